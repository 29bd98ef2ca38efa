// mod_type.hh -- field and 1D-array types of the host modules
// (src/mod_type.hh:16-316).
//
// The reference's Array3D owns shared host/device memory.  Here a field lives
// on the GPU inside the device context (include/pll.h); an Array3D names one
// of the context's field slots and carries a host mirror that is filled or
// flushed explicitly (to_host / to_device) for I/O and inspection.
#pragma once

#include <array>
#include <vector>

#include "pll.h"

using fp = double;

// 1-based array with one halo cell on each side: valid indices 0..n+1
// (src/mod_type.hh:16-67 with lhalo = rhalo = 1).
class Array1DH1 {
 public:
  Array1DH1() = default;
  explicit Array1DH1(int n) { allocate(n); }
  Array1DH1(const Array1DH1&) = delete;
  Array1DH1& operator=(const Array1DH1&) = delete;
  void allocate(int n) {
    sz_ = n;
    v_.assign(static_cast<size_t>(n) + 2, 0.0);
  }
  void deallocate() {
    v_.clear();
    v_.shrink_to_fit();
    sz_ = 0;
  }
  int size() const { return sz_; }
  fp* data() { return v_.data(); }
  const fp* data() const { return v_.data(); }
  fp& operator()(int k) { return v_[k]; }
  const fp& operator()(int k) const { return v_[k]; }

 private:
  std::vector<fp> v_;
  int sz_ = 0;
};

// A halo-padded (nx+2)(ny+2)(nz+2) field, x fastest, resident on the GPU.
class Array3D {
 public:
  static constexpr std::array<int, 6> nhalo = {1, 1, 1, 1, 1, 1};

  Array3D() = default;
  Array3D(const Array3D&) = delete;
  Array3D& operator=(const Array3D&) = delete;

  // Bind to a device-context field slot (pll_field) of interior size sz.
  void allocate(int slot, std::array<int, 3> sz);
  void deallocate();
  bool allocated() const { return slot_ >= 0; }
  int slot() const { return slot_; }

  // size without halos
  const std::array<int, 3>& size() const { return sz_; }
  // host mirror with halos (valid after to_host())
  fp* data() { return host_.data(); }
  const fp* data() const { return host_.data(); }
  // 1-based, halo at 0 and n+1, on the host mirror
  fp& operator()(int i, int j, int k) { return host_[index(i, j, k)]; }
  const fp& operator()(int i, int j, int k) const { return host_[index(i, j, k)]; }

  void to_host();    // device -> mirror
  void to_device();  // mirror -> device

 private:
  size_t index(int i, int j, int k) const {
    return static_cast<size_t>(i) +
           static_cast<size_t>(sz_[0] + 2) * (static_cast<size_t>(j) +
                                              static_cast<size_t>(sz_[1] + 2) * k);
  }
  int slot_ = -1;
  std::array<int, 3> sz_{0, 0, 0};
  std::vector<fp> host_;
};

using Array3DH3 = Array3D;
using Array3DH1 = Array3D;
