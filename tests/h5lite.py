"""Read and write the checkpoint files (one FP64 (nz, ny, nx) dataset plus
small int datasets) with the serial HDF5 C library through ctypes.

Test infrastructure: an independent path to the files pll_channel writes
(src/mod_hdf5.cc:87-338 layout), so that a wrong shape, order or timestamp in
its writer is not hidden by the same mistake in its reader.
"""
import ctypes as C
import os

import numpy as np

_CANDIDATES = [os.environ.get("PLL_HDF5_LIB"), "/opt/conda/lib/libhdf5.so.103",
               "/opt/conda/lib/libhdf5.so"]
_h5 = None

hid_t = C.c_int64
hsize_t = C.c_ulonglong


def available():
    try:
        lib()
        return True
    except OSError:
        return False


def lib():
    global _h5
    if _h5 is not None:
        return _h5
    err = None
    for p in _CANDIDATES:
        if not p:
            continue
        try:
            h = C.CDLL(p)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"HDF5 library not found: {err}")
    h.H5open()
    h.H5Fopen.restype = hid_t
    h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid_t]
    h.H5Fcreate.restype = hid_t
    h.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid_t, hid_t]
    h.H5Fclose.argtypes = [hid_t]
    h.H5Dopen2.restype = hid_t
    h.H5Dopen2.argtypes = [hid_t, C.c_char_p, hid_t]
    h.H5Dcreate2.restype = hid_t
    h.H5Dcreate2.argtypes = [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]
    h.H5Dclose.argtypes = [hid_t]
    h.H5Dget_space.restype = hid_t
    h.H5Dget_space.argtypes = [hid_t]
    h.H5Sget_simple_extent_ndims.argtypes = [hid_t]
    h.H5Sget_simple_extent_dims.argtypes = [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]
    h.H5Screate_simple.restype = hid_t
    h.H5Screate_simple.argtypes = [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]
    h.H5Sclose.argtypes = [hid_t]
    h.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
    h.H5Dwrite.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
    h.native_int = hid_t.in_dll(h, "H5T_NATIVE_INT_g").value
    h.native_double = hid_t.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    _h5 = h
    return h


def read(path):
    """{dataset name: array} for the given dataset names of one file."""
    h = lib()
    f = h.H5Fopen(path.encode(), 0, 0)
    if f < 0:
        raise OSError(f"cannot open {path}")
    return _File(h, f)


class _File:
    def __init__(self, h, f):
        self.h, self.f = h, f

    def dataset(self, name):
        h = self.h
        d = h.H5Dopen2(self.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        sp = h.H5Dget_space(d)
        nd = h.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        h.H5Sget_simple_extent_dims(sp, dims, None)
        h.H5Sclose(sp)
        shape = tuple(int(x) for x in dims[:nd])
        if name in ("nt", "nts", "nte", "nspl"):
            out = np.zeros(shape, dtype=np.int32)
            h.H5Dread(d, h.native_int, 0, 0, 0, out.ctypes.data)
        else:
            out = np.zeros(shape, dtype=np.float64)
            h.H5Dread(d, h.native_double, 0, 0, 0, out.ctypes.data)
        h.H5Dclose(d)
        return out

    def close(self):
        if self.f is not None:
            self.h.H5Fclose(self.f)
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def write_field(path, tag, nt, arr):
    """A checkpoint file as the reference writes it: <tag> (nz, ny, nx) FP64 and nt."""
    h = lib()
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    f = h.H5Fcreate(path.encode(), 2, 0, 0)
    for name, data, typ in (("nt", np.array([nt], dtype=np.int32), h.native_int),
                            (tag, arr, h.native_double)):
        dims = (hsize_t * data.ndim)(*data.shape)
        sp = h.H5Screate_simple(data.ndim, dims, None)
        d = h.H5Dcreate2(f, name.encode(), typ, sp, 0, 0, 0)
        h.H5Dwrite(d, typ, 0, 0, 0, data.ctypes.data)
        h.H5Dclose(d)
        h.H5Sclose(sp)
    h.H5Fclose(f)
