"""MI355X-native PowerLLEL channel-flow hot path.

The compute lives in ``libpll.so`` (hand-written HIP kernels for gfx950 behind
the C-ABI of ``include/pll.h``).  This module is the Python host side over that
C-ABI: a :class:`Channel` owns one device context and exposes one method per
routine of the reference's module interface, under the reference's names
(``timeIntVelRK1``, ``calcRHS``, ``execute_poisson_solver``, ``updateBoundVel``
...; src/mod_calcVel.hh, src/mod_calcRHS.hh, src/mod_poisson_solver.hh,
src/mod_updateBound.hh, src/mod_utils.hh).

Fields cross the boundary as float64 numpy arrays of shape (nz+2, ny+2, nx+2)
in the reference layout (src/mod_type.hh:107-117).  Failures raise
:class:`PllError` carrying the library's message; there is no CPU fallback --
if ``libpll.so`` is missing or no GPU is present, construction fails.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["Channel", "Closure", "Decomp", "PllError", "load_library", "mesh_by_func", "eigen_values", "FIELDS",
           "PRECISIONS", "PROFILES", "LIB_PATH", "HEADER_PATH"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# PLL_LIBRARY selects another build of the same C-ABI (e.g. to compare two builds)
LIB_PATH = os.environ.get("PLL_LIBRARY") or os.path.join(_HERE, "libpll.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "pll.h")

FIELDS = {"u": 0, "v": 1, "w": 2, "u1": 3, "v1": 4, "w1": 5, "p": 6}
PRECISIONS = {"fp64": 0, "fp32_mixed": 1}
PROFILES = {"uni": 0, "poi": 1, "log": 2}


class PllError(RuntimeError):
    """A pll_* call returned a non-zero status."""

    def __init__(self, code, what, msg):
        super().__init__(f"{what} failed with status {code}: {msg}")
        self.code = code


class PllConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("re", C.c_double), ("u_ref", C.c_double), ("l_ref", C.c_double),
        ("dt", C.c_double), ("u_crf", C.c_double),
        ("is_forced", C.c_int * 3), ("vel_force", C.c_double * 3),
        ("precision", C.c_int), ("pin_singular", C.c_int), ("device", C.c_int),
    ]


class PllValueIndex(C.Structure):
    _fields_ = [("value", C.c_double), ("rank", C.c_int), ("ig", C.c_int), ("jg", C.c_int),
                ("kg", C.c_int)]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None

# name -> (restype, argtypes); every entry point include/pll.h declares.
_SIGNATURES = {
    "pll_abi_version": (C.c_int, []),
    "pll_last_error": (C.c_char_p, []),
    "pll_config_default": (None, [C.POINTER(PllConfig)]),
    "pll_mesh_by_func": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D]),
    "pll_create": (C.c_int, [C.POINTER(PllConfig), _D, C.POINTER(C.c_void_p)]),
    "pll_destroy": (C.c_int, [C.c_void_p]),
    "pll_get_config": (C.c_int, [C.c_void_p, C.POINTER(PllConfig)]),
    "pll_device_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "pll_field_upload": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "pll_field_download": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "pll_field_device_ptr": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_int)]),
    "pll_spec_upload": (C.c_int, [C.c_void_p, _D]),
    "pll_spec_download": (C.c_int, [C.c_void_p, _D]),
    "pll_init_flow": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double,
                                C.POINTER(C.c_uint32)]),
    "pll_transform2crf": (C.c_int, [C.c_void_p, C.c_double]),
    "pll_time_int_vel_rk1": (C.c_int, [C.c_void_p]),
    "pll_time_int_vel_rk2": (C.c_int, [C.c_void_p]),
    "pll_correct_vel": (C.c_int, [C.c_void_p]),
    "pll_force_vel": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pll_calc_rhs": (C.c_int, [C.c_void_p]),
    "pll_poisson_execute": (C.c_int, [C.c_void_p]),
    "pll_fft_fwd_xpen": (C.c_int, [C.c_void_p]),
    "pll_fft_fwd_ypen": (C.c_int, [C.c_void_p]),
    "pll_solve_trid": (C.c_int, [C.c_void_p]),
    "pll_fft_bwd_ypen": (C.c_int, [C.c_void_p]),
    "pll_fft_bwd_xpen": (C.c_int, [C.c_void_p]),
    "pll_eigen_values": (C.c_int, [C.c_int, _D]),
    "pll_update_bound_vel": (C.c_int, [C.c_void_p, C.c_int]),
    "pll_impose_bc_vel": (C.c_int, [C.c_void_p, C.c_int]),
    "pll_update_bound_p": (C.c_int, [C.c_void_p]),
    "pll_mean": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "pll_calc_max_div": (C.c_int, [C.c_void_p, C.POINTER(PllValueIndex)]),
    "pll_calc_max_cfl": (C.c_int, [C.c_void_p, C.POINTER(PllValueIndex)]),
    "pll_check_nan": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "pll_stat_enable": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pll_stat_accumulate": (C.c_int, [C.c_void_p]),
    "pll_stat_download": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "pll_stat_upload": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "pll_stat_plane_sums": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "pll_closure_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "pll_closure_destroy": (C.c_int, [C.c_void_p]),
    "pll_closure_num_params": (C.c_int, []),
    "pll_closure_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pll_closure_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pll_closure_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                       C.POINTER(C.c_float), C.POINTER(C.c_double)]),
    "pll_closure_gradient_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                               C.POINTER(C.c_float), C.POINTER(C.c_double)]),
    "pll_closure_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                         C.c_double]),
    "pll_closure_predict": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p]),
    "pll_closure_synchronize": (C.c_int, [C.c_void_p]),
    "pll_closure_set_scales": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "pll_closure_get_scales": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
    "pll_closure_attach": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pll_closure_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pll_closure_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "pll_closure_nut": (C.c_int, [C.c_void_p, _D]),
    "pll_closure_filter": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pll_closure_samples": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pll_decomp_create": (C.c_int, [C.c_int] * 5 + [C.POINTER(C.c_void_p)]),
    "pll_decomp_destroy": (C.c_int, [C.c_void_p]),
    "pll_decomp_pencil": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_long)]),
    "pll_decomp_total": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_long)]),
    "pll_transpose_x_to_y_real": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pll_transpose_y_to_x_real": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pll_transpose_y_to_z_real": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pll_transpose_z_to_y_real": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pll_step": (C.c_int, [C.c_void_p, C.c_int]),
    "pll_synchronize": (C.c_int, [C.c_void_p]),
    "pll_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "pll_timing_reset": (C.c_int, [C.c_void_p]),
    "pll_timing_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pll_timing_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_char_p),
                                  C.POINTER(C.c_double), C.POINTER(C.c_long)]),
}


def load_library(path=LIB_PATH):
    """Load libpll.so and declare every entry point; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run make -C powerllel_channel_sycl_amd/csrc")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        msg = load_library().pll_last_error()
        raise PllError(rc, what, msg.decode() if msg else "")


def mesh_by_func(mesh_type, stretch_func, stretch_ratio, nz, lz):
    """initMeshByFunc (src/mod_mesh.cc:95-125): dzf_global[0..nz+1]."""
    out = np.zeros(nz + 2)
    _check(load_library().pll_mesh_by_func(mesh_type, stretch_func, stretch_ratio, nz, lz, out),
           "pll_mesh_by_func")
    return out


def eigen_values(n):
    """get_eigen_values (src/mod_fft.cc:196-222): perm-ordered, unscaled."""
    out = np.zeros(n)
    _check(load_library().pll_eigen_values(n, out), "pll_eigen_values")
    return out


class Closure:
    """The learned eddy-viscosity closure (no counterpart in the reference): a
    10 -> 64 -> 64 -> 1 perceptron with BF16 weights on MFMA.  Sample arrays
    are device addresses of float32 buffers: x[n][10], y[n], out[n]."""

    def __init__(self, seed=0):
        self._h = C.c_void_p()
        _check(load_library().pll_closure_create(int(seed), C.byref(self._h)), "pll_closure_create")
        self.nparam = load_library().pll_closure_num_params()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().pll_closure_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        _check(load_library().pll_closure_get_params(
            self._h, out.ctypes.data_as(C.POINTER(C.c_float))), "pll_closure_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _check(load_library().pll_closure_set_params(
            self._h, p.ctypes.data_as(C.POINTER(C.c_float))), "pll_closure_set_params")

    def gradient(self, x_ptr, y_ptr, n):
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double()
        _check(load_library().pll_closure_gradient(
            self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(n),
            g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)), "pll_closure_gradient")
        return g, loss.value

    def gradient_noalloc(self, x_ptr, y_ptr, n):
        """gradient() through pll_closure_gradient_noalloc (no allocation)."""
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double()
        _check(load_library().pll_closure_gradient_noalloc(
            self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(n),
            g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)), "pll_closure_gradient_noalloc")
        return g, loss.value

    def train_step(self, x_ptr, y_ptr, n, lr):
        _check(load_library().pll_closure_train_step(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                                     int(n), float(lr)), "pll_closure_train_step")

    def predict(self, x_ptr, n, out_ptr):
        _check(load_library().pll_closure_predict(self._h, C.c_void_p(x_ptr), int(n),
                                                  C.c_void_p(out_ptr)), "pll_closure_predict")

    def synchronize(self):
        _check(load_library().pll_closure_synchronize(self._h), "pll_closure_synchronize")

    def set_scales(self, gscale, yscale, nuscale):
        """Inputs: gradients * gscale, wall distance * yscale; nu_t = nuscale * softplus(o)."""
        _check(load_library().pll_closure_set_scales(self._h, float(gscale), float(yscale),
                                                     float(nuscale)), "pll_closure_set_scales")

    def scales(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(load_library().pll_closure_get_scales(self._h, C.byref(a), C.byref(b), C.byref(c)),
               "pll_closure_get_scales")
        return a.value, b.value, c.value

    def save(self, path):
        """Parameters and scales to a file (pll_closure_save)."""
        _check(load_library().pll_closure_save(self._h, os.fsencode(path)), "pll_closure_save")

    @classmethod
    def load(cls, path):
        """A closure with the parameters and scales of a saved file."""
        obj = cls.__new__(cls)
        obj._h = C.c_void_p()
        _check(load_library().pll_closure_load(os.fsencode(path), C.byref(obj._h)),
               "pll_closure_load")
        obj.nparam = load_library().pll_closure_num_params()
        return obj

    def samples(self, fine, coarse, x_ptr, y_ptr):
        """Filter fine into coarse and write its n = coarse cells training samples."""
        _check(load_library().pll_closure_samples(fine._h, coarse._h, self._h, C.c_void_p(x_ptr),
                                                  C.c_void_p(y_ptr)), "pll_closure_samples")


class Decomp:
    """decomp_2d for p_row x p_col ranks held on one GPU (src/decomp_2d.cc:41-589).

    Pencils: 0 = x, 1 = y, 2 = z.  Device buffers passed to the transposes hold
    every rank's pencil in rank order (rank = coords[0]*p_col + coords[1])."""

    def __init__(self, nx, ny, nz, p_row, p_col):
        self._h = C.c_void_p()
        _check(load_library().pll_decomp_create(nx, ny, nz, p_row, p_col, C.byref(self._h)),
               "pll_decomp_create")
        self.nranks = p_row * p_col

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().pll_decomp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pencil(self, pencil, rank):
        """(st[3], sz[3], element offset) of rank's pencil; st is 1-based."""
        st, sz, off = (C.c_int * 3)(), (C.c_int * 3)(), C.c_long()
        _check(load_library().pll_decomp_pencil(self._h, pencil, rank, st, sz, C.byref(off)),
               "pll_decomp_pencil")
        return tuple(st), tuple(sz), off.value

    def total(self, pencil):
        t = C.c_long()
        _check(load_library().pll_decomp_total(self._h, pencil, C.byref(t)), "pll_decomp_total")
        return t.value

    def transpose(self, kind, src_ptr, dst_ptr, elem_bytes=8):
        """kind in {'x_to_y', 'y_to_x', 'y_to_z', 'z_to_y'} on device addresses."""
        fn = getattr(load_library(), f"pll_transpose_{kind}_real")
        _check(fn(self._h, C.c_void_p(src_ptr), C.c_void_p(dst_ptr), elem_bytes),
               f"pll_transpose_{kind}_real")


class Channel:
    """One channel-flow problem on one GPU (the context of include/pll.h)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, re, dt, dzf_global, u_ref=1.0, l_ref=1.0,
                 u_crf=0.0, is_forced=(1, 0, 0), vel_force=(1.0, 0.0, 0.0), precision="fp64",
                 pin_singular=True, device=0):
        lib = load_library()
        cfg = PllConfig()
        lib.pll_config_default(C.byref(cfg))
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        cfg.lx, cfg.ly, cfg.lz = lx, ly, lz
        cfg.re, cfg.u_ref, cfg.l_ref = re, u_ref, l_ref
        cfg.dt, cfg.u_crf = dt, u_crf
        for d in range(3):
            cfg.is_forced[d] = int(bool(is_forced[d]))
            cfg.vel_force[d] = float(vel_force[d])
        cfg.precision = PRECISIONS[precision]
        cfg.pin_singular = int(bool(pin_singular))
        cfg.device = device
        self.nx, self.ny, self.nz = nx, ny, nz
        self.precision = precision
        self._h = C.c_void_p()
        _check(lib.pll_create(C.byref(cfg), np.ascontiguousarray(dzf_global, dtype=np.float64),
                              C.byref(self._h)), "pll_create")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().pll_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, name, *args):
        _check(getattr(load_library(), name)(self._h, *args), name)

    # ---- fields ---------------------------------------------------------------
    @property
    def field_shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    def upload(self, field, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        if arr.shape != self.field_shape:
            raise ValueError(f"field shape {arr.shape} != {self.field_shape}")
        self._call("pll_field_upload", FIELDS[field], arr)

    def download(self, field):
        out = np.empty(self.field_shape)
        self._call("pll_field_download", FIELDS[field], out)
        return out

    def upload_spec(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        if arr.size != self.nx * self.ny * self.nz:
            raise ValueError("work array must hold nx*ny*nz values")
        self._call("pll_spec_upload", arr)

    def download_spec(self):
        out = np.empty((self.nz, self.ny, self.nx))
        self._call("pll_spec_download", out)
        return out

    def device_bytes(self):
        b = C.c_size_t()
        self._call("pll_device_bytes", C.byref(b))
        return b.value

    # ---- mod_initFlow / mod_calcVel ----------------------------------------------
    def initFlow(self, profile="poi", u0=0.0, with_noise=True, noise_intensity=0.1,
                 seeds=(42, 43, 44)):
        s = (C.c_uint32 * 3)(*[int(x) for x in seeds])
        self._call("pll_init_flow", PROFILES[profile], float(u0), int(bool(with_noise)),
                   float(noise_intensity), s)

    def transform2CRF(self, vel_crf):
        self._call("pll_transform2crf", float(vel_crf))

    def timeIntVelRK1(self):
        self._call("pll_time_int_vel_rk1")

    def timeIntVelRK2(self):
        self._call("pll_time_int_vel_rk2")

    def correctVel(self):
        self._call("pll_correct_vel")

    def forceVel(self, return_mean=False):
        if return_mean:
            m = (C.c_double * 3)()
            self._call("pll_force_vel", m)
            return np.array(m[:])
        self._call("pll_force_vel", None)
        return None

    # ---- mod_calcRHS / poisson_solver -----------------------------------------------
    def calcRHS(self):
        self._call("pll_calc_rhs")

    def execute_poisson_solver(self):
        self._call("pll_poisson_execute")

    def execute_fft_fwd_xpen(self):
        self._call("pll_fft_fwd_xpen")

    def execute_fft_fwd_ypen(self):
        self._call("pll_fft_fwd_ypen")

    def solve_trid(self):
        self._call("pll_solve_trid")

    def execute_fft_bwd_ypen(self):
        self._call("pll_fft_bwd_ypen")

    def execute_fft_bwd_xpen(self):
        self._call("pll_fft_bwd_xpen")

    # ---- mod_updateBound -----------------------------------------------------------
    def updateBoundVel(self, which=0):
        self._call("pll_update_bound_vel", int(which))

    def imposeBCVel(self, which=0):
        """Wall conditions only (src/mod_updateBound.cc:183-227)."""
        self._call("pll_impose_bc_vel", int(which))

    def updateBoundP(self):
        self._call("pll_update_bound_p")

    # ---- mod_statistics ---------------------------------------------------------------
    STAT_NAMES = ("u", "v", "w", "u2", "v2", "w2", "uv", "uw", "vw", "p", "p2")

    def stat_enable(self, which=(1,) * 11):
        arr = (C.c_int * 11)(*[int(bool(x)) for x in which])
        self._call("pll_stat_enable", arr)

    def stat_accumulate(self):
        self._call("pll_stat_accumulate")

    def stat_download(self, var):
        out = np.zeros((self.nz, self.ny, self.nx))
        self._call("pll_stat_download", int(var), out)
        return out

    def stat_upload(self, var, arr):
        self._call("pll_stat_upload", int(var), np.ascontiguousarray(arr, dtype=np.float64))

    def stat_plane_sums(self, var):
        out = np.zeros(self.nz)
        self._call("pll_stat_plane_sums", int(var), out)
        return out

    # ---- mod_utils -------------------------------------------------------------------
    def Mean(self, field="u", nx_global=None, ny_global=None):
        out = C.c_double()
        self._call("pll_mean", FIELDS[field], int(nx_global or self.nx), int(ny_global or self.ny),
                   C.byref(out))
        return out.value

    def calcMaxDiv(self):
        r = PllValueIndex()
        self._call("pll_calc_max_div", C.byref(r))
        return r.value, (r.ig, r.jg, r.kg)

    def calcMaxCFL(self):
        r = PllValueIndex()
        self._call("pll_calc_max_cfl", C.byref(r))
        return r.value, (r.ig, r.jg, r.kg)

    def checkNaN(self, field="u"):
        ok = C.c_int()
        self._call("pll_check_nan", FIELDS[field], C.byref(ok))
        return bool(ok.value)

    # ---- learned closure --------------------------------------------------------------
    def attach_closure(self, closure):
        """Add the closure's eddy-viscosity stress to the RK stages (None detaches)."""
        self._call("pll_closure_attach", closure._h if closure is not None else None)
        self._closure = closure  # keep it alive while attached

    def closure_nut(self):
        out = np.zeros((self.nz, self.ny, self.nx))
        self._call("pll_closure_nut", out)
        return out

    def filter_from(self, fine):
        """This (coarse) context's u, v, w := the box filter of the fine context's."""
        _check(load_library().pll_closure_filter(fine._h, self._h), "pll_closure_filter")

    # ---- main loop ---------------------------------------------------------------------
    def step(self, nsteps=1):
        self._call("pll_step", int(nsteps))

    def synchronize(self):
        self._call("pll_synchronize")

    # ---- timers --------------------------------------------------------------------------
    def timing(self, on=True):
        self._call("pll_timing_enable", int(bool(on)))

    def timing_reset(self):
        self._call("pll_timing_reset")

    def timing_read(self):
        """{kernel name: (total ms, launches)} of the HIP-event timers."""
        n = C.c_int()
        self._call("pll_timing_count", C.byref(n))
        out = {}
        for k in range(n.value):
            name, ms, cnt = C.c_char_p(), C.c_double(), C.c_long()
            self._call("pll_timing_read", k, C.byref(name), C.byref(ms), C.byref(cnt))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out
