"""ctypes binding of the CPU FP64 oracle (oracle/pll_oracle.cc).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the CPU-baseline leg of bench.py, always as the
checker or the timed CPU baseline, never as part of the product path.

Fields are numpy float64 arrays of shape (nz+2, ny+2, nx+2) -- the reference
layout of src/mod_type.hh:107-117 (x fastest, one halo cell per side).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DEFAULT_LIB = os.path.join(_HERE, "build", "libpll_oracle.so")
# PLL_ORACLE_LIB names another build of pll_oracle.cc (the mutation tests load
# deliberately broken copies through it to show that the pins catch them).
_LIB_PATH = os.environ.get("PLL_ORACLE_LIB") or _DEFAULT_LIB


class OrcConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("re", C.c_double), ("u_ref", C.c_double), ("l_ref", C.c_double),
        ("dt", C.c_double), ("u_crf", C.c_double),
        ("is_forced", C.c_int * 3), ("vel_force", C.c_double * 3),
        ("pin_singular", C.c_int),
    ]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


def build():
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if _LIB_PATH == _DEFAULT_LIB and not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    L.orc_create.restype = vp
    L.orc_create.argtypes = [C.POINTER(OrcConfig), _D]
    L.orc_destroy.argtypes = [vp]
    L.orc_mesh_by_func.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D]
    L.orc_num_threads.restype = C.c_int
    for name in ("orc_rk1", "orc_rk2"):
        getattr(L, name).argtypes = [vp] + [_D] * 6
    L.orc_calc_rhs.argtypes = [vp, _D, _D, _D, _D]
    L.orc_correct_vel.argtypes = [vp, _D, _D, _D, _D]
    L.orc_update_bound_vel.argtypes = [vp, _D, _D, _D]
    L.orc_update_bound_p.argtypes = [vp, _D]
    L.orc_impose_bc_vel.argtypes = [vp, _D, _D, _D]
    L.orc_mean.restype = C.c_double
    L.orc_mean.argtypes = [vp, C.c_int, C.c_int, _D]
    L.orc_force_vel.argtypes = [vp, _D, _D, _D, _D]
    L.orc_fft_fwd_xpen.argtypes = [vp, _D, _D]
    L.orc_fft_bwd_xpen.argtypes = [vp, _D, _D]
    L.orc_fft_fwd_ypen.argtypes = [vp, _D]
    L.orc_fft_bwd_ypen.argtypes = [vp, _D]
    L.orc_solve_trid.argtypes = [vp, _D]
    L.orc_poisson.argtypes = [vp, _D]
    L.orc_get_spec.argtypes = [vp, _D]
    L.orc_step.argtypes = [vp, C.c_int] + [_D] * 7
    L.orc_max_div.argtypes = [vp, _D, _D, _D, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.orc_max_cfl.argtypes = [vp, _D, _D, _D, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.orc_check_nan.restype = C.c_int
    L.orc_check_nan.argtypes = [vp, _D]
    L.orc_noise_uniform.restype = C.c_double
    L.orc_noise_uniform.argtypes = [C.c_uint32, C.c_uint64]
    L.orc_init_flow.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_double, np.ctypeslib.ndpointer(dtype=np.uint32), _D, _D, _D]
    L.orc_dft.argtypes = [C.c_int, C.c_int, _D, _D]
    L.orc_rfft_perm.argtypes = [C.c_int, _D, _D]
    L.orc_irfft_perm.argtypes = [C.c_int, _D, _D]
    L.orc_eigen_values.argtypes = [C.c_int, _D]
    L.orc_trid_coeff.argtypes = [C.c_int, _D, _D, _D, _D]
    L.orc_get_mesh.argtypes = [vp] + [_D] * 6
    L.orc_get_trid.argtypes = [vp] + [_D] * 5
    L.orc_decomp_create.restype = vp
    L.orc_decomp_create.argtypes = [C.c_int] * 5
    L.orc_decomp_destroy.argtypes = [vp]
    L.orc_decomp_pencil.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_long)]
    L.orc_decomp_total.restype = C.c_long
    L.orc_decomp_total.argtypes = [vp, C.c_int]
    L.orc_transpose.argtypes = [vp, C.c_int, _D, _D]
    _lib = L
    return L


PROFILES = {"uni": 0, "poi": 1, "log": 2}


def mesh_by_func(mesh_type, stretch_func, stretch_ratio, nz, lz):
    out = np.zeros(nz + 2)
    lib().orc_mesh_by_func(mesh_type, stretch_func, stretch_ratio, nz, lz, out)
    return out


class Oracle:
    """One oracle context: mesh coefficients, Poisson setup and the step routines."""

    def __init__(self, nx, ny, nz, lx, ly, lz, re, dt, dzf_global, u_ref=1.0, l_ref=1.0,
                 u_crf=0.0, is_forced=(1, 0, 0), vel_force=(1.0, 0.0, 0.0), pin_singular=1):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.re, self.u_ref = re, u_ref
        cfg = OrcConfig()
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        cfg.lx, cfg.ly, cfg.lz = lx, ly, lz
        cfg.re, cfg.u_ref, cfg.l_ref = re, u_ref, l_ref
        cfg.dt, cfg.u_crf = dt, u_crf
        for d in range(3):
            cfg.is_forced[d] = int(is_forced[d])
            cfg.vel_force[d] = float(vel_force[d])
        cfg.pin_singular = int(pin_singular)
        self._h = lib().orc_create(C.byref(cfg), np.ascontiguousarray(dzf_global, dtype=np.float64))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    # fields -----------------------------------------------------------------
    def field(self):
        return np.zeros((self.nz + 2, self.ny + 2, self.nx + 2))

    def spec(self):
        return np.zeros((self.nz, self.ny, self.nx))

    def mesh(self):
        arrs = [np.zeros(self.nz + 2) for _ in range(6)]
        lib().orc_get_mesh(self._h, *arrs)
        return dict(zip(["dzf", "dzf_inv", "dzc", "dzc_inv", "dzflzi", "zc_global"], arrs))

    def trid(self):
        a, c = np.zeros(self.nz), np.zeros(self.nz)
        b = np.zeros((self.nz, self.ny, self.nx))
        lx, ly = np.zeros(self.nx), np.zeros(self.ny)
        lib().orc_get_trid(self._h, a, c, b, lx, ly)
        return dict(a=a, c=c, b=b, lambdax=lx, lambday=ly)

    # hot path ----------------------------------------------------------------
    def rk1(self, u, v, w, u1, v1, w1):
        lib().orc_rk1(self._h, u, v, w, u1, v1, w1)

    def rk2(self, u, v, w, u1, v1, w1):
        lib().orc_rk2(self._h, u, v, w, u1, v1, w1)

    def calc_rhs(self, u, v, w, p):
        lib().orc_calc_rhs(self._h, u, v, w, p)

    def correct_vel(self, p, u, v, w):
        lib().orc_correct_vel(self._h, p, u, v, w)

    def update_bound_vel(self, u, v, w):
        lib().orc_update_bound_vel(self._h, u, v, w)

    def impose_bc_vel(self, u, v, w):
        lib().orc_impose_bc_vel(self._h, u, v, w)

    def update_bound_p(self, p):
        lib().orc_update_bound_p(self._h, p)

    def mean(self, f, nx_global=None, ny_global=None):
        return lib().orc_mean(self._h, nx_global or self.nx, ny_global or self.ny, f)

    def force_vel(self, u, v, w):
        m = np.zeros(3)
        lib().orc_force_vel(self._h, u, v, w, m)
        return m

    def fft_fwd_xpen(self, p, spec):
        lib().orc_fft_fwd_xpen(self._h, p, spec)

    def fft_bwd_xpen(self, spec, p):
        lib().orc_fft_bwd_xpen(self._h, spec, p)

    def fft_fwd_ypen(self, spec):
        lib().orc_fft_fwd_ypen(self._h, spec)

    def fft_bwd_ypen(self, spec):
        lib().orc_fft_bwd_ypen(self._h, spec)

    def solve_trid(self, spec):
        lib().orc_solve_trid(self._h, spec)

    def poisson(self, p):
        lib().orc_poisson(self._h, p)

    def step(self, n, u, v, w, u1, v1, w1, p):
        lib().orc_step(self._h, int(n), u, v, w, u1, v1, w1, p)

    # checks ------------------------------------------------------------------
    def max_div(self, u, v, w):
        val, ijk = C.c_double(), (C.c_int * 3)()
        lib().orc_max_div(self._h, u, v, w, C.byref(val), ijk)
        return val.value, tuple(ijk)

    def max_cfl(self, u, v, w):
        val, ijk = C.c_double(), (C.c_int * 3)()
        lib().orc_max_cfl(self._h, u, v, w, C.byref(val), ijk)
        return val.value, tuple(ijk)

    def check_nan(self, f):
        return bool(lib().orc_check_nan(self._h, f))

    def init_flow(self, u, v, w, profile="poi", u0=0.0, with_noise=True, noise_intensity=0.1,
                  seeds=(42, 43, 44)):
        s = np.array(seeds, dtype=np.uint32)
        lib().orc_init_flow(self._h, PROFILES[profile], u0, self.u_ref, self.re, int(with_noise),
                            noise_intensity, s, u, v, w)


def dft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    n = x.size
    inp = np.empty(2 * n)
    inp[0::2], inp[1::2] = x.real, x.imag
    out = np.empty(2 * n)
    lib().orc_dft(n, sign, inp, out)
    return out[0::2] + 1j * out[1::2]


def rfft_perm(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_rfft_perm(x.size, x, out)
    return out


def irfft_perm(perm):
    perm = np.ascontiguousarray(perm, dtype=np.float64)
    out = np.empty_like(perm)
    lib().orc_irfft_perm(perm.size, perm, out)
    return out


def eigen_values(n):
    out = np.empty(n)
    lib().orc_eigen_values(n, out)
    return out


def trid_coeff(dzf0):
    n = dzf0.size - 2
    a, b, c = np.empty(n), np.empty(n), np.empty(n)
    lib().orc_trid_coeff(n, np.ascontiguousarray(dzf0, dtype=np.float64), a, b, c)
    return a, b, c


def noise_uniform(seed, offset):
    return lib().orc_noise_uniform(seed, offset)


TRANSPOSES = {"x_to_y": 0, "y_to_x": 1, "y_to_z": 2, "z_to_y": 3}


class OracleDecomp:
    """decomp_2d restated with mem_split -> Alltoallv -> mem_merge over virtual ranks."""

    def __init__(self, nx, ny, nz, p_row, p_col):
        self._h = lib().orc_decomp_create(nx, ny, nz, p_row, p_col)
        self.nranks = p_row * p_col

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_decomp_destroy(h)
            self._h = None

    def pencil(self, pencil, rank):
        st, sz, off = (C.c_int * 3)(), (C.c_int * 3)(), C.c_long()
        lib().orc_decomp_pencil(self._h, pencil, rank, st, sz, C.byref(off))
        return tuple(st), tuple(sz), off.value

    def total(self, pencil):
        return lib().orc_decomp_total(self._h, pencil)

    def transpose(self, kind, src):
        dst_pencil = {"x_to_y": 1, "y_to_x": 0, "y_to_z": 2, "z_to_y": 1}[kind]
        out = np.zeros(self.total(dst_pencil))
        lib().orc_transpose(self._h, TRANSPOSES[kind], np.ascontiguousarray(src, dtype=np.float64), out)
        return out
