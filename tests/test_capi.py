"""The C-ABI library on the CPU: it loads without a GPU, exports every entry
point include/pll.h declares, its host-only routines agree with the oracle
bit for bit, and invalid arguments come back as status codes (never aborts)."""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

import powerllel_channel_sycl_amd as pll


def _declared():
    text = open(pll.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pll_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pll.load_library()
    names = _declared()
    assert len(names) >= 39
    out = subprocess.run(["nm", "-D", "--defined-only", pll.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (pll_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)
    assert set(names) == set(pll._SIGNATURES)


def test_abi_version_and_defaults():
    lib = pll.load_library()
    assert lib.pll_abi_version() == 1
    cfg = pll.PllConfig()
    lib.pll_config_default(C.byref(cfg))
    # examples/template/param.json
    assert (cfg.nx, cfg.ny, cfg.nz) == (128, 128, 128)
    assert cfg.re == 2800.0 and cfg.dt == 1e-3 and cfg.is_forced[0] == 1


@pytest.mark.parametrize("r,nz", [(1.5, 64), (1.8, 384), (2.0, 512), (0.0, 16)])
def test_mesh_matches_oracle_bitwise(oracle_lib, r, nz):
    a = pll.mesh_by_func(1, 0, r, nz, 2.0)
    b = oracle_lib.mesh_by_func(1, 0, r, nz, 2.0)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [8, 24, 64, 96, 384, 768])
def test_eigenvalues_match_oracle_bitwise(oracle_lib, n):
    np.testing.assert_array_equal(pll.eigen_values(n), oracle_lib.eigen_values(n))


def test_invalid_arguments_return_status():
    dzf = pll.mesh_by_func(1, 0, 1.5, 9, 2.0)
    with pytest.raises(pll.PllError) as e:
        pll.Channel(63, 64, 9, 1.0, 1.0, 2.0, 100.0, 1e-3, dzf)
    assert e.value.code == 1 and "even" in str(e.value)
    dzf = pll.mesh_by_func(1, 0, 1.5, 8, 2.0)
    with pytest.raises(pll.PllError) as e:
        pll.Channel(14, 14, 8, 1.0, 1.0, 2.0, 100.0, 1e-3, dzf)  # nx/2 = 7
    assert e.value.code == 1 and "2, 3 and 5" in str(e.value)
    with pytest.raises(pll.PllError):
        pll.mesh_by_func(1, 0, 1.5, 0, 2.0)
    lib = pll.load_library()
    assert lib.pll_step(None, 1) == 1
    assert b"null" in lib.pll_last_error()
