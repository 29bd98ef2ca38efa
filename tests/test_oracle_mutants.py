"""Mutation check of the oracle pins (CPU only).

Each case plants one plausible mistake in a copy of oracle/pll_oracle.cc (a
wrong metric index, a flipped sign, a swapped operand, a changed constant),
builds the copy with the oracle's own flags, and runs the CPU pin suites
against it through PLL_ORACLE_LIB.  The case passes when at least one pin
fails on the mutant -- and the expected pin is among the failures -- so every
function the GPU path is compared against is shown to be pinned by checks that
would catch such a mistake.  The unmodified oracle passes the same suites
(they are part of the ordinary run).
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pll_oracle.cc")
PIN_SUITES = ["tests/test_oracle.py", "tests/test_oracle_momentum.py"]

# name: (text in pll_oracle.cc, replacement, a pin expected to fail)
MUTANTS = {
    # the four mistakes of the w tendency, the RK2 combination and the log
    # profile named as unpinned before these pins existed
    "w_visc_metric_index": (
        "dqt = (W(i, j, k + 1) - W(i, j, k)) * dzf_inv[k + 1];",
        "dqt = (W(i, j, k + 1) - W(i, j, k)) * dzf_inv[k];",
        "test_w_wall_normal_terms_on_stretched_grid"),
    "w_y_convection_sign": (
        "(vn * wn - vs * ws) * dy_inv",
        "(vs * ws - vn * wn) * dy_inv",
        "test_w_convection_and_diffusion_along_y"),
    "rk2_u_point_from_u1": (
        "u[id] = (u1[id] + dt * tu + u[id]) * 0.5;",
        "u[id] = (u1[id] + dt * tu + u1[id]) * 0.5;",
        "test_rk2_is_heun_over_rk1"),
    "log_profile_exponent": (
        "std::pow(re, 0.88)",
        "std::pow(re, 0.80)",
        "test_log_profile_is_the_law_of_the_wall"),
    # further mistakes of the same kinds
    "w_x_interp_weight": (
        "ue = (U(i, j, k) * dzf[k + 1] + U(i, j, k + 1) * dzf[k]) * dzc_inv[k];\n"
        "  uw = (U(i - 1, j, k)",
        "ue = (U(i, j, k) * dzf[k + 1] + U(i, j, k + 1) * dzf[k + 1]) * dzc_inv[k];\n"
        "  uw = (U(i - 1, j, k)",
        "test_w_convection_and_diffusion_along_x"),
    "w_z_convection_metric": (
        "(wt * wt - wb * wb) * dzc_inv[k]",
        "(wt * wt - wb * wb) * dzf_inv[k]",
        "test_w_wall_normal_terms_on_stretched_grid"),
    "w_y_visc_spacing": (
        "dqs = (W(i, j, k) - W(i, j - 1, k)) * dy_inv;",
        "dqs = (W(i, j, k) - W(i, j - 1, k)) * dx_inv;",
        "test_w_convection_and_diffusion_along_y"),
    "rk2_w_wrong_component": (
        "w[id] = (w1[id] + dt * tw + w[id]) * 0.5;",
        "w[id] = (w1[id] + dt * tv + w[id]) * 0.5;",
        "test_rk2_is_heun_over_rk1"),
    "log_law_intercept": (
        "2.5 * std::log(z) + 5.5",
        "2.5 * std::log(z) + 5.0",
        "test_log_profile_is_the_law_of_the_wall"),
    "u_bottom_metric_index": (
        "ub = (U(i, j, k) * dzf[k - 1] + U(i, j, k - 1) * dzf[k]) * dzc_inv[k - 1];",
        "ub = (U(i, j, k) * dzf[k - 1] + U(i, j, k - 1) * dzf[k]) * dzc_inv[k];",
        "test_rk1_conserves_momentum_up_to_wall_shear"),
}

CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
            "-w", "-shared"]


def _build(name, workdir):
    old, new, _ = MUTANTS[name]
    text = open(SRC).read()
    assert text.count(old) == 1, f"{name}: the planted text must occur exactly once"
    src = os.path.join(workdir, f"{name}.cc")
    lib = os.path.join(workdir, f"lib_{name}.so")
    with open(src, "w") as f:
        f.write(text.replace(old, new))
    subprocess.run(["g++", *CXXFLAGS, "-o", lib, src, "-lm"], check=True)
    return lib


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    work = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = dict(zip(MUTANTS, ex.map(lambda n: _build(n, work), MUTANTS)))
    return libs


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_planted_mistake(mutant_libs, name):
    env = dict(os.environ, PLL_ORACLE_LIB=mutant_libs[name], OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu", "-rf", *PIN_SUITES],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    failed = set(re.findall(r"^FAILED \S+::(\w+)", r.stdout, re.M))
    # exit code 1: tests ran and some failed (2+ would be a collection/usage error)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert MUTANTS[name][2] in failed, (name, sorted(failed))
