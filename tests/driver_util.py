"""Run-parameter files and runs of the host driver (pll_channel)."""
import copy
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "powerllel_channel_sycl_amd")
DRIVER = os.path.join(PKG, "pll_channel")

# The sections and keys of the reference's param.json, at a small channel.
BASE = {
    "PARA": {"p_row": 1, "p_col": 1},
    "MESH": {"read_mesh": False, "mesh_type": 1, "stretch_func": 0, "stretch_ratio": 1.5,
             "nx": 32, "ny": 24, "nz": 16, "lx": 6.283185307179586, "ly": 3.141592653589793,
             "lz": 2.0},
    "TIME": {"nt_end": 10, "dt": 2e-3, "nt_check": 5, "cfl_limit": 1.0, "div_limit": 1e-9},
    "RESTART": {"is_restart": False, "fn_prefix_input_inst": "save",
                "fn_prefix_input_stat": "save_stat"},
    "PHYSICAL": {"re": 2800, "u_ref": 1.0, "l_ref": 1.0, "initial_field": "poi", "u0": 1.0,
                 "init_with_noise": True, "noise_intensity": 0.1, "smooth_wall_visc": False,
                 "u_crf": 0.0},
    "FORCE": {"is_forced": [True, False, False], "vel_force": [1.0, 0.0, 0.0]},
    "STATISTICS": {"nt_init_stat": 100000, "sample_interval": 1,
                   "stat_which_var": [True] * 11},
    "OUTPUT": {"nt_out_scrn": 5, "nt_out_inst": 100000, "nt_out_stat": 100000,
               "nt_out_save": 10, "overwrite_save": True, "auto_cleanup": True,
               "num_retained": 0, "nt_out_moni": 100000},
    "MONITOR": {"out_forcing": False, "out_probe_point": False, "probe_ijk": [1, 1, 1],
                "out_skf_z": [False, False], "out_region": False, "nt_out_region": 100000,
                "region_ijk": [1, 1, 1, 1, 1, 1]},
}


def ensure_built():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.join(PKG, "host")], check=True)
    return DRIVER


def param(**sections):
    """BASE with per-section overrides: param(MESH={"nx": 16}, DEVICE={...})."""
    p = copy.deepcopy(BASE)
    for sec, kv in sections.items():
        if kv is None:
            p.pop(sec, None)
        else:
            p.setdefault(sec, {}).update(kv)
    return p


def write(dirpath, p, name="param.json"):
    path = os.path.join(dirpath, name)
    with open(path, "w") as f:
        json.dump(p, f, indent=1)
    return path


def run(dirpath, *args, timeout=300):
    ensure_built()
    return subprocess.run([DRIVER, *args], cwd=dirpath, capture_output=True, text=True,
                          timeout=timeout)
