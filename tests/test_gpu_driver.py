"""The C++ host driver (src/main.cc:24-240 on one GPU) end to end: param.json
in, checkpoint files out, compared with the oracle's run of the same steps;
restart from a checkpoint; the statistics files."""
import os

import numpy as np
import pytest

import driver_util as du
import h5lite

pytestmark = pytest.mark.gpu


def _oracle_fields(p, nsteps, sample_at=()):
    from oracle import pll_oracle as po
    m, ph, t = p["MESH"], p["PHYSICAL"], p["TIME"]
    dzf = po.mesh_by_func(m["mesh_type"], m["stretch_func"], m["stretch_ratio"], m["nz"], m["lz"])
    o = po.Oracle(m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"], float(ph["re"]), t["dt"],
                  dzf)
    u, v, w, u1, v1, w1, pp = (o.field() for _ in range(7))
    o.init_flow(u, v, w, ph["initial_field"], u0=ph["u0"], with_noise=ph["init_with_noise"],
                noise_intensity=ph["noise_intensity"], seeds=(123, 456, 789))
    o.update_bound_vel(u, v, w)
    o.update_bound_p(pp)
    samples = []
    for nt in range(1, nsteps + 1):
        o.step(1, u, v, w, u1, v1, w1, pp)
        if nt in sample_at:
            samples.append([a.copy() for a in (u, v, w, pp)])
    return (u, v, w), samples


def _read(path, tag):
    with h5lite.read(path) as f:
        return int(f.dataset("nt")[0]), f.dataset(tag)


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _run_ok(tmp_path, p):
    du.write(tmp_path, p)
    print(f"[driver] {tmp_path} DEVICE={p.get('DEVICE')}", flush=True)
    r = du.run(tmp_path, "param.json")
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Simulation ends successfully" in r.stdout
    return r


@pytest.mark.parametrize("fused", [True, False])
def test_driver_checkpoint_matches_oracle(tmp_path, fused):
    p = du.param(DEVICE={"fused_step": fused})
    r = _run_ok(tmp_path, p)
    # screen lines at nt = 5 and 10 with the divergence at round-off
    lines = [l.split() for l in r.stdout.splitlines() if l.strip().startswith(("5 ", "10 "))]
    assert len(lines) == 2 and all(float(l[-1]) < 1e-12 for l in lines), r.stdout
    (ou, ov, ow), _ = _oracle_fields(p, 10)
    for tag, ref in (("u", ou), ("v", ov), ("w", ow)):
        nt, got = _read(str(tmp_path / f"save_{tag}.h5"), tag)
        assert nt == 10
        assert got.shape == (16, 24, 32)
        assert _rel(got, ref[1:-1, 1:-1, 1:-1]) < 1e-11, tag


def test_driver_fp32_mixed_runs_close_to_oracle(tmp_path):
    # FP32 fields: the divergence sits at FP32 round-off (~1e-6 here), so the
    # run's div_limit is set for that precision
    p = du.param(DEVICE={"precision": "fp32_mixed"}, TIME={"div_limit": 1e-4})
    r = _run_ok(tmp_path, p)
    lines = [l.split() for l in r.stdout.splitlines() if l.strip().startswith(("5 ", "10 "))]
    assert len(lines) == 2 and all(float(l[-1]) < 1e-5 for l in lines), r.stdout
    (ou, ov, ow), _ = _oracle_fields(p, 10)
    for tag, ref in (("u", ou), ("v", ov), ("w", ow)):
        _, got = _read(str(tmp_path / f"save_{tag}.h5"), tag)
        assert np.max(np.abs(got - ref[1:-1, 1:-1, 1:-1])) < 1e-5, tag


def test_restart_continues_bit_for_bit(tmp_path):
    a, b = tmp_path / "straight", tmp_path / "restarted"
    a.mkdir()
    b.mkdir()
    _run_ok(a, du.param())
    _run_ok(b, du.param(TIME={"nt_end": 5}, OUTPUT={"nt_out_save": 5}))
    assert _read(str(b / "save_u.h5"), "u")[0] == 5
    _run_ok(b, du.param(RESTART={"is_restart": True}))
    for tag in "uvw":
        nta, ga = _read(str(a / f"save_{tag}.h5"), tag)
        ntb, gb = _read(str(b / f"save_{tag}.h5"), tag)
        assert nta == ntb == 10
        assert np.array_equal(ga, gb), tag


def test_restart_rejects_mismatched_timestamps(tmp_path):
    shape = (16, 24, 32)
    h5lite.write_field(str(tmp_path / "save_u.h5"), "u", 5, np.zeros(shape))
    h5lite.write_field(str(tmp_path / "save_v.h5"), "v", 6, np.zeros(shape))
    h5lite.write_field(str(tmp_path / "save_w.h5"), "w", 5, np.zeros(shape))
    du.write(tmp_path, du.param(RESTART={"is_restart": True}))
    r = du.run(tmp_path, "param.json")
    assert r.returncode == 2
    assert "The timestamp of <v> does not match with that of <u>!" in r.stderr


def test_statistics_files_match_oracle_samples(tmp_path):
    # samples at nt = 6, 8, 10 (nt_init_stat 4, every 2 steps), written at nt = 10
    p = du.param(STATISTICS={"nt_init_stat": 4, "sample_interval": 2},
                 OUTPUT={"nt_out_stat": 6})
    _run_ok(tmp_path, p)
    _, samples = _oracle_fields(p, 10, sample_at=(6, 8, 10))
    sums = np.zeros((11, 16, 24, 32))
    for u, v, w, pp in samples:
        uc = (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 1:-1]) * 0.5 + 0.0
        vc = (v[1:-1, :-2, 1:-1] + v[1:-1, 1:-1, 1:-1]) * 0.5
        wc = (w[:-2, 1:-1, 1:-1] + w[1:-1, 1:-1, 1:-1]) * 0.5
        pc = pp[1:-1, 1:-1, 1:-1]
        for t, q in enumerate((uc, vc, wc, uc * uc, vc * vc, wc * wc, uc * vc, uc * wc, vc * wc,
                               pc, pc * pc)):
            sums[t] += q
    names = ("u", "v", "w", "u2", "v2", "w2", "uv", "uw", "vw")
    for t, name in enumerate(names):  # velocity statistics; p carries an arbitrary constant
        path = str(tmp_path / f"stat_00000005-00000010_{name}.h5")
        with h5lite.read(path) as f:
            assert (int(f.dataset("nts")[0]), int(f.dataset("nte")[0]),
                    int(f.dataset("nspl")[0])) == (5, 10, 3)
            got = f.dataset(f"{name}_stat")
        assert _rel(got, sums[t]) < 1e-10, name
    prof = np.loadtxt(str(tmp_path / "stat_00000005-00000010_profiles.txt"))
    assert prof.shape == (16, 19)
    np.testing.assert_allclose(prof[:, 2], sums[0].mean(axis=(1, 2)) / 3, rtol=1e-10)


def test_driver_runs_with_a_saved_closure(tmp_path):
    """DEVICE.closure attaches a saved learned closure to the run's context."""
    import powerllel_channel_sycl_amd as pll
    cl = pll.Closure(seed=3)
    cl.set_scales(0.05, 2.0, 1.0 / 2800.0)
    cl.save(str(tmp_path / "les.closure"))
    cl.close()
    p = du.param(DEVICE={"closure": "les.closure"})
    r = _run_ok(tmp_path, p)
    assert "Learned closure attached from les.closure" in r.stdout
    lines = [l.split() for l in r.stdout.splitlines() if l.strip().startswith(("5 ", "10 "))]
    assert len(lines) == 2 and all(float(l[-1]) < 1e-12 for l in lines), r.stdout
    # the eddy viscosity changes the flow: the run differs from the plain one
    (ou, _, _), _ = _oracle_fields(p, 10)
    nt, got = _read(str(tmp_path / "save_u.h5"), "u")
    assert nt == 10 and _rel(got, ou[1:-1, 1:-1, 1:-1]) > 1e-9
