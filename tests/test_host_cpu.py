"""The C++ host modules on the CPU: param.json reading with the reference's
sections and checks (src/mod_parameters.cc:160-260), and the host mesh of
initMesh (src/mod_mesh.cc:13-125), through `pll_channel --check`."""
import json

import numpy as np
import pytest

import driver_util as du


def _check(tmp_path, p):
    du.write(tmp_path, p)
    return du.run(tmp_path, "--check", "param.json")


def test_check_reports_parameters_and_mesh(tmp_path, oracle_lib):
    r = _check(tmp_path, du.param())
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    m = du.BASE["MESH"]
    assert (out["nx"], out["ny"], out["nz"]) == (m["nx"], m["ny"], m["nz"])
    assert out["re_inv"] == 1.0 * 1.0 / 2800
    assert out["dx"] == m["lx"] / m["nx"] and out["dy"] == m["ly"] / m["ny"]
    dzf = oracle_lib.mesh_by_func(1, 0, 1.5, m["nz"], m["lz"])
    np.testing.assert_array_equal(np.array(out["dzf_global"]), dzf)
    o = oracle_lib.Oracle(m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"], 2800.0, 2e-3, dzf)
    mesh = o.mesh()
    np.testing.assert_array_equal(np.array(out["zc_global"]), mesh["zc_global"])
    np.testing.assert_array_equal(np.array(out["dzflzi"]), mesh["dzflzi"])
    np.testing.assert_array_equal(np.array(out["dzc"])[:-1], mesh["dzc"][:-1])
    assert out["precision"] == 0 and out["fused_step"] == 1


def test_read_mesh_from_mesh_in(tmp_path):
    nz = du.BASE["MESH"]["nz"]
    dz = np.linspace(0.05, 0.2, nz)
    (tmp_path / "mesh.in").write_text("\n".join(repr(float(x)) for x in dz))
    r = _check(tmp_path, du.param(MESH={"read_mesh": True}))
    assert r.returncode == 0, r.stderr
    got = np.array(json.loads(r.stdout)["dzf_global"])
    np.testing.assert_array_equal(got[1:-1], dz)
    assert got[0] == dz[0] and got[-1] == dz[-1]


def test_device_section(tmp_path):
    r = _check(tmp_path, du.param(DEVICE={"precision": "fp32_mixed", "fused_step": False,
                                          "closure": "les.closure"}))
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["precision"] == 1 and out["fused_step"] == 0 and out["closure"] == "les.closure"


@pytest.mark.parametrize("sections, message", [
    (dict(TIME=None), "Problem with TIME in the input file!"),
    (dict(MESH={"nx": 31}), "<nx>, <ny>, <nz> should be even integers!"),
    (dict(PARA={"p_row": 2}), "<p_row> and <p_col> must be 1"),
    (dict(MESH={"nx": 32.0}), "<nx> must be an integer"),
    (dict(PHYSICAL={"init_with_noise": 1}), "<init_with_noise> must be true or false"),
    (dict(DEVICE={"precision": "fp16"}), "<precision> must be"),
])
def test_parameter_errors(tmp_path, sections, message):
    r = _check(tmp_path, du.param(**sections))
    assert r.returncode == 2
    assert message in r.stdout + r.stderr


def test_unreadable_file(tmp_path):
    (tmp_path / "param.json").write_text("{\"PARA\": {\"p_row\": 1,,}}")
    r = du.run(tmp_path, "--check", "param.json")
    assert r.returncode == 2 and "readInputParam" in r.stderr
    r = du.run(tmp_path, "--check", "missing.json")
    assert r.returncode == 2 and "Fail to open file missing.json" in r.stderr


def test_mesh_files(tmp_path, oracle_lib):
    """outputMesh (src/mod_mesh.cc:127-145): mesh_<nx>-<ny>-<nz>.h5 holds the
    interior cell centres xc, yc, zc; mesh_<nz>.out lists k, dzf, zc."""
    import h5lite
    if not h5lite.available():
        pytest.skip("no HDF5 library")
    r = du.run(tmp_path, "--mesh", du.write(tmp_path, du.param()))
    assert r.returncode == 0, r.stderr
    m = du.BASE["MESH"]
    nx, ny, nz = m["nx"], m["ny"], m["nz"]
    name = f"mesh_{nx}-{ny}-{nz}.h5"
    assert f"Finish writing file {name}" in r.stdout
    with h5lite.read(str(tmp_path / name)) as f:
        xc, yc, zc = (f.dataset(t) for t in ("xc", "yc", "zc"))
    dx, dy = m["lx"] / nx, m["ly"] / ny
    np.testing.assert_allclose(xc, (np.arange(1, nx + 1) - 0.5) * dx, rtol=0, atol=1e-14)
    np.testing.assert_allclose(yc, (np.arange(1, ny + 1) - 0.5) * dy, rtol=0, atol=1e-14)
    dzf = oracle_lib.mesh_by_func(1, 0, m["stretch_ratio"], nz, m["lz"])
    o = oracle_lib.Oracle(nx, ny, nz, m["lx"], m["ly"], m["lz"], 2800.0, 2e-3, dzf)
    np.testing.assert_array_equal(zc, o.mesh()["zc_global"][1:nz + 1])
    rows = (tmp_path / f"mesh_{nz:05d}.out").read_text().splitlines()
    assert len(rows) == nz + 1
    k, dz, z = rows[-1].split()
    assert int(k) == nz and abs(float(dz) - dzf[nz]) < 1e-14 and abs(float(z) - zc[-1]) < 1e-14
