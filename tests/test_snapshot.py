"""HDF5 snapshots in the reference's layout (src/Particles.jl:711-787,
src/TimeIntegration.jl:520-549, src/Driver.jl:79-96): round trips as in
test/test_read_h5.jl, the on-disk layout read back by the HDF5 tools'
h5dump (independent of our reader), and restart from a snapshot continuing
bit-identically to an uninterrupted run (the oracle as the simulator)."""
import shutil
import subprocess

import numpy as np
import pytest

from conftest import grasph as M, oracle_scene

SN = M.snapshot


def filled(ps, seed):
    g = np.random.default_rng(seed)
    for f in ["x"] + SN.sim_field_names(ps):
        a = getattr(ps, f)
        if f == "id":
            a[:] = g.permutation(ps.n) + 1
        else:
            a[...] = g.uniform(-1, 1, a.shape)
    return ps


@pytest.mark.parametrize("make", [
    lambda: M.BasicParticleSystem("b", 7, 2, 1.5, 3.0),
    lambda: M.FluidParticleSystem("f", 9, 3, 0.5, 10.0),
    lambda: M.StressParticleSystem("s", 5, 2, 4, 2.0, 7.0),
    lambda: M.ElastoPlasticParticleSystem("e", 6, 3, 6, 2.0, 7.0),
], ids=["basic", "fluid", "stress", "ep"])
def test_write_and_read_back(tmp_path, make):  # test_read_h5.jl:8-88
    a = filled(make(), 1)
    p = tmp_path / "s.h5"
    SN.write_h5(a, str(p))
    b = make()
    SN.read_h5(b, str(p))
    for f in ["x"] + SN.sim_field_names(a):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_read_mismatch_errors(tmp_path):  # test_read_h5.jl:90-101
    a = filled(M.FluidParticleSystem("f", 9, 3, 0.5, 10.0), 2)
    p = str(tmp_path / "s.h5")
    SN.write_h5(a, p)
    with pytest.raises(ValueError, match="particle count"):
        SN.read_h5(M.FluidParticleSystem("f", 8, 3, 0.5, 10.0), p)
    with pytest.raises(ValueError, match="ndims"):
        SN.read_h5(M.FluidParticleSystem("f", 9, 2, 0.5, 10.0), p)


H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if shutil.os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_layout_seen_by_h5dump(tmp_path):
    """Groups, attribute types and dataset shapes as Julia's HDF5.jl writes
    them: (n, ndims) C-order vectors, (n,) scalars, Int64 ids and counts."""
    ps = filled(M.FluidParticleSystem("fluid", 4, 3, 0.5, 10.0), 3)
    p = str(tmp_path / "snap_0010.h5")
    SN.save_snapshot(p, [ps], [], 10, 0.25)
    out = subprocess.run([H5DUMP, "-H", p], capture_output=True, text=True, check=True).stdout
    flat = " ".join(out.split())
    assert 'ATTRIBUTE "step" { DATATYPE H5T_STD_I64LE DATASPACE SCALAR }' in flat
    assert 'ATTRIBUTE "sim_time" { DATATYPE H5T_IEEE_F64LE DATASPACE SCALAR }' in flat
    assert 'GROUP "fluid"' in flat
    for name, shape, ty in (("x", "( 4, 3 )", "H5T_IEEE_F64LE"), ("v", "( 4, 3 )", "H5T_IEEE_F64LE"),
                            ("dvdt", "( 4, 3 )", "H5T_IEEE_F64LE"), ("rho", "( 4 )", "H5T_IEEE_F64LE"),
                            ("p", "( 4 )", "H5T_IEEE_F64LE"), ("id", "( 4 )", "H5T_STD_I64LE")):
        assert f'DATASET "{name}" {{ DATATYPE {ty} DATASPACE SIMPLE {{ {shape} / {shape} }} }}' in flat, name
    vals = subprocess.run([H5DUMP, "-a", "/step", "-a", "/sim_time", p], capture_output=True,
                          text=True, check=True).stdout
    assert "(0): 10" in vals and "(0): 2.5" in vals
    x = subprocess.run([H5DUMP, "-d", "/fluid/x", "-m", "%.17g", p], capture_output=True,
                       text=True, check=True).stdout
    import re
    got = np.zeros((4, 3))
    for i, j, v in re.findall(r"\((\d+),(\d+)\):\s*([-+0-9.eE]+)", x):
        got[int(i), int(j)] = float(v)
    assert np.array_equal(got, ps.x)  # row i = particle i, columns = components


def _ep_cfg():
    return M.configs.ep_column_2d(dx=0.01, width=0.1, height=0.06)


def test_restart_continues_bit_identically(tmp_path):
    """Oracle: 12 steps straight vs 7 steps, snapshot (with ghosts), restart
    into fresh objects on a fresh context, 5 more steps."""
    def run(cfg, nsteps, sc=None):
        lf = cfg["integrator"]
        sc = sc or oracle_scene(2)
        h = sc.integrator(lf)
        sc.upload_all(lf)
        dt = sc.be.integrator_dt(h, cfg["cfl"])
        sc.be.step(h, nsteps, dt)
        sc.download_all(lf)
        return sc, dt
    ref = _ep_cfg()
    run(ref, 12)
    a = _ep_cfg()
    _, dt = run(a, 7)
    path = str(tmp_path / "out" / "col_07.h5")
    SN.save_snapshot(path, a["integrator"].systems, a["integrator"].ghosts, 7, dt)
    b = _ep_cfg()
    step = SN.load_restart(path, b["integrator"].systems)
    assert step == 7
    run(b, 5)
    for f in ["x"] + SN.sim_field_names(ref["fluid"]):
        assert np.array_equal(getattr(b["fluid"], f), getattr(ref["fluid"], f)), f
