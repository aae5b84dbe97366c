"""ProbeParticleSystem in the oracle and the host mirror (src/ProbeParticles.jl;
_measure_probes!, src/TimeIntegration.jl:454-502; NeighborCountFn and the
probe method of InterpolateFieldFn, src/PairwiseFunctors.jl:436-460).  The
cases of test/test_probe_particles.jl (construction, auto-zero, sort by id,
counts on a line, constant-field recovery, prescribed motion, HDF5 round
trip) plus counts and Shepard sums computed here by brute force."""
import numpy as np
import pytest

import types

from conftest import grasph as M, oracle_lib, oracle_scene

SN = M.snapshot


def line_fluid(n=8, dx=0.1, rho0=1000.0, c=100.0):
    """_make_line_fluid (test/test_probe_particles.jl:10-22)."""
    h = 1.2 * dx
    ps = M.FluidParticleSystem("fluid", n, 2, rho0 * dx * dx, c,
                               state_updater=M.TaitEOSUpdater(rho0))
    ps.x[:, 0] = np.arange(n) * dx
    ps.rho[:] = rho0
    return ps, M.CubicSplineKernel(h, ndims=2), h


def w_cubic(r, h):
    return oracle_lib().orc_kernel_w(0, 2, h, r / h) if r < 2 * h else 0.0


def run(lf, nsteps, cfl=0.1, measure=True):
    sc = oracle_scene(lf.ndims)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfl)
    sc.be.step(h, nsteps, dt)
    if measure:
        sc.measure_probes(lf)
    sc.download_all(lf)
    return dt


def test_construction():
    pos = [(0.1 * i, 0.0) for i in range(1, 6)]
    p = M.ProbeParticleSystem("p", pos, extras={"nbr": np.zeros(5, dtype=np.int64)})
    assert p.n == 5 and p.ndims == 2 and np.array_equal(p.id, np.arange(1, 6))
    assert np.allclose(p.x, np.array(pos)) and np.all(p.w_sum == 0) and np.all(p.nbr == 0)
    ps, _, _ = line_fluid(n=6)
    pm = M.ProbeParticleSystem("pm", ps, extras={"cnt": np.zeros(6, dtype=np.int64)})
    assert pm.mirror_target is ps and np.array_equal(pm.x, ps.x)
    assert pm.mass == ps.mass and pm.c == ps.c  # forwarded to the mirror target
    with pytest.raises(AttributeError):
        p.mass  # no mirror target
    with pytest.raises(ValueError, match="length"):
        M.ProbeParticleSystem("p", pos, extras={"cnt": np.zeros(3)})
    with pytest.raises(ValueError, match="empty"):
        M.ProbeParticleSystem("p", [])


def test_neighbour_counts_on_a_line():
    """test_probe_particles.jl:259-295 plus the exact counts: probe i sees
    the source particles within 2h of x_i."""
    ps, kern, h = line_fluid(n=8)
    probe = M.ProbeParticleSystem("probe", ps, extras={"nbr": np.zeros(8, dtype=np.int64)})
    si = M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, h), ps)
    sp = M.SystemInteraction(kern, M.NeighborCountFn("nbr"), ps, probe)
    lf = M.LeapFrogTimeIntegrator([ps], [si], probes=(probe,), probe_interactions=(sp,))
    run(lf, 1)
    nbr = probe.nbr
    assert nbr[0] < nbr[3] and nbr[7] < nbr[3] and nbr.max() in (nbr[3], nbr[4])
    # brute force on the mirrored (current) positions, probe slot = particle id
    xs = np.empty_like(ps.x)
    xs[ps.id - 1] = ps.x
    exact = [(np.linalg.norm(xs - xs[i], axis=1) < 2 * h).sum() for i in range(8)]
    assert np.array_equal(nbr, exact)
    assert np.array_equal(probe.id, np.arange(1, 9))
    np.testing.assert_array_equal(probe.x, xs)


def test_constant_field_recovered_and_shepard_sum():
    ps, kern, h = line_fluid(n=8)
    ps.v[:] = [1.0, 0.5]
    px = [(0.35, 0.0), (0.45, 0.0), (0.55, 0.0)]
    probe = M.ProbeParticleSystem("probe", px, extras={"v": np.zeros((3, 2))},
                                  state_updater=M.VirtualNormUpdater((1.0, 1.0), "v"))
    si = M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, h), ps)
    sp = M.SystemInteraction(kern, M.InterpolateFieldFn("v"), ps, probe)
    lf = M.LeapFrogTimeIntegrator([ps], [si], probes=(probe,), probe_interactions=(sp,))
    run(lf, 1)
    assert np.all(np.abs(probe.v[:, 0] - 1.0) < 0.01) and np.all(np.abs(probe.v[:, 1] - 0.5) < 0.01)
    # the exact normalised Shepard sum at the probes, from the final source state
    for k, xp in enumerate(np.array(px)):
        w = np.array([w_cubic(np.linalg.norm(xp - xi), h) * ps.mass / ri
                      for xi, ri in zip(ps.x, ps.rho)])
        assert probe.w_sum[k] == pytest.approx(w.sum(), rel=1e-14)
        np.testing.assert_allclose(probe.v[k], (w @ ps.v) / w.sum(), rtol=1e-13)


def test_scalar_rho_roundtrip():
    ps, kern, h = line_fluid(n=6)
    ps.rho[:] = 1234.0
    probe = M.ProbeParticleSystem("probe", [(0.2, 0.0), (0.3, 0.0)],
                                  extras={"rho": np.zeros(2)},
                                  state_updater=M.VirtualNormUpdater((1.0, 1.0), "rho"))
    si = M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, h), ps)
    sp = M.SystemInteraction(kern, M.InterpolateFieldFn("rho"), ps, probe)
    lf = M.LeapFrogTimeIntegrator([ps], [si], probes=(probe,), probe_interactions=(sp,))
    run(lf, 1)
    assert np.all(np.abs(probe.rho - 1234.0) < 10.0)


def test_prescribed_velocity_moves_probes_every_step():
    ps, kern, h = line_fluid(n=4)
    probe = M.ProbeParticleSystem("probe", [(0.0, 0.0)], extras={"cnt": np.zeros(1)},
                                  prescribed_v=(1.0, 0.0))
    si = M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, h), ps)
    sp = M.SystemInteraction(kern, M.NeighborCountFn("cnt"), ps, probe)
    lf = M.LeapFrogTimeIntegrator([ps], [si], probes=(probe,), probe_interactions=(sp,))
    dt = run(lf, 3, measure=False)
    assert probe.x[0, 0] == pytest.approx(3 * dt, abs=1e-12) and probe.x[0, 1] == 0.0


def test_zero_and_sort_by_id_after_measurement():
    """A second measurement starts from zero (auto_zero_probe!) and the
    probes come back in id order however the cell sort permuted them."""
    rng = np.random.default_rng(3)
    ps, kern, h = line_fluid(n=8)
    pos = rng.uniform(0.0, 0.7, size=(6, 2))
    pos[:, 1] = rng.uniform(-0.1, 0.1, size=6)
    probe = M.ProbeParticleSystem("probe", pos, extras={"cnt": np.zeros(6)})
    si = M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, h), ps)
    sp = M.SystemInteraction(kern, M.NeighborCountFn("cnt"), ps, probe)
    lf = M.LeapFrogTimeIntegrator([ps], [si], probes=(probe,), probe_interactions=(sp,))
    sc = oracle_scene(2)
    hh = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(hh, 0.1)
    sc.be.step(hh, 1, dt)
    sc.measure_probes(lf)
    first = probe.cnt.copy()
    sc.measure_probes(lf)
    sc.download_all(lf)
    np.testing.assert_array_equal(probe.cnt, first)
    np.testing.assert_array_equal(probe.x, pos)
    np.testing.assert_array_equal(probe.id, np.arange(1, 7))
    exact = [(np.linalg.norm(ps.x - p, axis=1) < 2 * h).sum() for p in pos]
    np.testing.assert_array_equal(probe.cnt, exact)


def test_time_integrate_writes_probe_groups(tmp_path):
    """test_probe_particles.jl:296-317 through the host driver: the snapshot
    holds the probe group (x, id, extras; no w_sum), read back with the
    probe reader (id reset to 1..n)."""
    ps, kern, h = line_fluid(n=5)
    probe = M.ProbeParticleSystem("probe", ps, extras={"nbr": np.zeros(5, dtype=np.int64)})
    si = M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, h), ps)
    sp = M.SystemInteraction(kern, M.NeighborCountFn("nbr"), ps, probe)
    lf = M.LeapFrogTimeIntegrator([ps], [si], probes=(probe,), probe_interactions=(sp,))
    sc = oracle_scene(2)
    sess = types.SimpleNamespace(backend=sc.be, scene=sc)
    M.time_integrate(lf, 1, 1, 1, 0.1, str(tmp_path / "out"), print_timer=False, session=sess)
    back = M.ProbeParticleSystem("probe", [(9.0, 9.0)] * 5,
                                 extras={"nbr": np.zeros(5, dtype=np.int64)})
    back.id[:] = [5, 4, 3, 2, 1]
    h5 = SN._lib()
    f = h5.H5Fopen(SN._bytes(str(tmp_path / "out_1.h5")), SN._F_ACC_RDONLY, SN._P_DEFAULT)
    try:
        g = h5.H5Gopen2(f, b"probe", SN._P_DEFAULT)
        assert not SN._exists(g, "w_sum")
        SN.read_group(g, back)
        h5.H5Gclose(g)
    finally:
        h5.H5Fclose(f)
    nbr = back.nbr
    assert nbr[0] == nbr[4] and nbr[1] == nbr[3] and nbr[2] >= nbr[1]
    np.testing.assert_array_equal(nbr, probe.nbr)
    np.testing.assert_array_equal(back.x, probe.x)
    np.testing.assert_array_equal(back.id, np.arange(1, 6))
