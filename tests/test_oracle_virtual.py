"""Virtual particle boundaries in the oracle (src/Particles.jl:476-528,
src/PairwiseFunctors.jl:405-434, src/StateUpdaters.jl:58-92,
src/TimeIntegration.jl:373-397,401-419): interpolation weights and values
against sums computed here from the kernel closed form, normalisation with
v_mult and prescribed_v, zero weight → zero, fields outside zero_fields
accumulating across steps, w_sum untouched when accumulate_wsum = false,
prescribed motion and the virtual sort."""
import numpy as np
import pytest

from conftest import grasph as M, oracle_lib, oracle_scene

H = 0.05
WENDLAND = 1


def w_of(r):
    lib = oracle_lib()
    return lib.orc_kernel_w(WENDLAND, 2, H, r / H)


def scene(zero_fields=("v", "rho"), interp2=None, pv=(0.0, 0.0), vnorm=("v", "rho", "p"),
          shuffle_virtual=False):
    real = M.FluidParticleSystem("real", 3, 2, 2.0, 10.0)
    real.x[:] = [[0.00, 0.00], [0.03, 0.01], [0.05, -0.02]]
    real.v[:] = [[1.0, 2.0], [-0.5, 0.25], [0.75, -1.5]]
    real.rho[:] = [1000.0, 1010.0, 990.0]
    real.p[:] = [100.0, 250.0, -50.0]
    src = M.FluidParticleSystem("vsrc", 3, 2, 2.0, 10.0)
    vx = np.array([[0.02, -0.03], [0.01, 0.04], [1.0, 1.0]])  # the last sees no real particle
    order = [2, 0, 1] if shuffle_virtual else [0, 1, 2]
    src.x[:] = vx[order]
    src.rho[:] = 1.0
    vps = M.VirtualParticleSystem(
        src, "virt", 3, 2, 2.0, 10.0, zero_fields=zero_fields, prescribed_v=pv,
        state_updater=(None, M.VirtualNormUpdater((-1.0, 0.5), *vnorm)))
    kern = M.WenlandC2Kernel(H, ndims=2)
    pf2 = interp2 if interp2 is not None else None
    si = M.SystemInteraction(kern, (M.InterpolateFieldFn("v", "rho", "p"), pf2), real, vps)
    lf = M.LeapFrogTimeIntegrator([real], [si], virtual_systems=(vps,))
    return real, vps, lf, vx


def run(lf, nsteps, dt=1e-4):
    sc = oracle_scene(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    sc.be.step(h, nsteps, dt)
    sc.download_all(lf)
    return sc


def shepard(real, xv):
    ws, fv, fr, fp = [], [], [], []
    for x in xv:
        kw = np.array([w_of(np.linalg.norm(x - xi)) * (real.mass / ri) if np.linalg.norm(x - xi) < 2 * H
                       else 0.0 for xi, ri in zip(real.x, real.rho)])
        ws.append(kw.sum())
        fv.append(kw @ real.v)
        fr.append(kw @ real.rho)
        fp.append(kw @ real.p)
    return np.array(ws), np.array(fv), np.array(fr), np.array(fp)


def test_interpolation_and_normalisation():
    real, vps, lf, vx = scene()
    x0, v0, r0, p0 = real.x.copy(), real.v.copy(), real.rho.copy(), real.p.copy()
    run(lf, 1)
    ref = M.FluidParticleSystem("r", 3, 2, 2.0, 10.0)
    ref.x[:], ref.v[:], ref.rho[:], ref.p[:] = x0, v0, r0, p0
    ws, fv, fr, fp = shepard(ref, vx)
    assert ws[2] == 0.0 and ws[0] > 0 and ws[1] > 0
    o = np.argsort(vps.id)
    np.testing.assert_allclose(vps.w_sum[o], ws, rtol=1e-14, atol=0)
    vm = np.array([-1.0, 0.5])
    for k in (0, 1):
        np.testing.assert_allclose(vps.v[o][k], fv[k] / ws[k] * vm, rtol=1e-13)
        assert vps.rho[o][k] == pytest.approx(fr[k] / ws[k], rel=1e-14)
        assert vps.p[o][k] == pytest.approx(fp[k] / ws[k], rel=1e-13)
    assert np.all(vps.v[o][2] == 0.0) and vps.rho[o][2] == 0.0 and vps.p[o][2] == 0.0


def test_fields_outside_zero_fields_accumulate():
    # p is interpolated but neither zeroed nor normalised: it grows by the raw
    # sum every step
    real, vps, lf, vx = scene(zero_fields=("v", "rho"), vnorm=("v", "rho"))
    dt = 1e-4
    run(lf, 2, dt)
    r1 = _ref()
    _, _, _, fp1 = shepard(r1, vx)           # step 1: initial positions
    r1.x[:] = r1.x + dt * r1.v                # no forces: the reals drift at v
    ws2, _, _, fp2 = shepard(r1, vx)         # step 2
    o = np.argsort(vps.id)
    np.testing.assert_allclose(vps.p[o][:2], fp1[:2] + fp2[:2], rtol=1e-13)
    np.testing.assert_allclose(vps.w_sum[o], ws2, rtol=1e-13)  # w_sum is re-zeroed


def _ref():
    real, _, _, _ = scene()
    return real


def test_wsum_not_accumulated_when_disabled():
    real, vps, lf, vx = scene(interp2=M.InterpolateFieldFn("v", accumulate_wsum=False))
    run(lf, 1)
    ws, fv, fr, fp = shepard(_ref(), vx)
    o = np.argsort(vps.id)
    np.testing.assert_allclose(vps.w_sum[o], ws, rtol=1e-14)
    # stage 2 adds the raw v sum on top of the normalised value
    vm = np.array([-1.0, 0.5])
    np.testing.assert_allclose(vps.v[o][:2], (fv[:2] / ws[:2, None]) * vm + fv[:2], rtol=1e-13)


def test_prescribed_velocity_moves_and_offsets():
    pv = (0.2, -0.1)
    real, vps, lf, vx = scene(pv=pv)
    dt = 1e-4
    run(lf, 3, dt)
    o = np.argsort(vps.id)
    exp = vx.copy()
    for _ in range(3):
        exp = exp + np.array(pv) * dt
    np.testing.assert_allclose(vps.x[o], exp, rtol=0, atol=1e-15)
    # the far particle has w_sum = 0: v = 0·v_mult + prescribed_v
    np.testing.assert_array_equal(vps.v[o][2], np.array(pv))


def test_virtual_systems_are_sorted():
    real, vps, lf, vx = scene(shuffle_virtual=True)
    run(lf, 1)
    keys = np.floor(vps.x / (2 * H)).astype(np.int64)
    order = np.lexsort((keys[:, 1], keys[:, 0]))
    assert np.array_equal(order, np.arange(3))  # stored in sorted order
    assert sorted(vps.id.tolist()) == [1, 2, 3]


def test_virtual_column_runs():
    cfg = M.configs.ep_column_virtual_2d(dx=0.01, width=0.1, height=0.06)
    lf = cfg["integrator"]
    sc = oracle_scene(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfg["cfl"])
    sc.be.step(h, 20, dt)
    sc.download_all(lf)
    f = cfg["fluid"]
    b, left = cfg["virtual"]
    assert np.all(np.isfinite(f.stress)) and np.all(np.isfinite(f.v))
    assert np.any(b.w_sum > 0) and np.any(left.w_sum > 0)
    # dt uses the real systems only (src/TimeIntegration.jl:63-64)
    assert dt == pytest.approx(cfg["cfl"] * lf.h / f.c, rel=1e-15)
