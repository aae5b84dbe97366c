"""RK4 time integration in the oracle (src/TimeIntegration.jl:712-838).

Pins: a damped particle (dv/dt = −(Γ/dt)·v, no other force) is multiplied
every step by the classical RK4 stability polynomial R(−Γ) = 1 − Γ + Γ²/2 −
Γ³/6 + Γ⁴/24, which fixes the advance coefficients (0, ½, ½, 1) and the
weights (1/6, 1/3, 1/3, 1/6); free fall gives v = n·dt·g and
x = x0 + g·dt²·n(n+1)/2 (positions move once per step with the new
velocity); and a small dam break stays close to the leapfrog trajectory."""
import numpy as np
import pytest

from conftest import grasph as M, oracle_scene


def lone(gamma=0.0, g=0.0, v0=(1.5, -0.5), cls=M.RK4TimeIntegrator):
    ps = M.FluidParticleSystem("p", 1, 2, 1.0, 10.0, source_v=[0.0, g])
    ps.x[0] = (0.1, 0.2)
    ps.v[0] = v0
    ps.rho[0] = 1000.0
    si = M.SystemInteraction(M.WenlandC2Kernel(0.01, ndims=2), M.FluidPfn(0.1, 0.0, 0.01), ps)
    return ps, cls([ps], [si], gamma=gamma)


def run(lf, nsteps, dt):
    sc = oracle_scene(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    sc.be.step(h, nsteps, dt)
    sc.download_all(lf)


@pytest.mark.parametrize("gamma", [0.3, 0.05])
def test_damping_follows_rk4_polynomial(gamma):
    ps, lf = lone(gamma=gamma)
    v0 = ps.v[0].copy()
    run(lf, 5, 1e-3)
    z = -gamma
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(ps.v[0], v0 * R ** 5, rtol=1e-13)


def test_leapfrog_damping_differs_from_rk4():
    """The same damped particle under the leapfrog follows its own factor
    (1 − Γ) per step (kick-drift with the half-step velocity in the damping
    term: v1 = v0 − Γ·(v0 + 0) = v0·(1 − Γ)), so the two schemes are told
    apart."""
    ps, lf = lone(gamma=0.3, cls=M.LeapFrogTimeIntegrator)
    v0 = ps.v[0].copy()
    run(lf, 1, 1e-3)
    np.testing.assert_allclose(ps.v[0], v0 * (1 - 0.3), rtol=1e-14)


def test_free_fall():
    g, dt, n = -9.81, 1e-3, 7
    ps, lf = lone(g=g, v0=(0.0, 0.0))
    x0 = ps.x[0].copy()
    run(lf, n, dt)
    assert ps.v[0][1] == pytest.approx(n * dt * g, rel=1e-13)
    assert ps.x[0][1] == pytest.approx(x0[1] + g * dt * dt * n * (n + 1) / 2, rel=1e-12)
    assert ps.x[0][0] == x0[0]


def test_rk4_close_to_leapfrog_on_a_dam_break():
    out = {}
    for cls in (M.LeapFrogTimeIntegrator, M.RK4TimeIntegrator):
        cfg = M.configs.dambreak_2d(dx=0.04)
        lf0 = cfg["integrator"]
        lf = cls(lf0.systems, lf0.interactions)
        sc = oracle_scene(2)
        h = sc.integrator(lf)
        sc.upload_all(lf)
        dt = sc.be.integrator_dt(h, cfg["cfl"])
        sc.be.step(h, 20, dt)
        sc.download_all(lf)
        f = cfg["fluid"]
        o = np.argsort(f.id)
        out[cls.__name__] = (f.x[o].copy(), f.rho[o].copy())
    (xl, rl), (xr, rr) = out["LeapFrogTimeIntegrator"], out["RK4TimeIntegrator"]
    assert not np.array_equal(xl, xr)
    assert np.max(np.abs(xl - xr)) < 1e-3 * 0.04  # well inside a particle spacing
    assert np.max(np.abs(rl - rr)) / np.max(rl) < 5e-4  # acoustic start-up differs slightly
