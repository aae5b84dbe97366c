"""Pins the oracle's fluid–solid coupling (FluidSolidPfn,
src/PairwiseFunctors.jl:463-505) and linear-elastic stress update
(HookeLawStressUpdater, src/StateUpdaters.jl:337-394) against closed forms
and against FluidPfn, which FluidSolidPfn equals once the solid's pressure is
replaced by the fluid's.  The reference has no tests of either.  Nothing here
runs on a GPU.
"""
import math

import numpy as np
import pytest

from conftest import A, grasph, oracle_lib, oracle_scene

M = grasph
WEND = A.GR_KERNEL_WENDLAND_C2


@pytest.fixture(scope="module")
def orc():
    return oracle_lib()


# ---------------------------------------------------------------------------
# HookeLawStressUpdater: Δσ = D_e·ε̇·dt, Jaumann correction, p = −tr σ / 3
# ---------------------------------------------------------------------------


def _hooke(E, nu, stress, sr, vort, dt, ns=4):
    ps = M.ElastoPlasticParticleSystem(
        "e", 1, 2, ns, 1.0, 100.0, state_updater=(M.HookeLawStressUpdater(E, nu),))
    ps.stress[0] = stress
    ps.strain_rate[0] = sr
    ps.vorticity[0] = vort
    sc = oracle_scene(2)
    sc.upload(ps)
    sc.be.update_state(sc.system(ps), 1, dt)
    sc.download(ps)
    return ps


def test_hooke_uniaxial_increment_and_pressure():
    E, nu, dt = 5e8, 0.2, 1e-3
    ps = _hooke(E, nu, (0, 0, 0, 0), (0.1, 0, 0, 0), 0.0, dt)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))  # Lamé λ, μ: D_e = λ·I⊗I + 2μ·I
    mu = E / (2 * (1 + nu))
    de = 0.1 * dt
    assert ps.stress[0, 0] == pytest.approx((lam + 2 * mu) * de, rel=1e-14)
    assert ps.stress[0, 1] == pytest.approx(lam * de, rel=1e-14)
    assert ps.stress[0, 2] == pytest.approx(lam * de, rel=1e-14)
    assert ps.stress[0, 3] == 0.0
    # p = −tr σ / 3 = −(3λ + 2μ)·ε / 3 = −K·ε (bulk modulus K)
    K = E / (3 * (1 - 2 * nu))
    assert ps.p[0] == pytest.approx(-K * de, rel=1e-14)


def test_hooke_shear_uses_engineering_half():
    """Δσxy = D0·(1 − 2ν)·ε̇xy·dt = 2μ·ε̇xy·dt (tensor shear component)."""
    E, nu, dt = 2e7, 0.3, 1e-4
    ps = _hooke(E, nu, (0, 0, 0, 0), (0, 0, 0, 3.0), 0.0, dt)
    mu = E / (2 * (1 + nu))
    assert ps.stress[0, 3] == pytest.approx(2 * mu * 3.0 * dt, rel=1e-14)
    assert ps.stress[0, :3].tolist() == [0.0, 0.0, 0.0]
    assert ps.p[0] == 0.0


def test_hooke_jaumann_rotation():
    """With no strain rate only the spin terms act: dσxx = −2σxy·ωdt,
    dσyy = +2σxy·ωdt, dσxy = (σxx − σyy)·ωdt; the trace is unchanged."""
    s0 = (1e5, -2e4, 3e3, 4e4)
    w, dt = 10.0, 1e-3
    ps = _hooke(1e6, 0.3, s0, (0, 0, 0, 0), w, dt)
    wdt = w * dt
    assert ps.stress[0, 0] == pytest.approx(s0[0] - 2 * s0[3] * wdt, rel=1e-14)
    assert ps.stress[0, 1] == pytest.approx(s0[1] + 2 * s0[3] * wdt, rel=1e-14)
    assert ps.stress[0, 2] == s0[2]
    assert ps.stress[0, 3] == pytest.approx(s0[3] + (s0[0] - s0[1]) * wdt, rel=1e-14)
    assert ps.p[0] == pytest.approx(-(s0[0] + s0[1] + s0[2]) / 3, rel=1e-14)


def test_hooke_has_no_yield():
    """Linear in the strain increment however large (the elasto-plastic
    updater would return such a state to its yield surface)."""
    a = _hooke(1e6, 0.3, (0, 0, 0, 0), (50.0, -20.0, 0, 30.0), 0.0, 1.0)
    b = _hooke(1e6, 0.3, (0, 0, 0, 0), (100.0, -40.0, 0, 60.0), 0.0, 1.0)
    assert np.allclose(b.stress[0], 2 * a.stress[0], rtol=1e-14, atol=0)
    assert abs(a.stress[0, 3]) > 1e7


def test_hooke_needs_four_stress_components():  # StateUpdaters.jl:369
    sc = oracle_scene(2)
    ps = M.ElastoPlasticParticleSystem(
        "e", 1, 2, 3, 1.0, 100.0, state_updater=(M.HookeLawStressUpdater(1e6, 0.3),))
    with pytest.raises((AssertionError, ValueError)):
        sc.upload(ps)
        sc.system(ps)


# ---------------------------------------------------------------------------
# FluidSolidPfn: FluidPfn with the fluid's pressure on both sides
# ---------------------------------------------------------------------------


def _pair_sweep(fluid, solid, pfn, h):
    sc = oracle_scene(fluid.ndims)
    for ps in (fluid, solid):
        sc.upload(ps)
        sc.be.sort(sc.system(ps), 2 * h)
    k = M.WenlandC2Kernel(h, ndims=fluid.ndims)
    hh = sc.interaction(M.SystemInteraction(k, pfn, fluid, solid))
    sc.be.create_grid(hh)
    sc.be.sweep(hh, 1)
    for ps in (fluid, solid):
        sc.download(ps)
        o = np.argsort(ps.id)
        for f in ("x", "v", "rho", "p", "dvdt", "drhodt", "id"):
            getattr(ps, f)[:] = getattr(ps, f)[o]
    return fluid, solid


def _solid(n, D, mass, c):
    return M.ElastoPlasticParticleSystem("s", n, D, 4, mass, c)


def test_fluid_solid_two_particles_at_rest(orc):
    """At rest with α = β = 0 only the pressure term acts, with p_f on both
    sides: a_f = −m_s·(p_f/ρ_f² + p_f/ρ_s²)·∇W, a_s = +m_f·(…)·∇W."""
    h, r = 0.1, 0.07
    pf, rf, rs = 2000.0, 1000.0, 2400.0
    mf, ms = 0.3, 0.72
    fluid = M.FluidParticleSystem("f", 1, 2, mf, 10.0)
    fluid.x[:] = [(0.5, 0.5)]
    fluid.rho[:] = rf
    fluid.p[:] = pf
    solid = _solid(1, 2, ms, 50.0)
    solid.x[:] = [(0.5 + r, 0.5)]
    solid.rho[:] = rs
    solid.p[:] = 7.5e5  # the solid's own pressure takes no part
    _pair_sweep(fluid, solid, M.FluidSolidPfn(0.0, 0.0, h), h)
    dwdr = orc.orc_kernel_dw_dq(WEND, 2, h, r / h) / h
    coef = pf / rf ** 2 + pf / rs ** 2
    assert fluid.dvdt[0, 0] == pytest.approx(-ms * coef * abs(dwdr), rel=1e-12)
    assert solid.dvdt[0, 0] == pytest.approx(mf * coef * abs(dwdr), rel=1e-12)
    assert fluid.dvdt[0, 1] == 0.0 and solid.dvdt[0, 1] == 0.0
    assert fluid.drhodt[0] == 0.0 and solid.drhodt[0] == 0.0


def test_fluid_solid_each_pair_uses_its_fluid_pressure(orc):
    """One solid particle among fluid particles of distinct pressures, all at
    rest: a_s = Σ_f m_f·(p_f/ρ_f² + p_f/ρ_s²)·|∇W_fs|·(x_s − x_f)/r."""
    h = 0.1
    rs, ms = 2400.0, 0.5
    xs = np.array([0.5, 0.5])
    xf = xs + np.array([(0.06, 0.0), (-0.03, 0.05), (0.0, -0.08), (-0.07, -0.02)])
    pf = np.array([1500.0, -300.0, 4200.0, 800.0])
    rf = np.array([1000.0, 990.0, 1010.0, 1003.0])
    fluid = M.FluidParticleSystem("f", 4, 2, 0.25, 10.0)
    fluid.x[:] = xf
    fluid.rho[:] = rf
    fluid.p[:] = pf
    solid = _solid(1, 2, ms, 50.0)
    solid.x[:] = [xs]
    solid.rho[:] = rs
    solid.p[:] = -9e5
    _pair_sweep(fluid, solid, M.FluidSolidPfn(0.0, 0.0, h), h)
    exp = np.zeros(2)
    for k in range(4):
        d = xs - xf[k]
        r = math.hypot(*d)
        dwdr = orc.orc_kernel_dw_dq(WEND, 2, h, r / h) / h
        exp += 0.25 * (pf[k] / rf[k] ** 2 + pf[k] / rs ** 2) * abs(dwdr) * d / r
    assert np.allclose(solid.dvdt[0], exp, rtol=1e-12, atol=0)


@pytest.mark.parametrize("delta", [None, 0.1])
def test_fluid_solid_equals_fluid_pfn_with_fluid_pressure(delta):
    """With every fluid particle at one pressure p0, FluidSolidPfn equals the
    two-sided fluid–fluid FluidPfn (PairwiseFunctors.jl:172-190) with the
    solid replaced by a fluid system of the same mass, sound speed and state
    at pressure p0 (viscosity, continuity and density diffusion are shared).
    FluidPfn on a fluid–solid pair would take the one-sided generic form
    instead, which writes nothing to the solid."""
    rng = np.random.default_rng(11)
    h, dx = 0.13, 0.1
    g = np.array([(i * dx, j * dx) for i in range(8) for j in range(8)], float)
    xf = g + rng.uniform(-0.02, 0.02, g.shape)
    xs = g[:24] + np.array([0.05, 0.8]) + rng.uniform(-0.02, 0.02, (24, 2))
    p0 = 1234.5

    def make(p_solid, as_fluid=False):
        fluid = M.FluidParticleSystem("f", len(xf), 2, 0.01, 12.0)
        fluid.x[:] = xf
        fluid.v[:] = rng_v[0]
        fluid.rho[:] = rho_f
        fluid.p[:] = p0
        solid = (M.FluidParticleSystem("s", len(xs), 2, 0.024, 40.0) if as_fluid
                 else _solid(len(xs), 2, 0.024, 40.0))
        solid.x[:] = xs
        solid.v[:] = rng_v[1]
        solid.rho[:] = rho_s
        solid.p[:] = p_solid
        return fluid, solid

    rng_v = (rng.normal(0, 0.3, xf.shape), rng.normal(0, 0.1, xs.shape))
    rho_f = rng.uniform(990, 1010, len(xf))
    rho_s = rng.uniform(2390, 2410, len(xs))
    fa, sa = _pair_sweep(*make(rng.uniform(-1e6, 1e6, len(xs))),
                         M.FluidSolidPfn(0.1, 0.05, h, delta=delta), h)
    fb, sb = _pair_sweep(*make(p0, as_fluid=True), M.FluidPfn(0.1, 0.05, h, delta=delta), h)
    assert np.any(sa.dvdt != 0)
    for a, b in ((fa, fb), (sa, sb)):
        assert np.allclose(a.dvdt, b.dvdt, rtol=1e-13, atol=1e-13 * np.abs(b.dvdt).max())
        assert np.allclose(a.drhodt, b.drhodt, rtol=1e-13,
                           atol=1e-13 * np.abs(b.drhodt).max())


def test_dambreak_wall_scene_keeps_hooke_pressure():
    """DambreakWall.jl's scene (fluid, Hooke wall, four dynamic boundaries)
    steps in the oracle with the wall's p = −tr σ / 3 after every stage and
    the wall under load from its own weight (σyy < 0 at its foot)."""
    from conftest import oracle_scene as osc
    cfg = M.configs.dambreak_wall_2d()
    lf = cfg["integrator"]
    sc = osc(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfg["cfl"])
    sc.be.step(h, 50, dt)
    sc.download_all(lf)
    w = cfg["wall"]
    assert np.all(np.isfinite(w.stress)) and np.all(np.isfinite(cfg["fluid"].x))
    assert np.allclose(w.p, -w.stress[:, :3].sum(axis=1) / 3, rtol=1e-13,
                       atol=1e-13 * np.abs(w.stress).max())
    foot = w.x[:, 1] < 1.0
    assert np.all(w.stress[foot, 1] < 0)
