"""Closed-form pins of the oracle's pair functors (src/PairwiseFunctors.jl,
src/PairwisePhysics.jl).

Every case is a two-particle scene whose expected rates are worked out by
hand for round numbers, chosen so that each plausible mistake changes the
answer by far more than the tolerance:

- artificial viscosity (PairwisePhysics.jl:29-34): the denominator is the
  product 0.5·ρi·ρj, not the mean density; the softening is 0.01·h², not h²;
  the sound speed is the mean ½(ci + cj); only approaching pairs count;
- δ-SPH diffusion (:238-243): the factor 2, the sign of ρi − ρj, the division
  by r² and by the *partner's* density on each side;
- artificial surface tension (:286-290): the absolute values of both pressures;
- Lennard-Jones wall force (:258-263): exponents 12 and 6, the 1/r² factor,
  0.01·c² with the fluid's c, zero beyond lj_cutoff;
- Cauchy stress force (:125-151): every Voigt index, checked through the full
  symmetric tensor times the gradient (a matrix product, not the component
  formulas);
- strain rate and spin (:169-220): symmetric and antisymmetric parts of
  dv ⊗ ∇W, again through the full tensor;
- dynamic-wall partners (PairwiseFunctors.jl:53-68, 222-242, 334-355):
  v_j = −min(β, |db/da|)·v_i, ρ_j = ρ_i, p_j = p_i, σ_j = σ_i and the *fluid's*
  mass, with the wall particle's own mass, density, pressure and velocity set
  to values that would show if they were read.

The kernel enters only through g = dW/dq(q)/(r·h), with dW/dq taken from the
oracle's kernel, which test_oracle_known_answers.py pins (norms, values,
derivative by finite differences, normalisation).

The ghost cases port test/test_ghost_updates.jl:58-132 (wall repulsion of
CauchyFluidPfn and FluidPfn through a mirrored ghost) and add their closed
forms.
"""
import math

import numpy as np
import pytest

from conftest import A, grasph, oracle_lib, oracle_scene

M = grasph
CUBIC = A.GR_KERNEL_CUBIC
H = 0.1
REL = 1e-12


@pytest.fixture(scope="module")
def orc():
    return oracle_lib()


def _g(orc, r, D, h=H):
    """∇W = g·dx with g = dW/dq(r/h)/(r·h) (src/Interaction.jl:546-553)."""
    return orc.orc_kernel_dw_dq(CUBIC, D, h, r / h) / (r * h)


def _sweep(pfn, a, b=None, ghost_entry=None, h=H):
    """One sweep of stage 1 on the oracle; results are mapped back to the
    order the particles were given in (by id)."""
    D = a.ndims
    sc = oracle_scene(D)
    sc.upload(a)
    partner = b
    if ghost_entry is not None:
        gh = sc.ghost(ghost_entry)
        sc.be.generate_ghosts(gh)
        sc.be.update_ghost_kinematics(gh)
        sc.be.update_ghost(gh, 1)
        partner = ghost_entry.ghost
    elif b is not None:
        sc.upload(b)
    si = M.SystemInteraction(M.CubicSplineKernel(h, ndims=D), pfn, a, partner)
    hh = sc.interaction(si)
    sc.be.sort(sc.system(a), 2 * h)
    if b is not None:
        sc.be.sort(sc.system(b), 2 * h)
    sc.be.create_grid(hh)
    sc.be.sweep(hh, 1)
    out = []
    for ps in (a, b):
        if ps is None:
            continue
        sc.download(ps)
        inner = ps.inner if hasattr(ps, "inner") else ps
        order = np.argsort(inner.id)
        out.append({f: np.asarray(getattr(inner, f))[order] for f in inner.field_names()})
    return out


def _fluid(pts, mass, c, rho, v=None, p=None, name="f"):
    pts = np.asarray(pts, float)
    ps = M.FluidParticleSystem(name, len(pts), pts.shape[1], mass, c)
    ps.x[:] = pts
    ps.rho[:] = rho
    if v is not None:
        ps.v[:] = v
    if p is not None:
        ps.p[:] = p
    return ps


def _stress(pts, ns, mass, c, rho, stress, v=None, name="s"):
    pts = np.asarray(pts, float)
    ps = M.StressParticleSystem(name, len(pts), pts.shape[1], ns, mass, c)
    ps.x[:] = pts
    ps.rho[:] = rho
    ps.stress[:] = stress
    if v is not None:
        ps.v[:] = v
    return ps


def _tensor(s, ns):
    """Voigt (σxx, σyy, σxy) / (σxx, σyy, σzz, σxy) / (σxx, σyy, σzz, σxy,
    σxz, σyz) as the full symmetric tensor of the force plane (2D) or space."""
    s = np.asarray(s, float)
    if ns == 3:
        return np.array([[s[0], s[2]], [s[2], s[1]]])
    if ns == 4:
        return np.array([[s[0], s[3]], [s[3], s[1]]])
    return np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])


def _close(a, b, rel=REL):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(np.max(np.abs(b)), 1e-300)
    assert np.max(np.abs(a - b)) <= rel * scale, (a, b)


# ---------------------------------------------------------------------------
# Artificial viscosity (PairwisePhysics.jl:29-34) through FluidPfn (self, :151-169)
# ---------------------------------------------------------------------------


def test_art_visc_closed_form_self(orc):
    """r = 0.1h, so r² + 0.01h² = 2·10⁻⁴; particle a closes on b at 1 m/s:
    dx·dv = −0.01, μ = h·(−0.01)/(2·10⁻⁴) = −5.  With α = 1, β = 2, c = 3:
    (βμ − αc)μ = (−13)(−5) = 65, and 0.5·ρa·ρb = 0.5·2·5 = 5, so Π = 13.
    (Mean density 3.5 would give 18.57; h² instead of 0.01h² gives μ ≈ −0.099.)
    With p = 0: dvdt_a = −m·Π·g·(xa − xb), dvdt_b the opposite; drhodt = m·dv·∇W
    on both sides (S = 2)."""
    m = 0.3
    a = _fluid([(0.5, 0.5), (0.51, 0.5)], m, 3.0, [2.0, 5.0], v=[(1.0, 0.0), (0.0, 0.0)])
    (r,) = _sweep(M.FluidPfn(1.0, 2.0, H), a)
    g = _g(orc, 0.01, 2)
    dxa = np.array([-0.01, 0.0])
    _close(r["dvdt"][0], -m * 13.0 * g * dxa)
    _close(r["dvdt"][1], m * 13.0 * g * dxa)
    _close(r["drhodt"], [m * (-0.01) * g] * 2)


def test_art_visc_receding_pair_is_zero(orc):
    """Receding pair: min(dx·dv, 0) = 0, so μ = 0 and Π = 0; with p = 0 the
    momentum rate vanishes while the continuity term does not."""
    m = 0.3
    a = _fluid([(0.5, 0.5), (0.51, 0.5)], m, 3.0, [2.0, 5.0], v=[(-1.0, 0.0), (0.0, 0.0)])
    (r,) = _sweep(M.FluidPfn(1.0, 2.0, H), a)
    g = _g(orc, 0.01, 2)
    assert np.all(r["dvdt"] == 0.0)
    _close(r["drhodt"], [m * 0.01 * g] * 2)


def test_art_visc_mean_sound_speed_two_sided(orc):
    """Fluid–fluid coupled (two-sided, :171-190): c̄ = ½(10 + 30) = 20.  α = 1,
    β = 0, μ = −5 (same geometry), 0.5·ρa·ρb = 5: Π = −20·(−5)/5 = 20.
    dvdt_a = −m_b·Π·∇W, dvdt_b = +m_a·Π·∇W (the partner's mass on each side);
    drhodt_a = m_b·dv·∇W, drhodt_b = m_a·dv·∇W."""
    ma, mb = 0.2, 0.7
    a = _fluid([(0.5, 0.5)], ma, 10.0, 2.0, v=[(1.0, 0.0)], name="a")
    b = _fluid([(0.51, 0.5)], mb, 30.0, 5.0, name="b")
    ra, rb = _sweep(M.FluidPfn(1.0, 0.0, H), a, b)
    g = _g(orc, 0.01, 2)
    gx = g * np.array([-0.01, 0.0])
    _close(ra["dvdt"][0], -mb * 20.0 * gx)
    _close(rb["dvdt"][0], ma * 20.0 * gx)
    _close(ra["drhodt"], [mb * (-0.01) * g])
    _close(rb["drhodt"], [ma * (-0.01) * g])


# ---------------------------------------------------------------------------
# Pressure coefficients and continuity factors (:58-63, :99-102)
# ---------------------------------------------------------------------------


def test_pressure_sigma1_and_continuity_ratio(orc):
    """S = 1 (Colagrossi–Landrini): dh = −(pa + pb)/(ρa·ρb) = −(4 + 8)/(2·4) = −1.5;
    continuity factors ρa/ρb = 0.5 for a and ρb/ρa = 2 for b.  α = β = 0."""
    m = 0.25
    a = _fluid([(0.5, 0.5), (0.53, 0.54)], m, 3.0, [2.0, 4.0], v=[(0.2, -0.1), (0.0, 0.3)],
               p=[4.0, 8.0])
    (r,) = _sweep(M.FluidPfn(0.0, 0.0, H, sigma=1), a)
    rr = 0.05
    g = _g(orc, rr, 2)
    gx = g * np.array([-0.03, -0.04])
    dvr = np.dot(np.array([0.2, -0.4]), gx)
    _close(r["dvdt"][0], m * (-1.5) * gx)
    _close(r["dvdt"][1], -m * (-1.5) * gx)
    _close(r["drhodt"], [m * dvr * 0.5, m * dvr * 2.0])


def test_pressure_sigma2(orc):
    """S = 2: dh = −(pa/ρa² + pb/ρb²) = −(4/4 + 16/16) = −2 (the S = 1 form would
    give −(4 + 16)/(2·4) = −2.5)."""
    m = 0.25
    a = _fluid([(0.5, 0.5), (0.53, 0.54)], m, 3.0, [2.0, 4.0], p=[4.0, 16.0])
    (r,) = _sweep(M.FluidPfn(0.0, 0.0, H), a)
    gx = _g(orc, 0.05, 2) * np.array([-0.03, -0.04])
    _close(r["dvdt"][0], m * (-2.0) * gx)
    _close(r["dvdt"][1], -m * (-2.0) * gx)


# ---------------------------------------------------------------------------
# δ-SPH density diffusion (:238-243)
# ---------------------------------------------------------------------------


def test_delta_sph_closed_form(orc):
    """ψ = 2δ·c̄·h̄·(ρa − ρb)·(dx·∇W)/r², and dx·∇W = g·r², so ψ = 2δ·c·h·(ρa − ρb)·g
    = 2·0.1·3·0.1·(2 − 5)·g = −0.18·g.  drhodt_a = m·ψ/ρb, drhodt_b = −m·ψ/ρa.
    At rest with p = 0 the momentum rate is zero."""
    m = 0.3
    a = _fluid([(0.5, 0.5), (0.51, 0.5)], m, 3.0, [2.0, 5.0])
    (r,) = _sweep(M.FluidPfn(0.1, 0.1, H, delta=0.1), a)
    g = _g(orc, 0.01, 2)
    psi = -0.18 * g
    _close(r["drhodt"], [m * psi / 5.0, -m * psi / 2.0])
    assert np.all(r["dvdt"] == 0.0)


def test_delta_sph_two_sided_mean_c(orc):
    """Two-sided coupled: c̄ = ½(ca + cb) = 20, ψ = 2·0.1·20·0.1·(2 − 5)·g = −1.2·g;
    drhodt_a = m_b·ψ/ρb, drhodt_b = −m_a·ψ/ρa."""
    ma, mb = 0.2, 0.7
    a = _fluid([(0.5, 0.5)], ma, 10.0, 2.0, name="a")
    b = _fluid([(0.51, 0.5)], mb, 30.0, 5.0, name="b")
    ra, rb = _sweep(M.FluidPfn(0.0, 0.0, H, delta=0.1), a, b)
    g = _g(orc, 0.01, 2)
    psi = -1.2 * g
    _close(ra["drhodt"], [mb * psi / 5.0])
    _close(rb["drhodt"], [-ma * psi / 2.0])


# ---------------------------------------------------------------------------
# Artificial surface tension (:286-290), two-sided FluidPfn only
# ---------------------------------------------------------------------------


def test_surface_tension_uses_absolute_pressures(orc):
    """pa = −3, pb = 5, ρa = 2, ρb = 4, S = 2: dh = −(−3/4 + 5/16) = 0.4375;
    ast = ε·(−(|pa| + |pb|)/(ρa·ρb)) = 0.5·(−8/8) = −0.5 (without the absolute
    values it would be −0.125).  dvdt_a = m_b·(dh + ast)·∇W, dvdt_b = −m_a·(…)·∇W."""
    ma, mb = 0.2, 0.7
    a = _fluid([(0.5, 0.5)], ma, 10.0, 2.0, p=-3.0, name="a")
    b = _fluid([(0.53, 0.54)], mb, 30.0, 4.0, p=5.0, name="b")
    ra, rb = _sweep(M.FluidPfn(0.0, 0.0, H, epsilon=0.5), a, b)
    gx = _g(orc, 0.05, 2) * np.array([-0.03, -0.04])
    _close(ra["dvdt"][0], mb * (0.4375 - 0.5) * gx)
    _close(rb["dvdt"][0], -ma * (0.4375 - 0.5) * gx)


# ---------------------------------------------------------------------------
# Lennard-Jones static wall (:258-263; FluidPfn :210-219, CauchyFluidPfn :322-331)
# ---------------------------------------------------------------------------


def _lj_case(pfn_for, r_wall, va=(0.0, 0.0), vwall=(0.0, 0.0), cutoff=0.1):
    a_pts = [(0.3, 0.05 + r_wall)]
    if pfn_for == "cauchy":
        a = _stress(a_pts, 4, 0.2, 10.0, 2.0, (7.0, 9.0, 1.0, 3.0), v=[va], name="a")
        pfn = M.CauchyFluidPfn(1.0, 0.0, H)
    else:
        a = _fluid(a_pts, 0.2, 10.0, 2.0, v=[va], p=6.0, name="a")
        pfn = M.FluidPfn(1.0, 0.0, H)
    inner = _fluid([(0.3, 0.05)], 0.7, 99.0, 5.0, v=[vwall], p=123.0, name="wall")
    b = M.StaticBoundarySystem(inner, cutoff)
    ra, rb = _sweep(pfn, a, b)
    return ra, rb


@pytest.mark.parametrize("kind", ["fluid", "cauchy"])
def test_lennard_jones_inside_cutoff(orc, kind):
    """r = 0.05, lj_cutoff = 0.1: (0.1/0.05)^12 − (0.1/0.05)^6 = 4032, /r² → 1612800,
    ×0.01·c² with the fluid's c = 10 → 1612800; the force is that times dx = (0, 0.05):
    (0, 80640).  At rest Π = 0; the wall particle is not updated."""
    ra, rb = _lj_case(kind, 0.05)
    _close(ra["dvdt"][0], [0.0, 80640.0])
    assert np.all(rb["dvdt"] == 0.0) and np.all(rb["drhodt"] == 0.0)


@pytest.mark.parametrize("kind", ["fluid", "cauchy"])
def test_lennard_jones_outside_cutoff(orc, kind):
    """r = 0.15 > lj_cutoff (still inside 2h): no force at rest."""
    ra, _ = _lj_case(kind, 0.15)
    assert np.all(ra["dvdt"] == 0.0)


@pytest.mark.parametrize("kind", ["fluid", "cauchy"])
def test_static_wall_viscosity_uses_fluid_c(orc, kind):
    """Fluid moving down at 1 m/s, the wall particle up at 1 m/s (a static wall
    reads its particles' own velocity, :212): dv = (0, −2), dx·dv = −0.1,
    μ = 0.1·(−0.1)/(0.0025 + 0.0001) = −50/13; Π = (−α·c_a)·μ/(0.5·ρa·ρwall)
    = (−10)(−50/13)/5 = 100/13 (the wall's own c = 99 is not used).
    dvdt_a = −m_wall·Π·∇W + LJ·dx."""
    ra, _ = _lj_case(kind, 0.05, va=(0.0, -1.0), vwall=(0.0, 1.0))
    gx = _g(orc, 0.05, 2) * np.array([0.0, 0.05])
    _close(ra["dvdt"][0], -0.7 * (100.0 / 13.0) * gx + np.array([0.0, 80640.0]))
    assert np.all(ra["drhodt"] == 0.0)  # the static-wall functors leave drhodt alone


# ---------------------------------------------------------------------------
# Dynamic wall (FluidPfn :222-242, CauchyFluidPfn :334-355, StrainRate(Vorticity)Pfn
# :53-68, :111-126)
# ---------------------------------------------------------------------------


def _dyn_wall(a, beta=1.5, wall_y=-0.02):
    inner = _fluid([(0.3, wall_y)], 0.7, 99.0, 5.0, v=[(7.0, 7.0)], p=123.0, name="wall")
    return M.DynamicBoundarySystem(inner, (0.0, 1.0), (0.0, 0.0), beta)


@pytest.mark.parametrize("wall_y,factor", [(-0.02, 0.5), (-0.08, 1.5)])
def test_dynamic_wall_fluid(orc, wall_y, factor):
    """Fluid at da = 0.04 above the plane y = 0, wall particle at db = wall_y:
    |db/da| = 0.5 (below β = 1.5) or 2 (capped at β).  v_wall = −factor·v_a,
    so dv = (1 + factor)·v_a; ρ_j = ρ_a = 2, p_j = p_a = 6 so dh = −3; mass = m_a.
    α = 1, β_visc = 0: Π = −c_a·μ/(0.5·ρa²) with μ = h·(dx·dv)/(r² + 0.01h²)."""
    ma = 0.2
    va = np.array([0.5, -1.0])
    a = _fluid([(0.3, 0.04)], ma, 10.0, 2.0, v=[va], p=6.0, name="a")
    ra, rb = _sweep(M.FluidPfn(1.0, 0.0, H), a, _dyn_wall(a, wall_y=wall_y))
    dx = np.array([0.0, 0.04 - wall_y])
    r = np.linalg.norm(dx)
    g = _g(orc, r, 2)
    gx = g * dx
    dv = (1.0 + factor) * va
    mu = H * min(np.dot(dx, dv), 0.0) / (r * r + 0.01 * H * H)
    piv = (-10.0) * mu / (0.5 * 2.0 * 2.0)
    _close(ra["dvdt"][0], ma * (-3.0 - piv) * gx)
    _close(ra["drhodt"], [ma * np.dot(dv, gx)])
    assert np.all(rb["dvdt"] == 0.0)


def test_dynamic_wall_fluid_numbers(orc):
    """The first case in plain numbers: dx = (0, 0.06), dv = 1.5·(0.5, −1) =
    (0.75, −1.5), dx·dv = −0.09, μ = 0.1·(−0.09)/0.0037 = −90/37,
    Π = (−10)(−90/37)/2 = 450/37; drhodt = 0.2·(−1.5·0.06)·g = −0.018·g."""
    a = _fluid([(0.3, 0.04)], 0.2, 10.0, 2.0, v=[(0.5, -1.0)], p=6.0, name="a")
    ra, _ = _sweep(M.FluidPfn(1.0, 0.0, H), a, _dyn_wall(a))
    g = _g(orc, 0.06, 2)
    gx = g * np.array([0.0, 0.06])
    _close(ra["dvdt"][0], 0.2 * (-3.0 - 450.0 / 37.0) * gx)
    _close(ra["drhodt"], [-0.018 * g])


def test_dynamic_wall_cauchy(orc):
    """σ_j = σ_a, ρ_j = ρ_a, mass = m_a: at rest dvdt_a = m_a·(2·S_a/ρa²)·∇W,
    with the full tensor S_a of (σxx, σyy, σzz, σxy) = (−40, −24, −30, 8)."""
    ma = 0.2
    s = (-40.0, -24.0, -30.0, 8.0)
    a = _stress([(0.31, 0.04)], 4, ma, 10.0, 2.0, s, name="a")
    ra, _ = _sweep(M.CauchyFluidPfn(1.0, 0.0, H), a, _dyn_wall(a))
    dx = np.array([0.01, 0.06])
    r = np.linalg.norm(dx)
    gx = _g(orc, r, 2) * dx
    _close(ra["dvdt"][0], ma * (2.0 * _tensor(s, 4) / 4.0) @ gx)
    assert np.all(ra["drhodt"] == 0.0)


def test_dynamic_wall_strain_rate_vorticity(orc):
    """StrainRateVorticityPfn against a dynamic wall: dv = v_wall − v_a =
    −(1 + 0.5)·v_a, weight m_a/ρ_a; ε̇ = sym(dv ⊗ ∇W), W12 = ½(dv_x·g_y − dv_y·g_x)."""
    ma = 0.2
    va = np.array([0.5, -1.0])
    pts = [(0.31, 0.04)]
    a = M.ElastoPlasticParticleSystem("a", 1, 2, 4, ma, 10.0)
    a.x[:] = pts
    a.rho[:] = 2.0
    a.v[:] = va
    ra, _ = _sweep(M.StrainRateVorticityPfn(), a, _dyn_wall(a))
    dx = np.array([0.01, 0.06])
    gx = _g(orc, np.linalg.norm(dx), 2) * dx
    dv = -1.5 * va
    E = 0.5 * (np.outer(dv, gx) + np.outer(gx, dv)) * (ma / 2.0)
    Om = 0.5 * (np.outer(dv, gx) - np.outer(gx, dv)) * (ma / 2.0)
    _close(ra["strain_rate"][0], [E[0, 0], E[1, 1], 0.0, E[0, 1]])
    _close(ra["vorticity"], [Om[0, 1]])


# ---------------------------------------------------------------------------
# Cauchy stress force (:125-151) through CauchyFluidPfn (self, :265-283;
# two-sided :303-319)
# ---------------------------------------------------------------------------


CAUCHY_CASES = [
    (3, [(0.5, 0.5), (0.53, 0.54)], [(-10.0, -6.0, 2.0), (-7.0, -3.0, 5.0)]),
    (4, [(0.5, 0.5), (0.53, 0.54)], [(-10.0, -6.0, 99.0, 2.0), (-7.0, -3.0, -77.0, 5.0)]),
    (6, [(0.5, 0.5, 0.5), (0.52, 0.53, 0.56)],
     [(-10.0, -6.0, -4.0, 2.0, 3.0, 1.5), (-7.0, -3.0, -5.0, 5.0, -1.0, 0.5)]),
]


@pytest.mark.parametrize("ns,pts,stress", CAUCHY_CASES)
def test_cauchy_force_tensor_product(orc, ns, pts, stress):
    """At rest: dvdt_a = m·(S_a/ρa² + S_b/ρb²)·∇W_ab with the full symmetric
    tensors (a matrix product; σzz of NS = 4 never enters); dvdt_b the opposite."""
    m = 0.3
    rho = [2.0, 5.0]
    a = _stress(pts, ns, m, 3.0, rho, stress)
    (r,) = _sweep(M.CauchyFluidPfn(1.0, 2.0, H), a)
    D = len(pts[0])
    dx = np.subtract(pts[0], pts[1])
    gx = _g(orc, np.linalg.norm(dx), D) * dx
    Sa, Sb = _tensor(stress[0], ns), _tensor(stress[1], ns)
    h_vec = (Sa / rho[0] ** 2 + Sb / rho[1] ** 2) @ gx
    _close(r["dvdt"][0], m * h_vec)
    _close(r["dvdt"][1], -m * h_vec)
    assert np.all(r["drhodt"] == 0.0)


def test_cauchy_with_viscosity_and_continuity(orc):
    """Moving pair (NS = 4): dvdt_a = m·(h_vec − Π·∇W) with Π = 13 (the numbers of
    test_art_visc_closed_form_self) and drhodt = m·dv·∇W with no density factor."""
    m = 0.3
    s = [(-10.0, -6.0, 99.0, 2.0), (-7.0, -3.0, -77.0, 5.0)]
    a = _stress([(0.5, 0.5), (0.51, 0.5)], 4, m, 3.0, [2.0, 5.0], s,
                v=[(1.0, 0.0), (0.0, 0.0)])
    (r,) = _sweep(M.CauchyFluidPfn(1.0, 2.0, H), a)
    g = _g(orc, 0.01, 2)
    gx = g * np.array([-0.01, 0.0])
    h_vec = (_tensor(s[0], 4) / 4.0 + _tensor(s[1], 4) / 25.0) @ gx
    _close(r["dvdt"][0], m * (h_vec - 13.0 * gx))
    _close(r["drhodt"], [m * (-0.01) * g] * 2)


def test_cauchy_two_sided(orc):
    """Two stress systems: c̄ = ½(10 + 30) = 20 in Π, the partner's mass on
    each side, drhodt_a = m_b·dv·∇W, drhodt_b = m_a·dv·∇W."""
    ma, mb = 0.2, 0.7
    sa, sb = (-10.0, -6.0, 99.0, 2.0), (-7.0, -3.0, -77.0, 5.0)
    a = _stress([(0.5, 0.5)], 4, ma, 10.0, 2.0, sa, v=[(1.0, 0.0)], name="a")
    b = _stress([(0.51, 0.5)], 4, mb, 30.0, 5.0, sb, name="b")
    ra, rb = _sweep(M.CauchyFluidPfn(1.0, 0.0, H), a, b)
    g = _g(orc, 0.01, 2)
    gx = g * np.array([-0.01, 0.0])
    h_vec = (_tensor(sa, 4) / 4.0 + _tensor(sb, 4) / 25.0) @ gx
    piv = 20.0  # −α·c̄·μ/(0.5·ρa·ρb) = −20·(−5)/5
    _close(ra["dvdt"][0], mb * (h_vec - piv * gx))
    _close(rb["dvdt"][0], -ma * (h_vec - piv * gx))
    _close(ra["drhodt"], [mb * (-0.01) * g])
    _close(rb["drhodt"], [ma * (-0.01) * g])


# ---------------------------------------------------------------------------
# Strain rate and spin (:169-220) through the self functors (:18-34, :78-94)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D,ns", [(2, 3), (2, 4), (3, 6)])
def test_strain_rate_vorticity_tensor_parts(orc, D, ns):
    """dv = v_b − v_a, E = sym(dv ⊗ ∇W_ab), Ω = skew(dv ⊗ ∇W_ab); particle a gets
    (m/ρb)·E, particle b gets (m/ρa)·E (the pair term is the same from both
    sides); Voigt order (xx, yy[, zz], xy[, xz, yz]), εzz = 0 for NS = 4;
    spin W12 (2D) or (W12, W13, W23)."""
    m = 0.3
    rho = [2.0, 5.0]
    if D == 2:
        pts = [(0.5, 0.5), (0.53, 0.54)]
        v = [(0.2, -0.1), (-0.3, 0.7)]
    else:
        pts = [(0.5, 0.5, 0.5), (0.52, 0.53, 0.56)]
        v = [(0.2, -0.1, 0.4), (-0.3, 0.7, -0.2)]
    a = M.ElastoPlasticParticleSystem("e", 2, D, ns, m, 3.0)
    a.x[:] = pts
    a.v[:] = v
    a.rho[:] = rho
    (r,) = _sweep(M.StrainRateVorticityPfn(), a)
    dx = np.subtract(pts[0], pts[1])
    gx = _g(orc, np.linalg.norm(dx), D) * dx
    dv = np.subtract(v[1], v[0])
    E = 0.5 * (np.outer(dv, gx) + np.outer(gx, dv))
    Om = 0.5 * (np.outer(dv, gx) - np.outer(gx, dv))
    if ns == 3:
        ev = [E[0, 0], E[1, 1], E[0, 1]]
    elif ns == 4:
        ev = [E[0, 0], E[1, 1], 0.0, E[0, 1]]
    else:
        ev = [E[0, 0], E[1, 1], E[2, 2], E[0, 1], E[0, 2], E[1, 2]]
    om = [Om[0, 1]] if D == 2 else [Om[0, 1], Om[0, 2], Om[1, 2]]
    _close(r["strain_rate"][0], np.array(ev) * m / rho[1])
    _close(r["strain_rate"][1], np.array(ev) * m / rho[0])
    _close(np.reshape(r["vorticity"][0], -1), np.array(om) * m / rho[1])
    _close(np.reshape(r["vorticity"][1], -1), np.array(om) * m / rho[0])


def test_strain_rate_one_sided_coupled(orc):
    """StrainRatePfn, real partner system (one-sided, :37-51): only a is
    updated, with the partner's mass and density."""
    a = M.ElastoPlasticParticleSystem("a", 1, 2, 4, 0.2, 3.0)
    a.x[:] = [(0.5, 0.5)]
    a.v[:] = [(0.2, -0.1)]
    a.rho[:] = 2.0
    b = M.ElastoPlasticParticleSystem("b", 1, 2, 4, 0.7, 3.0)
    b.x[:] = [(0.53, 0.54)]
    b.v[:] = [(-0.3, 0.7)]
    b.rho[:] = 5.0
    ra, rb = _sweep(M.StrainRatePfn(), a, b)
    dx = np.array([-0.03, -0.04])
    gx = _g(orc, 0.05, 2) * dx
    dv = np.array([-0.5, 0.8])
    E = 0.5 * (np.outer(dv, gx) + np.outer(gx, dv)) * (0.7 / 5.0)
    _close(ra["strain_rate"][0], [E[0, 0], E[1, 1], 0.0, E[0, 1]])
    assert np.all(rb["strain_rate"] == 0.0)


# ---------------------------------------------------------------------------
# Ghost walls (test/test_ghost_updates.jl:58-132)
# ---------------------------------------------------------------------------


def test_cauchy_ghost_wall_repulsion(orc):  # test_ghost_updates.jl:58-97
    """One particle at (0.02, 0.05), σ = (−100, −100, 0), ghost wall x = 0:
    dvdt_x > 0 and dvdt_y ≈ 0.  Closed form: the ghost sits at (−0.02, 0.05)
    with the same σ and ρ (plain copy), so h_vec = −200/ρ²·∇W with ∇W = g·(0.04, 0)
    and the ghost's mass 1; at rest Π = 0."""
    ps = M.StressParticleSystem("test", 1, 2, 3, 1.0, 1000.0)
    ps.rho[:] = 1000.0
    ps.x[0] = (0.02, 0.05)
    ps.stress[0] = (-100.0, -100.0, 0.0)
    ghost = M.GhostParticleSystem(ps, M.GhostCopier("stress"))
    entry = M.GhostEntry(ghost, 0.1, ((1.0, 0.0), (0.0, 0.0)))
    (r,) = _sweep(M.CauchyFluidPfn(0.1, 0.1, H), ps, ghost_entry=entry)
    assert r["dvdt"][0, 0] > 0
    assert abs(r["dvdt"][0, 1]) <= 1e-10
    gx = _g(orc, 0.04, 2) * np.array([0.04, 0.0])
    _close(r["dvdt"][0], -200.0 / 1e6 * gx)


def test_fluid_ghost_wall_repulsion(orc):  # test_ghost_updates.jl:99-132
    """FluidPfn through a ghost with p copied (p = 100): dvdt_x > 0; closed form
    dh = −2·100/1000² with the ghost's mass 1."""
    ps = M.FluidParticleSystem("test", 1, 2, 1.0, 1000.0)
    ps.rho[:] = 1000.0
    ps.p[:] = 100.0
    ps.x[0] = (0.02, 0.05)
    ghost = M.GhostParticleSystem(ps, M.GhostCopier("p"))
    entry = M.GhostEntry(ghost, 0.1, ((1.0, 0.0), (0.0, 0.0)))
    (r,) = _sweep(M.FluidPfn(0.1, 0.1, H), ps, ghost_entry=entry)
    assert r["dvdt"][0, 0] > 0
    gx = _g(orc, 0.04, 2) * np.array([0.04, 0.0])
    _close(r["dvdt"][0], -2e-4 * gx)


def test_ghost_wall_viscosity_mirrored_velocity(orc):
    """Moving toward the wall, v = (−1, 0.5): the ghost's velocity is the
    reflection (1, 0.5), dv = (−2, 0), dx = (0.04, 0): dx·dv = −0.08,
    μ = 0.1·(−0.08)/(0.0016 + 0.0001) = −8/1.7; Π = (β·μ − α·c)·μ/(0.5·ρ²) with
    the source's c for both sides; drhodt = m·dv·∇W."""
    ps = M.FluidParticleSystem("test", 1, 2, 1.0, 10.0)
    ps.rho[:] = 1000.0
    ps.x[0] = (0.02, 0.05)
    ps.v[0] = (-1.0, 0.5)
    ghost = M.GhostParticleSystem(ps, M.GhostCopier("p"))
    entry = M.GhostEntry(ghost, 0.1, ((1.0, 0.0), (0.0, 0.0)))
    (r,) = _sweep(M.FluidPfn(1.0, 2.0, H), ps, ghost_entry=entry)
    g = _g(orc, 0.04, 2)
    gx = g * np.array([0.04, 0.0])
    mu = -8.0 / 1.7
    piv = (2.0 * mu - 10.0) * mu / (0.5 * 1e6)
    _close(r["dvdt"][0], -piv * gx)
    _close(r["drhodt"], [(-2.0) * 0.04 * g])
