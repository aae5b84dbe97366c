"""Pins the FP64 oracle against the reference's own known answers and
against closed forms, invariants and brute force.

Each test names the reference test (test/*.jl:line) or the mathematical fact
it restates.  Nothing here runs on a GPU.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import A, brute_counts, grasph, oracle_backend, oracle_lib, oracle_scene

M = grasph
CUBIC, WEND = A.GR_KERNEL_CUBIC, A.GR_KERNEL_WENDLAND_C2


@pytest.fixture(scope="module")
def orc():
    return oracle_lib()


# ---------------------------------------------------------------------------
# Kernels (test/test_kernels.jl:16-104; closed forms of src/Kernels.jl)
# ---------------------------------------------------------------------------


def test_cubic_norms(orc):  # test_kernels.jl:16-29
    assert orc.orc_kernel_norm(CUBIC, 2, 1.0) == pytest.approx(10 / (7 * math.pi), rel=1e-15)
    assert orc.orc_kernel_norm(CUBIC, 3, 1.0) == pytest.approx(1 / math.pi, rel=1e-15)
    assert orc.orc_kernel_norm(CUBIC, 1, 1.0) == pytest.approx(2 / 3, rel=1e-15)


def test_cubic_shape(orc):  # test_kernels.jl:49-104
    C = orc.orc_kernel_norm(CUBIC, 3, 1.0)
    assert orc.orc_kernel_w(CUBIC, 3, 1.0, 0.0) == pytest.approx(C)
    assert orc.orc_kernel_w(CUBIC, 3, 1.0, 1.0) == pytest.approx(C * 0.25)
    assert orc.orc_kernel_w(CUBIC, 3, 1.0, 2.0) == 0.0
    assert orc.orc_kernel_w(CUBIC, 3, 1.0, 2.5) == 0.0
    assert all(orc.orc_kernel_w(CUBIC, 3, 1.0, q) >= 0 for q in np.arange(0, 3.01, 0.1))
    assert orc.orc_kernel_dw_dq(CUBIC, 3, 1.0, 0.0) == 0.0
    assert orc.orc_kernel_dw_dq(CUBIC, 3, 1.0, 2.0) == 0.0
    assert orc.orc_kernel_dw_dq(CUBIC, 3, 1.0, 3.0) == 0.0
    for q in (0.5, 1.0, 1.5):
        assert orc.orc_kernel_dw_dq(CUBIC, 3, 1.0, q) < 0


@pytest.mark.parametrize("kk", [CUBIC, WEND])
@pytest.mark.parametrize("D", [2, 3])
def test_kernel_normalisation(orc, kk, D):
    """∫ W dV = 1 (radial quadrature, Simpson on [0, 2h])."""
    h = 0.37
    r = np.linspace(0, 2 * h, 4001)
    w = np.array([orc.orc_kernel_w(kk, D, h, x / h) for x in r])
    shell = 2 * math.pi * r if D == 2 else 4 * math.pi * r * r
    f = w * shell
    dr = r[1] - r[0]
    integral = dr / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    assert integral == pytest.approx(1.0, abs=1e-9)


@pytest.mark.parametrize("kk", [CUBIC, WEND])
def test_kernel_derivative_is_derivative(orc, kk):
    """dW/dq equals the central difference of W (pins the gradient formula)."""
    h = 1.0
    for q in np.linspace(0.05, 1.95, 39):
        e = 1e-6
        fd = (orc.orc_kernel_w(kk, 3, h, q + e) - orc.orc_kernel_w(kk, 3, h, q - e)) / (2 * e)
        assert orc.orc_kernel_dw_dq(kk, 3, h, q) == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_wendland_values(orc):
    """Wendland C2 at q = 0: C·2⁴ = 16C; at q = 1: C·1·3 = 3C."""
    for D in (2, 3):
        C = orc.orc_kernel_norm(WEND, D, 1.0)
        assert orc.orc_kernel_w(WEND, D, 1.0, 0.0) == pytest.approx(16 * C, rel=1e-15)
        assert orc.orc_kernel_w(WEND, D, 1.0, 1.0) == pytest.approx(3 * C, rel=1e-15)
        assert C == pytest.approx((7 / (64 * math.pi)) if D == 2 else 21 / (256 * math.pi), rel=1e-15)


def test_pow_int_correctly_rounded(orc):
    """Julia's compensated ^(Float64, Integer): within half an ulp of the exact
    rational power for the exponents the path uses (4 in Wendland, 6 and 12
    in Lennard-Jones, 7 in Tait)."""
    rng = np.random.default_rng(0)
    for n in (2, 4, 6, 7, 12):
        for x in rng.uniform(0.5, 2.0, 200):
            exact = Fraction(float(x)) ** n
            got = orc.orc_pow_int(float(x), n)
            err = abs(Fraction(got) - exact)
            assert err <= Fraction(math.ulp(got)) * Fraction(51, 100)


# ---------------------------------------------------------------------------
# Sort (test/test_sorting.jl:28-225)
# ---------------------------------------------------------------------------


def _sort(points, cutoff, tag=True):
    pts = np.asarray(points, dtype=np.float64)
    n, D = pts.shape
    ps = M.BasicParticleSystem("t", n, D, 1.0, 1.0)
    ps.x[:] = pts
    ps.rho[:] = np.arange(1, n + 1, dtype=np.float64)
    sc = oracle_scene(D)
    sc.upload(ps)
    sc.be.sort(sc.system(ps), cutoff)
    sc.download(ps)
    return ps


def test_sort_2d_row_major():  # test_sorting.jl:50-77
    ps = _sort([(0.35, 0.05), (0.05, 0.15), (0.25, 0.05), (0.05, 0.05)], 0.2)
    # keys (1,0),(0,0),(1,0),(0,0) → stable: 2,4,1,3
    assert list(ps.id) == [2, 4, 1, 3]
    assert list(ps.rho) == [2.0, 4.0, 1.0, 3.0]


def test_sort_3d_order():  # test_sorting.jl:192-225
    ps = _sort([(0.25, 0.05, 0.05), (0.05, 0.25, 0.05), (0.05, 0.05, 0.25), (0.05, 0.05, 0.05)], 0.2)
    assert list(ps.id) == [4, 3, 2, 1]  # (0,0,0)<(0,0,1)<(0,1,0)<(1,0,0)


def test_sort_negative_and_idempotent():  # test_sorting.jl:151-183
    ps = _sort([(0.05, 0.05), (-0.15, -0.15), (-0.35, 0.05)], 0.2)
    keys = np.floor(ps.x / 0.2)
    assert all(tuple(keys[i]) <= tuple(keys[i + 1]) for i in range(2))
    again = _sort(ps.x, 0.2)
    assert np.array_equal(again.x, ps.x)


def test_sort_single_and_sorted_noop():  # test_sorting.jl:106-130
    ps = _sort([(0.5, 0.5)], 0.2)
    assert list(ps.id) == [1]
    pts = [(0.05, 0.05), (0.05, 0.25), (0.25, 0.05), (0.45, 0.35)]
    ps = _sort(pts, 0.2)
    assert np.array_equal(ps.x, np.array(pts))


@pytest.mark.parametrize("D", [2, 3])
def test_sort_is_stable_lexicographic(D):
    """The permutation equals numpy's stable lexsort of floor(x/c) (an
    independent implementation of the unique stable order)."""
    rng = np.random.default_rng(D)
    x = rng.uniform(-1, 1, (3000, D))
    x[::3] = x[1::3]  # duplicates: ties must keep input order
    ps = _sort(x, 0.13)
    keys = np.floor(x / 0.13).astype(np.int64)
    order = np.lexsort(tuple(keys[:, d] for d in reversed(range(D))))
    assert np.array_equal(ps.id - 1, order)


# ---------------------------------------------------------------------------
# Grid (test/test_interaction_grid.jl:46-287, test_sorting.jl:340-367)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D", [2, 3])
def test_grid_invariants(D):
    rng = np.random.default_rng(10 + D)
    n = 2000
    h = 0.05
    ps = M.BasicParticleSystem("t", n, D, 1.0, 1.0)
    ps.x[:] = rng.uniform(-0.3, 0.7, (n, D))
    sc = oracle_scene(D)
    sc.upload(ps)
    sc.be.sort(sc.system(ps), 2 * h)
    sc.download(ps)
    k = M.CubicSplineKernel(h, ndims=D)
    si = M.SystemInteraction(k, M.NeighborCounterPfn(), ps)
    sc.be.create_grid(sc.interaction(si))
    ming, ng, cs, _ = sc.grid(si)
    c = 2 * h
    assert cs[-1] - 1 == n  # total count (:46-58)
    for d in range(D):  # origin: multiple of cutoff, at most min − 2c (:129-200)
        assert ming[d] <= ps.x[:, d].min() - 2 * c + 1e-12
        assert abs(ming[d] / c - round(ming[d] / c)) < 1e-9
    # every particle of each CSR range lies in that cell (test_sorting.jl:340-367)
    cells = np.zeros(n, dtype=np.int64)
    for d in range(D):
        cells = cells * ng[d] + np.floor((ps.x[:, d] - ming[d]) / c).astype(np.int64)
    for cell in np.unique(cells):
        s, e = cs[cell], cs[cell + 1]
        assert np.all(cells[s - 1:e - 1] == cell)
    # independent construction: start of each cell = 1 + #particles in lower cells
    expect = 1 + np.searchsorted(cells, np.arange(len(cs) - 1), side="left")
    assert np.array_equal(cs[:-1], expect)


# ---------------------------------------------------------------------------
# Sweep counts (test/test_interaction_sweep.jl:85-305, _3d.jl:69-148)
# ---------------------------------------------------------------------------


def _counts(xa, xb=None, h=0.1):
    xa = np.asarray(xa, float)
    D = xa.shape[1]
    a = M.FluidParticleSystem("a", len(xa), D, 1.0, 1.0)
    a.x[:] = xa
    sc = oracle_scene(D)
    sc.upload(a)
    sc.be.sort(sc.system(a), 2 * h)
    k = M.CubicSplineKernel(h, ndims=D)
    if xb is None:
        si = M.SystemInteraction(k, M.NeighborCounterPfn(), a)
        objs = [a]
    else:
        b = M.FluidParticleSystem("b", len(xb), D, 1.0, 1.0)
        b.x[:] = np.asarray(xb, float)
        sc.upload(b)
        sc.be.sort(sc.system(b), 2 * h)
        si = M.SystemInteraction(k, M.NeighborCounterPfn(), a, b)
        objs = [a, b]
    hh = sc.interaction(si)
    sc.be.create_grid(hh)
    sc.be.sweep(hh, 1)
    out = []
    for ps in objs:
        sc.download(ps)
        out.append(ps.drhodt[np.argsort(ps.id)])
    return out


def test_counts_pair_and_cutoff():  # :85-127
    (c,) = _counts([(0.05, 0.05), (0.10, 0.10)])
    assert list(c) == [1, 1]
    (c,) = _counts([(0.05, 0.05), (0.10, 0.10), (0.25, 0.25)])
    assert list(c) == [1, 1, 0]
    (c,) = _counts([(0.0, 0.0), (1.0, 1.0)])
    assert list(c) == [0, 0]


def test_counts_eight_clustered():  # :129-146 (LCG positions in a 0.05 box)
    rng = lambda i: (i * 6364136223846793005 + 1442695040888963407) % 2 ** 32
    vals = [(rng(i) % 100) / 2000.0 for i in range(1, 17)]
    (c,) = _counts([(vals[2 * i], vals[2 * i + 1]) for i in range(8)])
    assert list(c) == [7] * 8


def test_counts_coupled():  # :215-305
    ca, cb = _counts([(0.05, 0.05)], [(0.10, 0.10)])
    assert list(ca) == [1] and list(cb) == [1]
    ca, cb = _counts([(0.05, 0.05), (0.10, 0.10), (0.50, 0.50)],
                     [(0.06, 0.06), (0.11, 0.11), (0.55, 0.55)])
    assert list(ca) == [2, 2, 1] and list(cb) == [2, 2, 1]
    rng = lambda i: (i * 6364136223846793005 + 1442695040888963407) % 2 ** 32
    va = [(rng(i) % 100) / 2000.0 for i in range(1, 17)]
    vb = [(rng(i + 16) % 100) / 2000.0 for i in range(1, 17)]
    ca, cb = _counts([(va[2 * i], va[2 * i + 1]) for i in range(8)],
                     [(vb[2 * i], vb[2 * i + 1]) for i in range(8)])
    assert list(ca) == [8] * 8 and list(cb) == [8] * 8


@pytest.mark.parametrize("D", [2, 3])
def test_counts_equal_brute_force(D):
    """The half-shell colour sweep finds exactly the pairs with r² < (2h)²
    (brute force over all pairs), self and coupled."""
    rng = np.random.default_rng(30 + D)
    n = 1500 if D == 2 else 2500
    xa = rng.uniform(0, 0.6, (n, D))
    xb = rng.uniform(-0.1, 0.1, (n // 3, D))
    h = 0.03 if D == 2 else 0.05
    ca, cb = _counts(xa, xb, h)
    (cself,) = _counts(xa, None, h)
    sc = oracle_scene(D)
    a = M.BasicParticleSystem("a", n, D, 1, 1)
    b = M.BasicParticleSystem("b", len(xb), D, 1, 1)
    a.x[:] = xa
    b.x[:] = xb
    sc.upload(a)
    sc.upload(b)
    assert np.array_equal(cself, brute_counts(sc, a, a, 2 * h))
    assert np.array_equal(ca, brute_counts(sc, a, b, 2 * h))
    assert np.array_equal(cb, brute_counts(sc, b, a, 2 * h))


# ---------------------------------------------------------------------------
# Pair physics invariants (src/PairwisePhysics.jl)
# ---------------------------------------------------------------------------


def _fluid_cloud(D, rng, n_side):
    shape = (n_side,) * D
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, D)
    dx = 0.01
    x = (idx + 0.5) * dx + rng.uniform(-0.2, 0.2, idx.shape) * dx
    ps = M.FluidParticleSystem("f", len(x), D, 0.1, 20.0)
    ps.x[:] = x
    ps.rho[:] = 1000 * (1 + rng.uniform(-0.01, 0.01, len(x)))
    ps.p[:] = rng.uniform(-1e3, 1e3, len(x))
    ps.v[:] = rng.uniform(-0.5, 0.5, x.shape)
    return ps, dx


def _sweep_once(ps, pfn, h):
    sc = oracle_scene(ps.ndims)
    sc.upload(ps)
    sc.be.sort(sc.system(ps), 2 * h)
    k = M.WenlandC2Kernel(h, ndims=ps.ndims)
    si = M.SystemInteraction(k, pfn, ps)
    hh = sc.interaction(si)
    sc.be.create_grid(hh)
    sc.be.sweep(hh, 1)
    sc.download(ps)
    return ps


@pytest.mark.parametrize("D", [2, 3])
def test_fluid_momentum_conserved(D):
    """Self FluidPfn forces are antisymmetric: Σ m·dvdt = 0."""
    rng = np.random.default_rng(D)
    ps, dx = _fluid_cloud(D, rng, 24 if D == 2 else 10)
    ps.dvdt[:] = 0
    _sweep_once(ps, M.FluidPfn(0.1, 0.1, 1.3 * dx), 1.3 * dx)
    tot = ps.dvdt.sum(axis=0)
    scale = np.abs(ps.dvdt).sum(axis=0)
    assert np.all(np.abs(tot) <= 1e-12 * scale)


def test_uniform_velocity_no_density_rate():
    """Uniform v: dv = 0 for every pair, so continuity gives drhodt = 0 and
    the artificial viscosity vanishes (only the pressure term remains)."""
    rng = np.random.default_rng(5)
    ps, dx = _fluid_cloud(2, rng, 20)
    ps.v[:] = (0.3, -0.2)
    ps.drhodt[:] = 0
    ps.dvdt[:] = 0
    _sweep_once(ps, M.FluidPfn(0.5, 0.5, 1.3 * dx, delta=None), 1.3 * dx)
    assert np.all(ps.drhodt == 0)


def test_two_particle_pressure_repulsion(orc):
    """Two particles at rest with equal p > 0 repel along dx with magnitude
    m·2p/ρ²·|dW/dr| (pressure term of FluidPfn, S = 2)."""
    h = 0.1
    r = 0.07
    ps = M.FluidParticleSystem("f", 2, 2, 0.3, 10.0)
    ps.x[:] = [(0.5, 0.5), (0.5 + r, 0.5)]
    ps.rho[:] = 1000.0
    ps.p[:] = 2000.0
    _sweep_once(ps, M.FluidPfn(0.0, 0.0, h), h)
    order = np.argsort(ps.id)
    a = ps.dvdt[order]
    dwdr = orc.orc_kernel_dw_dq(WEND, 2, h, r / h) / h
    mag = 0.3 * 2 * 2000.0 / 1000.0 ** 2 * abs(dwdr)
    assert a[0, 0] == pytest.approx(-mag, rel=1e-12)  # pushed to −x
    assert a[1, 0] == pytest.approx(mag, rel=1e-12)
    assert abs(a[0, 1]) < 1e-15


def test_vorticity_known_values(orc):  # test_vorticity.jl:8-85
    """Two EP particles, v2 − v1 = (0, 1): vorticity_tensor(dv, gx) with
    gx = dW/dq·dx/(r·h) and dx = x1 − x2 = (−0.05, 0): W12 = 0.5·(0·gy − 1·gx)."""
    h = 0.1
    mass = 1000.0 * h * h
    ps = M.ElastoPlasticParticleSystem("e", 2, 2, 4, mass, 10.0)
    ps.x[:] = [(0.0, 0.0), (0.05, 0.0)]
    ps.v[:] = [(0.0, 0.0), (0.0, 1.0)]
    ps.rho[:] = 1000.0
    sc = oracle_scene(2)
    sc.upload(ps)
    k = M.CubicSplineKernel(h, ndims=2)
    si = M.SystemInteraction(k, M.StrainRateVorticityPfn(), ps)
    hh = sc.interaction(si)
    sc.be.create_grid(hh)
    sc.be.sweep(hh, 1)
    sc.download(ps)
    r = 0.05
    gxx = orc.orc_kernel_dw_dq(CUBIC, 2, h, r / h) * (-0.05) / (r * h)
    expect = 0.5 * (0.0 * 0.0 - 1.0 * gxx) * (mass / 1000.0)
    assert ps.vorticity[0] == pytest.approx(expect, rel=1e-14)
    assert ps.vorticity[1] == pytest.approx(expect, rel=1e-14)  # same sign from j's side
    assert ps.strain_rate[0, 2] == 0.0  # εzz pinned (NS = 4)


# ---------------------------------------------------------------------------
# Elasto-plastic updater (test/test_ep_updater.jl:18-85)
# ---------------------------------------------------------------------------


def _ep_call(E, nu, phi, psi, coh, stress, sr, vort, dt):
    ps = M.ElastoPlasticParticleSystem(
        "e", 1, 2, 4, 1.0, 100.0,
        state_updater=(M.ElastoPlasticStressUpdater(E, nu, phi, psi, coh),))
    ps.stress[0] = stress
    ps.strain_rate[0] = sr
    ps.vorticity[0] = vort
    sc = oracle_scene(2)
    sc.upload(ps)
    sc.be.update_state(sc.system(ps), 1, dt)
    sc.download(ps)
    return ps


def test_ep_elastic_increment():  # :18-41
    E, nu, dt = 1e6, 0.3, 0.001
    ps = _ep_call(E, nu, 30 * math.pi / 180, 0.0, 1e6, (0, 0, 0, 0), (0.1, 0, 0, 0), 0.0, dt)
    D0 = E / ((1 + nu) * (1 - 2 * nu))
    de = 0.1 * dt
    assert ps.stress[0, 0] == pytest.approx(D0 * (1 - nu) * de, rel=1e-14)
    assert ps.stress[0, 1] == pytest.approx(D0 * nu * de, rel=1e-14)
    assert ps.stress[0, 2] == pytest.approx(D0 * nu * de, rel=1e-14)
    assert ps.stress[0, 3] == 0.0
    assert np.allclose(ps.strain[0], np.array([0.1, 0, 0, 0]) * dt, rtol=1e-15)
    assert not np.any(ps.strain_p[0])


def test_ep_jaumann_rotation():  # :43-57
    ps = _ep_call(1e6, 0.3, 30 * math.pi / 180, 0.0, 1e6, (1e5, 0, 0, 0), (0, 0, 0, 0), 10.0, 0.001)
    assert ps.stress[0, 3] == pytest.approx(1000.0, rel=1e-14)
    assert ps.stress[0, 0] == pytest.approx(1e5, rel=1e-14)


def test_ep_plastic_return():  # :59-85
    phi = 30 * math.pi / 180
    ps = _ep_call(1e6, 0.3, phi, 0.0, 10.0, (0, 0, 0, 0), (0, 0, 0, 100.0), 0.0, 0.001)
    a_phi = 2 * math.sin(phi) / (math.sqrt(3) * (3 - math.sin(phi)))
    k_c = 6 * 10.0 * math.cos(phi) / (math.sqrt(3) * (3 - math.sin(phi)))
    s = ps.stress[0]
    I1 = s[0] + s[1] + s[2]
    dev = s[:3] - I1 / 3
    J2 = 0.5 * (dev @ dev + 2 * s[3] ** 2)
    assert a_phi * I1 + math.sqrt(J2) - k_c == pytest.approx(0.0, abs=1e-7)
    assert np.any(ps.strain_p[0])


def test_ep_apex_cap():
    """A tensile trial state beyond the apex (tr σ > k_c/α_φ after the return)
    is set to the apex σ = (k_c/α_φ/3)(1,1,1,0) (src/StateUpdaters.jl:330-332);
    with c = 0 the apex is the origin."""
    ps = _ep_call(1e6, 0.3, 0.5, 0.0, 0.0, (30.0, 10.0, 10.0, 5.0), (0, 0, 0, 0), 0.0, 1e-3)
    assert np.all(ps.stress[0] == 0.0)
    ps = _ep_call(1e6, 0.3, 0.5, 0.0, 100.0, (3000.0, 1000.0, 1000.0, 50.0), (0, 0, 0, 0), 0.0, 1e-3)
    a_phi = 2 * math.sin(0.5) / (math.sqrt(3) * (3 - math.sin(0.5)))
    k_c = 6 * 100.0 * math.cos(0.5) / (math.sqrt(3) * (3 - math.sin(0.5)))
    assert np.allclose(ps.stress[0], (k_c / a_phi / 3,) * 3 + (0.0,), rtol=1e-14)


def test_ep_hydrostatic_tension_is_nan():
    """A purely hydrostatic tensile trial state has J2 = 0, so the reference's
    inv(2·sqrt(J2)) is Inf and the return produces NaN
    (src/StateUpdaters.jl:276-311); the final apex test compares NaN and
    leaves it.  Restated as is."""
    ps = _ep_call(1e6, 0.3, 0.5, 0.0, 0.0, (10.0, 10.0, 10.0, 0.0), (0, 0, 0, 0), 0.0, 1e-3)
    assert np.all(np.isnan(ps.stress[0]))


# ---------------------------------------------------------------------------
# Visco-plastic Mohr-Coulomb updater (src/StateUpdaters.jl:144-189): closed
# forms.  No reference test covers it; these pin the rate magnitude (factor 2
# on the shear terms, which components are shear), the 2(c + p·tanφ) scale,
# the pressure shift and the apex zeroing.
# ---------------------------------------------------------------------------


def _vpmc_call(D, ns, rho, sr, phi, coh, eos="tait"):
    upd = M.TaitEOSUpdater(1000.0) if eos == "tait" else M.LinearEOSUpdater(1000.0)
    ps = M.StressParticleSystem("v", 1, D, ns, 1.0, 10.0,
                                state_updater=(M.ViscoPlasticMCStressUpdater(upd, phi, coh),))
    ps.rho[0] = rho
    ps.strain_rate[0] = sr
    sc = oracle_scene(D)
    sc.upload(ps)
    sc.be.update_state(sc.system(ps), 1, 1e-4)
    sc.download(ps)
    return ps


def _tait(rho, rho0=1000.0, c=10.0, gamma=7.0):
    return rho0 * c * c / gamma * ((rho / rho0) ** 7 - 1.0)


def test_vpmc_pure_shear_3d():
    g, phi, coh = 3.0, 0.5, 20.0
    ps = _vpmc_call(3, 6, 1010.0, (0, 0, 0, g, 0, 0), phi, coh)
    p = _tait(1010.0)
    assert ps.p[0] == pytest.approx(p, rel=1e-14)
    tau = 2 * (coh + p * math.tan(phi)) * g / math.sqrt(2 * g * g)
    assert np.allclose(ps.stress[0], (-p, -p, -p, tau, 0, 0), rtol=1e-14, atol=0)


def test_vpmc_plane_rate_2d():
    a, b, phi, coh = 2.0, 1.5, 0.6, 5.0
    ps = _vpmc_call(2, 4, 1005.0, (a, -a, 0.0, b), phi, coh)
    p = _tait(1005.0)
    mag = math.sqrt(a * a + a * a + 2 * b * b)
    sc = 2 * (coh + p * math.tan(phi)) / mag
    expect = (sc * a - p, -sc * a - p, -p, sc * b)
    assert np.allclose(ps.stress[0], expect, rtol=1e-14, atol=0)


def test_vpmc_zero_rate_is_hydrostatic():
    ps = _vpmc_call(3, 6, 1010.0, (0, 0, 0, 0, 0, 0), 0.5, 20.0, eos="linear")
    p = 100.0 * 10.0  # c²(ρ − ρ0)
    assert ps.p[0] == pytest.approx(p, rel=1e-15)
    assert np.array_equal(ps.stress[0], np.array([-p, -p, -p, 0.0, 0.0, 0.0]))


def test_vpmc_tension_beyond_apex_is_zeroed():
    ps = _vpmc_call(3, 6, 990.0, (0, 0, 0, 1.0, 0, 0), 0.5, 0.0)
    assert ps.p[0] < 0
    assert np.all(ps.stress[0] == 0.0)


# ---------------------------------------------------------------------------
# Ghosts (test/test_ghost_particles.jl:112-193, test_ghost_copier.jl:186-305)
# ---------------------------------------------------------------------------


def _ghost(points, stress, entries, ns=4):
    pts = np.asarray(points, float)
    D = pts.shape[1]
    ps = M.StressParticleSystem("s", len(pts), D, ns, 1.0, 10.0)
    ps.x[:] = pts
    ps.rho[:] = 1000.0
    ps.v[:] = np.arange(pts.size, dtype=float).reshape(pts.shape) * 0.1 + 0.05
    ps.stress[:] = stress
    g = M.GhostParticleSystem(ps, M.GhostCopier(("stress", M.HouseholderReflect())))
    ge = M.GhostEntry(g, *entries)
    sc = oracle_scene(D)
    sc.upload(ps)
    gh = sc.ghost(ge)
    sc.be.generate_ghosts(gh)
    sc.be.update_ghost_kinematics(gh)
    sc.be.update_ghost(gh, 1)
    sc.download(g)
    return ps, g


def test_ghost_floor_reflection():  # test_ghost_copier.jl:186-221
    ps, g = _ghost([(0.1, 0.04), (0.2, 0.06)],
                   [(10, 20, 30, 5), (-5, -5, 0, -3)], (0.1, ((0.0, 1.0), (0.0, 0.0))))
    assert g.n == 2
    for k in range(g.n):
        src = g.idx_original[k] - 1
        assert np.allclose(g.x[k], ps.x[src] * (1, -1))  # mirrored position
        assert np.allclose(g.v[k], ps.v[src] * (1, -1))  # free-slip velocity
        assert np.allclose(g.stress[k], ps.stress[src] * (1, 1, 1, -1))


def test_ghost_45_degree_swaps_normals():  # :236-248
    n = (1 / math.sqrt(2), 1 / math.sqrt(2))
    ps, g = _ghost([(0.05, 0.05)], [(10, 20, 30, 7)], (0.2, (n, (0.0, 0.0))))
    assert np.allclose(g.stress[0], (20, 10, 30, 7))


def test_ghost_3d_flips_shear():  # :250-262
    ps, g = _ghost([(0.05, 0.05, 0.05)], [(1, 2, 3, 4, 5, 6)],
                   (0.2, ((0.0, 0.0, 1.0), (0.0, 0.0, 0.0))), ns=6)
    assert np.allclose(g.stress[0], (1, 2, 3, 4, -5, -6))


def test_ghost_qualification_and_order():  # test_ghost_particles.jl:112-193
    """Ghosts only for 0 < da ≤ cutoff, generated boundary-major then by
    source index (src/GhostParticles.jl:350-389)."""
    pts = [(0.3, 0.05), (0.02, 0.5), (0.6, 0.2), (0.01, 0.01), (0.5, -0.01)]
    ps, g = _ghost(pts, [(0, 0, 0, 0)] * 5,
                   (0.1, ((0.0, 1.0), (0.0, 0.0)), ((1.0, 0.0), (0.0, 0.0))))
    assert list(g.idx_boundary) == [1, 1, 2, 2]
    assert list(g.idx_original) == [1, 4, 2, 4]


def test_ghost_normals_keep_generation_order():
    """_particle_arrays(::GhostParticleSystem) (src/Sorting.jl:33-40) does not
    include `normals`: after a ghost sort the normal of slot k is still the one
    of the k-th generated ghost."""
    pts = [(0.5, 0.03), (0.02, 0.5)]  # floor ghost first, wall ghost second
    ps = M.StressParticleSystem("s", 2, 2, 4, 1.0, 10.0)
    ps.x[:] = pts
    g = M.GhostParticleSystem(ps)
    ge = M.GhostEntry(g, 0.1, ((0.0, 1.0), (0.0, 0.0)), ((1.0, 0.0), (0.0, 0.0)))
    sc = oracle_scene(2)
    sc.upload(ps)
    gh = sc.ghost(ge)
    sc.be.generate_ghosts(gh)
    sc.be.sort(gh, 0.2)
    sc.download(g)
    # the wall ghost (x < 0) sorts first but slot 0 keeps the floor normal
    assert list(g.idx_boundary) == [2, 1]
    assert np.allclose(g.normals, [(0.0, 1.0), (1.0, 0.0)])


# ---------------------------------------------------------------------------
# Leapfrog closed forms (test/test_time_integration.jl:79-197)
# ---------------------------------------------------------------------------


def _lf_run(ps, nsteps, cfl, h=0.1):
    k = M.CubicSplineKernel(h, ndims=ps.ndims)
    si = M.SystemInteraction(k, None, ps)
    lf = M.LeapFrogTimeIntegrator(ps, si)
    sc = oracle_scene(ps.ndims)
    hh = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(hh, cfl)
    sc.be.step(hh, nsteps, dt)
    sc.download_all(lf)
    return dt


def test_constant_velocity_drift():  # :81-95
    ps = M.BasicParticleSystem("f", 1, 2, 1.0, 10.0)
    ps.v[:] = (1.0, 0.0)
    ps.rho[:] = 1000.0
    dt = _lf_run(ps, 3, 0.5)
    assert dt == 0.5 * 0.1 / 10.0
    assert ps.x[0, 0] == pytest.approx(3 * dt, abs=1e-15)
    assert ps.x[0, 1] == 0.0


def test_rest_and_density_unchanged():  # :97-125
    ps = M.BasicParticleSystem("f", 2, 2, 1.0, 10.0)
    ps.x[:] = [(0.2, 0.2), (0.4, 0.4)]
    ps.rho[:] = [800.0, 1200.0]
    x0 = ps.x.copy()
    _lf_run(ps, 5, 0.5)
    assert np.array_equal(ps.x, x0) and list(ps.rho) == [800.0, 1200.0]


def test_free_fall_closed_form():
    """Kick-drift leapfrog with a constant source g (derived from
    src/TimeIntegration.jl:645-685): v_n = n·dt·g, x_n = x_0 + g·dt²·n(n+1)/2."""
    ps = M.BasicParticleSystem("f", 1, 2, 1.0, 10.0, source_v=[0.0, -1.0])
    ps.x[:] = (0.5, 0.5)
    ps.rho[:] = 1000.0
    n = 40
    dt = _lf_run(ps, n, 0.5)
    assert ps.v[0, 1] == pytest.approx(-n * dt, rel=1e-13)
    assert ps.x[0, 1] == pytest.approx(0.5 - dt * dt * n * (n + 1) / 2, rel=1e-13)
    assert ps.v[0, 0] == 0.0
    assert list(ps.dvdt[0]) == [0.0, -1.0]  # accumulator reset to the source (:167-195)


def test_damping_closed_form():
    """With Γ ≠ 0 and no forces, v_{n+1} = v_n − Γ·v_{n+1/2} where the
    half-step value is v_n + dt/2·dvdt_n (src/TimeIntegration.jl:673)."""
    ps = M.BasicParticleSystem("f", 1, 2, 1.0, 10.0)
    ps.v[:] = (1.0, 0.0)
    ps.rho[:] = 1000.0
    k = M.CubicSplineKernel(0.1, ndims=2)
    lf = M.LeapFrogTimeIntegrator(ps, M.SystemInteraction(k, None, ps), gamma=0.1)
    sc = oracle_scene(2)
    hh = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(hh, 0.5)
    v, a = 1.0, 0.0
    for _ in range(5):
        sc.be.step(hh, 1, dt)
        vh = v + dt / 2 * a
        a = 0.0 - (0.1 / dt) * vh
        v = v + dt * a
    sc.download_all(lf)
    assert ps.v[0, 0] == pytest.approx(v, rel=1e-14)


def test_stage_count_mismatch_raises():  # :229-288
    ps = M.BasicParticleSystem("f", 2, 2, 1.0, 10.0)
    k = M.CubicSplineKernel(0.1, ndims=2)
    a = M.SystemInteraction(k, (None, None), ps)
    b = M.SystemInteraction(k, None, ps)
    lf = M.LeapFrogTimeIntegrator(ps, [a, b])
    sc = oracle_scene(2)
    with pytest.raises(AssertionError):
        sc.integrator(lf)


def test_constructor_errors():  # src/Particles.jl:73-76,287; Interaction.jl:69-72
    with pytest.raises(ValueError):
        M.FluidParticleSystem("f", 0, 2, 1.0, 1.0)
    with pytest.raises(ValueError):
        M.StressParticleSystem("f", 2, 2, 5, 1.0, 1.0)
    with pytest.raises(ValueError):
        M.BasicParticleSystem("f", 2, 2, 1.0, 1.0, source_v=[0.0])
    ps = M.BasicParticleSystem("f", 2, 2, 1.0, 1.0)
    with pytest.raises(ValueError):
        M.SystemInteraction(M.CubicSplineKernel(0.1, ndims=3), None, ps)
    with pytest.raises(ValueError):
        M.LeapFrogTimeIntegrator([], [])
    sc = oracle_scene(2)
    ep = M.StressParticleSystem("s", 2, 2, 3, 1.0, 1.0,
                                state_updater=(M.ElastoPlasticStressUpdater(1, 0.3, 0.5, 0, 0),))
    with pytest.raises((AssertionError, ValueError)):
        sc.system(ep)
