"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle on
identical seeded inputs.

Integer and position data (cell keys → permutations, CSR starts, ghost
indices, mirrored positions) must be bit-exact.  Floating-point results are
compared with max|Δ| / max|ref| per field:
  * GR_PREC_F64 ("f64"): 1e-11 for one call, 1e-9 after short trajectories
    (only the summation order differs from the oracle);
  * GR_PREC_F32 ("f32"): 2e-5 for one call (FP32 pair math, FP64 sums of
    the FP32 pair terms).
"""
import math

import numpy as np
import pytest

from conftest import A, brute_counts, grasph, oracle_scene

M = grasph
pytestmark = pytest.mark.gpu

TOL1 = {"f64": 1e-11, "f32": 2e-5}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if b.size else 0.0


def lattice(shape, dx, rng, jitter=0.3, origin=0.0):
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, len(shape))
    x = (idx + 0.5) * dx + origin
    x += rng.uniform(-jitter, jitter, x.shape) * dx
    return x


def fill_state(ps, rng, v=0.3, rho0=1000.0, p=5e3):
    ps.v[:] = rng.uniform(-v, v, ps.v.shape)
    ps.rho[:] = rho0 * (1 + rng.uniform(-0.01, 0.01, ps.n))
    if hasattr(ps, "p"):
        ps.p[:] = rng.uniform(-p, p, ps.n)
    if hasattr(ps, "stress"):
        ps.stress[:] = rng.uniform(-p, p, ps.stress.shape)


# ---------------------------------------------------------------------------
# sort_particles! — bit-exact permutation (src/Sorting.jl:157-195)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D,n", [(2, 40000), (3, 150000)])
def test_sort_bit_exact(D, n):
    rng = np.random.default_rng(11 + D)
    h = 0.013
    cutoff = 2 * h
    span = 0.6
    x = rng.uniform(-0.1, span, (n, D))
    # adversarial: exact cell faces and many equal keys
    x[: n // 10] = np.round(x[: n // 10] / cutoff) * cutoff
    x[n // 10: n // 5] = x[n // 5: n // 5 + n // 10]
    outs = []
    for make in ("dev", "orc"):
        ps = M.FluidParticleSystem("s", n, D, 1.0, 1.0)
        ps.x[:] = x
        ps.rho[:] = np.arange(n, dtype=np.float64)
        sc = M.Session(D, "f64").scene if make == "dev" else oracle_scene(D)
        sc.upload(ps)
        sc.be.sort(sc.system(ps), cutoff)
        sc.download(ps)
        outs.append((ps.x.copy(), ps.id.copy(), ps.rho.copy()))
    (xg, ig, rg), (xo, io, ro) = outs
    assert np.array_equal(ig, io)
    assert np.array_equal(xg, xo)
    assert np.array_equal(rg, ro)
    # independent check: stable lexicographic order of floor(x / cutoff)
    keys = np.floor(x / cutoff).astype(np.int64)
    order = np.lexsort(tuple(keys[:, d] for d in reversed(range(D))))
    assert np.array_equal(ig - 1, order)


# ---------------------------------------------------------------------------
# create_grid! — bit-exact CSR (src/Interaction.jl:123-236)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("sorted_first", [True, False])
def test_grid_bit_exact(D, sorted_first):
    rng = np.random.default_rng(5 + D)
    h = 0.02
    na = 20000 if D == 2 else 60000
    xa = rng.uniform(0.0, 0.5, (na, D))
    xb = rng.uniform(-0.1, 0.05, (na // 4, D))
    res = []
    for make in ("dev", "orc"):
        a = M.FluidParticleSystem("a", na, D, 1.0, 1.0)
        b = M.BasicParticleSystem("b", len(xb), D, 1.0, 1.0)
        a.x[:] = xa
        b.x[:] = xb
        k = (M.CubicSplineKernel if D == 2 else M.WenlandC2Kernel)(h, ndims=D)
        si_self = M.SystemInteraction(k, M.NeighborCounterPfn(), a)
        si_cpl = M.SystemInteraction(k, M.NeighborCounterPfn(), a, b)
        sc = M.Session(D, "f64").scene if make == "dev" else oracle_scene(D)
        sc.upload(a)
        sc.upload(b)
        if sorted_first:
            sc.be.sort(sc.system(a), 2 * h)
            sc.be.sort(sc.system(b), 2 * h)
        out = []
        for si in (si_self, si_cpl):
            sc.be.create_grid(sc.interaction(si))
            out.append(sc.grid(si))
        res.append(out)
    for gdev, gorc in zip(res[0], res[1]):
        assert np.array_equal(gdev[0], gorc[0])  # mingridx
        assert np.array_equal(gdev[1], gorc[1])  # ngridx
        assert np.array_equal(gdev[2], gorc[2])  # cell_start
        if gorc[3] is not None:
            assert np.array_equal(gdev[3], gorc[3])  # cell_start_a


def test_grid_bit_exact_long_empty_run():
    """Two clusters 2000 apart along the slowest grid axis: ~400k empty cells
    (hundreds of fill blocks) between them, so an empty cell's start comes
    from a block far to its right."""
    rng = np.random.default_rng(11)
    h = 0.02
    xa = rng.uniform(0.0, 0.3, (6000, 2))
    xa[3000:, 0] += 2000.0
    res = []
    for make in ("dev", "orc"):
        a = M.FluidParticleSystem("a", len(xa), 2, 1.0, 1.0)
        a.x[:] = xa
        si = M.SystemInteraction(M.CubicSplineKernel(h, ndims=2), M.NeighborCounterPfn(), a)
        sc = M.Session(2, "f64").scene if make == "dev" else oracle_scene(2)
        sc.upload(a)
        sc.be.sort(sc.system(a), 2 * h)
        sc.be.create_grid(sc.interaction(si))
        res.append(sc.grid(si))
    gdev, gorc = res
    assert len(gorc[2]) > 300000
    assert np.array_equal(gdev[1], gorc[1])
    assert np.array_equal(gdev[2], gorc[2])


# ---------------------------------------------------------------------------
# Neighbour counts (test/test_interaction_sweep.jl:85-164, 215-305 analogues
# at scale): device == oracle == brute force
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D", [2, 3])
def test_counts_match_brute_force(D):
    rng = np.random.default_rng(21 + D)
    h = 0.02 if D == 2 else 0.03
    shape = (60, 60) if D == 2 else (20, 20, 20)
    xa = lattice(shape, 0.7 * h, rng)
    xb = lattice(shape[:-1] + (3,), 0.7 * h, rng, origin=-0.05)
    got = {}
    for make in ("dev", "orc"):
        a = M.FluidParticleSystem("a", len(xa), D, 1.0, 1.0)
        b = M.FluidParticleSystem("b", len(xb), D, 1.0, 1.0)
        a.x[:] = xa
        b.x[:] = xb
        k = M.WenlandC2Kernel(h, ndims=D)
        sc = M.Session(D, "f64").scene if make == "dev" else oracle_scene(D)
        for ps in (a, b):
            ps.dvdt[:] = 0
            ps.drhodt[:] = 0
            sc.upload(ps)
            sc.be.sort(sc.system(ps), 2 * h)  # create_grid! needs sorted input
        for si in (M.SystemInteraction(k, M.NeighborCounterPfn(), a),
                   M.SystemInteraction(k, M.NeighborCounterPfn(), a, b)):
            h_ = sc.interaction(si)
            sc.be.create_grid(h_)
            sc.be.sweep(h_, 1)
        sc.download(a)
        sc.download(b)
        got[make] = (a.drhodt.copy(), b.drhodt.copy(), a.dvdt[:, 0].copy())
        if make == "orc":
            ca = brute_counts(sc, a, a, 2 * h) + brute_counts(sc, a, b, 2 * h)
            assert np.array_equal(a.drhodt, ca)
    for u, v in zip(got["dev"], got["orc"]):
        assert np.array_equal(u, v)


# ---------------------------------------------------------------------------
# Pair functors through one sweep, each partner kind
# ---------------------------------------------------------------------------

PARTNERS = ["self", "dynamic", "static", "two_sided", "ghost"]


def _functor_scene(D, pfn_kind, partner, rng_seed):
    rng = np.random.default_rng(rng_seed)
    h = 0.02
    shape = (40, 40) if D == 2 else (14, 14, 14)
    dxl = 0.75 * h
    xa = lattice(shape, dxl, rng, origin=0.01)
    xb = lattice(shape[:-1] + (3,), dxl, rng, origin=-0.045)
    ns = 4 if D == 2 else 6
    kern = (M.CubicSplineKernel if D == 2 else M.WenlandC2Kernel)(h, ndims=D)
    if pfn_kind in ("fluid",):
        a = M.FluidParticleSystem("a", len(xa), D, 0.4, 30.0)
    else:
        a = M.ElastoPlasticParticleSystem("a", len(xa), D, ns, 0.4, 30.0)
    a.x[:] = xa
    fill_state(a, rng)
    pfn = {"fluid": M.FluidPfn(0.1, 0.05, h, delta=0.1),
           "fluid_s1": M.FluidPfn(0.1, 0.05, h, sigma=1),
           "cauchy": M.CauchyFluidPfn(0.1, 0.05, h, delta=0.1),
           "sr": M.StrainRatePfn(), "srv": M.StrainRateVorticityPfn()}[pfn_kind]
    if pfn_kind == "fluid_s1":
        a = M.FluidParticleSystem("a", len(xa), D, 0.4, 30.0)
        a.x[:] = xa
        fill_state(a, rng)
    objs = [a]
    entries = []
    if partner == "self":
        si = M.SystemInteraction(kern, pfn, a)
    elif partner in ("dynamic", "static"):
        b = M.BasicParticleSystem("b", len(xb), D, 0.5, 20.0)
        b.x[:] = xb
        fill_state(b, rng)
        objs.append(b)
        nrm = np.zeros(D)
        nrm[-1] = 1.0
        w = (M.DynamicBoundarySystem(b, nrm, np.zeros(D), 2.5) if partner == "dynamic"
             else M.StaticBoundarySystem(b, 0.5 * h))
        si = M.SystemInteraction(kern, pfn, a, w)
    elif partner == "two_sided":
        b = (M.FluidParticleSystem("b", len(xb), D, 0.5, 20.0) if pfn_kind.startswith("fluid")
             else M.ElastoPlasticParticleSystem("b", len(xb), D, ns, 0.5, 20.0))
        b.x[:] = xb
        fill_state(b, rng)
        objs.append(b)
        if pfn_kind == "fluid":
            pfn = M.FluidPfn(0.1, 0.05, h, delta=0.1, epsilon=0.2)
        si = M.SystemInteraction(kern, pfn, a, b)
    else:  # ghost
        cop = M.GhostCopier("p") if pfn_kind.startswith("fluid") else \
            M.GhostCopier(("stress", M.HouseholderReflect()))
        g = M.GhostParticleSystem(a, cop)
        nrm = np.zeros(D)
        nrm[-1] = 1.0
        nrm2 = np.zeros(D)
        nrm2[0] = 1.0
        ge = M.GhostEntry(g, 3 * h, (nrm, np.zeros(D) + 0.01), (nrm2, np.zeros(D) + 0.01))
        entries.append(ge)
        si = M.SystemInteraction(kern, pfn, a, g)
    return objs, entries, si


def _run_functor(sc, objs, entries, si):
    for ps in objs:
        sc.upload(ps)
        sc.be.sort(sc.system(ps), 2 * si.kernel.h)  # create_grid! needs sorted input
    for ge in entries:
        gh = sc.ghost(ge)
        sc.be.generate_ghosts(gh)
        sc.be.sort(gh, 2 * si.kernel.h)
        sc.be.update_ghost_kinematics(gh)
        sc.be.update_ghost(gh, 1)
    h_ = sc.interaction(si)
    sc.be.create_grid(h_)
    sc.be.sweep(h_, 1)
    for ps in objs:
        sc.download(ps)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("pfn_kind,partner", [
    ("fluid", p) for p in PARTNERS] + [("fluid_s1", "self")] + [
    ("cauchy", p) for p in PARTNERS] + [("sr", "self"), ("sr", "dynamic"), ("sr", "ghost"),
                                        ("srv", "self"), ("srv", "dynamic"), ("srv", "ghost")])
def test_functor_sweep(prec, D, pfn_kind, partner):
    res = {}
    for make in ("dev", "orc"):
        objs, entries, si = _functor_scene(D, pfn_kind, partner, 100 + D)
        sc = M.Session(D, prec).scene if make == "dev" else oracle_scene(D)
        _run_functor(sc, objs, entries, si)
        fields = {}
        for k, ps in enumerate(objs):
            order = np.argsort(ps.id)
            for f in ("dvdt", "drhodt", "strain_rate", "vorticity"):
                if hasattr(ps, f):
                    fields[f"{k}.{f}"] = getattr(ps, f)[order].copy()
        res[make] = fields
    for key, ref in res["orc"].items():
        if not np.any(ref):
            assert not np.any(res["dev"][key]), key
            continue
        e = rel(res["dev"][key], ref)
        assert e < TOL1[prec], f"{key}: rel err {e:.3e}"


# ---------------------------------------------------------------------------
# State updaters (src/StateUpdaters.jl) — one call, FP64 bit-level
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D", [2, 3])
def test_updaters(D):
    rng = np.random.default_rng(77)
    n = 5000
    ns = 4 if D == 2 else 6
    res = {}
    for make in ("dev", "orc"):
        rng = np.random.default_rng(77)
        if D == 2:
            ps = M.ElastoPlasticParticleSystem(
                "ep", n, 2, 4, 1.0, 24.7,
                state_updater=(M.ZeroFieldUpdater("strain_rate"),
                               M.ElastoPlasticStressUpdater(0.84e6, 0.3, 0.35, 0.05, 50.0)))
        else:
            ps = M.StressParticleSystem(
                "vp", n, 3, 6, 1.0, 24.7,
                state_updater=(M.ViscoPlasticMCStressUpdater(M.TaitEOSUpdater(1500.0), 0.5, 10.0),
                               M.ViscoPlasticMCStressUpdater(M.LinearEOSUpdater(1500.0), 0.6, 0.0)))
        ps.rho[:] = 1500 * (1 + rng.uniform(-0.02, 0.02, n))
        ps.stress[:] = rng.uniform(-5e3, 2e3, (n, ns))
        ps.strain_rate[:] = rng.uniform(-50, 50, (n, ns))
        if D == 2:
            ps.strain_rate[:, 2] = 0
            ps.vorticity[:] = rng.uniform(-20, 20, n)
        sc = M.Session(D, "f64").scene if make == "dev" else oracle_scene(D)
        sc.upload(ps)
        sc.be.update_state(sc.system(ps), 2, 1e-4)
        if D == 3:
            sc.be.update_state(sc.system(ps), 1, 1e-4)
        sc.download(ps)
        res[make] = {f: getattr(ps, f).copy() for f in ps.field_names()}
    for f, ref in res["orc"].items():
        assert rel(res["dev"][f], ref) < 1e-14, f


# ---------------------------------------------------------------------------
# Ghosts: generation, sort, kinematics, copier (src/GhostParticles.jl)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D", [2, 3])
def test_ghosts(D):
    res = {}
    for make in ("dev", "orc"):
        rng = np.random.default_rng(3 + D)
        ns = 4 if D == 2 else 6
        shape = (80, 50) if D == 2 else (24, 20, 16)
        ps = M.StressParticleSystem("s", int(np.prod(shape)), D, ns, 1.0, 10.0)
        ps.x[:] = lattice(shape, 0.01, rng, jitter=0.2)
        fill_state(ps, rng)
        g = M.GhostParticleSystem(ps, M.GhostCopier(("stress", M.HouseholderReflect()), "p"))
        nrms = [np.eye(D)[-1], np.eye(D)[0], -np.eye(D)[0]]
        pts = [np.zeros(D), np.zeros(D), np.eye(D)[0] * shape[0] * 0.01]
        ge = M.GhostEntry(g, 0.035, *zip(nrms, pts))
        sc = M.Session(D, "f64").scene if make == "dev" else oracle_scene(D)
        sc.upload(ps)
        gh = sc.ghost(ge)
        sc.be.generate_ghosts(gh)
        sc.be.sort(gh, 0.026)
        sc.be.update_ghost_kinematics(gh)
        sc.be.update_ghost(gh, 1)
        sc.download(g)
        res[make] = {f: getattr(g, f).copy() for f in
                     ("x", "v", "rho", "idx_original", "idx_boundary", "normals", "stress", "p")}
    assert res["dev"]["x"].shape == res["orc"]["x"].shape and res["orc"]["x"].shape[0] > 100
    for f in ("x", "idx_original", "idx_boundary", "normals", "rho", "p"):
        assert np.array_equal(res["dev"][f], res["orc"][f]), f
    for f in ("v", "stress"):
        assert rel(res["dev"][f], res["orc"][f]) < 1e-14, f


# ---------------------------------------------------------------------------
# Whole leapfrog steps (src/TimeIntegration.jl:645-693)
# ---------------------------------------------------------------------------


def _as_rk4(cfg):
    lf = cfg["integrator"]
    cfg["integrator"] = M.RK4TimeIntegrator(lf.systems, lf.interactions, ghosts=lf.ghosts,
                                            virtual_systems=lf.virtual_systems, gamma=lf.gamma)
    return cfg


def _run_steps(builder, kw, D, prec, nsteps, make):
    cfg = builder(**kw)
    lf = cfg["integrator"]
    sc = M.Session(D, prec).scene if make == "dev" else oracle_scene(D)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfg["cfl"])
    sc.be.step(h, nsteps, dt)
    sc.download_all(lf)
    out = {}
    for ps in lf.systems:
        order = np.argsort(ps.id)
        for f in ps.field_names():
            out[f"{ps.name}.{f}"] = getattr(ps, f)[order].copy()
    return out


def _bubble3_xy(**kw):
    """bubble3 with XSPH also on the two-sided Y–X interaction."""
    cfg = M.configs.bubble_2d(**kw)
    cfg["integrator"].interactions[2].vadjust_pfn = M.XSPHPfn(0.3)
    return cfg


STEP_CASES = [
    ("dambreak_2d", M.configs.dambreak_2d, dict(dx=0.04), 2, 20),
    ("ep_column_2d", M.configs.ep_column_2d, dict(dx=0.01, width=0.2, height=0.1), 2, 20),
    ("dambreak_3d", M.configs.dambreak_3d, dict(dx=0.04, water=(0.24, 0.2, 0.24),
                                                tank=(0.48, 0.2, 0.32)), 3, 10),
    ("granular_column_3d", M.configs.granular_column_3d, dict(dx=0.01), 3, 20),
    ("ep_column_virtual_2d", M.configs.ep_column_virtual_2d,
     dict(dx=0.01, width=0.1, height=0.06), 2, 20),
    ("bubble_2d_rk4", M.configs.bubble_2d, dict(dx=0.2), 2, 10),
    ("bubble3_2d_rk4", M.configs.bubble_2d, dict(dx=0.2, variant="bubble3"), 2, 10),
    ("bubble3_xy_xsph_2d_leapfrog", _bubble3_xy, dict(dx=0.2, variant="bubble3", rk4=False), 2,
     10),
    ("ep_column_2d_rk4", lambda **kw: _as_rk4(M.configs.ep_column_2d(**kw)),
     dict(dx=0.01, width=0.2, height=0.1), 2, 10),
    ("dambreak_wall_2d", M.configs.dambreak_wall_2d, dict(), 2, 20),
]


# f64 state tolerance per case: the XSPH scenes sum the velocity adjustment in
# another order than the oracle's pairwise sweep, and over ten steps of the
# multiphase flow the last-bit differences grow; a single XSPH step is held to
# 1e-12 below
F64_TOL = {"bubble3_2d_rk4": 1e-9, "bubble3_xy_xsph_2d_leapfrog": 1e-9}


@pytest.mark.parametrize("name,builder,kw,D,nsteps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_steps_f64(name, builder, kw, D, nsteps):
    dev = _run_steps(builder, kw, D, "f64", nsteps, "dev")
    orc = _run_steps(builder, kw, D, "f64", nsteps, "orc")
    base = F64_TOL.get(name, 1e-10)
    for key, ref in orc.items():
        if key.endswith(".id"):
            continue
        if not np.any(ref):
            assert not np.any(dev[key]), key
            continue
        e = rel(dev[key], ref)
        tol = 100 * base if key.endswith(("dvdt", "drhodt", "strain_rate", "vorticity")) else base
        assert e < tol, f"{key}: {e:.3e}"


@pytest.mark.parametrize("rk4", [False, True])
def test_xsph_one_step_f64(rk4):
    """One step of bubble3 (XSPH on X, Y, X–ghost and Y–X): v_adjustment,
    v and x equal the oracle's to the last bits."""
    kw = dict(dx=0.2, variant="bubble3", rk4=rk4)
    dev = _run_steps(_bubble3_xy, kw, 2, "f64", 1, "dev")
    orc = _run_steps(_bubble3_xy, kw, 2, "f64", 1, "orc")
    for key in (k for k in orc if k.endswith((".x", ".v", ".v_adjustment", ".rho"))):
        assert np.any(orc[key]), key
        e = rel(dev[key], orc[key])
        assert e < 1e-12, f"{key}: {e:.3e}"


@pytest.mark.parametrize("name,builder,kw,D,nsteps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_steps_f32(name, builder, kw, D, nsteps):
    dev = _run_steps(builder, kw, D, "f32", nsteps, "dev")
    orc = _run_steps(builder, kw, D, "f64", nsteps, "orc")
    for key in (k for k in orc if k.endswith((".x", ".rho", ".v", ".stress", ".p"))):
        ref = orc[key]
        if not np.any(ref):
            continue
        e = rel(dev[key], ref)
        tol = 2e-4 if key.endswith((".v", ".stress", ".p")) else 1e-5
        assert e < tol, f"{key}: {e:.3e}"


def test_step_many_equals_alone():
    """Two simulations stepped in the same launches equal each stepped alone
    (bit-identical: every particle's sums depend only on its own simulation)."""
    cfgs = [M.configs.ep_column_2d(dx=0.01, width=w, height=0.08, name=f"c{k}")
            for k, w in enumerate((0.1, 0.16))]
    alone = []
    for k in range(2):
        cfg = M.configs.ep_column_2d(dx=0.01, width=(0.1, 0.16)[k], height=0.08, name=f"c{k}")
        s = M.Session(2, "f32")
        h = s.scene.integrator(cfg["integrator"])
        s.scene.upload_all(cfg["integrator"])
        dt = s.backend.integrator_dt(h, cfg["cfl"])
        s.backend.step(h, 15, dt)
        s.scene.download_all(cfg["integrator"])
        alone.append(cfg["fluid"])
    s = M.Session(2, "f32")
    hs = []
    for cfg in cfgs:
        h = s.scene.integrator(cfg["integrator"])
        s.scene.upload_all(cfg["integrator"])
        s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
        hs.append(h)
    s.backend.step_many(hs, 15)
    for cfg, ref in zip(cfgs, alone):
        s.scene.download_all(cfg["integrator"])
        for f in ("x", "v", "rho", "stress", "dvdt", "strain_p"):
            assert np.array_equal(getattr(cfg["fluid"], f), getattr(ref, f)), f


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 2e-4)])
def test_ensemble_members_match_the_oracle_alone(prec, tol):
    """Eight granular columns of the ensemble workload (aspect, width, φ and ρ
    drawn per simulation, dx = 5 mm) advanced together in shared launches
    (gr_step_many, graph replay); four members with different φ and ρ each
    equal the oracle running that column alone (src/TimeIntegration.jl:
    645-693): particle order bit for bit, state to the step tolerance."""
    nsteps = 25
    sims = M.configs.ensemble_columns(n_sims=8, dx=0.005, seed=3)
    s = M.Session(2, prec)
    s.backend.set_graph_mode(True)
    hs = []
    for cfg in sims:
        h = s.scene.integrator(cfg["integrator"])
        s.scene.upload_all(cfg["integrator"])
        s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
        hs.append(h)
    s.backend.step_many(hs, 10)
    s.backend.step_many(hs, nsteps - 10)
    phis, rhos = [], []
    for k in (0, 2, 5, 7):
        cfg = sims[k]
        s.scene.download_all(cfg["integrator"])
        dev = cfg["fluid"]
        ref = M.configs.ensemble_columns(n_sims=8, dx=0.005, seed=3)[k]
        sc = oracle_scene(2)
        h = sc.integrator(ref["integrator"])
        sc.upload_all(ref["integrator"])
        sc.be.step(h, nsteps, sc.be.integrator_dt(h, ref["cfl"]))
        sc.download_all(ref["integrator"])
        o = ref["fluid"]
        upd = o.state_updater[1]
        phis.append(upd.phi)
        rhos.append(float(o.rho.max()))
        assert np.array_equal(dev.id, o.id), k
        assert np.array_equal(dev.x, o.x) or rel(dev.x, o.x) <= tol, k
        for f in ("x", "v", "rho", "stress", "strain_p"):
            assert rel(getattr(dev, f), getattr(o, f)) <= tol, (k, f, rel(getattr(dev, f), getattr(o, f)))
    assert len(set(np.round(phis, 6))) == 4 and len(set(np.round(rhos, 3))) == 4


def test_graph_mode_equals_eager():
    res = []
    for graph in (False, True):
        cfg = M.configs.dambreak_2d(dx=0.04)
        s = M.Session(2, "f32")
        s.backend.set_graph_mode(graph)
        lf = cfg["integrator"]
        h = s.scene.integrator(lf)
        s.scene.upload_all(lf)
        s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
        for _ in range(3):  # replays with host work in between
            s.backend.step_many([h], 5)
            s.scene.download_all(lf)
        res.append({f: getattr(cfg["fluid"], f).copy() for f in ("x", "v", "rho", "p")})
    for f in res[0]:
        assert np.array_equal(res[0][f], res[1][f]), f


def _translated_dambreak(offset, **kw):
    """The 2D dam break moved by `offset` (particles and wall points)."""
    cfg = M.configs.dambreak_2d(**kw)
    lf = cfg["integrator"]
    off = np.asarray(offset, float)
    for ps in lf.systems:
        ps.x[:] = ps.x + off
    for si in lf.interactions:
        b = si.system_b
        if isinstance(b, M.DynamicBoundarySystem):
            b.boundary_point = b.boundary_point + off
    return cfg


@pytest.mark.parametrize("offset", [(0.0, 0.0), (1.0e5, -3.0e4)])
def test_fp32_steps_translation_invariant(offset):
    """FP32 steps far from the origin: the fluid-only force sweep takes pair
    offsets from 32-bit fixed-point positions kept modulo 2^32 (here wrapping
    about a million times), and must match the oracle as closely as at the
    origin."""
    kw = dict(dx=0.04)
    dev = _run_steps(lambda **k: _translated_dambreak(offset, **k), kw, 2, "f32", 20, "dev")
    orc = _run_steps(lambda **k: _translated_dambreak(offset, **k), kw, 2, "f64", 20, "orc")
    for key in (k for k in orc if k.endswith((".rho", ".v", ".p"))):
        ref = orc[key]
        if not np.any(ref):
            continue
        e = rel(dev[key], ref)
        tol = 2e-4 if key.endswith((".v", ".p")) else 1e-5
        assert e < tol, f"{key}: {e:.3e}"
    f = "fluid.x"
    assert np.max(np.abs(dev[f] - orc[f])) < 1e-9 * (1.0 + np.max(np.abs(orc[f])))


def test_step_skips_unwalked_cells_and_create_grid_builds_them():
    """A step builds no fluid cells for the fluid–wall grids (the dynamic
    wall functor writes only the fluid); reading them then fails with a
    state error instead of returning stale starts, and create_grid builds
    them."""
    cfg = M.configs.dambreak_2d(dx=0.04)
    lf = cfg["integrator"]
    sc = M.Session(2, "f32").scene
    h = sc.integrator(lf)
    sc.upload_all(lf)
    sc.be.step(h, 2, sc.be.integrator_dt(h, cfg["cfl"]))
    si = lf.interactions[1]
    assert si.system_b is not None
    hi = sc.interaction(si)
    _, _, ncells = sc.be.get_grid(hi, 2)
    assert len(sc.be.get_cell_start(hi, 0, ncells + 1)) == ncells + 1
    with pytest.raises(Exception, match="not built"):
        sc.be.get_cell_start(hi, 1, ncells + 1)
    sc.be.create_grid(hi)
    _, _, ncells = sc.be.get_grid(hi, 2)
    cs = sc.be.get_cell_start(hi, 1, ncells + 1)
    assert cs[-1] == cfg["fluid"].n + 1


LIST_CASES = [
    ("ep_column_2d", M.configs.ep_column_2d, dict(dx=0.01, width=0.2, height=0.1), 2, "f32"),
    ("dambreak_3d", M.configs.dambreak_3d, dict(dx=0.04, water=(0.24, 0.2, 0.24),
                                                tank=(0.48, 0.2, 0.32)), 3, "f32"),
    ("dambreak_2d_f64", M.configs.dambreak_2d, dict(dx=0.04), 2, "f64"),
    ("granular_column_3d", M.configs.granular_column_3d, dict(dx=0.01), 3, "f32"),
    ("ep_column_virtual_2d", M.configs.ep_column_virtual_2d,
     dict(dx=0.01, width=0.1, height=0.06), 2, "f32"),
    ("dambreak_wall_2d", M.configs.dambreak_wall_2d, dict(), 2, "f32"),
]


@pytest.mark.parametrize("name,builder,kw,D,prec", LIST_CASES, ids=[c[0] for c in LIST_CASES])
def test_neighbour_lists_equal_cell_walk(monkeypatch, name, builder, kw, D, prec):
    """Steps through the neighbour lists, through the direct cell walk, and
    with lists so short that most targets overflow and fall back to the walk
    give bit-identical states: the lists keep the partners in the walk's
    order and the distance test is the same FP64 expression."""
    res = {}
    for mode, env in (("list", {}), ("walk", {"GRASPH_NO_NLIST": "1"}),
                      ("overflow", {"GRASPH_NLIST_MAXNB": "6"})):
        for k in ("GRASPH_NO_NLIST", "GRASPH_NLIST_MAXNB"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        res[mode] = _run_steps(builder, kw, D, prec, 8, "dev")
    for mode in ("walk", "overflow"):
        for key, ref in res["list"].items():
            assert np.array_equal(res[mode][key], ref, equal_nan=True), f"{mode} {key}"


# ---------------------------------------------------------------------------
# Golden fixtures (tests/golden/make_golden.py): the device in FP64 gives the
# fixture's permutation, CSR and counts bit for bit and its one-step state to
# rounding.
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("D", [2, 3])
def test_device_reproduces_golden(D):
    import sys
    from pathlib import Path
    here = Path(__file__).resolve().parent
    sys.path.insert(0, str(here / "golden"))
    import make_golden as MG
    with np.load(here / "golden" / f"golden_{D}d.npz", allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    dev = MG.make(D, scene_factory=lambda nd: M.Session(nd, "f64").scene)
    for k in ("perm", "mingridx", "ngridx", "cell_start", "counts"):
        assert np.array_equal(dev[k], g[k]), k
    for k in ("step_x", "step_v", "step_rho", "step_p", "step_dvdt", "step_drhodt"):
        assert rel(dev[k], g[k]) < 1e-12, k


def test_release_reuses_segments_and_keeps_results():
    """Integrators admitted and retired on one context (the serving pattern):
    after A is released, C takes A's pool segments; B and C stepped together
    (graph replay on) equal each stepped alone on a fresh context."""
    def col(w, name):
        return M.configs.ep_column_2d(dx=0.01, width=w, height=0.08, name=name)

    def alone(w, name, nsteps):
        cfg = col(w, name)
        s = M.Session(2, "f32")
        h = s.scene.integrator(cfg["integrator"])
        s.scene.upload_all(cfg["integrator"])
        s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
        s.backend.step_many([h], nsteps)
        s.scene.download_all(cfg["integrator"])
        return cfg["fluid"]

    s = M.Session(2, "f32")
    s.backend.set_graph_mode(True)
    cfgs = {k: col(w, k) for k, w in (("a", 0.12), ("b", 0.1), ("c", 0.12))}
    hs = {}
    for k in ("a", "b"):
        lf = cfgs[k]["integrator"]
        hs[k] = s.scene.integrator(lf)
        s.scene.upload_all(lf)
        s.backend.integrator_set_dt(hs[k], s.backend.integrator_dt(hs[k], cfgs[k]["cfl"]))
    s.backend.step_many([hs["a"], hs["b"]], 5)
    s.scene.release(cfgs["a"]["integrator"])
    lf = cfgs["c"]["integrator"]
    hs["c"] = s.scene.integrator(lf)
    s.scene.upload_all(lf)
    s.backend.integrator_set_dt(hs["c"], s.backend.integrator_dt(hs["c"], cfgs["c"]["cfl"]))
    s.backend.step_many([hs["c"]], 5)        # c catches up with b
    s.backend.step_many([hs["b"], hs["c"]], 5)
    ref_b = alone(0.1, "b", 10)
    ref_c = alone(0.12, "c", 10)
    for k, ref in (("b", ref_b), ("c", ref_c)):
        s.scene.download_all(cfgs[k]["integrator"])
        for f in ("x", "v", "rho", "stress"):
            assert np.array_equal(getattr(cfgs[k]["fluid"], f), getattr(ref, f)), (k, f)
    with pytest.raises(RuntimeError):
        s.backend.step_many([hs["a"]], 1)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_device_snapshot_restart_continues_bit_identically(tmp_path, prec):
    """Device: 12 steps straight vs 7 steps, HDF5 snapshot in the reference
    layout (with dvdt/drhodt and ghosts), restart into fresh objects on a new
    context, 5 more steps (src/Driver.jl:79-96)."""
    SN = M.snapshot

    def cfg():
        return M.configs.ep_column_2d(dx=0.01, width=0.1, height=0.06)

    def run(c, nsteps):
        lf = c["integrator"]
        s = M.Session(2, prec)
        h = s.scene.integrator(lf)
        s.scene.upload_all(lf)
        dt = s.backend.integrator_dt(h, c["cfl"])
        s.backend.step(h, nsteps, dt)
        s.scene.download_all(lf)
        return dt

    ref = cfg()
    run(ref, 12)
    a = cfg()
    dt = run(a, 7)
    path = str(tmp_path / "col_07.h5")
    SN.save_snapshot(path, a["integrator"].systems, a["integrator"].ghosts, 7, dt)
    b = cfg()
    assert SN.load_restart(path, b["integrator"].systems) == 7
    run(b, 5)
    for f in ["x"] + SN.sim_field_names(ref["fluid"]):
        assert np.array_equal(getattr(b["fluid"], f), getattr(ref["fluid"], f)), f


@pytest.mark.parametrize("writer", ["orc", "dev"])
def test_snapshot_cross_restart_device_oracle(tmp_path, writer):
    """One backend runs 7 steps and writes an HDF5 snapshot in the reference
    layout (src/Particles.jl:711-735, src/TimeIntegration.jl:520-549); the
    other restarts from that file (src/Particles.jl:747-787,
    src/Driver.jl:79-96) and runs 5 more steps, while the writer continues
    its own run.  The two continuations agree to the f64 step tolerance, the
    particle order bit for bit; the file's state equals what the reader's own
    run reaches at step 7 to the same tolerance."""
    SN = M.snapshot
    reader = "dev" if writer == "orc" else "orc"

    def cfg():
        return M.configs.ep_column_2d(dx=0.01, width=0.1, height=0.06)

    def session(make):
        return M.Session(2, "f64").scene if make == "dev" else oracle_scene(2)

    def start(make, c):
        lf = c["integrator"]
        sc = session(make)
        h = sc.integrator(lf)
        sc.upload_all(lf)
        return sc, h, sc.be.integrator_dt(h, c["cfl"])

    def state(c):
        ps = c["fluid"]
        return {f: getattr(ps, f).copy() for f in ["x", "id"] + SN.sim_field_names(ps)}

    # writer: 7 steps, snapshot, 5 more steps
    w = cfg()
    sc, h, dt = start(writer, w)
    sc.be.step(h, 7, dt)
    sc.download_all(w["integrator"])
    path = str(tmp_path / f"{writer}_07.h5")
    SN.save_snapshot(path, w["integrator"].systems, w["integrator"].ghosts, 7, dt)
    at7_writer = state(w)
    sc.be.step(h, 5, dt)
    sc.download_all(w["integrator"])
    cont_writer = state(w)
    # the reader's own run to step 7 agrees with the file
    r7 = cfg()
    sc7, h7, dt7 = start(reader, r7)
    assert dt7 == dt
    sc7.be.step(h7, 7, dt7)
    sc7.download_all(r7["integrator"])
    own7 = state(r7)
    assert np.array_equal(own7["id"], at7_writer["id"])
    for f, ref in at7_writer.items():
        if f != "id":
            assert rel(own7[f], ref) <= 1e-10, (f, rel(own7[f], ref))
    # reader: restart from the writer's file, 5 steps
    r = cfg()
    assert SN.load_restart(path, r["integrator"].systems) == 7
    scr, hr, dtr = start(reader, r)
    scr.be.step(hr, 5, dtr)
    scr.download_all(r["integrator"])
    cont_reader = state(r)
    assert np.array_equal(cont_reader["id"], cont_writer["id"])
    for f, ref in cont_writer.items():
        if f != "id":
            assert rel(cont_reader[f], ref) <= 1e-10, (f, rel(cont_reader[f], ref))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_virtual_interpolation_matches_oracle(prec, tol):
    """The small virtual scene of the oracle tests (interpolation, zero
    weight, normalisation with v_mult and prescribed velocity, accumulation
    of fields outside zero_fields, prescribed motion) on the device."""
    import test_oracle_virtual as TV
    res = {}
    for make in ("dev", "orc"):
        real, vps, lf, vx = TV.scene(zero_fields=("v", "rho"), vnorm=("v", "rho"),
                                     pv=(0.2, -0.1), shuffle_virtual=True,
                                     interp2=M.InterpolateFieldFn("v", accumulate_wsum=False))
        sc = M.Session(2, prec).scene if make == "dev" else oracle_scene(2)
        h = sc.integrator(lf)
        sc.upload_all(lf)
        sc.be.step(h, 3, 1e-4)
        sc.download_all(lf)
        res[make] = {f: getattr(vps, f).copy() for f in ("x", "v", "rho", "p", "w_sum", "id")}
        res[make].update({"real_" + f: getattr(real, f).copy() for f in ("x", "v", "rho")})
    assert np.array_equal(res["dev"]["id"], res["orc"]["id"])
    for k, ref in res["orc"].items():
        if k == "id":
            continue
        assert rel(res["dev"][k], ref) <= tol, (k, rel(res["dev"][k], ref))


def test_cell_arrays_follow_a_spreading_flow():
    """Two fluid clusters fly apart at 20 m/s each (the sound speed): their
    common grid spreads to ten times its first extent over the run, far beyond the cell arrays
    planned at creation.  The device grows them between chunks of steps
    and matches the oracle (which has no capacity) at the f64 tolerance,
    with no fault."""
    def scene():
        dx, h = 0.01, 0.013
        n1 = 12
        xs = (np.stack(np.meshgrid(np.arange(n1), np.arange(n1), indexing="ij"), -1)
              .reshape(-1, 2) + 0.5) * dx
        x = np.concatenate([xs, xs + [0.3, 0.0]])
        f = M.FluidParticleSystem("f", len(x), 2, 1000 * dx * dx, 20.0,
                                  state_updater=M.TaitEOSUpdater(1000.0))
        f.x[:] = x
        rng = np.random.default_rng(4)
        f.rho[:] = 1000.0 * (1 + rng.uniform(-1e-3, 1e-3, len(x)))
        f.v[:] = rng.uniform(-0.2, 0.2, (len(x), 2))
        f.v[: len(xs), 0] -= 20.0
        f.v[len(xs):, 0] += 20.0
        si = M.SystemInteraction(M.WenlandC2Kernel(h, ndims=2), M.FluidPfn(0.1, 0.0, h), f)
        return f, M.LeapFrogTimeIntegrator([f], [si])

    res = {}
    for make in ("dev", "orc"):
        f, lf = scene()
        sc = M.Session(2, "f64").scene if make == "dev" else oracle_scene(2)
        h = sc.integrator(lf)
        sc.upload_all(lf)
        dt = sc.be.integrator_dt(h, 0.2)
        for _ in range(4):
            sc.be.step(h, 200, dt)
        sc.download_all(lf)
        if make == "dev":
            assert sc.be.integrator_status(h) == 0
        res[make] = (f.id.copy(), f.x.copy(), f.v.copy(), f.rho.copy())
    assert np.array_equal(res["dev"][0], res["orc"][0])
    span = res["orc"][1][:, 0].max() - res["orc"][1][:, 0].min()
    assert span > 10 * 0.42  # the flow spread far beyond its first extent
    for a, b in zip(res["dev"][1:], res["orc"][1:]):
        assert rel(a, b) <= 1e-10, rel(a, b)
