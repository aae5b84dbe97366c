"""Synthetic, seeded scenes of the benchmark workloads.

Built with the same object model a reference user writes in a script
(patterns of dambreak.jl, DambreakWall.jl:109-171, EP_ColumnCollapse.jl,
Trapdoor.jl:165-176, bubble.jl:94-98).  Lattice particles sit at (i+½)dx.
Every builder takes the spacing as a parameter so the test-suite can run the
same scene at a size the CPU oracle finishes in seconds.
"""
from __future__ import annotations

import math

import numpy as np

from . import model as M

G = 9.81


def _lattice(nx, ny, dx, nz=None, origin=(0.0, 0.0, 0.0)):
    if nz is None:
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        return np.stack([(i.ravel() + 0.5) * dx + origin[0],
                         (j.ravel() + 0.5) * dx + origin[1]], axis=1)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([(i.ravel() + 0.5) * dx + origin[0], (j.ravel() + 0.5) * dx + origin[1],
                     (k.ravel() + 0.5) * dx + origin[2]], axis=1)


def _hydrostatic_rho(height_above, rho0, c0, gamma=7.0):
    """Tait inverse for p = ρ0·g·depth (bubble.jl:94-98)."""
    p = height_above * rho0 * G
    return (p * gamma / (c0 * c0 * rho0) + 1.0) ** (1.0 / gamma) * rho0


def _boundary(name, pts, mass, c, normal, point, beta=3.0):
    inner = M.BasicParticleSystem(name, len(pts), pts.shape[1], mass, c)
    inner.x[:] = pts
    inner.rho[:] = 1000.0
    return inner, M.DynamicBoundarySystem(inner, normal, point, beta)


def _round(v):
    return int(round(v))


def dambreak_2d(dx=0.01, water=(1.0, 2.0), tank=(4.0, 3.0), layers=3, alpha=0.1,
                h_fac=1.3, cfl=0.2):
    """2D weakly-compressible dam break with 3-layer dynamic walls
    (floor, left, right): Wendland C2 h = 1.3dx, Tait ρ0 = 1000, γ = 7,
    c0 = 10·sqrt(2·g·H), α = 0.1, gravity −9.81, hydrostatic ρ, v = 0."""
    rho0 = 1000.0
    H = water[1]
    c0 = 10.0 * math.sqrt(2.0 * G * H)
    h = h_fac * dx
    nwx, nwy = _round(water[0] / dx), _round(water[1] / dx)
    ntx, nty = _round(tank[0] / dx), _round(tank[1] / dx)
    mass = rho0 * dx * dx
    fluid = M.FluidParticleSystem("fluid", nwx * nwy, 2, mass, c0, source_v=[0.0, -G],
                                  state_updater=M.TaitEOSUpdater(rho0))
    fluid.x[:] = _lattice(nwx, nwy, dx)
    fluid.rho[:] = _hydrostatic_rho(H - fluid.x[:, 1], rho0, c0)
    L = layers
    fl = np.array([((i + 0.5) * dx, -(l + 0.5) * dx) for l in range(L) for i in range(-L, ntx + L)])
    le = np.array([(-(l + 0.5) * dx, (j + 0.5) * dx) for l in range(L) for j in range(nty)])
    ri = np.array([(tank[0] + (l + 0.5) * dx, (j + 0.5) * dx) for l in range(L) for j in range(nty)])
    floor, floor_w = _boundary("floor", fl, mass, c0, (0.0, 1.0), (0.0, 0.0))
    left, left_w = _boundary("left", le, mass, c0, (1.0, 0.0), (0.0, 0.0))
    right, right_w = _boundary("right", ri, mass, c0, (-1.0, 0.0), (tank[0], 0.0))
    kern = M.WenlandC2Kernel(h, ndims=2)
    pfn = M.FluidPfn(alpha, 0.0, h)
    ints = [M.SystemInteraction(kern, pfn, fluid)] + [
        M.SystemInteraction(kern, pfn, fluid, w) for w in (floor_w, left_w, right_w)]
    lf = M.LeapFrogTimeIntegrator([fluid, floor, left, right], ints)
    return {"integrator": lf, "fluid": fluid, "cfl": cfl, "name": "dambreak_2d",
            "n_particles": sum(ps.n for ps in lf.systems), "fluid_n": fluid.n}


def dambreak_3d(dx=0.006, water=(0.6, 0.6, 0.6), tank=(1.6, 0.6, 0.8), layers=3,
                alpha=0.1, h_fac=1.3, cfl=0.2, seed=20261202):
    """3D dam break (z vertical): water block in the x-low corner of the
    tank, 3-layer dynamic floor and four walls (no lid), Wendland C2
    h = 1.3dx, Tait, c0 = 10·sqrt(2·g·H), α = 0.1; fluid order shuffled by
    `seed` before the first sort."""
    rho0 = 1000.0
    H = water[2]
    c0 = 10.0 * math.sqrt(2.0 * G * H)
    h = h_fac * dx
    nw = [_round(w / dx) for w in water]
    nt = [_round(t / dx) for t in tank]
    mass = rho0 * dx ** 3
    fluid = M.FluidParticleSystem("fluid", nw[0] * nw[1] * nw[2], 3, mass, c0,
                                  source_v=[0.0, 0.0, -G], state_updater=M.TaitEOSUpdater(rho0))
    x = _lattice(nw[0], nw[1], dx, nw[2])
    perm = np.random.default_rng(seed).permutation(len(x))
    fluid.x[:] = x[perm]
    fluid.rho[:] = _hydrostatic_rho(H - fluid.x[:, 2], rho0, c0)
    L = layers
    r = np.arange
    # floor covers the corners under the walls
    fi, fj, fl = np.meshgrid(r(-L, nt[0] + L), r(-L, nt[1] + L), r(L), indexing="ij")
    floor_pts = np.stack([(fi + 0.5) * dx, (fj + 0.5) * dx, -(fl + 0.5) * dx], -1).reshape(-1, 3)
    wj, wk, wl = np.meshgrid(r(nt[1]), r(nt[2]), r(L), indexing="ij")
    left_pts = np.stack([-(wl + 0.5) * dx, (wj + 0.5) * dx, (wk + 0.5) * dx], -1).reshape(-1, 3)
    right_pts = np.stack([tank[0] + (wl + 0.5) * dx, (wj + 0.5) * dx, (wk + 0.5) * dx], -1).reshape(-1, 3)
    bi, bk, bl = np.meshgrid(r(-L, nt[0] + L), r(nt[2]), r(L), indexing="ij")
    front_pts = np.stack([(bi + 0.5) * dx, -(bl + 0.5) * dx, (bk + 0.5) * dx], -1).reshape(-1, 3)
    back_pts = np.stack([(bi + 0.5) * dx, tank[1] + (bl + 0.5) * dx, (bk + 0.5) * dx], -1).reshape(-1, 3)
    walls = [("floor", floor_pts, (0, 0, 1.0), (0, 0, 0)),
             ("left", left_pts, (1.0, 0, 0), (0, 0, 0)),
             ("right", right_pts, (-1.0, 0, 0), (tank[0], 0, 0)),
             ("front", front_pts, (0, 1.0, 0), (0, 0, 0)),
             ("back", back_pts, (0, -1.0, 0), (0, tank[1], 0))]
    kern = M.WenlandC2Kernel(h, ndims=3)
    pfn = M.FluidPfn(alpha, 0.0, h)
    inners, ints = [], [M.SystemInteraction(kern, pfn, fluid)]
    for name, pts, nrm, pnt in walls:
        inner, w = _boundary(name, pts, mass, c0, nrm, pnt)
        inners.append(inner)
        ints.append(M.SystemInteraction(kern, pfn, fluid, w))
    lf = M.LeapFrogTimeIntegrator([fluid] + inners, ints)
    return {"integrator": lf, "fluid": fluid, "cfl": cfl, "name": "dambreak_3d",
            "n_particles": sum(ps.n for ps in lf.systems), "fluid_n": fluid.n,
            "boundary_n": sum(ps.n for ps in inners)}


def ep_column_2d(dx=0.0025, width=0.4, height=0.2, rho0=1850.0, E=0.84e6, nu=0.3,
                 phi_deg=19.8, psi=0.0, cohesion=0.0, alpha=0.1, beta=0.1, h_fac=1.5,
                 cfl=0.1, k0=None, name="soil"):
    """2D elasto-plastic soil column (plane strain, NS = 4): cubic spline
    h = 1.5dx, Drucker–Prager (ψ = 0, c = 0), stages
    (StrainRateVorticityPfn, CauchyFluidPfn) as in EP_ColumnCollapse.jl, and
    free-slip ghost floor and left wall with Householder stress reflection
    (the Trapdoor.jl:165-176 pattern; ghost cutoff 3h).  Initial stress is
    geostatic: σyy = −ρ0·g·(H − y), σxx = σzz = K0·σyy with K0 = 1 − sin φ
    (Jaky) unless given."""
    phi = phi_deg * math.pi / 180.0
    c = math.sqrt(E * (1 - nu) / (rho0 * (1 + nu) * (1 - 2 * nu)))
    h = h_fac * dx
    nx, ny = _round(width / dx), _round(height / dx)
    mass = rho0 * dx * dx
    soil = M.ElastoPlasticParticleSystem(
        name, nx * ny, 2, 4, mass, c, source_v=[0.0, -G],
        state_updater=(M.ZeroFieldUpdater("strain_rate", "vorticity"),
                       M.ElastoPlasticStressUpdater(E, nu, phi, psi, cohesion)))
    soil.x[:] = _lattice(nx, ny, dx)
    soil.rho[:] = rho0
    K0 = (1.0 - math.sin(phi)) if k0 is None else k0
    syy = -rho0 * G * (height - soil.x[:, 1])
    soil.stress[:, 0] = K0 * syy
    soil.stress[:, 1] = syy
    soil.stress[:, 2] = K0 * syy
    floor_g = M.GhostParticleSystem(soil, M.GhostCopier(("stress", M.HouseholderReflect())),
                                    name=f"{name}_floor_ghost")
    wall_g = M.GhostParticleSystem(soil, M.GhostCopier(("stress", M.HouseholderReflect())),
                                   name=f"{name}_wall_ghost")
    fe = M.GhostEntry(floor_g, 3.0 * h, ((0.0, 1.0), (0.0, 0.0)))
    we = M.GhostEntry(wall_g, 3.0 * h, ((1.0, 0.0), (0.0, 0.0)))
    kern = M.CubicSplineKernel(h, ndims=2)
    pfns = (M.StrainRateVorticityPfn(), M.CauchyFluidPfn(alpha, beta, h))
    ints = [M.SystemInteraction(kern, pfns, soil),
            M.SystemInteraction(kern, pfns, soil, floor_g),
            M.SystemInteraction(kern, pfns, soil, wall_g)]
    lf = M.LeapFrogTimeIntegrator([soil], ints, ghosts=(fe, we))
    return {"integrator": lf, "fluid": soil, "cfl": cfl, "name": "ep_column_2d",
            "n_particles": soil.n, "fluid_n": soil.n}


def ensemble_columns(n_sims=256, dx=0.0025, seed=7, aspect=(0.5, 4.0), width=(0.1, 0.4),
                     phi_deg=(20.0, 40.0), rho=(1500.0, 2000.0)):
    """Independent 2D granular columns with parameters drawn per simulation
    (one PRNG stream per simulation from `seed`)."""
    sims = []
    for k in range(n_sims):
        g = np.random.default_rng([seed, k])
        a = g.uniform(*aspect)
        w = g.uniform(*width)
        ph = g.uniform(*phi_deg)
        r0 = g.uniform(*rho)
        hgt = a * w
        sims.append(ep_column_2d(dx=dx, width=w, height=hgt, rho0=r0, phi_deg=ph,
                                 name=f"col{k}"))
    return sims


def granular_column_3d(dx=0.005, width=0.2, depth=0.05, height=0.1, rho0=1850.0, c=20.0,
                       phi_deg=19.8, alpha=0.1, beta=0.1, h_fac=1.2, cfl=0.1, box_x=None,
                       name="soil3d"):
    """3D granular column (GranularColumnCollapse3D.jl): StressParticleSystem
    NS = 6 with (ZeroFieldUpdater(:strain_rate),
    ViscoPlasticMCStressUpdater(LinearEOSUpdater(ρ0), φ, 0)), cubic spline
    h = 1.2dx, stages (StrainRatePfn, CauchyFluidPfn(α, β, h)); a 3-layer
    dynamic floor sheet (β = 3) under a box 4× the column width, and ghost
    mirrors (cutoff 3h) on the left, front and back walls and the two
    vertical corners, copying stress at stage 2.  z is vertical."""
    phi = phi_deg * math.pi / 180.0
    h = h_fac * dx
    nfx, nfy, nfz = (max(1, int(math.floor(v / dx + 1e-9))) for v in (width, depth, height))
    mass = rho0 * dx ** 3
    soil = M.StressParticleSystem(
        name, nfx * nfy * nfz, 3, 6, mass, c, source_v=[0.0, 0.0, -G],
        state_updater=(M.ZeroFieldUpdater("strain_rate"),
                       M.ViscoPlasticMCStressUpdater(M.LinearEOSUpdater(rho0), phi, 0.0)))
    soil.x[:] = _lattice(nfx, nfy, dx, nfz)
    soil.rho[:] = rho0
    nbx = int(math.floor((4.0 * width if box_x is None else box_x) / dx + 1e-9))
    nby = nfy
    i, j, m = np.meshgrid(np.arange(1, nbx + 4), np.arange(1, nby + 7), np.arange(1, 4),
                          indexing="ij")
    bot = np.stack([(i - 3.5) * dx, (j - 2.5) * dx, -(m - 0.5) * dx], -1).reshape(-1, 3)
    bottom, bottom_w = _boundary(f"{name}_bottom", bot, mass, c, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0))
    bottom.rho[:] = rho0
    y_max = nfy * dx
    r2 = math.sqrt(2.0)
    ghost = M.GhostParticleSystem(soil, None, M.GhostCopier("stress"), name=f"{name}_ghost")
    entry = M.GhostEntry(ghost, 3.0 * h,
                         ((1.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
                         ((0.0, 1.0, 0.0), (0.0, 0.0, 0.0)),
                         ((0.0, -1.0, 0.0), (0.0, y_max, 0.0)),
                         ((1.0 / r2, -1.0 / r2, 0.0), (0.0, y_max, 0.0)),
                         ((1.0 / r2, 1.0 / r2, 0.0), (0.0, 0.0, 0.0)))
    kern = M.CubicSplineKernel(h, ndims=3)
    pfns = (M.StrainRatePfn(), M.CauchyFluidPfn(alpha, beta, h))
    ints = [M.SystemInteraction(kern, pfns, soil),
            M.SystemInteraction(kern, pfns, soil, bottom_w),
            M.SystemInteraction(kern, pfns, soil, ghost)]
    lf = M.LeapFrogTimeIntegrator([soil, bottom], ints, ghosts=(entry,))
    return {"integrator": lf, "fluid": soil, "cfl": cfl, "name": "granular_column_3d",
            "n_particles": soil.n + bottom.n, "fluid_n": soil.n, "boundary_n": bottom.n}


def ep_column_virtual_2d(dx=0.0025, width=0.2, height=0.1, rho0=1850.0, E=0.84e6, nu=0.3,
                         phi_deg=19.8, psi=0.0, cohesion=0.0, alpha=0.1, beta=0.1, h_fac=1.5,
                         cfl=0.1, name="soil"):
    """EP column collapse with VirtualParticleSystem boundaries
    (EP_ColumnCollapse2.jl): four stages per step, (InterpolateFieldFn(:v,
    :rho), StrainRateVorticityPfn, InterpolateFieldFn(:stress; no w_sum),
    CauchyFluidPfn) against a no-slip virtual floor (v_mult (−1, −1)) and a
    free-slip virtual left wall (v_mult (−1, 1)), 3 layers each; soil
    updaters (nothing, Zero(:strain_rate, :vorticity), Drucker–Prager,
    nothing); initial stress from the stage-3 updater at rest (zero)."""
    phi = phi_deg * math.pi / 180.0
    c = math.sqrt(E * (1 - nu) / (rho0 * (1 + nu) * (1 - 2 * nu)))
    h = h_fac * dx
    nfx, nfy = _round(width / dx), _round(height / dx)
    nbx = _round(4.0 * width / dx)
    mass = rho0 * dx * dx
    soil = M.ElastoPlasticParticleSystem(
        name, nfx * nfy, 2, 4, mass, c, source_v=[0.0, -G],
        state_updater=(None, M.ZeroFieldUpdater("strain_rate", "vorticity"),
                       M.ElastoPlasticStressUpdater(E, nu, phi, psi, cohesion), None))
    soil.x[:] = _lattice(nfx, nfy, dx)
    soil.rho[:] = rho0
    i, j = np.meshgrid(np.arange(1, nbx + 4), np.arange(1, 4), indexing="ij")
    bot_src = M.StressParticleSystem(f"{name}_bottom_boundary", i.size, 2, 4, mass, c)
    bot_src.x[:] = np.stack([(i - 3.5) * dx, -(j - 0.5) * dx], -1).reshape(-1, 2)
    bot_src.rho[:] = rho0
    jj, col = np.meshgrid(np.arange(nfy), np.arange(1, 4), indexing="ij")
    left_src = M.StressParticleSystem(f"{name}_left_boundary", jj.size, 2, 4, mass, c)
    left_src.x[:] = np.stack([-(col - 0.5) * dx, (jj + 0.5) * dx], -1).reshape(-1, 2)
    left_src.rho[:] = rho0
    bottom = M.VirtualParticleSystem(
        bot_src, f"{name}_bottom_virt", bot_src.n, 2, mass, c, zero_fields=("v", "rho", "stress"),
        state_updater=(None, M.VirtualNormUpdater((-1.0, -1.0), "v", "rho"), None,
                       M.VirtualNormUpdater((-1.0, -1.0), "stress")))
    left = M.VirtualParticleSystem(
        left_src, f"{name}_left_virt", left_src.n, 2, mass, c, zero_fields=("v", "rho", "stress"),
        state_updater=(None, M.VirtualNormUpdater((-1.0, 1.0), "v", "rho"), None,
                       M.VirtualNormUpdater((-1.0, 1.0), "stress")))
    kern = M.CubicSplineKernel(h, ndims=2)
    sr = M.StrainRateVorticityPfn()
    kin = M.CauchyFluidPfn(alpha, beta, h)
    iv = M.InterpolateFieldFn("v", "rho", accumulate_wsum=True)
    ist = M.InterpolateFieldFn("stress", accumulate_wsum=False)
    ints = [M.SystemInteraction(kern, (None, sr, None, kin), soil),
            M.SystemInteraction(kern, (iv, sr, ist, kin), soil, bottom),
            M.SystemInteraction(kern, (iv, sr, ist, kin), soil, left)]
    lf = M.LeapFrogTimeIntegrator([soil], ints, virtual_systems=(bottom, left))
    return {"integrator": lf, "fluid": soil, "cfl": cfl, "name": "ep_column_virtual_2d",
            "n_particles": soil.n + bot_src.n + left_src.n, "fluid_n": soil.n,
            "virtual": (bottom, left)}


def bubble_2d(dx=0.05, R=1.0, variant="bubble", cfl=1.5, rk4=True, xsph=None):
    """Rising bubble (bubble.jl / bubble3.jl, after Colagrossi & Landrini
    2003): a 6R × 10R column of heavy fluid X with a light fluid Y (half the
    density) filling x² + y² < R, hydrostatic densities, Wendland C2
    h = 1.2dx, ghost walls on all four sides and corners copying p (cutoff
    3h), RK4.  variant "bubble": FluidPfn(0.01, 0, h) everywhere, Tait γ = 7
    for both fluids.  variant "bubble3": FluidPfn with S = 1, surface tension
    ε = 0.1 between the phases, γ = 1.4 for Y, and XSPHPfn(0.5) on the X, Y
    and X–ghost interactions (xsph overrides ε_XSPH; 0 switches it off)."""
    height, width = 10.0 * R, 6.0 * R
    y_min, x_min = -2.0 * R, -3.0 * R
    y_max, x_max = y_min + height, x_min + width
    h = 1.2 * dx
    rho_x, rho_y = 1000.0, 500.0
    g = 9.81
    c_x = math.sqrt(800.0 * g * R)
    c_y = 40.0 * math.sqrt(g * R)
    alpha, beta = 0.01, 0.0
    nx, ny = int(math.floor(width / dx)), int(math.floor(height / dx))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = x_min + (i.ravel() + 0.5) * dx
    y = y_min + (j.ravel() + 0.5) * dx
    inb = x * x + y * y < R
    gam_y = 1.4 if variant == "bubble3" else 7.0
    fx = M.FluidParticleSystem("fluid X", int((~inb).sum()), 2, dx * dx * rho_x, c_x,
                               source_v=[0.0, -g], state_updater=M.TaitEOSUpdater(rho_x))
    fy = M.FluidParticleSystem("fluid Y", int(inb.sum()), 2, dx * dx * rho_y, c_y,
                               source_v=[0.0, -g],
                               state_updater=M.TaitEOSUpdater(rho_y, gamma=gam_y))
    fx.x[:] = np.stack([x[~inb], y[~inb]], 1)
    fy.x[:] = np.stack([x[inb], y[inb]], 1)
    px = (y_min + height - fx.x[:, 1]) * rho_x * g
    fx.rho[:] = (px * 7.0 / (c_x * c_x * rho_x) + 1.0) ** (1.0 / 7.0) * rho_x
    py = (y_min + height - fy.x[:, 1]) * rho_x * g
    fy.rho[:] = (py * gam_y / (c_y * c_y * rho_y) + 1.0) ** (1.0 / gam_y) * rho_y
    ghost = M.GhostParticleSystem(fx, M.GhostCopier("p"), name="ghost[fluid_X]")
    r2 = math.sqrt(2.0)
    entry = M.GhostEntry(ghost, 3.0 * h,
                         ((1.0, 0.0), (x_min, 0.0)), ((-1.0, 0.0), (x_max, 0.0)),
                         ((0.0, 1.0), (0.0, y_min)), ((0.0, -1.0), (0.0, y_max)),
                         ((1.0 / r2, 1.0 / r2), (x_min, y_min)),
                         ((-1.0 / r2, 1.0 / r2), (x_max, y_min)),
                         ((1.0 / r2, -1.0 / r2), (x_min, y_max)),
                         ((-1.0 / r2, -1.0 / r2), (x_max, y_max)))
    kern = M.WenlandC2Kernel(h, ndims=2)
    if variant == "bubble3":
        pf = M.FluidPfn(alpha, beta, h, sigma=1)
        pxy = M.FluidPfn(alpha, beta, h, sigma=1, epsilon=0.1)
    else:
        pf = pxy = M.FluidPfn(alpha, beta, h)
    if xsph is None:
        xsph = 0.5 if variant == "bubble3" else 0.0
    va = M.XSPHPfn(xsph) if xsph else None
    ints = [M.SystemInteraction(kern, pf, fx, velocity_adjust_pairwise_fn=va),
            M.SystemInteraction(kern, pf, fy, velocity_adjust_pairwise_fn=va),
            M.SystemInteraction(kern, pxy, fy, fx),
            M.SystemInteraction(kern, pf, fx, ghost, velocity_adjust_pairwise_fn=va)]
    cls = M.RK4TimeIntegrator if rk4 else M.LeapFrogTimeIntegrator
    lf = cls([fx, fy], ints, ghosts=(entry,))
    return {"integrator": lf, "fluid": fx, "light": fy, "cfl": cfl, "name": f"{variant}_2d",
            "n_particles": fx.n + fy.n, "fluid_n": fx.n}


def dambreak_wall_2d(dx=0.5, alpha=0.02, cfl=0.05):
    """Dam break against an elastic concrete wall (DambreakWall.jl): a
    25 m × 25 m water column (Tait at stage 2, c = 10·sqrt(2·g·25)) in a
    75 m × 40 m box of 3-layer dynamic boundaries (floor, left, right, top),
    a 1.5 m × 10 m wall at x = 60 m (ElastoPlastic NS = 4, ρ = 2400,
    E = 5e8, ν = 0.2; stages Zero(strain_rate, vorticity) then
    HookeLawStressUpdater), cubic spline h = 1.2dx; interactions (nothing,
    FluidPfn) for the fluid, (StrainRateVorticityPfn, CauchyFluidPfn) for the
    wall and wall–floor, (nothing, FluidSolidPfn) fluid–wall.  The reference
    geometry is tied to dx = 0.5 (particle counts scale with 0.5/dx)."""
    s = 0.5 / dx
    h = 1.2 * dx
    rho_w, g = 1000.0, 9.81
    c_s = 10.0 * math.sqrt(2.0 * g * 25.0)
    rho_wall, E, nu = 2400.0, 5.0e8, 0.2
    c_wall = math.sqrt(E * (1 - nu) / (rho_wall * (1 + nu) * (1 - 2 * nu)))
    nf = _round(50 * s)
    fmass = rho_w * dx * dx
    fluid = M.FluidParticleSystem("fluid", nf * nf, 2, fmass, c_s, source_v=[0.0, -g],
                                  state_updater=(None, M.TaitEOSUpdater(rho_w)))
    fluid.x[:] = _lattice(nf, nf, dx)
    fluid.rho[:] = rho_w
    nwx, nwy = _round(3 * s), _round(20 * s)
    wall = M.ElastoPlasticParticleSystem(
        "wall", nwx * nwy, 2, 4, rho_wall * dx * dx, c_wall, source_v=[0.0, -g],
        state_updater=(M.ZeroFieldUpdater("strain_rate", "vorticity"),
                       M.HookeLawStressUpdater(E, nu)))
    wall.x[:] = _lattice(nwx, nwy, dx) + np.array([60.0, 0.0])
    wall.rho[:] = rho_wall
    L = 3
    nfx = _round(156 * s)
    fl = np.array([(-1.25 * s * dx / 0.5 + ix * dx, -(l + 0.5) * dx)
                   for l in range(L) for ix in range(nfx)])
    nly = _round(80 * s)
    le = np.array([(-(l + 0.5) * dx, (iy + 0.5) * dx) for l in range(L) for iy in range(nly)])
    ri = np.array([(75.0 + (l + 0.5) * dx, (iy + 0.5) * dx) for l in range(L) for iy in range(nly)])
    tp = np.array([(-1.25 * s * dx / 0.5 + ix * dx, 40.0 + (l + 0.5) * dx)
                   for l in range(L) for ix in range(nfx)])
    floor, floor_w = _boundary("floor", fl, fmass, c_s, (0.0, 1.0), (0.0, 0.0))
    left, left_w = _boundary("left", le, fmass, c_s, (1.0, 0.0), (0.0, 0.0))
    right, right_w = _boundary("right", ri, fmass, c_s, (-1.0, 0.0), (75.0, 0.0))
    top, top_w = _boundary("top", tp, fmass, c_s, (0.0, -1.0), (0.0, 40.0))
    kern = M.CubicSplineKernel(h, ndims=2)
    fp = M.FluidPfn(alpha, 0.0, h)
    cp = M.CauchyFluidPfn(alpha, 0.0, h)
    fs = M.FluidSolidPfn(alpha, 0.0, h)
    sr = M.StrainRateVorticityPfn()
    ints = [M.SystemInteraction(kern, (None, fp), fluid),
            M.SystemInteraction(kern, (sr, cp), wall),
            M.SystemInteraction(kern, (None, fs), fluid, wall),
            M.SystemInteraction(kern, (None, fp), fluid, floor_w),
            M.SystemInteraction(kern, (None, fp), fluid, left_w),
            M.SystemInteraction(kern, (None, fp), fluid, right_w),
            M.SystemInteraction(kern, (None, fp), fluid, top_w),
            M.SystemInteraction(kern, (sr, cp), wall, floor_w)]
    lf = M.LeapFrogTimeIntegrator([fluid, wall, floor, left, right, top], ints)
    # update_state!(fluid, 2): the initial pressure
    fluid.p[:] = (rho_w * c_s * c_s / 7.0) * ((fluid.rho / rho_w) ** 7 - 1.0)
    return {"integrator": lf, "fluid": fluid, "wall": wall, "cfl": cfl,
            "name": "dambreak_wall_2d", "n_particles": sum(ps.n for ps in lf.systems),
            "fluid_n": fluid.n}
