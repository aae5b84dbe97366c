"""XSPH velocity correction in the oracle (XSPHPfn, src/PairwiseFunctors.jl:
360-390; xsph_veladjust, src/PairwisePhysics.jl:268-270; _xsph_correction!,
src/TimeIntegration.jl:423-436): after the full step every pair adds
du·m_other to each particle's v_adjustment with du = −ε(v_i − v_j)W/(ρ_i+ρ_j)
(pair positions of the step's start), v += v_adjustment, then positions move
with the corrected v.  Checked against the same step without XSPH plus the
closed-form adjustment; total momentum is unchanged."""
import numpy as np
import pytest

from conftest import grasph as M, oracle_lib, oracle_scene

H = 0.05


def pair_scene(eps, coupled=False, gamma=0.0):
    a = M.FluidParticleSystem("a", 3, 2, 2.0, 10.0)
    a.x[:] = [[0.0, 0.0], [0.04, 0.01], [0.02, 0.05]]
    a.v[:] = [[1.0, 0.5], [-0.5, 0.25], [0.3, -0.7]]
    a.rho[:] = [1000.0, 1010.0, 990.0]
    kern = M.WenlandC2Kernel(H, ndims=2)
    vadj = M.XSPHPfn(eps) if eps is not None else None
    if not coupled:
        si = [M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, H), a,
                                  velocity_adjust_pairwise_fn=vadj)]
        return [a], si, M.LeapFrogTimeIntegrator([a], si, gamma=gamma)
    b = M.FluidParticleSystem("b", 2, 2, 3.0, 10.0)
    b.x[:] = [[0.05, 0.04], [-0.03, 0.02]]
    b.v[:] = [[0.1, 0.2], [-0.4, 0.6]]
    b.rho[:] = [995.0, 1005.0]
    si = [M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, H), a, b,
                              velocity_adjust_pairwise_fn=vadj)]
    return [a, b], si, M.LeapFrogTimeIntegrator([a, b], si)


def run(lf, n, dt=1e-4):
    sc = oracle_scene(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    sc.be.step(h, n, dt)
    sc.download_all(lf)


def W(r):
    return oracle_lib().orc_kernel_w(1, 2, H, r / H) if r < 2 * H else 0.0


@pytest.mark.parametrize("coupled", [False, True])
def test_one_step_adds_the_closed_form_adjustment(coupled):
    eps, dt = 0.5, 1e-4
    sysp, _, lfp = pair_scene(None, coupled)
    x0 = [s.x.copy() for s in sysp]
    ids0 = [s.id.copy() for s in sysp]
    run(lfp, 1, dt)
    sysx, _, lfx = pair_scene(eps, coupled)
    run(lfx, 1, dt)
    # undo the sort on both runs: rows by id
    def by_id(s, arr):
        return arr[np.argsort(s.id)]
    vp = [by_id(s, s.v) for s in sysp]
    rp = [by_id(s, s.rho) for s in sysp]
    x0 = [x[np.argsort(i)] for x, i in zip(x0, ids0)]
    adj = [np.zeros_like(v) for v in vp]
    pairs = []
    if not coupled:
        for i in range(3):
            for j in range(i + 1, 3):
                pairs.append((0, i, 0, j))
    else:
        for i in range(3):
            for j in range(2):
                pairs.append((0, i, 1, j))
    for sa, i, sb, j in pairs:
        r = np.linalg.norm(x0[sa][i] - x0[sb][j])
        if r >= 2 * H:
            continue
        du = ((-eps) * (vp[sa][i] - vp[sb][j])) * W(r) / (rp[sa][i] + rp[sb][j])
        adj[sa][i] += du * sysp[sb].mass
        adj[sb][j] -= du * sysp[sa].mass
    for k, s in enumerate(sysx):
        v = by_id(s, s.v)
        np.testing.assert_allclose(v, vp[k] + adj[k], rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(by_id(s, s.v_adjustment), adj[k], rtol=1e-13, atol=1e-18)
        np.testing.assert_allclose(by_id(s, s.x), x0[k] + dt * v, rtol=0, atol=1e-15)
    assert any(np.any(a != 0) for a in adj)
    mom = sum(s.mass * by_id(s, s.v).sum(0) for s in sysx)
    mom_p = sum(s.mass * v.sum(0) for s, v in zip(sysp, vp))
    np.testing.assert_allclose(mom, mom_p, rtol=1e-13, atol=1e-14)


def test_previous_adjustment_is_removed_first():
    """Over two steps the stored v_adjustment is the second step's only."""
    sysx, _, lfx = pair_scene(0.5)
    run(lfx, 2)
    a = sysx[0]
    assert np.any(a.v_adjustment != 0)
    # it is antisymmetric in the pair sums: Σ m·v_adjustment = 0
    np.testing.assert_allclose((a.mass * a.v_adjustment).sum(0), 0.0, atol=1e-15)


def test_ghost_partner_aliases_the_source_index():
    """bubble3.jl puts XSPHPfn on the fluid–ghost interaction.  The coupled
    method writes ps_b.v_adjustment[j] -= du·m_i with ps_b the ghost system,
    whose v_adjustment is the source's array (src/GhostParticles.jl:226-236):
    ghost j's share lands on the source particle with the same index j."""
    eps, dt, dx = 0.5, 1e-4, 0.02
    gx, gy = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")
    n = gx.size
    a = M.FluidParticleSystem("a", n, 2, 0.4, 10.0)
    a.x[:] = np.stack([(gx.ravel() + 0.5) * dx, (gy.ravel() + 0.5) * dx], 1)
    rng = np.random.default_rng(5)
    a.v[:] = rng.normal(size=(n, 2))
    a.rho[:] = 1000.0 + rng.normal(size=n)
    ghost = M.GhostParticleSystem(a, M.GhostCopier("p"))
    entry = M.GhostEntry(ghost, 0.06, ((0.0, 1.0), (0.0, 0.0)))
    kern = M.WenlandC2Kernel(H, ndims=2)
    ints = [M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, H), a),
            M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, H), a, ghost,
                                velocity_adjust_pairwise_fn=M.XSPHPfn(eps))]
    lf = M.LeapFrogTimeIntegrator([a], ints, ghosts=(entry,))
    x0 = a.x.copy()
    sc = oracle_scene(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    sc.be.step(h, 1, dt)
    sc.download_all(lf)
    sc.download(ghost)
    assert 0 < ghost.n < n
    xs = x0[a.id - 1]                      # step-start positions, stored order
    v_pre = a.v - a.v_adjustment           # velocity the sweep saw
    adj = np.zeros_like(a.v)
    for i in range(n):
        for j in range(ghost.n):
            r = np.linalg.norm(xs[i] - ghost.x[j])
            if r >= 2 * H:
                continue
            du = ((-eps) * (v_pre[i] - ghost.v[j])) * W(r) / (a.rho[i] + ghost.rho[j])
            adj[i] += du * a.mass
            adj[j] -= du * a.mass          # the source entry with the ghost's index
    assert np.any(adj[:ghost.n] != 0)
    np.testing.assert_allclose(a.v_adjustment, adj, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(a.x, xs + dt * a.v, rtol=0, atol=1e-15)


def test_ghost_partner_more_ghosts_than_sources_is_an_error():
    a = M.FluidParticleSystem("a", 2, 2, 0.4, 10.0)
    a.x[:] = [[0.01, 0.01], [0.03, 0.01]]
    a.rho[:] = 1000.0
    ghost = M.GhostParticleSystem(a, M.GhostCopier("p"))
    # two walls: every particle is mirrored twice
    entry = M.GhostEntry(ghost, 0.06, ((0.0, 1.0), (0.0, 0.0)), ((1.0, 0.0), (0.0, 0.0)))
    kern = M.WenlandC2Kernel(H, ndims=2)
    ints = [M.SystemInteraction(kern, M.FluidPfn(0.1, 0.0, H), a, ghost,
                                velocity_adjust_pairwise_fn=M.XSPHPfn(0.5))]
    lf = M.LeapFrogTimeIntegrator([a], ints, ghosts=(entry,))
    sc = oracle_scene(2)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    with pytest.raises(ValueError, match="more ghosts than source"):
        sc.be.step(h, 1, 1e-4)
