"""Writes the golden fixtures golden_2d.npz and golden_3d.npz.

Provenance: the oracle (oracle/libgrasph_oracle.so, the FP64 C++ restatement
of the reference algorithm), NOT a Julia run of the reference (no Julia is
available here).  Each fixture holds a small seeded scene and, for it:

  * `x_in`            input positions (shuffled lattice with jitter)
  * `perm`            1-based ids after sort_particles! (src/Sorting.jl:157)
  * `mingridx`, `ngridx`, `cell_start`
                      grid of the self interaction after the sort
                      (create_grid!, src/Interaction.jl:123-236)
  * `counts`          neighbours with r < 2h of each particle, in sorted
                      order (counting functor of test_interaction_sweep.jl)
  * `step_*`          state after one leapfrog step of a FluidPfn + Tait
                      scene (src/TimeIntegration.jl:580-697), rows in id order

Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT))

from conftest import grasph as M, oracle_scene  # noqa: E402

PROVENANCE = ("oracle: FP64 C++ restatement of the reference algorithm "
              "(not Julia-verified)")
CASES = {2: dict(shape=(20, 20), seed=202612021), 3: dict(shape=(8, 8, 8), seed=202612023)}
DX = 0.05
H = 1.3 * DX
C0 = 20.0
DT = 1e-4


def scene_input(D):
    c = CASES[D]
    rng = np.random.default_rng(c["seed"])
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in c["shape"]], indexing="ij"),
                   -1).reshape(-1, D)
    x = (idx + 0.5) * DX + rng.uniform(-0.3, 0.3, idx.shape) * DX
    order = rng.permutation(len(x))
    x = x[order]
    v = rng.uniform(-0.5, 0.5, x.shape)
    rho = 1000.0 * (1.0 + rng.uniform(-0.01, 0.01, len(x)))
    return x, v, rho


def fluid(D, x, v, rho, updater=True):
    kw = {"state_updater": M.TaitEOSUpdater(1000.0)} if updater else {}
    ps = M.FluidParticleSystem("f", len(x), D, 1000.0 * DX ** D, C0, **kw)
    ps.x[:] = x
    ps.v[:] = v
    ps.rho[:] = rho
    return ps


def kernel(D):
    return M.WenlandC2Kernel(H, ndims=D)


def make(D, scene_factory=oracle_scene, precision=None):
    """Run the fixture recipe on a scene factory (the oracle by default)."""
    x, v, rho = scene_input(D)
    out = {"x_in": x, "v_in": v, "rho_in": rho, "h": np.float64(H), "dt": np.float64(DT),
           "provenance": np.array(PROVENANCE)}
    # sort + grid + counts
    ps = fluid(D, x, v, rho, updater=False)
    si = M.SystemInteraction(kernel(D), M.NeighborCounterPfn(), ps)
    sc = scene_factory(D)
    sc.upload(ps)
    sc.be.sort(sc.system(ps), 2 * H)
    sc.be.create_grid(sc.interaction(si))
    ming, ng, cs, _ = sc.grid(si)
    sc.be.sweep(sc.interaction(si), 1)
    sc.download(ps)
    out.update(perm=ps.id.copy(), mingridx=ming, ngridx=ng, cell_start=cs,
               counts=ps.drhodt.astype(np.int64))
    # one leapfrog step
    ps = fluid(D, x, v, rho)
    si = M.SystemInteraction(kernel(D), M.FluidPfn(0.1, 0.0, H), ps)
    lf = M.LeapFrogTimeIntegrator([ps], [si])
    sc = scene_factory(D)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    sc.be.step(h, 1, DT)
    sc.download_all(lf)
    o = np.argsort(ps.id)
    for f in ("x", "v", "rho", "p", "dvdt", "drhodt"):
        out["step_" + f] = getattr(ps, f)[o].copy()
    return out


def main():
    for D in (2, 3):
        out = make(D)
        np.savez(HERE / f"golden_{D}d.npz", **out)
        print(f"golden_{D}d.npz: n={len(out['x_in'])} cells={len(out['cell_start'])} "
              f"pairs={int(out['counts'].sum()) // 2}")


if __name__ == "__main__":
    main()
