"""FP32 trajectory accuracy: the device in f32 mode (FP32 pair math with
Kahan sums, FP64 positions, velocity, density and return map) against the
FP64 oracle on the same seeded scene after many steps.  Per field: max |Δ|
over max |ref| (range-normalised) and the L2 relative error, particles
matched by id.  The 2D dam break runs the full configuration (20k fluid
particles); the 3D dam break a reduced-resolution copy of its geometry.
Writes the table to gpurun_out/accuracy.json when that folder exists."""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import grasph as M, oracle_scene

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def run(builder, kw, D, prec, nsteps, make, threads=0):
    cfg = builder(**kw)
    lf = cfg["integrator"]
    if make == "dev":
        sc = M.Session(D, prec).scene
    else:
        sc = oracle_scene(D)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfg["cfl"])
    sc.be.step(h, nsteps, dt)
    sc.download_all(lf)
    f = cfg["fluid"]
    o = np.argsort(f.id)
    return {k: getattr(f, k)[o].copy() for k in ("x", "v", "rho", "p")}


def errors(dev, ref):
    out = {}
    for k in ref:
        d = (dev[k] - ref[k]).ravel()
        r = ref[k].ravel()
        out[k] = {"max_rel": float(np.max(np.abs(d)) / max(np.max(np.abs(r)), 1e-300)),
                  "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(r), 1e-300))}
    return out


CASES = [
    # name, builder, kwargs, D, steps, bound on range-normalised rho and p error
    ("dambreak_2d_full_2000", M.configs.dambreak_2d, dict(dx=0.01), 2, 2000, 1e-3),
    ("dambreak_3d_dx0.03_1000", M.configs.dambreak_3d, dict(dx=0.03), 3, 1000, 1e-3),
]


@pytest.mark.parametrize("name,builder,kw,D,nsteps,bound", CASES, ids=[c[0] for c in CASES])
def test_fp32_trajectory_accuracy(name, builder, kw, D, nsteps, bound):
    dev = run(builder, kw, D, "f32", nsteps, "dev")
    ref = run(builder, kw, D, "f64", nsteps, "orc")
    err = errors(dev, ref)
    print(name, json.dumps(err))
    out = ROOT / "gpurun_out"
    if out.is_dir():
        p = out / "accuracy.json"
        allr = json.loads(p.read_text()) if p.exists() else {}
        allr[name] = err
        p.write_text(json.dumps(allr, indent=1))
    for k in ("rho", "p"):
        assert err[k]["max_rel"] < bound, (k, err[k])
    assert err["x"]["max_rel"] < bound
