"""FP32 trajectory accuracy: the device in f32 mode (FP32 pair math with
FP64 sums, FP64 positions, velocity, density and return map) against the
FP64 oracle on the same seeded scene after many steps.  Per field: max |Δ|
over max |ref| (range-normalised) and the L2 relative error, particles
matched by id, at several step counts.

A dam break is chaotic: a difference in the last bits of the state grows by
orders of magnitude over a few thousand steps.  Each case therefore also
runs the FP64 oracle from the same initial state with every density rounded
once to FP32 (`floor`): the distance that one FP32 rounding of the initial
state, and nothing else, puts between two FP64 runs.  The FP32 device must
stay within twice that distance at the last step count, and the density
within the 1e-5 target after 1000 steps in 3D.

Writes the table to gpurun_out/accuracy.json when that folder exists."""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import grasph as M, oracle_scene

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
FIELDS = ("x", "v", "rho", "p")


def run(builder, kw, D, prec, checkpoints, make, round_rho=False):
    """States at each step count of `checkpoints`, ordered by id."""
    cfg = builder(**kw)
    lf = cfg["integrator"]
    if round_rho:
        f = cfg["fluid"]
        f.rho[:] = f.rho.astype(np.float32).astype(np.float64)
    sc = M.Session(D, prec).scene if make == "dev" else oracle_scene(D)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfg["cfl"])
    out, done = {}, 0
    for n in checkpoints:
        sc.be.step(h, n - done, dt)
        done = n
        sc.download_all(lf)
        f = cfg["fluid"]
        o = np.argsort(f.id)
        out[n] = {k: getattr(f, k)[o].copy() for k in FIELDS}
    return out


def errors(dev, ref):
    out = {}
    for k in ref:
        d = (dev[k] - ref[k]).ravel()
        r = ref[k].ravel()
        out[k] = {"max_rel": float(np.max(np.abs(d)) / max(np.max(np.abs(r)), 1e-300)),
                  "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(r), 1e-300))}
    return out


def record(name, table):
    out = ROOT / "gpurun_out"
    if out.is_dir():
        p = out / "accuracy.json"
        allr = json.loads(p.read_text()) if p.exists() else {}
        allr[name] = table
        p.write_text(json.dumps(allr, indent=1))


CASES = [
    # name, builder, kwargs, D, checkpoints, {step: {field: absolute bound}}
    ("dambreak_2d_full_2000", M.configs.dambreak_2d, dict(dx=0.01), 2, (100, 500, 1000, 2000),
     {}),
    ("dambreak_3d_dx0.03_1000", M.configs.dambreak_3d, dict(dx=0.03), 3, (100, 500, 1000),
     {1000: {"rho": 1e-5}}),
]


@pytest.mark.parametrize("name,builder,kw,D,checkpoints,target", CASES,
                         ids=[c[0] for c in CASES])
def test_fp32_trajectory_accuracy(name, builder, kw, D, checkpoints, target):
    dev = run(builder, kw, D, "f32", checkpoints, "dev")
    ref = run(builder, kw, D, "f64", checkpoints, "orc")
    flo = run(builder, kw, D, "f64", checkpoints, "orc", round_rho=True)
    table = {str(n): {"device": errors(dev[n], ref[n]), "floor": errors(flo[n], ref[n])}
             for n in checkpoints}
    print(name, json.dumps(table))
    record(name, table)
    last = str(checkpoints[-1])
    for k in FIELDS:
        e, fl = table[last]["device"][k]["max_rel"], table[last]["floor"][k]["max_rel"]
        assert e <= 2.0 * fl, (k, e, fl)
    for n, bounds in target.items():
        for k, b in bounds.items():
            assert table[str(n)]["device"][k]["max_rel"] <= b, (n, k, table[str(n)]["device"][k])
