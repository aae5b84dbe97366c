"""The 3D dam break at its full size (1.0M fluid particles, dx = 6 mm, plus
0.384M particles in five 3-layer dynamic walls) on the device against the
FP64 oracle.

* First step, integer data, bit for bit: the sort of the seeded shuffled
  order (src/Sorting.jl:157-195), the CSR cell starts of all six
  interactions (src/Interaction.jl:123-236) and the per-particle neighbour
  counts of the fluid and of every wall (the counting functor over the same
  pairs the physics sweeps visit).
* 100 leapfrog steps in f32 mode against the FP64 oracle: per field, max
  |Δ|/max|ref| and the L2 relative error, particles matched by id; the
  density within the 1e-5 target, and every field within twice the distance
  by which one FP32 rounding of the initial densities moves the FP64 run.

Writes its figures to gpurun_out/fullsize.json when that folder exists."""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import grasph as M, oracle_scene

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def record(name, table):
    out = ROOT / "gpurun_out"
    if out.is_dir():
        p = out / "fullsize.json"
        allr = json.loads(p.read_text()) if p.exists() else {}
        allr[name] = table
        p.write_text(json.dumps(allr, indent=1))


def scene(make, prec="f64"):
    return M.Session(3, prec).scene if make == "dev" else oracle_scene(3)


def test_dambreak3d_full_first_step_integer_data():
    out = {}
    for make in ("dev", "orc"):
        cfg = M.configs.dambreak_3d()
        lf = cfg["integrator"]
        fluid = cfg["fluid"]
        walls = [ps for ps in lf.systems if ps is not fluid]
        h = lf.interactions[0].kernel.h
        sc = scene(make)
        for ps in lf.systems:
            sc.upload(ps)
            sc.be.sort(sc.system(ps), 2.0 * h)
        res = {"id": [], "grid": [], "counts": []}
        for ps in lf.systems:
            sc.download(ps, ["x"])
            res["id"].append(sc.be.system_download_index(sc.system(ps), M._abi.GR_I_ID, ps.n))
            res["x"] = res.get("x", []) + [ps.x.copy()]
        for si in lf.interactions:
            sc.be.create_grid(sc.interaction(si))
            res["grid"].append(sc.grid(si))
        # neighbour counts over the same pairs: fluid self and fluid-wall
        kern = lf.interactions[0].kernel
        for ps in lf.systems:
            ps.dvdt[:] = 0.0
            ps.drhodt[:] = 0.0
            sc.upload(ps, ["dvdt", "drhodt"])
        for w in [None] + walls:
            si = M.SystemInteraction(kern, M.NeighborCounterPfn(), fluid, w)
            hi = sc.interaction(si)
            sc.be.create_grid(hi)
            sc.be.sweep(hi, 1)
        for ps in lf.systems:
            sc.download(ps, ["drhodt"])
            res["counts"].append(ps.drhodt.copy())
        out[make] = res
    dev, orc = out["dev"], out["orc"]
    for a, b in zip(dev["id"], orc["id"]):
        assert np.array_equal(a, b)
    for a, b in zip(dev["x"], orc["x"]):
        assert np.array_equal(a, b)
    # the shuffled fluid was actually reordered
    assert not np.array_equal(orc["id"][0], np.arange(1, len(orc["id"][0]) + 1))
    ncells = []
    for gd, go in zip(dev["grid"], orc["grid"]):
        for k in range(3):
            assert np.array_equal(gd[k], go[k])
        if go[3] is not None:
            assert np.array_equal(gd[3], go[3])
        ncells.append(int(len(go[2]) - 1))
    for a, b in zip(dev["counts"], orc["counts"]):
        assert np.array_equal(a, b)
    fl = orc["counts"][0]
    record("first_step_integer", {"particles": int(sum(len(i) for i in orc["id"])),
                                  "cells": ncells, "fluid_counts_mean": float(fl.mean()),
                                  "fluid_counts_max": int(fl.max()),
                                  "pairs_fluid_self_plus_walls": int(fl.sum())})


def _run(prec, make, nsteps, round_rho=False):
    cfg = M.configs.dambreak_3d()
    lf = cfg["integrator"]
    if round_rho:
        f = cfg["fluid"]
        f.rho[:] = f.rho.astype(np.float32).astype(np.float64)
    sc = scene(make, prec)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = sc.be.integrator_dt(h, cfg["cfl"])
    sc.be.step(h, nsteps, dt)
    sc.download_all(lf)
    f = cfg["fluid"]
    o = np.argsort(f.id)
    return {k: getattr(f, k)[o].copy() for k in ("x", "v", "rho", "p")}


def _errors(dev, ref):
    out = {}
    for k in ref:
        d = (dev[k] - ref[k]).ravel()
        r = ref[k].ravel()
        out[k] = {"max_rel": float(np.max(np.abs(d)) / max(np.max(np.abs(r)), 1e-300)),
                  "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(r), 1e-300))}
    return out


def test_dambreak3d_full_100_steps_fp32_vs_oracle():
    nsteps = 100
    dev = _run("f32", "dev", nsteps)
    ref = _run("f64", "orc", nsteps)
    flo = _run("f64", "orc", nsteps, round_rho=True)
    table = {"steps": nsteps, "device": _errors(dev, ref), "floor": _errors(flo, ref)}
    print(json.dumps(table))
    record("fp32_100_steps", table)
    assert table["device"]["rho"]["max_rel"] <= 1e-5, table["device"]["rho"]
    for k in ("x", "v", "rho", "p"):
        e, f = table["device"][k]["max_rel"], table["floor"][k]["max_rel"]
        assert e <= 2.0 * f, (k, e, f)
