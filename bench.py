"""Benchmark of the particle step on one MI355X.

Workload: the 3D dam break — 1.0M fluid particles (0.6 m cube, dx = 0.006 m)
in a 1.6 × 0.6 × 0.8 m tank with 3-layer dynamic floor and walls (~0.38M
boundary particles), Wendland C2 h = 1.3dx, Tait EOS, FluidPfn α = 0.1,
fluid order shuffled by seed.  Synthetic, seeded input built in-process.

A "step" is one leapfrog step of the whole scene (sort, grids, half step,
fused sweep, full step).  The timed region holds inputs already resident in
HBM; the library synchronises its stream on return of every gr_step call.

Prints one JSON line (see the fields below).  --gpus N > 1 is not supported:
this build runs on one GPU.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
import __graft_entry__  # noqa: E402

METRIC = "particle-pair updates/s & % HBM roofline; served sims/h + p50/p99 vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def progress(msg):
    """One progress line on stderr (the JSON result stays alone on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dx", type=float, default=0.006)
    ap.add_argument("--start-step", type=int, default=300,
                    help="untimed steps of the dam break before tuning, warm-up and timing: "
                         "from rest no particle crosses a cell face and the sort skips its "
                         "radix passes, which a flowing dam break does not")
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--cpu-steps", type=int, default=60,
                    help="timed oracle steps on the host (0 disables the CPU baseline)")
    ap.add_argument("--graph", type=int, default=1, help="replay the step as a hipGraph")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run the plain copy kernel on this many MiB (counter calibration)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick the launch shapes with the autotuner before timing (0: defaults)")
    ap.add_argument("--shapes", default="",
                    help="fixed launch shapes instead of the autotuner's, e.g. "
                         "'nl_win=256,nl_span=4,sweep_block=128,maxnb=32'")
    ap.add_argument("--ensemble", type=int, default=256,
                    help="also step an ensemble of this many 2D columns in shared launches "
                         "and the same columns one by one (0 disables)")
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="timed steps of the batched ensemble")
    ap.add_argument("--ensemble-serial-steps", type=int, default=20,
                    help="timed steps of each column run alone (0: the batched run only)")
    ap.add_argument("--serve", type=int, default=100,
                    help="also serve this many column-collapse jobs (continuous batching) "
                         "and the same jobs serially (0 disables)")
    ap.add_argument("--serve-step-scale", type=float, default=0.25,
                    help="step counts of the served sample are scaled by this factor")
    ap.add_argument("--serve-size-scale", type=float, default=1.0,
                    help="particle counts of the served sample are scaled by this factor")
    return ap.parse_args()


def ensemble_leg(n_sims, nsteps, serial_steps, precision, tune, graph=True):
    """`n_sims` 2D granular columns (the ensemble workload: aspect, width, φ
    and ρ drawn per simulation, dx = 2.5 mm) advanced together (gr_step_many:
    one graph replay per step, one launch per phase over the packed pool)
    vs each column alone, one after another (graph replay too).  The
    batched run times `nsteps` steps, each column alone `serial_steps` (0:
    the batched run only).  graph=False launches the phases one by one (for
    a kernel trace)."""
    g = __graft_entry__.package()
    t0 = time.perf_counter()
    sims = g.configs.ensemble_columns(n_sims=n_sims)
    npart = sum(c["n_particles"] for c in sims)
    s = g.Session(2, precision)
    s.backend.set_graph_mode(graph)
    hs = []
    for cfg in sims:
        h = s.scene.integrator(cfg["integrator"])
        s.scene.upload_all(cfg["integrator"])
        s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
        hs.append(h)
    build_s = time.perf_counter() - t0
    progress(f"ensemble: {n_sims} columns, {npart} particles built in {build_s:.1f} s")
    rec = (g.autotune.autotune(s, [c["integrator"] for c in sims], steps=2, reps=1)
           if tune else None)
    progress("ensemble: tuned; batched run")
    s.backend.step_many(hs, 2)  # plan, capture
    t0 = time.perf_counter()
    s.backend.step_many(hs, nsteps)
    batched = time.perf_counter() - t0
    progress(f"ensemble: batched {nsteps} steps in {batched:.2f} s")
    serial = None
    if serial_steps > 0:
        progress("ensemble: columns alone")
        for h in hs:  # plan and capture of each column alone, untimed
            s.backend.step_many([h], 1)
        t0 = time.perf_counter()
        for h in hs:
            s.backend.step_many([h], serial_steps)
        serial = time.perf_counter() - t0
    s.close()
    b_rate = n_sims * nsteps / batched
    s_rate = n_sims * serial_steps / serial if serial_steps > 0 else None
    out = {"workload": f"{n_sims} 2D granular columns (EP, Drucker-Prager, ghost walls), "
                       f"aspect U[0.5,4], width U[0.1,0.4] m, phi U[20,40] deg, "
                       f"rho U[1500,2000], dx 2.5 mm",
           "sims": n_sims, "particles": npart, "steps_batched": nsteps,
           "steps_serial_each": serial_steps, "batched_s": batched, "serial_s": serial,
           "sim_steps_per_s_batched": b_rate, "sim_steps_per_s_serial": s_rate,
           "particle_steps_per_s_batched": npart * nsteps / batched,
           "ms_per_batched_step": batched / nsteps * 1e3,
           "speedup": b_rate / s_rate if serial_steps > 0 else None, "build_s": build_s,
           "projected_5000_steps_s": 5000 * batched / nsteps}
    if rec:
        out["autotune"] = rec
    return out


def serve_leg(n_jobs, step_scale, size_scale, precision):
    """The served stream: Poisson arrivals at 20 jobs/s, 80 % 2D elasto-plastic
    columns (the ensemble's parameter ranges) and 20 % 3D visco-plastic
    columns (50k-500k particles, dx 5 mm, phi U[25, 40] deg), step counts
    log-uniform in [500, 20000] times `step_scale`, particle counts times
    `size_scale`, an HBM budget of 200 GB.  Continuous batching (one 2D and
    three 3D contexts, shared launches, admission by footprint, retirement)
    against the same jobs run one at a time, each alone on the GPU, in
    arrival order (graph replay and scene building overlapped in both)."""
    g = __graft_entry__.package()
    S = g.serving
    mk = lambda: S.column_jobs(n_jobs, step_scale=step_scale, size_scale=size_scale)  # noqa: E731
    progress(f"serving: {n_jobs} jobs, continuous batching")
    # three 3D lanes: the 3D jobs are the long ones, and a late long job then
    # shares its steps with a third of them (a shorter makespan and lower
    # latencies than one or two 3D lanes in this stream; four contexts use
    # the four hardware queues)
    served = S.Server(precision=precision, quantum=50, graph=True, hbm_budget=200e9,
                      lanes={2: 1, 3: 3}).run(mk())
    progress(f"serving: served in {served['makespan_s']:.1f} s; the same jobs one at a time")
    plain = S.run_plain(mk(), precision=precision, graph=True,
                        progress=lambda m: progress("serving (plain): " + m))
    progress(f"serving: plain in {plain['makespan_s']:.1f} s")
    return {"lanes": {"2d": 1, "3d": 3},
            "workload": "served stream of 2D/3D column-collapse jobs (Poisson 20/s, 80% 2D / "
                        "20% 3D, steps log-U[500, 20000], HBM budget 200 GB)",
            "jobs": n_jobs, "step_scale": step_scale, "size_scale": size_scale,
            "served": served, "plain": plain,
            "jobs_per_h_served": served["jobs_per_h"], "jobs_per_h_plain": plain["jobs_per_h"],
            "jobs_per_h_ratio": served["jobs_per_h"] / plain["jobs_per_h"],
            "latency_p50_s": {"served": served["latency_p50_s"], "plain": plain["latency_p50_s"]},
            "latency_p99_s": {"served": served["latency_p99_s"], "plain": plain["latency_p99_s"]}}


def sweep_bytes(cfg, grids):
    """Algorithmic bytes of the pair sweep of one step (FP32 state, FP64
    positions): fluid reads x (8D), v (4D), ρ (4), p (4) and writes dvdt
    (4D), drhodt (4); each dynamic-boundary particle is read for x (8D);
    every cell array is read once (4 B per entry).  The neighbour search
    (k_nlist) and the functor pass (k_sweep_list) together do this work."""
    D = 3
    nf = cfg["fluid_n"]
    nb = cfg["boundary_n"]
    cells = sum(nc + 1 for nc in grids)
    return nf * (8 * D + 4 * D + 4 + 4) + nf * (4 * D + 4) + nb * 8 * D + 4 * cells


def sort_bytes(cfg, key_cells, precision):
    """Algorithmic bytes of the sort phase of one step: the key pass reads x
    (8D) and writes a key (4), the ranking reads and writes a pool index (8),
    the histogram and its scan touch 8 B per key cell, and the permutation
    reads and writes every per-particle array once (2·S_state: x 8D, v 8D,
    ρ 8, id 4, dvdt and drhodt in the state precision, p for fluids).  The
    leapfrog half step and stage 0's partner records are written by the same
    permutation: the half-step shadows (v_s in the state precision, ρ_h 8,
    ρ_s) and, in FP32, the 32-byte compact record of every particle."""
    D = 3
    r = 4 if precision == "f32" else 8
    nf, nb = cfg["fluid_n"], cfg["boundary_n"]
    n = nf + nb
    s_fluid = 8 * D + 8 * D + 8 + 4 + r * D + r + r
    s_bnd = 8 * D + 8 * D + 8 + 4 + r * D + r
    half = n * (r * D + 8 + r) + (n * 32 if precision == "f32" else 0)
    return (n * (8 * D + 4) + 8 * n + 8 * key_cells + 2 * (nf * s_fluid + nb * s_bnd) + half)


def load_traffic():
    """HBM bytes of the sweep per step (the k_nlist and k_sweep_list
    dispatches of one step) from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, calibrated on a plain copy
    kernel of known size), or None if no summary is committed."""
    p = ROOT / "profiles" / "sweep_dambreak3d_pmc.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return float(d["traffic_bytes_per_step"])
    except Exception:
        return None


# FP32 operations of one FluidPfn pair (S = 2, Wendland C2, artificial
# viscosity, the records' 1/ρ and p/ρ²), counted from the functor: x_i − x_j
# from the fixed-point records (3), r² (5), 1/r and q (3), dW/dq and the
# gradient factor (7), the gradient (3), v_ij and v_ij·x_ij (8), μ (4), Π (7),
# the pressure term (3), the momentum factor (2), its FP64 sums (6), v_ij·∇W
# and the density sum (8).
FLOP_PER_PAIR = 59
VALU_CLOCK_GHZ = 2.4  # peak engine clock: a wave64 VALU instruction issues in 4 cycles
SIMDS = 1024          # 256 CUs x 4 SIMDs


def load_sq():
    """SQ counters of the pair kernels from the committed rocprofv3 --pmc
    summary (per dispatch of the timed steps), or None."""
    p = ROOT / "profiles" / "sweep_dambreak3d_pmc.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())["sq"]["kernels"]
    except Exception:
        return None


def valu_block(sq, t_nlist_ms, t_sweep_ms, ordered_pairs):
    """VALU view of the two pair kernels: instructions per wavefront and per
    ordered pair (counters), their issue rate against one wave64 instruction
    per 4 cycles per SIMD at the peak clock (this run's kernel times), and the
    pair functor's FLOP rate against the FP32 vector peak."""
    if not sq:
        return None
    out = {"source": "SQ_INSTS_VALU / SQ_WAVES of the committed --pmc pass "
                     "(profiles/sweep_dambreak3d_pmc.json); kernel times of this run",
           "clock_ghz": VALU_CLOCK_GHZ, "simds": SIMDS}
    for key, t in (("k_nlist", t_nlist_ms), ("k_sweep_list", t_sweep_ms)):
        k = next((v for n, v in sq.items() if key + "<" in n), None)
        if not k or not t:
            continue
        insts = k["SQ_INSTS_VALU"]
        out[key] = {"valu_insts_per_wave": k["valu_insts_per_wave"],
                    "lane_insts_per_ordered_pair": insts * 64 / ordered_pairs,
                    "issue_frac": insts * 4 / (SIMDS * t * 1e-3 * VALU_CLOCK_GHZ * 1e9),
                    "wait_frac": k["wait_fraction_of_wave_cycles"]}
    if t_sweep_ms:
        tf = FLOP_PER_PAIR * ordered_pairs / (t_sweep_ms * 1e-3) / 1e12
        out["pair_flop"] = {"flop_per_ordered_pair": FLOP_PER_PAIR, "achieved_tflops": tf,
                            "peak_tflops": FP32_PEAK_TFLOPS, "frac": tf / FP32_PEAK_TFLOPS,
                            "kernel": "k_sweep_list"}
    return out


def cpu_baseline(cfg_builder, kw, nsteps, pairs_per_step):
    """The FP64 oracle (reference-algorithm restatement, OpenMP over cell
    colours) on the same workload, timed on this host's cores."""
    if nsteps <= 0:
        return None
    g = __graft_entry__.package()
    A = g._abi
    lib_path = ROOT / "oracle" / "libgrasph_oracle.so"
    lib = C.CDLL(str(lib_path))
    lib.orc_scene_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    lib.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
    ctx = C.c_void_p()
    assert lib.orc_scene_create(C.byref(ctx), 3) == 0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    lib.orc_set_threads(ctx, threads)
    be = A.CBackend(lib, "orc_", ctx, "orc_last_error", "orc_scene_destroy")
    cfg = cfg_builder(**kw)
    lf = cfg["integrator"]
    sc = g.DeviceScene(be, 3)
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = be.integrator_dt(h, cfg["cfl"])
    t0 = time.perf_counter()
    be.step(h, 1, dt)  # first step: sorts the shuffled order
    t1 = time.perf_counter()
    be.step(h, nsteps, dt)
    t2 = time.perf_counter()
    el = t2 - t1
    be.close()
    return {"value": pairs_per_step * nsteps / el, "unit": "pair-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle (FP64 C++ restatement of the reference algorithm) on the same "
                      f"3D dam break, {nsteps} steps after 1 warm-up step "
                      f"({t1 - t0:.2f} s incl. the first sort); {el / nsteps * 1e3:.1f} ms/step"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(1)
    g = __graft_entry__.package()
    # The served stream runs first.  In a process where another context had
    # run and been closed before, the same stream took 9.1 s instead of 7.2 s
    # (two of the 3D lanes stepped about 30 % slower); reusing that context's
    # stream for a lane did not change it, and eight hardware queues instead of
    # four made it slower still.
    srv = (serve_leg(args.serve, args.serve_step_scale, args.serve_size_scale, args.precision)
           if args.serve else None)
    kw = dict(dx=args.dx)
    progress("3D dam break: building the scene")
    cfg = g.configs.dambreak_3d(**kw)
    lf = cfg["integrator"]
    s = g.Session(3, args.precision)
    be, sc = s.backend, s.scene
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = be.integrator_dt(h, cfg["cfl"])
    be.integrator_set_dt(h, dt)
    be.set_graph_mode(bool(args.graph))
    info = be.ctx_info()
    # first sort of the shuffled order and the start of the flow, then the
    # launch shapes (the tuner restores the state it found)
    be.step_many([h], max(1, args.start_step))
    if args.shapes:
        knob = {v: k for k, v in g.autotune.KNOB_NAMES.items()}
        for kv in args.shapes.split(","):
            k, v = kv.split("=")
            be.tune_set(knob[k.strip()], int(v))
    tune = (g.autotune.autotune(s, [lf], steps=5, reps=3, passes=2)
            if args.autotune and not args.shapes else None)
    progress("3D dam break: tuned; timed steps")
    # warm-up (graph capture)
    if args.warmup:
        be.step_many([h], args.warmup)
    # pair count of the current configuration (diagnostic counting kernels)
    be.set_timing(2)
    be.step_many([h], 1)
    pairs = be.get_pair_count(h)
    self_ordered, coupled = be.get_pair_stats()
    # timed region: HIP events at every phase, no extra synchronisation
    be.set_timing(1)
    t0 = time.perf_counter()
    be.step_many([h], args.steps)
    t1 = time.perf_counter()
    phases = be.get_timing()
    be.set_timing(0)
    # graph replay without events (the production path), same step count
    t2 = time.perf_counter()
    be.step_many([h], args.steps)
    t3 = time.perf_counter()
    el_events, el_graph = t1 - t0, t3 - t2
    if args.calibrate_copy:
        be.bench_copy(args.calibrate_copy << 20, 3)
    el = min(el_events, el_graph) if args.graph else el_events
    ms_step = el / args.steps * 1e3
    grids = []
    for si in lf.interactions:
        _, _, nc = be.get_grid(sc.interaction(si), 3)
        grids.append(nc)
    bytes_sweep = sweep_bytes(cfg, grids)
    sweep_ms = phases[5] + phases[8]
    achieved = bytes_sweep / (sweep_ms * 1e-3) / 1e9
    # the sort's keys span the fluid's cells of side 2h (its self grid)
    bytes_sort = sort_bytes(cfg, grids[0] if grids else 0, args.precision)
    sort_gbs = bytes_sort / (phases[0] * 1e-3) / 1e9 if phases[0] > 0 else None
    # the committed counter passes are of the FP32 kernels
    traffic = load_traffic() if args.precision == "f32" else None
    # the dam break's context (its stream, pool and graphs) is not needed by
    # the legs below
    s.close()
    progress(f"3D dam break: {el / args.steps * 1e3:.3f} ms/step; CPU baseline")
    cpu = cpu_baseline(g.configs.dambreak_3d, kw, args.cpu_steps, pairs)
    ens = (ensemble_leg(args.ensemble, args.ensemble_steps, args.ensemble_serial_steps,
                        args.precision, bool(args.autotune), graph=bool(args.graph))
           if args.ensemble else None)
    total_pairs = pairs * args.steps
    value = total_pairs / el
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair math + fp64 positions/velocity/density" if args.precision == "f32"
        else "fp64",
        "data": "synthetic (seeded lattice, shuffled order)",
        "config": {"workload": "dambreak_3d", "timed_from_step": max(1, args.start_step),
                   "n_fluid": cfg["fluid_n"],
                   "n_boundary": cfg["boundary_n"], "dx": args.dx, "kernel": "WendlandC2",
                   "h_over_dx": 1.3, "boundary": "3-layer DynamicBoundarySystem x5",
                   "pairs_per_step": pairs, "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": "pair sweep = k_nlist<3,float> + k_sweep_list<3,float,FORCE "
                               "fluid-only> (fused gather: self + 5 dynamic walls), one launch "
                               "each per step",
                     "algorithmic_bytes_per_step": bytes_sweep,
                     "launch_ms": sweep_ms,
                     "k_nlist_ms": phases[8], "k_sweep_list_ms": phases[5],
                     "ordered_pairs": self_ordered + coupled,
                     "valu": valu_block(load_sq() if args.precision == "f32" else None,
                                        phases[8], phases[5], self_ordered + coupled)},
        "sort_roofline": {"bound": "hbm", "achieved": sort_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s",
                          "frac": sort_gbs / HBM_PEAK_GBS if sort_gbs else None,
                          "algorithmic_bytes_per_step": bytes_sort, "phase_ms": phases[0],
                          "kernels": "key min/max, key write (with the first pass's tile "
                                     "counts), LSD radix (a scan and a scatter kernel per "
                                     "8-bit digit), permutation of every per-particle array "
                                     "with the leapfrog half step and stage 0's partner "
                                     "records"},
        "cpu_baseline": cpu,
        "steps_per_s": 1e3 / ms_step,
        "phase_ms": {k: round(v, 4) for k, v in zip(
            ("sort", "ghosts", "grid", "halfstep", "updaters", "sweep", "fullstep", "step",
             "nlist"),
            phases)},
        "ms_per_step_events": el_events / args.steps * 1e3,
        "ms_per_step_graph": el_graph / args.steps * 1e3,
        "device": info,
    }
    if tune:
        res["autotune"] = tune
    if ens:
        res["ensemble"] = ens
    if srv:
        res["serving"] = srv
    print(json.dumps(res))


if __name__ == "__main__":
    main()
