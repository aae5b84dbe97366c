"""Autotuner of the launch shapes of the neighbour search and functor gather.

The reference has no counterpart: its hot loop (src/Interaction.jl:477-541)
runs on Polyester threads with no launch shape.  On the device the knobs
that matter per scene are (include/grasph_hip.h, GR_TUNE_*):

  * GR_TUNE_SWEEP_BLOCK — target particles per workgroup of the functor
    gather (64, 128, 256);
  * GR_TUNE_NL_WIN — candidates staged in LDS per window of the neighbour
    search (the staging tile: 64, 128, 256, 512);
  * GR_TUNE_NL_SPAN — how far apart the lanes' candidate ranges may start
    and still share one staging pass (a cell-column grouping);
  * GR_TUNE_MAXNB — list entries kept per target before it walks its cells
    directly (memory against overflow walks).

None of them changes a result (bit-identical lists, sums and states), so
the tuner only times.  It runs a coordinate descent: for each knob in turn,
every candidate value is timed over a few steps of the caller's own
integrators (the wall time of the steps as they will run), the best
value is kept, and the next knob is tried with it.  `passes` > 1 runs the
descent again from the values found, since the knobs interact (the best
workgroup size of the gather depends on the lists the search writes).  The
simulations' state
is saved to the host before the trials, restored before each trial (so
every shape is timed on the same state) and afterwards, so tuning does not
advance them.
"""
from __future__ import annotations

import time

from . import _abi as A

KNOB_NAMES = {A.GR_TUNE_SWEEP_BLOCK: "sweep_block", A.GR_TUNE_NL_WIN: "nl_win",
              A.GR_TUNE_NL_SPAN: "nl_span", A.GR_TUNE_MAXNB: "maxnb"}


def default_space(ndims):
    """Candidate values per knob, in the order the descent visits them."""
    mx = (32, 48, 64) if ndims == 3 else (16, 24, 32)
    return [(A.GR_TUNE_SWEEP_BLOCK, (256, 128, 64)),
            (A.GR_TUNE_NL_WIN, (128, 256, 512)),
            (A.GR_TUNE_NL_SPAN, (2, 4, 8)),
            (A.GR_TUNE_MAXNB, mx)]


def _time_steps(be, integs, steps, reps):
    """Mean ms per step, best of `reps`: the wall time of `steps` steps in
    the context's own mode (graph replay when it is on; the call returns
    with the stream synchronised), as the steps will run.  Timing each phase
    with events instead launches eagerly and ranked the shapes differently."""
    best = float("inf")
    be.step_many(integs, 1)  # plan, lists and graph for the current shape
    for _ in range(reps):
        t0 = time.perf_counter()
        be.step_many(integs, steps)
        best = min(best, (time.perf_counter() - t0) / steps * 1e3)
    return best


def current(be):
    return {KNOB_NAMES[k]: be.tune_get(k) for k in KNOB_NAMES}


def autotune(session, integrators, steps=3, reps=2, space=None, restore=True, passes=1):
    """Pick the launch shapes for `integrators` (host LeapFrogTimeIntegrator
    objects registered on `session`'s scene) and leave them set on the
    context.  Returns a record: the chosen values, the default's and the
    tuned time per step, and every trial."""
    be, sc = session.backend, session.scene
    hs = [sc.integrator(lf) for lf in integrators]
    t0 = time.perf_counter()
    if restore:
        for lf in integrators:
            sc.download_all(lf)
    space = space or default_space(sc.ndims)
    start = current(be)

    def timed():
        # every trial starts from the state the tuner found (the flow would
        # otherwise move on between trials and change what is timed)
        if restore:
            for lf in integrators:
                sc.upload_all(lf)
        return _time_steps(be, hs, steps, reps)

    ms_start = timed()
    trials = []
    for _ in range(max(1, int(passes))):
        for knob, values in space:
            best_v, best_ms = be.tune_get(knob), None
            for v in values:
                be.tune_set(knob, v)
                ms = timed()
                trials.append([KNOB_NAMES[knob], int(v), ms])
                if best_ms is None or ms < best_ms:
                    best_v, best_ms = v, ms
            be.tune_set(knob, best_v)
    ms_tuned = timed()
    if restore:
        for lf in integrators:
            sc.upload_all(lf)
    return {"chosen": current(be), "start": start, "ms_per_step_start": ms_start,
            "ms_per_step_tuned": ms_tuned, "trials": trials, "steps_per_trial": steps,
            "reps": reps, "passes": max(1, int(passes)), "tuning_s": time.perf_counter() - t0}
