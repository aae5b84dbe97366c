"""Continuous-batching simulation server and the plain serial baseline.

Jobs (a scene builder, a step count, an arrival time) are served by device
contexts ("lanes"), one or more per spatial dimension, each driven by its own
host thread so their work overlaps on the GPU (ctypes releases the GIL inside
the library calls).  A job goes to the lane of its dimension with the fewest
steps assigned so far: a batch advances at the pace of all its particles, so
spreading long jobs over lanes keeps a late long job from sharing every step
with all the others.  A context loop:

  1. admits every queued job that has arrived and whose estimated HBM
     footprint fits the shared budget: builds the scene, registers it on the
     context, uploads it;
  2. advances all running jobs of the context together, `quantum` steps at a
     time (gr_step_many: one launch per phase over the packed pool, each job
     with its own dt), a quantum ending early at the next job completion;
  3. retires finished jobs: downloads their state if asked, releases them
     (gr_integrator_release, so later jobs reuse their pool segments) and
     returns their footprint to the budget.

A failed job does not disturb the others:
  * a job whose footprint can never fit the budget is rejected at arrival;
  * a job whose scene cannot be built, registered or uploaded fails alone;
  * a job whose simulation faults on the device (positions turning NaN or
    infinite, or particles leaving the cell range) is isolated by the
    library inside the shared launches (GR_ERR_FAULT, gr_integrator_status):
    the other jobs of the batch complete the quantum exactly as they would
    alone, and the faulted job is retired as failed;
  * only an error of the context itself (a HIP failure) ends that context's
    loop, failing the jobs it holds; the other context goes on.
The reference's only recovery is a restart from a snapshot
(src/Driver.jl:79-96).

`run_plain` serves the same jobs one at a time, each alone on the GPU, in
arrival order: the "plain single-sim kernel run serially" baseline.

A job's result does not depend on what it was batched with: every particle's
sums use only its own simulation's particles, in the same order.  The
device tests check this bit for bit.
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _abi as A
from .integration import Session


@dataclass
class Job:
    jid: int
    ndims: int
    build: Callable[[], dict]
    steps: int
    arrival: float = 0.0
    n_particles: int = 0
    bytes: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    done_steps: int = 0
    result: Optional[dict] = None
    meta: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def footprint_bytes(cfg, precision="f32", maxnb=None):
    """Upper estimate of the device bytes a scene holds while resident: two
    pool families of every per-particle array, the shared shadows and
    partner records, the neighbour lists and the ghost capacity."""
    lf = cfg["integrator"]
    D = lf.systems[0].ndims
    r = 4 if precision == "f32" else 8
    maxnb = maxnb or (144 if D == 3 else 72)  # list words: 48 / 24 entries of 12 B
    per = 0
    n_real = 0
    for ps in lf.systems:
        ns = getattr(ps, "ns", 0) or 0
        fam = 8 * D * 2 + 8 + r * D + r + r + 4 + ns * (8 + r + r + 16) + 3 * r
        per += ps.n * (2 * fam + r * D + r + 8 + 12 + 8 * D + 48)
        n_real += ps.n
    lists = n_real * (maxnb * 4 + 8 * 4)
    ghosts = 0
    for ge in lf.ghosts:
        src = ge.ghost.source
        ghosts += len(ge.normals) * src.n * (2 * (8 * D * 2 + 8 + 8 * r) + 64)
    return int(per + lists + ghosts + (1 << 20))


class _Budget:
    def __init__(self, total):
        self.total = float(total)
        self.used = 0.0
        self.lock = threading.Lock()

    def try_take(self, b):
        with self.lock:
            if self.used + b <= self.total:
                self.used += b
                return True
            return False

    def give(self, b):
        with self.lock:
            self.used -= b


class Server:
    """Continuous batching over `lanes` contexts per spatial dimension (an
    int for every dimension, or a dict D -> count)."""

    def __init__(self, precision="f32", hbm_budget=200e9, quantum=50, graph=False,
                 keep_results=False, session_factory=None, builders=2, lanes=1):
        self.precision = precision
        self._make_session = session_factory or (lambda D: Session(D, precision))
        self.budget = _Budget(hbm_budget)
        self.quantum = int(quantum)
        self.graph = bool(graph)
        self.keep_results = keep_results
        self.sessions = {}
        self.errors = []
        # per context: seconds in admission (register, upload; the scene is
        # built on a builder thread as soon as the job arrives), in the shared
        # steps, in retirement, and the number of shared calls
        self.stats = {}
        self.builders = max(1, int(builders))
        self.lanes = lanes
        self._pool = None

    def _nlanes(self, D):
        n = self.lanes.get(D, 1) if isinstance(self.lanes, dict) else self.lanes
        return max(1, int(n))

    def _session(self, key):
        """The context of lane key = (D, lane)."""
        if key not in self.sessions:
            s = self._make_session(key[0])
            s.backend.set_graph_mode(self.graph)
            self.sessions[key] = s
        return self.sessions[key]

    def _fail(self, job, t0, why, taken=True):
        job.meta["failed"] = why
        job.meta.pop("cfg", None)
        job.finished = time.perf_counter() - t0
        if taken:
            self.budget.give(job.bytes)

    def _retire(self, sc, job, cfg, t0):
        lf = cfg["integrator"]
        if self.keep_results:
            sc.download_all(lf)
            job.result = {f: getattr(cfg["fluid"], f).copy() for f in cfg["fluid"].field_names()}
        sc.release(lf)
        job.finished = time.perf_counter() - t0
        self.budget.give(job.bytes)

    def _admit(self, sc, be, job, t0):
        """Build (if not yet), register and upload one job; None on failure
        (the job is marked failed and its budget returned)."""
        cfg = job.meta.pop("cfg", None)
        lf = None
        try:
            if cfg is None:
                cfg = job.build()
            lf = cfg["integrator"]
            h = sc.integrator(lf)
            sc.upload_all(lf)
            be.integrator_set_dt(h, be.integrator_dt(h, cfg["cfl"]))
        except (ValueError, AssertionError, KeyError, TypeError) as e:
            if lf is not None:
                try:
                    sc.release(lf)
                except Exception:
                    pass
            self._fail(job, t0, f"admission: {e}")
            return None
        job.admitted = time.perf_counter() - t0
        return cfg, h

    def _loop(self, key, jobs, t0):
        queue = sorted(jobs, key=lambda j: (j.arrival, j.jid))
        running = []  # (job, cfg, handle)
        st = self.stats.setdefault(key, {"admit_s": 0.0, "step_s": 0.0, "retire_s": 0.0,
                                         "calls": 0, "max_running": 0, "jobs": len(jobs)})
        try:
            s = self._session(key)
            be, sc = s.backend, s.scene
            while queue or running:
                now = time.perf_counter() - t0
                # the scene of every arrived job is built on a builder thread
                for job in queue:
                    if job.arrival > now:
                        break
                    if not job.bytes and "fut" not in job.meta:
                        job.meta["fut"] = self._pool.submit(job.build)
                # admit arrived jobs in order while they fit
                while queue and queue[0].arrival <= now:
                    job = queue[0]
                    if not job.bytes:
                        fut = job.meta["fut"]
                        if not fut.done() and running:
                            break  # step the running jobs meanwhile
                        try:
                            cfg = fut.result()
                        except (ValueError, AssertionError, KeyError, TypeError) as e:
                            queue.pop(0)
                            job.meta.pop("fut", None)
                            self._fail(job, t0, f"build: {e}", taken=False)
                            continue
                        job.meta.pop("fut", None)
                        job.meta["cfg"] = cfg
                        job.n_particles = cfg["n_particles"]
                        job.bytes = footprint_bytes(cfg, self.precision)
                    if job.bytes > self.budget.total:  # can never fit: reject
                        queue.pop(0)
                        job.meta.pop("cfg", None)
                        job.meta["rejected"] = True
                        continue
                    if not self.budget.try_take(job.bytes):
                        if not running:
                            time.sleep(0.001)  # the other context will free memory
                        break
                    queue.pop(0)
                    ta = time.perf_counter()
                    got = self._admit(sc, be, job, t0)
                    st["admit_s"] += time.perf_counter() - ta
                    if got is not None:
                        running.append((job,) + got)
                if not running:
                    if queue:
                        wait = queue[0].arrival - (time.perf_counter() - t0)
                        if wait > 0:
                            time.sleep(min(wait, 0.05))
                    continue
                q = min(self.quantum, min(j.steps - j.done_steps for j, _, _ in running))
                faulted = set()
                ts = time.perf_counter()
                st["max_running"] = max(st["max_running"], len(running))
                try:
                    be.step_many([h for _, _, h in running], q)
                except A.FaultError:
                    # the library isolated the faulted simulations: the
                    # others advanced all q steps
                    faulted = {h for _, _, h in running if be.integrator_status(h)}
                st["step_s"] += time.perf_counter() - ts
                st["calls"] += 1
                tr = time.perf_counter()
                keep = []
                for job, cfg, h in running:
                    if h in faulted:
                        bits = be.integrator_status(h)
                        why = ("non-finite positions" if bits & A.GR_FAULT_NONFINITE
                               else "particles beyond the cell range")
                        sc.release(cfg["integrator"])
                        self._fail(job, t0, f"device fault after step <= {job.done_steps + q}: "
                                            f"{why}")
                        continue
                    job.done_steps += q
                    if job.done_steps < job.steps:
                        keep.append((job, cfg, h))
                        continue
                    self._retire(sc, job, cfg, t0)
                running = keep
                st["retire_s"] += time.perf_counter() - tr
        except Exception as e:  # an error of the context itself
            self.errors.append(e)
            for job, _, _ in running:
                self._fail(job, t0, f"context error: {e}")
            for job in queue:
                self._fail(job, t0, f"context error: {e}", taken=False)

    def _assign(self, jobs):
        """Lane of every job, in arrival order, from the jobs before it only."""
        by_lane = {}
        assigned = {}
        for j in sorted(jobs, key=lambda j: (j.arrival, j.jid)):
            keys = [(j.ndims, k) for k in range(self._nlanes(j.ndims))]
            key = min(keys, key=lambda q: (assigned.get(q, 0), q[1]))
            assigned[key] = assigned.get(key, 0) + j.steps
            by_lane.setdefault(key, []).append(j)
        return by_lane

    def run(self, jobs):
        """Serve `jobs` (arrival times in seconds from now); returns a report."""
        t0 = time.perf_counter()
        by_lane = self._assign(jobs)
        for key in by_lane:  # create contexts before the clock matters
            self._session(key)
        self._pool = ThreadPoolExecutor(max_workers=self.builders)
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._loop, args=(key, js, t0))
                   for key, js in by_lane.items()]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        self._pool.shutdown(wait=True)
        rep = report(jobs, time.perf_counter() - t0)
        rep["errors"] = [str(e) for e in self.errors]
        rep["contexts"] = {(f"{D}d" if self._nlanes(D) == 1 else f"{D}d.{k}"): v
                           for (D, k), v in self.stats.items()}
        return rep


def run_plain(jobs, precision="f32", keep_results=False, session_factory=None, graph=False,
              progress=None):
    """Each job alone on the GPU, one after another in arrival order (a job
    starts when it has arrived and the previous one has finished); `graph`
    replays each job's step as a captured graph, as the server can.  The
    next job's scene is built on a helper thread while the current job runs
    (as soon as it has arrived), so the baseline loses no GPU time to scene
    building that the server would hide."""
    make = session_factory or (lambda D: Session(D, precision))
    sessions = {}
    t0 = time.perf_counter()
    for D in sorted({j.ndims for j in jobs}):
        sessions[D] = make(D)
        sessions[D].backend.set_graph_mode(bool(graph))
    order = sorted(jobs, key=lambda j: (j.arrival, j.jid))
    pool = ThreadPoolExecutor(max_workers=1)

    def build_after_arrival(job, t_start):
        wait = job.arrival - (time.perf_counter() - t_start)
        if wait > 0:
            time.sleep(wait)
        return job.build()

    t0 = time.perf_counter()
    fut = pool.submit(build_after_arrival, order[0], t0) if order else None
    for k, job in enumerate(order):
        cfg = fut.result()
        fut = pool.submit(build_after_arrival, order[k + 1], t0) if k + 1 < len(order) else None
        s = sessions[job.ndims]
        job.n_particles = cfg["n_particles"]
        lf = cfg["integrator"]
        h = s.scene.integrator(lf)
        s.scene.upload_all(lf)
        s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
        job.admitted = time.perf_counter() - t0
        try:
            s.backend.step_many([h], job.steps)
        except A.FaultError:  # the same failure as in the server: retired alone
            bits = s.backend.integrator_status(h)
            s.scene.release(lf)
            job.meta["failed"] = ("device fault: " + ("non-finite positions"
                                  if bits & A.GR_FAULT_NONFINITE
                                  else "particles beyond the cell range"))
            job.finished = time.perf_counter() - t0
            continue
        job.done_steps = job.steps
        if keep_results:
            s.scene.download_all(lf)
            job.result = {f: getattr(cfg["fluid"], f).copy() for f in cfg["fluid"].field_names()}
        s.scene.release(lf)
        job.finished = time.perf_counter() - t0
        if progress and (k + 1) % 20 == 0:
            progress(f"{k + 1}/{len(order)} jobs done")
    pool.shutdown(wait=True)
    return report(jobs, time.perf_counter() - t0)


def report(jobs, makespan):
    done = [j for j in jobs if not j.meta.get("rejected") and not j.meta.get("failed")]
    failed = [j for j in jobs if j.meta.get("failed")]
    lat = np.array([j.latency for j in done]) if done else np.array([math.nan])
    psteps = sum(j.n_particles * j.steps for j in done)
    return {"jobs": len(done), "rejected": sum(1 for j in jobs if j.meta.get("rejected")),
            "failed": len(failed), "makespan_s": makespan,
            "jobs_per_h": len(done) / makespan * 3600.0,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "particle_steps_per_s": psteps / makespan}


# ---------------------------------------------------------------------------
# Served workload: Poisson arrivals of 2D and 3D granular column collapses
# ---------------------------------------------------------------------------


def column_jobs(n_jobs, rate=20.0, frac_3d=0.2, steps=(500, 20000), step_scale=1.0, seed=11,
                dx2=0.0025, dx3=0.005, n3=(50_000, 500_000), size_scale=1.0):
    """Job stream: Poisson arrivals at `rate` jobs/s; 2D elasto-plastic
    columns (aspect U[0.5, 4], width U[0.1, 0.4] m, φ U[20°, 40°],
    ρ U[1500, 2000], dx = 2.5 mm) and, with probability `frac_3d`, 3D
    visco-plastic columns (target N log-uniform in `n3`, dx = 5 mm,
    φ U[25°, 40°]); step counts log-uniform in `steps`.  `step_scale` and
    `size_scale` shrink a sample for short runs (stated with the results)."""
    from . import configs as CF
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        nsteps = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))
                           * step_scale))
        nsteps = max(nsteps, 1)
        if rng.uniform() < frac_3d:
            n = math.exp(rng.uniform(math.log(n3[0]), math.log(n3[1]))) * size_scale
            a = rng.uniform(0.5, 1.0)
            w = (2.0 * n * dx3 ** 3 / a) ** (1.0 / 3.0)
            ph = rng.uniform(25.0, 40.0)
            build = (lambda w=w, a=a, ph=ph, k=k: CF.granular_column_3d(
                dx=dx3, width=w, depth=w / 2, height=a * w, phi_deg=ph, name=f"job{k}"))
            jobs.append(Job(k, 3, build, nsteps, t, meta={"kind": "3d"}))
        else:
            a = rng.uniform(0.5, 4.0)
            w = rng.uniform(0.1, 0.4) * math.sqrt(size_scale)
            ph = rng.uniform(20.0, 40.0)
            r0 = rng.uniform(1500.0, 2000.0)
            build = (lambda w=w, a=a, ph=ph, r0=r0, k=k: CF.ep_column_2d(
                dx=dx2, width=w, height=a * w, rho0=r0, phi_deg=ph, name=f"job{k}"))
            jobs.append(Job(k, 2, build, nsteps, t, meta={"kind": "2d"}))
    return jobs
