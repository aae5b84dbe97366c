"""Host logic of the continuous-batching server (grasph.jl_amd/serving.py)
on the CPU: the oracle stands in for the device through a thin adapter that
adds the batching calls the oracle lacks (each integrator stepped with its
own dt, release as a no-op)."""
import types

import numpy as np
import pytest

from conftest import grasph as M, oracle_backend


class OracleDevice:
    """The oracle context plus the device-only batching calls."""

    def __init__(self, D):
        self.be = oracle_backend(D)
        self.dt = {}
        self.calls = []

    def __getattr__(self, name):
        return getattr(self.be, name)

    def set_graph_mode(self, on):
        pass

    def integrator_set_dt(self, h, dt):
        self.dt[h] = dt

    def step_many(self, hs, n):
        self.calls.append((tuple(hs), n))
        for h in hs:
            self.be.step(h, n, self.dt[h])

    def integrator_release(self, h):
        self.dt.pop(h)


def session(D):
    dev = OracleDevice(D)
    return types.SimpleNamespace(backend=dev, scene=M.DeviceScene(dev, D))


def small_jobs(n=6):
    S = M.serving
    jobs = []
    for k in range(n):
        if k % 3 == 2:
            b = (lambda k=k: M.configs.granular_column_3d(dx=0.02, width=0.1 + 0.02 * k,
                                                          depth=0.06, height=0.08, name=f"j{k}"))
            jobs.append(S.Job(k, 3, b, 7 + k, arrival=0.0))
        else:
            b = (lambda k=k: M.configs.ep_column_2d(dx=0.01, width=0.06 + 0.01 * k, height=0.05,
                                                    name=f"j{k}"))
            jobs.append(S.Job(k, 2, b, 5 + 2 * k, arrival=0.0))
    return jobs


def test_server_finishes_every_job_and_matches_alone():
    S = M.serving
    served = small_jobs()
    srv = S.Server(quantum=4, keep_results=True, session_factory=session)
    rep = srv.run(served)
    assert rep["jobs"] == len(served) and rep["rejected"] == 0
    assert all(j.done_steps == j.steps for j in served)
    alone = small_jobs()
    S.run_plain(alone, keep_results=True, session_factory=session)
    for a, b in zip(served, alone):
        for f, v in b.result.items():
            assert np.array_equal(a.result[f], v), (a.jid, f)
    # quanta never overshoot a job's step count and batch several jobs
    calls = srv.sessions[2].backend.calls
    assert max(len(hs) for hs, _ in calls) > 1
    assert all(n <= 4 for _, n in calls)


def test_budget_limits_residency_and_rejects_oversized():
    S = M.serving
    jobs = small_jobs(4)
    one = max(S.footprint_bytes(j.build()) for j in jobs)
    srv = S.Server(quantum=3, hbm_budget=one * 1.01, session_factory=session)
    srv.run(jobs)
    # at every admission, the resident footprints fit the budget
    for j in jobs:
        resident = [k for k in jobs if k.admitted <= j.admitted < k.finished]
        assert sum(k.bytes for k in resident) <= one * 1.01
    assert max(j.bytes for j in jobs) <= one * 1.01
    huge = small_jobs(2)
    rep = S.Server(hbm_budget=1000, session_factory=session).run(huge)
    assert rep["rejected"] == 2


def test_column_jobs_stream_statistics():
    S = M.serving
    jobs = S.column_jobs(400, rate=20.0, seed=3)
    arr = np.array([j.arrival for j in jobs])
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.01
    frac3 = np.mean([j.ndims == 3 for j in jobs])
    assert 0.13 < frac3 < 0.27
    st = np.array([j.steps for j in jobs])
    assert st.min() >= 500 and st.max() <= 20000
