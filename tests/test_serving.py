"""Host logic of the continuous-batching server (grasph.jl_amd/serving.py)
on the CPU: the oracle stands in for the device through a thin adapter that
adds the batching calls the oracle lacks (each integrator stepped with its
own dt, release as a no-op, and the device's fault contract: a simulation
the oracle refuses to step is reported through GR_ERR_FAULT and
integrator_status while the others complete the call)."""
import types

import numpy as np
import pytest

from conftest import grasph as M, oracle_backend


class OracleDevice:
    """The oracle context plus the device-only batching calls."""

    def __init__(self, D):
        self.be = oracle_backend(D)
        self.dt = {}
        self.status = {}
        self.calls = []

    def __getattr__(self, name):
        return getattr(self.be, name)

    def set_graph_mode(self, on):
        pass

    def integrator_set_dt(self, h, dt):
        self.dt[h] = dt

    def step_many(self, hs, n):
        self.calls.append((tuple(hs), n))
        bad = []
        for h in hs:
            if self.status.get(h):
                bad.append(h)
                continue
            try:
                self.be.step(h, n, self.dt[h])
            except ValueError:  # the oracle refuses a non-finite position
                self.status[h] = M._abi.GR_FAULT_NONFINITE
                bad.append(h)
        if bad:
            raise M._abi.FaultError(f"integrator(s) {bad} faulted")

    def integrator_status(self, h):
        return self.status.get(h, 0)

    def integrator_release(self, h):
        self.dt.pop(h)
        self.status.pop(h, None)


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
    calls = srv.sessions[(2, 0)].backend.calls
    assert max(len(hs) for hs, _ in calls) > 1
    assert all(n <= 4 for _, n in calls)


def test_server_lanes_split_the_jobs_and_match_alone():
    """Two contexts per dimension: each job goes to the lane with the fewest
    steps assigned, both lanes step, and every result equals the job alone."""
    S = M.serving
    served = small_jobs(6)
    srv = S.Server(quantum=4, keep_results=True, session_factory=session, lanes=2)
    rep = srv.run(served)
    assert rep["jobs"] == len(served) and rep["failed"] == 0
    assert set(rep["contexts"]) == {"2d.0", "2d.1", "3d.0", "3d.1"}
    assert all(srv.sessions[(2, k)].backend.calls for k in (0, 1))
    alone = small_jobs(6)
    S.run_plain(alone, keep_results=True, session_factory=session)
    for a, b in zip(served, alone):
        for f, v in b.result.items():
            assert np.array_equal(a.result[f], v), (a.jid, f)


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


def _nan_rho_column(k):
    cfg = M.configs.ep_column_2d(dx=0.01, width=0.07, height=0.05, name=f"nan{k}")
    cfg["fluid"].rho[7] = float("nan")  # NaN density: positions turn NaN a step later
    return cfg


def _bad_column(k):
    raise ValueError(f"job {k}: width must be positive")


def test_failed_jobs_do_not_disturb_the_others():
    """A job whose simulation turns NaN mid-run, a job whose scene cannot be
    built and a job that can never fit the budget fail or are rejected
    alone; every other job finishes with the state it reaches alone."""
    S = M.serving
    served = small_jobs()
    healthy = {j.jid for j in served}
    one = max(S.footprint_bytes(j.build()) for j in served)
    served.append(S.Job(10, 2, lambda: _nan_rho_column(10), 9, arrival=0.0))
    served.append(S.Job(11, 2, lambda: _bad_column(11), 9, arrival=0.0))
    big = S.Job(12, 2, lambda: M.configs.ep_column_2d(dx=0.01, width=0.06, height=0.05), 5,
                arrival=0.0)
    big.bytes = int(one * 100)  # declared footprint beyond the budget
    served.append(big)
    srv = S.Server(quantum=4, keep_results=True, hbm_budget=one * 20, session_factory=session)
    rep = srv.run(served)
    assert rep["errors"] == []
    assert rep["jobs"] == len(healthy) and rep["failed"] == 2 and rep["rejected"] == 1
    by = {j.jid: j for j in served}
    assert "non-finite" in by[10].meta["failed"]
    assert by[10].done_steps < by[10].steps
    assert "build" in by[11].meta["failed"]
    assert by[12].meta.get("rejected")
    # the NaN job ran batched with healthy jobs before it failed
    calls = srv.sessions[(2, 0)].backend.calls
    assert any(len(hs) > 1 for hs, _ in calls)
    alone = small_jobs()
    S.run_plain(alone, keep_results=True, session_factory=session)
    for b in alone:
        a = by[b.jid]
        assert a.done_steps == a.steps
        for f, v in b.result.items():
            assert np.array_equal(a.result[f], v), (a.jid, f)
    # the budget is whole again
    assert abs(srv.budget.used) < 1e-6
