"""Continuous batching on the device: jobs admitted while others run, stepped
together in shared launches and retired with their segments reused, give the
same state as each job run alone (bit for bit)."""
import numpy as np
import pytest

from conftest import grasph as M

pytestmark = pytest.mark.gpu


def small_jobs():
    S = M.serving
    jobs = []
    for k in range(6):
        if k % 3 == 2:
            b = (lambda k=k: M.configs.granular_column_3d(dx=0.02, width=0.1 + 0.02 * k,
                                                          depth=0.06, height=0.08, name=f"j{k}"))
            jobs.append(S.Job(k, 3, b, 9 + 3 * k, arrival=0.05 * k))
        else:
            b = (lambda k=k: M.configs.ep_column_2d(dx=0.01, width=0.06 + 0.01 * k, height=0.05,
                                                    name=f"j{k}"))
            jobs.append(S.Job(k, 2, b, 11 + 4 * k, arrival=0.05 * k))
    return jobs


@pytest.mark.parametrize("graph", [False, True])
def test_served_results_equal_alone(graph):
    S = M.serving
    served = small_jobs()
    rep = S.Server(quantum=5, graph=graph, keep_results=True).run(served)
    assert rep["jobs"] == 6 and all(j.done_steps == j.steps for j in served)
    alone = small_jobs()
    S.run_plain(alone, keep_results=True)
    for a, b in zip(served, alone):
        assert set(a.result) == set(b.result)
        for f, v in b.result.items():
            assert np.array_equal(a.result[f], v, equal_nan=True), (a.jid, f)


def _state(cfg):
    f = cfg["fluid"]
    return {k: getattr(f, k).copy() for k in f.field_names()}


@pytest.mark.parametrize("graph", [False, True])
def test_faulted_simulations_are_isolated_in_shared_launches(graph):
    """Three columns stepped together; one turns NaN (a NaN density), one
    throws a particle far out of its domain.  gr_step_many completes, raises
    GR_ERR_FAULT, reports each faulted integrator's bits, and the healthy
    column's state equals its state alone, bit for bit, through further
    shared calls."""
    A = M._abi

    def col(w, name):
        return M.configs.ep_column_2d(dx=0.01, width=w, height=0.05, name=name)

    s = M.Session(2, "f32")
    s.backend.set_graph_mode(graph)
    cfgs = {"ok": col(0.08, "ok"), "nan": col(0.07, "nan"), "far": col(0.06, "far")}
    cfgs["nan"]["fluid"].rho[5] = float("nan")
    cfgs["far"]["fluid"].v[3] = (4.0e7, 0.0)
    hs = {}
    for k, c in cfgs.items():
        lf = c["integrator"]
        hs[k] = s.scene.integrator(lf)
        s.scene.upload_all(lf)
        s.backend.integrator_set_dt(hs[k], s.backend.integrator_dt(hs[k], c["cfl"]))
    allh = [hs["ok"], hs["nan"], hs["far"]]
    with pytest.raises(A.FaultError):
        s.backend.step_many(allh, 6)
    assert s.backend.integrator_status(hs["ok"]) == 0
    assert s.backend.integrator_status(hs["nan"]) & A.GR_FAULT_NONFINITE
    assert s.backend.integrator_status(hs["far"]) & A.GR_FAULT_DOMAIN
    with pytest.raises(A.FaultError):  # still batched with the faulted ones
        s.backend.step_many(allh, 4)
    s.scene.release(cfgs["nan"]["integrator"])
    s.backend.step_many([hs["ok"], hs["far"]], 0)  # nothing stepped, nothing raised
    s.scene.release(cfgs["far"]["integrator"])
    s.backend.step_many([hs["ok"]], 3)
    s.scene.download_all(cfgs["ok"]["integrator"])
    got = _state(cfgs["ok"])
    # alone
    s2 = M.Session(2, "f32")
    s2.backend.set_graph_mode(graph)
    ref = col(0.08, "ok")
    h = s2.scene.integrator(ref["integrator"])
    s2.scene.upload_all(ref["integrator"])
    s2.backend.integrator_set_dt(h, s2.backend.integrator_dt(h, ref["cfl"]))
    s2.backend.step_many([h], 13)
    s2.scene.download_all(ref["integrator"])
    for f, v in _state(ref).items():
        assert np.array_equal(got[f], v, equal_nan=True), f


def test_server_survives_a_failed_job():
    """The served stream with a NaN job, a job that cannot be built and one
    that can never fit the budget: the others finish with their alone state."""
    S = M.serving
    served = small_jobs()
    nan_cfg = lambda: M.configs.ep_column_2d(dx=0.01, width=0.07, height=0.05, name="nan")  # noqa: E731

    def nan_job():
        c = nan_cfg()
        c["fluid"].rho[4] = float("nan")
        return c

    def bad_job():
        raise ValueError("width must be positive")

    served.append(S.Job(20, 2, nan_job, 15, arrival=0.0))
    served.append(S.Job(21, 2, bad_job, 15, arrival=0.02))
    big = S.Job(22, 2, nan_cfg, 5, arrival=0.0)
    big.bytes = int(300e9)
    served.append(big)
    rep = S.Server(quantum=5, graph=True, keep_results=True).run(served)
    assert rep["errors"] == [] and rep["failed"] == 2 and rep["rejected"] == 1
    assert rep["jobs"] == 6
    by = {j.jid: j for j in served}
    assert "non-finite" in by[20].meta["failed"]
    alone = small_jobs()
    S.run_plain(alone, keep_results=True)
    for b in alone:
        a = by[b.jid]
        for f, v in b.result.items():
            assert np.array_equal(a.result[f], v, equal_nan=True), (a.jid, f)
