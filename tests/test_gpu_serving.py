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
