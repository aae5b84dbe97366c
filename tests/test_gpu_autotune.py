"""Launch-shape tuning (GR_TUNE_*, grasph.jl_amd/autotune.py): every
choice of the knobs gives bit-identical states, and the autotuner leaves the
simulations where it found them."""
import numpy as np
import pytest

from conftest import grasph as M

pytestmark = pytest.mark.gpu
A = M._abi


def _run(builder, kw, D, nsteps, knobs=(), tune=False, passes=1):
    cfg = builder(**kw)
    lf = cfg["integrator"]
    s = M.Session(D, "f32")
    s.backend.set_graph_mode(True)
    h = s.scene.integrator(lf)
    s.scene.upload_all(lf)
    s.backend.integrator_set_dt(h, s.backend.integrator_dt(h, cfg["cfl"]))
    for k, v in knobs:
        s.backend.tune_set(k, v)
    rec = None
    if tune:
        rec = M.autotune.autotune(s, [lf], steps=2, reps=1, passes=passes)
    s.backend.step_many([h], nsteps)
    s.scene.download_all(lf)
    f = cfg["fluid"]
    return {k: getattr(f, k).copy() for k in ["id"] + f.field_names()}, rec


CASES = [(M.configs.dambreak_3d, dict(dx=0.03), 3),
         (M.configs.ep_column_2d, dict(dx=0.005, width=0.1, height=0.06), 2),
         (M.configs.granular_column_3d, dict(dx=0.02, width=0.12, depth=0.06, height=0.08), 3)]


@pytest.mark.parametrize("builder,kw,D", CASES, ids=["dambreak3d", "ep_column2d", "vpmc3d"])
def test_every_launch_shape_gives_the_same_state(builder, kw, D):
    ref, _ = _run(builder, kw, D, 12)
    shapes = [((A.GR_TUNE_SWEEP_BLOCK, 64),), ((A.GR_TUNE_SWEEP_BLOCK, 128),),
              ((A.GR_TUNE_NL_WIN, 64), (A.GR_TUNE_NL_SPAN, 1)),
              ((A.GR_TUNE_NL_WIN, 256), (A.GR_TUNE_NL_SPAN, 4)),
              ((A.GR_TUNE_MAXNB, 4),), ((A.GR_TUNE_INV_RECORDS, 0),),
              ((A.GR_TUNE_MAXNB, 6), (A.GR_TUNE_SWEEP_BLOCK, 64), (A.GR_TUNE_NL_WIN, 64)),
              ((A.GR_TUNE_FUSE_SORT, 0),), ((A.GR_TUNE_FUSE_SORT, 0), (A.GR_TUNE_INV_RECORDS, 0))]
    for knobs in shapes:
        got, _ = _run(builder, kw, D, 12, knobs)
        for f, v in ref.items():
            assert np.array_equal(got[f], v, equal_nan=True), (knobs, f)


def test_autotune_records_a_choice_and_leaves_the_state_alone():
    ref, _ = _run(M.configs.dambreak_3d, dict(dx=0.03), 3, 10)
    got, rec = _run(M.configs.dambreak_3d, dict(dx=0.03), 3, 10, tune=True)
    for f, v in ref.items():
        assert np.array_equal(got[f], v), f
    assert set(rec["chosen"]) == {"sweep_block", "nl_win", "nl_span", "maxnb"}
    assert len(rec["trials"]) == 12 and rec["ms_per_step_tuned"] > 0
    with pytest.raises(ValueError):
        M.Session(3, "f32").backend.tune_set(A.GR_TUNE_NL_WIN, 100)


def test_autotune_second_pass_leaves_the_state_alone():
    ref, _ = _run(M.configs.ep_column_2d, dict(dx=0.005, width=0.1, height=0.06), 2, 10)
    got, rec = _run(M.configs.ep_column_2d, dict(dx=0.005, width=0.1, height=0.06), 2, 10,
                    tune=True, passes=2)
    for f, v in ref.items():
        assert np.array_equal(got[f], v), f
    assert rec["passes"] == 2 and len(rec["trials"]) == 24
