"""gr_sort_static: the sort entry point that allocates nothing (its scratch
is a static device array of the library) against gr_sort, through the C
ABI: bit-identical particle order, positions and payload, on two calls in a
row on one GPU (the second with a smaller system, so the scratch holds the
first call's larger contents), on a shuffled and on an already sorted
input, and the size limit refused with GR_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

from conftest import grasph as M

pytestmark = pytest.mark.gpu
A = M._abi

MAXN = 1 << 21


def _static_sort(sess, h, cutoff):
    lib = sess.backend.lib
    fn = lib.gr_sort_static
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_double]
    return fn(sess.backend.ctx, h, cutoff)


def _systems(seed, sizes, D=3):
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        ps = M.FluidParticleSystem(f"s{k}", n, D, 1.0, 1.0)
        ps.x[:] = rng.uniform(-0.1, 0.7, (n, D))
        ps.x[: n // 8] = np.round(ps.x[: n // 8] / 0.026) * 0.026  # on cell faces
        ps.rho[:] = np.arange(n, dtype=np.float64)
        ps.v[:] = rng.uniform(-1, 1, (n, D))
        out.append(ps)
    return out


def _state(sc, ps):
    sc.download(ps)
    return {k: getattr(ps, k).copy() for k in ("id", "x", "v", "rho")}


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_static_sort_equals_sort_on_two_calls(prec):
    cutoff = 0.026
    sizes = (150_000, 40_000)
    ref = {}
    s1 = M.Session(3, prec)
    for ps in _systems(5, sizes):
        s1.scene.upload(ps)
        s1.backend.sort(s1.scene.system(ps), cutoff)
        ref[ps.name] = _state(s1.scene, ps)
    s2 = M.Session(3, prec)
    systems = _systems(5, sizes)
    for ps in systems:
        s2.scene.upload(ps)
    for ps in systems:  # larger first, then smaller: stale scratch contents
        assert _static_sort(s2, s2.scene.system(ps), cutoff) == A.GR_OK
        got = _state(s2.scene, ps)
        for k, v in ref[ps.name].items():
            assert np.array_equal(got[k], v), (ps.name, k)
    # a second pass over sorted input (the early-out), larger system again
    for ps in systems:
        before = _state(s2.scene, ps)
        assert _static_sort(s2, s2.scene.system(ps), cutoff) == A.GR_OK
        after = _state(s2.scene, ps)
        for k, v in before.items():
            assert np.array_equal(after[k], v), (ps.name, k)


def test_static_sort_refuses_beyond_its_limit():
    s = M.Session(3, "f32")
    ps = M.BasicParticleSystem("big", MAXN + 1, 3, 1.0, 1.0)
    ps.x[:] = np.random.default_rng(1).uniform(0, 1, ps.x.shape)
    s.scene.upload(ps, ["x"])
    rc = _static_sort(s, s.scene.system(ps), 0.05)
    assert rc == A.GR_ERR_CAPACITY
    assert b"static" in s.backend._last_error(s.backend.ctx)
    # the regular sort of the same system still works
    s.backend.sort(s.scene.system(ps), 0.05)
