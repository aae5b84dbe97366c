import ctypes as C
import importlib.util
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

grasph = __graft_entry__.package()
A = grasph._abi

ORACLE_LIB = ROOT / "oracle" / "libgrasph_oracle.so"
# tests/test_oracle_mutants.py points this at oracles built with a planted
# mistake, to show that the pins fail on them
if os.environ.get("GRASPH_ORACLE_LIB"):
    ORACLE_LIB = Path(os.environ["GRASPH_ORACLE_LIB"])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def _oracle_lib():
    if not ORACLE_LIB.exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return C.CDLL(str(ORACLE_LIB))


_ORC = None


def oracle_backend(ndims, threads=0):
    """A context of the CPU oracle behind the same ABI as the HIP library."""
    global _ORC
    if _ORC is None:
        _ORC = _oracle_lib()
        _ORC.orc_scene_create.restype = C.c_int
        _ORC.orc_scene_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        _ORC.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        for name in ("orc_kernel_w", "orc_kernel_dw_dq", "orc_kernel_norm"):
            getattr(_ORC, name).restype = C.c_double
        _ORC.orc_kernel_norm.argtypes = [C.c_int, C.c_int, C.c_double]
        _ORC.orc_kernel_w.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        _ORC.orc_kernel_dw_dq.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        _ORC.orc_pow_int.restype = C.c_double
        _ORC.orc_pow_int.argtypes = [C.c_double, C.c_int64]
        _ORC.orc_brute_force_counts.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double,
                                                C.POINTER(C.c_int64)]
    ctx = C.c_void_p()
    assert _ORC.orc_scene_create(C.byref(ctx), ndims) == 0
    if threads:
        _ORC.orc_set_threads(ctx, threads)
    return A.CBackend(_ORC, "orc_", ctx, "orc_last_error", "orc_scene_destroy")


def oracle_lib():
    oracle_backend(2).close()
    return _ORC


def oracle_scene(ndims):
    return grasph.DeviceScene(oracle_backend(ndims), ndims)


def brute_counts(scene, ps_a, ps_b, cutoff):
    n = ps_a.n
    out = np.zeros(n, dtype=np.int64)
    rc = _ORC.orc_brute_force_counts(scene.be.ctx, scene.system(ps_a), scene.system(ps_b),
                                     cutoff, out.ctypes.data_as(C.POINTER(C.c_int64)))
    assert rc == 0
    return out


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
