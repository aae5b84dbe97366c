"""C-ABI checks that need no GPU: the HIP library and the oracle load, and
every entry point declared in their headers is exported.  No compute call is
made (a context needs a HIP device)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from conftest import A, grasph

ROOT = Path(__file__).resolve().parents[1]


def declared(header, prefix):
    text = Path(header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(rf"\b({prefix}[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_step_api():
    names = declared(ROOT / "include" / "grasph_hip.h", "gr_")
    for core in ("gr_ctx_create", "gr_system_create", "gr_interaction_create",
                 "gr_integrator_create", "gr_step", "gr_step_many", "gr_sort",
                 "gr_create_grid", "gr_sweep", "gr_update_state", "gr_generate_ghosts",
                 "gr_update_ghost", "gr_update_ghost_kinematics"):
        assert core in names, core


def test_hip_library_exports_every_declared_symbol():
    lib = A.load_hip_library()
    missing = [n for n in declared(ROOT / "include" / "grasph_hip.h", "gr_")
               if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    """Every declared entry point is bound by the ctypes layer (common or
    device-only signature table)."""
    names = declared(ROOT / "include" / "grasph_hip.h", "gr_")
    bound = {"gr_" + k for k in list(A._SIGS) + list(A._GPU_SIGS)}
    bound |= {"gr_ctx_create", "gr_ctx_destroy", "gr_last_error"}
    assert sorted(set(names) - bound) == []
    # and each bound entry point has a method of the same name
    methods = {n for n in list(A._SIGS) + list(A._GPU_SIGS) if callable(getattr(A.CBackend, n, None))}
    assert sorted(set(A._SIGS) | set(A._GPU_SIGS)) == sorted(methods)


def test_oracle_exports_every_declared_symbol():
    lib = C.CDLL(str(ROOT / "oracle" / "libgrasph_oracle.so"))
    missing = [n for n in declared(ROOT / "oracle" / "grasph_oracle.h", "orc_")
               if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_mirrors_the_device_api():
    """The oracle has a counterpart for every device entry point that is not
    device-only (streams, graphs, timing, batching)."""
    dev = {n[3:] for n in declared(ROOT / "include" / "grasph_hip.h", "gr_")}
    orc = {n[4:] for n in declared(ROOT / "oracle" / "grasph_oracle.h", "orc_")}
    device_only = {"ctx_create", "ctx_destroy", "ctx_info", "last_error", "integrator_set_dt",
                   "integrator_release",
                   "step_many", "set_graph_mode", "set_timing", "get_timing",
                   "get_pair_count", "get_pair_stats", "bench_copy", "integrator_status",
                   "tune_set", "tune_get", "sort_static"}
    assert sorted(dev - orc - device_only) == []


def test_missing_hip_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        A.load_hip_library(tmp_path / "libgrasph_hip.so")


def test_product_package_does_not_reach_the_oracle():
    """The package under grasph.jl_amd/ never loads the oracle library."""
    pkg = ROOT / "grasph.jl_amd"
    for p in pkg.rglob("*"):
        if p.suffix in (".py", ".hip", ".h", ".cpp", ".jl") and p.is_file():
            text = p.read_text(errors="replace")
            assert "libgrasph_oracle" not in text, p
            assert "orc_" not in text, p
