"""ctypes binding of the C ABI declared in include/grasph_hip.h.

The constants below restate the enums of that header.  `CBackend` wraps one
context of a shared library exporting the ABI with a given symbol prefix
("gr_" for the HIP library).  Every call checks the int status and raises
the exception the reference raises for the same condition (ArgumentError →
ValueError, AssertionError → AssertionError, anything else → RuntimeError),
with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# ---- status codes -------------------------------------------------------
(GR_OK, GR_ERR_ARGUMENT, GR_ERR_ASSERTION, GR_ERR_HIP, GR_ERR_CAPACITY, GR_ERR_STATE,
 GR_ERR_FAULT) = range(7)
# ---- per-integrator fault bits ---------------------------------------------
GR_FAULT_NONFINITE, GR_FAULT_DOMAIN = 1, 2
# ---- precision ------------------------------------------------------------
GR_PREC_F32, GR_PREC_F64 = 0, 1
# ---- system kinds -----------------------------------------------------------
GR_SYS_BASIC, GR_SYS_FLUID, GR_SYS_STRESS, GR_SYS_EP, GR_SYS_GHOST = range(5)
# ---- fields -------------------------------------------------------------------
(GR_F_X, GR_F_V, GR_F_RHO, GR_F_DVDT, GR_F_DRHODT, GR_F_P, GR_F_STRESS,
 GR_F_STRAIN_RATE, GR_F_VORTICITY, GR_F_STRAIN, GR_F_STRAIN_P,
 GR_F_V_ADJUSTMENT, GR_F_NORMALS, GR_F_W_SUM, GR_F_COUNT) = range(15)
GR_I_ID, GR_I_IDX_ORIGINAL, GR_I_IDX_BOUNDARY = range(3)
FIELD_IDS = {
    "x": GR_F_X, "v": GR_F_V, "rho": GR_F_RHO, "dvdt": GR_F_DVDT,
    "drhodt": GR_F_DRHODT, "p": GR_F_P, "stress": GR_F_STRESS,
    "strain_rate": GR_F_STRAIN_RATE, "vorticity": GR_F_VORTICITY,
    "strain": GR_F_STRAIN, "strain_p": GR_F_STRAIN_P,
    "v_adjustment": GR_F_V_ADJUSTMENT, "normals": GR_F_NORMALS, "w_sum": GR_F_W_SUM,
}
INDEX_IDS = {"id": GR_I_ID, "idx_original": GR_I_IDX_ORIGINAL,
             "idx_boundary": GR_I_IDX_BOUNDARY}
# ---- kernels / functors / wrappers / updaters / copier modes ------------------
GR_KERNEL_CUBIC, GR_KERNEL_WENDLAND_C2 = 0, 1
(GR_PFN_NONE, GR_PFN_FLUID, GR_PFN_CAUCHY, GR_PFN_SR, GR_PFN_SRV, GR_PFN_COUNT,
 GR_PFN_INTERP, GR_PFN_XSPH, GR_PFN_NCOUNT, GR_PFN_FLUIDSOLID) = range(10)
GR_WRAP_NONE, GR_WRAP_STATIC, GR_WRAP_DYNAMIC = range(3)
(GR_UPD_NOTHING, GR_UPD_ZERO, GR_UPD_TAIT, GR_UPD_LINEAR, GR_UPD_EP, GR_UPD_VPMC, GR_UPD_VNORM,
 GR_UPD_PRESCRIBED_V, GR_UPD_HOOKE) = range(9)
GR_COPY_PLAIN, GR_COPY_HOUSEHOLDER = 0, 1
GR_SCHEME_LEAPFROG, GR_SCHEME_RK4 = 0, 1
(GR_TUNE_NL_WIN, GR_TUNE_NL_SPAN, GR_TUNE_SWEEP_BLOCK, GR_TUNE_MAXNB, GR_TUNE_INV_RECORDS,
 GR_TUNE_FUSE_SORT) = range(6)


class gr_pfn_desc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("sigma", C.c_int32), ("alpha", C.c_double),
                ("beta", C.c_double), ("h", C.c_double), ("delta", C.c_double),
                ("epsilon", C.c_double)]


_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_ip = C.POINTER(C.c_int)

# name → argtypes after the leading context pointer
_SIGS = {
    "system_create": [_i, _i64, _i, _d, _d, _dp, _d, _ip],
    "system_add_updater": [_i, _i, _i, _dp, _i],
    "system_upload": [_i, _i, _dp, _i64],
    "system_download": [_i, _i, _dp, _i64],
    "system_upload_index": [_i, _i, _i64p, _i64],
    "system_download_index": [_i, _i, _i64p, _i64],
    "system_get_n": [_i, _i64p],
    "ghost_create": [_i, _d, _i, _dp, _dp, _ip],
    "virtual_create": [_i, _i, _ip, _dp],
    "probe_create": [_i, _dp, _i, _i, _ip],
    "integrator_set_probes": [_i, _i, _ip, _i, _ip],
    "measure_probes": [_i],
    "ghost_set_copier": [_i, _i, _i, _ip, _ip],
    "interaction_create": [_i, _d, _i, _i, _i, _dp, _dp, _d, _d, _i,
                           C.POINTER(gr_pfn_desc), _ip],
    "integrator_create": [_i, _ip, _i, _ip, _i, _ip, _d, _ip],
    "interaction_set_vadjust": [_i, _d],
    "integrator_set_scheme": [_i, _i],
    "integrator_dt": [_i, _d, _dp],
    "step": [_i, _i, _d],
    "sort": [_i, _d],
    "create_grid": [_i],
    "sweep": [_i, _i],
    "adjust_v": [_i],
    "update_state": [_i, _i, _d],
    "generate_ghosts": [_i],
    "update_ghost_kinematics": [_i],
    "update_ghost": [_i, _i],
    "get_grid": [_i, _dp, _i64p, _i64p],
    "get_cell_start": [_i, _i, _i64p, _i64],
}
# entry points only the device library has
_GPU_SIGS = {
    "integrator_set_dt": [_i, _d],
    "integrator_release": [_i],
    "step_many": [_i, _ip, _i],
    "set_graph_mode": [_i],
    "set_timing": [_i],
    "get_timing": [_dp, _i],
    "get_pair_count": [_i, _i64p],
    "get_pair_stats": [_i64p, _i64p],
    "ctx_info": [C.c_char_p, _i64],
    "bench_copy": [_i64, _i],
    "integrator_status": [_i, _ip],
    "tune_set": [_i, _i],
    "sort_static": [_i, _d],
    "tune_get": [_i, _ip],
}

PKG_DIR = Path(__file__).resolve().parent
HIP_LIB_PATH = PKG_DIR / "csrc" / "libgrasph_hip.so"


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _i64ptr(a):
    return a.ctypes.data_as(_i64p)


def _iptr(a):
    return a.ctypes.data_as(_ip)


class FaultError(RuntimeError):
    """gr_step / gr_step_many completed, but some of the integrators stepped
    faulted (GR_ERR_FAULT); `integrator_status` tells which."""


class CBackend:
    """One context of a library exporting the ABI of include/grasph_hip.h.

    `make_ctx(lib)` returns the opaque context pointer; `prefix` selects the
    symbol family; `last_error` is the name of the error-string getter.
    """

    def __init__(self, lib, prefix, ctx, last_error, destroy, extra=()):
        self.lib = lib
        self.prefix = prefix
        self.ctx = ctx
        self._last_error = getattr(lib, last_error)
        self._last_error.restype = C.c_char_p
        self._last_error.argtypes = [_vp]
        self._destroy = getattr(lib, destroy)
        self._destroy.argtypes = [_vp]
        self._fns = {}
        sigs = dict(_SIGS)
        for name in extra:
            sigs[name] = _GPU_SIGS[name]
        for name, args in sigs.items():
            fn = getattr(lib, prefix + name)
            fn.restype = C.c_int
            fn.argtypes = [_vp] + args
            self._fns[name] = fn

    def close(self):
        if self.ctx:
            self._destroy(self.ctx)
            self.ctx = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def has(self, name):
        return name in self._fns

    def call(self, name, *args):
        rc = self._fns[name](self.ctx, *args)
        if rc != GR_OK:
            msg = (self._last_error(self.ctx) or b"").decode(errors="replace")
            if rc == GR_ERR_ARGUMENT:
                raise ValueError(msg)
            if rc == GR_ERR_ASSERTION:
                raise AssertionError(msg)
            if rc == GR_ERR_FAULT:
                raise FaultError(msg)
            raise RuntimeError(f"{self.prefix}{name} failed ({rc}): {msg}")
        return rc

    # ---- typed helpers --------------------------------------------------
    def system_create(self, kind, n, ns, mass, c, source_v, source_rho):
        sv = np.ascontiguousarray(source_v, dtype=np.float64)
        out = C.c_int()
        self.call("system_create", kind, n, ns, mass, c, _dptr(sv), source_rho, C.byref(out))
        return out.value

    def system_add_updater(self, sys, stage, kind, params):
        p = np.ascontiguousarray(params if len(params) else [0.0], dtype=np.float64)
        self.call("system_add_updater", sys, stage, kind, _dptr(p), len(params))

    def system_upload(self, sys, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self.call("system_upload", sys, field, _dptr(a), a.size)

    def system_download(self, sys, field, count):
        a = np.empty(count, dtype=np.float64)
        self.call("system_download", sys, field, _dptr(a), count)
        return a

    def system_upload_index(self, sys, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.int64).reshape(-1)
        self.call("system_upload_index", sys, field, _i64ptr(a), a.size)

    def system_download_index(self, sys, field, count):
        a = np.empty(count, dtype=np.int64)
        self.call("system_download_index", sys, field, _i64ptr(a), count)
        return a

    def system_get_n(self, sys):
        out = C.c_int64()
        self.call("system_get_n", sys, C.byref(out))
        return out.value

    def virtual_create(self, sys, zero_fields, prescribed_v):
        z = np.ascontiguousarray(zero_fields, dtype=np.int32)
        pv = np.ascontiguousarray(prescribed_v, dtype=np.float64)
        self.call("virtual_create", sys, len(z), _iptr(z), _dptr(pv))

    def probe_create(self, sys, prescribed_v, mirror, zero_fields):
        z = np.ascontiguousarray(zero_fields, dtype=np.int32)
        pv = np.ascontiguousarray(prescribed_v, dtype=np.float64)
        self.call("probe_create", sys, _dptr(pv), int(mirror), len(z), _iptr(z))

    def integrator_set_probes(self, integ, probes, inters):
        p = np.ascontiguousarray(probes, dtype=np.int32)
        t = np.ascontiguousarray(inters, dtype=np.int32)
        self.call("integrator_set_probes", integ, len(p), _iptr(p), len(t), _iptr(t))

    def measure_probes(self, integ):
        self.call("measure_probes", integ)

    def ghost_create(self, src, cutoff, normals, points):
        nn = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1)
        pp = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
        nb = len(normals)
        out = C.c_int()
        self.call("ghost_create", src, cutoff, nb, _dptr(nn), _dptr(pp), C.byref(out))
        return out.value

    def ghost_set_copier(self, ghost, stage, fields, modes):
        f = np.ascontiguousarray(fields if fields else [0], dtype=np.int32)
        m = np.ascontiguousarray(modes if modes else [0], dtype=np.int32)
        self.call("ghost_set_copier", ghost, stage, len(fields), _iptr(f), _iptr(m))

    def interaction_create(self, kernel_kind, h, a, partner, wrap, normal, point, beta,
                           ljcut, pfns):
        nrm = np.zeros(3) if normal is None else np.ascontiguousarray(normal, dtype=np.float64)
        pnt = np.zeros(3) if point is None else np.ascontiguousarray(point, dtype=np.float64)
        arr = (gr_pfn_desc * max(1, len(pfns)))()
        for k, pd in enumerate(pfns):
            arr[k] = gr_pfn_desc(*pd)
        out = C.c_int()
        self.call("interaction_create", kernel_kind, h, a, partner, wrap, _dptr(nrm), _dptr(pnt),
                  beta, ljcut, len(pfns), arr, C.byref(out))
        return out.value

    def integrator_create(self, systems, inters, ghosts, gamma):
        s = np.ascontiguousarray(systems if len(systems) else [0], dtype=np.int32)
        i = np.ascontiguousarray(inters if len(inters) else [0], dtype=np.int32)
        g = np.ascontiguousarray(ghosts if len(ghosts) else [0], dtype=np.int32)
        out = C.c_int()
        self.call("integrator_create", len(systems), _iptr(s), len(inters), _iptr(i),
                  len(ghosts), _iptr(g), gamma, C.byref(out))
        return out.value

    def interaction_set_vadjust(self, inter, epsilon):
        self.call("interaction_set_vadjust", inter, float(epsilon))

    def integrator_set_scheme(self, integ, scheme):
        self.call("integrator_set_scheme", integ, scheme)

    def integrator_dt(self, integ, cfl):
        out = C.c_double()
        self.call("integrator_dt", integ, cfl, C.byref(out))
        return out.value

    def step(self, integ, nsteps, dt):
        self.call("step", integ, nsteps, dt)

    def sort(self, sys, cutoff):
        self.call("sort", sys, cutoff)

    def create_grid(self, inter):
        self.call("create_grid", inter)

    def sweep(self, inter, stage):
        self.call("sweep", inter, stage)

    def adjust_v(self, inter):
        self.call("adjust_v", inter)

    def update_state(self, sys, stage, dt):
        self.call("update_state", sys, stage, dt)

    def generate_ghosts(self, g):
        self.call("generate_ghosts", g)

    def update_ghost_kinematics(self, g):
        self.call("update_ghost_kinematics", g)

    def update_ghost(self, g, stage):
        self.call("update_ghost", g, stage)

    def get_grid(self, inter, ndims):
        m = np.zeros(3)
        n = np.zeros(3, dtype=np.int64)
        nc = C.c_int64()
        self.call("get_grid", inter, _dptr(m), _i64ptr(n), C.byref(nc))
        return m[:ndims].copy(), n[:ndims].copy(), nc.value

    def get_cell_start(self, inter, which, count):
        a = np.empty(count, dtype=np.int64)
        self.call("get_cell_start", inter, which, _i64ptr(a), count)
        return a

    # ---- device-only ------------------------------------------------------
    def integrator_release(self, integ):
        self.call("integrator_release", integ)

    def integrator_set_dt(self, integ, dt):
        self.call("integrator_set_dt", integ, dt)

    def sort_static(self, sys, cutoff):
        self.call("sort_static", sys, cutoff)

    def tune_set(self, param, value):
        self.call("tune_set", param, int(value))

    def tune_get(self, param):
        out = C.c_int()
        self.call("tune_get", param, C.byref(out))
        return out.value

    def integrator_status(self, integ):
        out = C.c_int()
        self.call("integrator_status", integ, C.byref(out))
        return out.value

    def step_many(self, integs, nsteps):
        a = np.ascontiguousarray(integs, dtype=np.int32)
        self.call("step_many", len(integs), _iptr(a), nsteps)

    def set_graph_mode(self, on):
        self.call("set_graph_mode", int(on))

    def set_timing(self, mode):
        """mode bit 0: per-phase HIP events; bit 1: pair counting."""
        self.call("set_timing", int(mode))

    def bench_copy(self, nbytes, reps):
        self.call("bench_copy", int(nbytes), int(reps))

    def get_timing(self):
        a = np.zeros(9)
        self.call("get_timing", _dptr(a), 9)
        return a

    def get_pair_stats(self):
        a, b = C.c_int64(), C.c_int64()
        self.call("get_pair_stats", C.byref(a), C.byref(b))
        return a.value, b.value

    def get_pair_count(self, integ):
        out = C.c_int64()
        self.call("get_pair_count", integ, C.byref(out))
        return out.value

    def ctx_info(self):
        buf = C.create_string_buffer(512)
        self.call("ctx_info", buf, 512)
        return buf.value.decode()


_HIP_LIB = None


def load_hip_library(path=None):
    """Load libgrasph_hip.so; raise if it has not been built.

    There is no fallback: without the HIP library there is no product path.
    """
    global _HIP_LIB
    if _HIP_LIB is not None and path is None:
        return _HIP_LIB
    p = Path(path or os.environ.get("GRASPH_HIP_LIB", HIP_LIB_PATH))
    if not p.exists():
        raise ImportError(
            f"{p} is missing: the HIP library has not been built "
            "(run __graft_entry__.build() or make -C grasph.jl_amd/csrc)")
    lib = C.CDLL(str(p))
    if path is None:
        _HIP_LIB = lib
    return lib


def hip_backend(ndims, precision=GR_PREC_F32, device=0, lib=None):
    """Create a device context (one HIP stream, one particle pool)."""
    lib = lib or load_hip_library()
    lib.gr_ctx_create.restype = C.c_int
    lib.gr_ctx_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_int, C.c_int]
    lib.gr_last_error.restype = C.c_char_p
    lib.gr_last_error.argtypes = [_vp]
    ctx = _vp()
    rc = lib.gr_ctx_create(C.byref(ctx), device, ndims, precision)
    if rc != GR_OK:
        msg = (lib.gr_last_error(ctx) or b"").decode(errors="replace") if ctx else ""
        raise RuntimeError(f"gr_ctx_create failed ({rc}): {msg or 'no HIP device'}")
    return CBackend(lib, "gr_", ctx, "gr_last_error", "gr_ctx_destroy",
                    extra=tuple(_GPU_SIGS))
