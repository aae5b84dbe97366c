"""HDF5 snapshots and restart in the reference's file layout.

Writer: `_maybe_save!` (src/TimeIntegration.jl:520-549) with
`_write_to_group!` (src/Particles.jl:721-735) and the ghost writer
(src/GhostParticles.jl:434-459):

    file attrs   step::Int64, sim_time::Float64
    /<ps.name>   attrs n::Int64, ndims::Int64, mass::Float64, c::Float64
                 x (n × ndims) and the fields of _sim_field_names
                 (src/Particles.jl:699-702), vectors as n × components,
                 scalars as n, id as Int64
    /<ghost>     attrs n, ndims, mass; when n > 0: x, v, rho, idx_original,
                 idx_boundary and the copier's extra fields

Julia's HDF5.jl stores a Julia (components × n) array as a C-order (n ×
components) dataset, which is how the host mirror already holds its arrays,
so rows are particles here too.  Reader: `read_h5!` (src/Particles.jl:
747-787: n and ndims must match, ArgumentError otherwise; x and every field
present are read) and `_load_restart!` (src/Driver.jl:79-96).

The HDF5 C library is used directly through ctypes (no Python HDF5 package
is required); GRASPH_HDF5_LIB names it when it is not on the loader path.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
from pathlib import Path

import numpy as np

from . import model as M

_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_uint64
_P_DEFAULT = 0
_S_ALL = 0
_S_SCALAR = 0
_F_ACC_RDONLY = 0
_F_ACC_TRUNC = 2
_T_INTEGER = 0
_T_FLOAT = 1

_H5 = None


def _lib():
    global _H5
    if _H5 is not None:
        return _H5
    cands = [os.environ.get("GRASPH_HDF5_LIB"), ctypes.util.find_library("hdf5"),
             "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            lib = C.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise ImportError(f"the HDF5 C library was not found ({err}); set GRASPH_HDF5_LIB")
    sig = {
        "H5open": (_herr, []),
        "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
        "H5Gopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Gclose": (_herr, [_hid]),
        "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.c_void_p]),
        "H5Screate": (_hid, [C.c_int]),
        "H5Sclose": (_herr, [_hid]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.c_void_p]),
        "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dget_type": (_hid, [_hid]),
        "H5Dclose": (_herr, [_hid]),
        "H5Tget_class": (C.c_int, [_hid]),
        "H5Tclose": (_herr, [_hid]),
        "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
        "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Awrite": (_herr, [_hid, _hid, C.c_void_p]),
        "H5Aread": (_herr, [_hid, _hid, C.c_void_p]),
        "H5Aexists": (C.c_int, [_hid, C.c_char_p]),
        "H5Aclose": (_herr, [_hid]),
        "H5Lexists": (C.c_int, [_hid, C.c_char_p, _hid]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.H5open() < 0:
        raise RuntimeError("H5open failed")
    lib.T_F64 = _hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    lib.T_I64 = _hid.in_dll(lib, "H5T_NATIVE_INT64_g").value
    _H5 = lib
    return lib


def _ok(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def _bytes(s):
    return str(s).encode()


# ---- attributes and datasets --------------------------------------------


def _attr_write(loc, name, value):
    h = _lib()
    sp = _ok(h.H5Screate(_S_SCALAR), "H5Screate")
    if isinstance(value, (int, np.integer)):
        t, buf = h.T_I64, C.c_int64(int(value))
    else:
        t, buf = h.T_F64, C.c_double(float(value))
    a = _ok(h.H5Acreate2(loc, _bytes(name), t, sp, _P_DEFAULT, _P_DEFAULT), f"attr {name}")
    _ok(h.H5Awrite(a, t, C.byref(buf)), f"write attr {name}")
    h.H5Aclose(a)
    h.H5Sclose(sp)


def _attr_read(loc, name, kind):
    h = _lib()
    if h.H5Aexists(loc, _bytes(name)) <= 0:
        return None
    a = _ok(h.H5Aopen(loc, _bytes(name), _P_DEFAULT), f"open attr {name}")
    if kind is int:
        buf = C.c_int64()
        _ok(h.H5Aread(a, h.T_I64, C.byref(buf)), f"read attr {name}")
    else:
        buf = C.c_double()
        _ok(h.H5Aread(a, h.T_F64, C.byref(buf)), f"read attr {name}")
    h.H5Aclose(a)
    return buf.value


def _dset_write(loc, name, arr):
    h = _lib()
    arr = np.ascontiguousarray(arr)
    if arr.dtype.kind in "iu":
        arr = arr.astype(np.int64)
        t = h.T_I64
    else:
        arr = arr.astype(np.float64)
        t = h.T_F64
    dims = (_hsize * arr.ndim)(*arr.shape)
    sp = _ok(h.H5Screate_simple(arr.ndim, dims, None), "H5Screate_simple")
    d = _ok(h.H5Dcreate2(loc, _bytes(name), t, sp, _P_DEFAULT, _P_DEFAULT, _P_DEFAULT),
            f"dataset {name}")
    _ok(h.H5Dwrite(d, t, _S_ALL, _S_ALL, _P_DEFAULT, arr.ctypes.data), f"write {name}")
    h.H5Dclose(d)
    h.H5Sclose(sp)


def _exists(loc, name):
    return _lib().H5Lexists(loc, _bytes(name), _P_DEFAULT) > 0


def _dset_read(loc, name):
    h = _lib()
    d = _ok(h.H5Dopen2(loc, _bytes(name), _P_DEFAULT), f"open {name}")
    sp = h.H5Dget_space(d)
    nd = h.H5Sget_simple_extent_ndims(sp)
    dims = (_hsize * max(nd, 1))()
    h.H5Sget_simple_extent_dims(sp, dims, None)
    shape = tuple(int(dims[k]) for k in range(nd))
    ty = h.H5Dget_type(d)
    is_int = h.H5Tget_class(ty) == _T_INTEGER
    h.H5Tclose(ty)
    out = np.empty(shape, dtype=np.int64 if is_int else np.float64)
    _ok(h.H5Dread(d, h.T_I64 if is_int else h.T_F64, _S_ALL, _S_ALL, _P_DEFAULT,
                  out.ctypes.data), f"read {name}")
    h.H5Sclose(sp)
    h.H5Dclose(d)
    return out


# ---- particle systems -----------------------------------------------------


def sim_field_names(ps):
    """_sim_field_names (src/Particles.jl:699-702)."""
    names = ["id", "v", "rho", "dvdt", "drhodt"]
    if isinstance(ps, M.FluidParticleSystem) or isinstance(ps, M.StressParticleSystem):
        names.append("p")
    if isinstance(ps, M.StressParticleSystem):
        names += ["stress", "strain_rate"]
    if isinstance(ps, M.ElastoPlasticParticleSystem):
        names += ["vorticity", "strain", "strain_p"]
    return names


def _inner(ps):
    if isinstance(ps, (M.StaticBoundarySystem, M.DynamicBoundarySystem)):
        return ps.inner
    if isinstance(ps, M.VirtualParticleSystem):  # fields forwarded to the source
        return ps.source
    return ps


def write_group(loc, ps):
    """_write_to_group! (src/Particles.jl:721-735) / ghost write_h5
    (src/GhostParticles.jl:434-459) / probe write_h5 (src/ProbeParticles.jl:
    166-185: attrs n, ndims; x, id and the extras, no w_sum)."""
    ps = _inner(ps)
    if isinstance(ps, M.ProbeParticleSystem):
        _attr_write(loc, "n", ps.n)
        _attr_write(loc, "ndims", ps.ndims)
        if ps.n > 0:
            _dset_write(loc, "x", ps.x)
            _dset_write(loc, "id", ps.id)
            for k, a in ps.extras.items():
                _dset_write(loc, k, a)
        return
    if isinstance(ps, M.GhostParticleSystem):
        _attr_write(loc, "n", ps.n)
        _attr_write(loc, "ndims", ps.ndims)
        _attr_write(loc, "mass", ps.mass)
        if ps.n > 0:
            for f in ("x", "v", "rho", "idx_original", "idx_boundary") + tuple(ps.extras):
                _dset_write(loc, f, getattr(ps, f))
        return
    _attr_write(loc, "n", ps.n)
    _attr_write(loc, "ndims", ps.ndims)
    _attr_write(loc, "mass", ps.mass)
    _attr_write(loc, "c", ps.c)
    _dset_write(loc, "x", ps.x)
    for f in sim_field_names(ps):
        _dset_write(loc, f, getattr(ps, f))


def read_group(loc, ps):
    """_read_from_group! (src/Particles.jl:759-787)."""
    ps = _inner(ps)
    if isinstance(ps, M.ProbeParticleSystem):  # read_h5! (src/ProbeParticles.jl:187-204)
        n = int(_attr_read(loc, "n", int))
        if n != ps.n:
            raise ValueError(f"Saved probe count ({n}) does not match probe.n={ps.n}")
        if _exists(loc, "x"):
            ps.x[...] = _dset_read(loc, "x").reshape(ps.x.shape)
        ps.id[:] = np.arange(1, ps.n + 1)
        for k, a in ps.extras.items():
            if _exists(loc, k):
                a[...] = _dset_read(loc, k).reshape(a.shape)
        return ps
    n = int(_attr_read(loc, "n", int))
    nd = int(_attr_read(loc, "ndims", int))
    if n != ps.n:
        raise ValueError(f"Saved particle count ({n}) does not match system count ({ps.n})")
    if nd != ps.ndims:
        raise ValueError(f"Saved ndims ({nd}) does not match system ndims ({ps.ndims})")
    for f in ["x"] + sim_field_names(ps):
        if _exists(loc, f):
            arr = getattr(ps, f)
            arr[...] = _dset_read(loc, f).reshape(arr.shape)
    return ps


def write_h5(ps, path):
    """write_h5(ps, path) (src/Particles.jl:711-715): one system, file mode "w"."""
    h = _lib()
    f = _ok(h.H5Fcreate(_bytes(path), _F_ACC_TRUNC, _P_DEFAULT, _P_DEFAULT), f"create {path}")
    try:
        write_group(f, ps)
    finally:
        h.H5Fclose(f)


def read_h5(ps, path):
    """read_h5!(ps, path) (src/Particles.jl:747-752)."""
    h = _lib()
    f = _ok(h.H5Fopen(_bytes(path), _F_ACC_RDONLY, _P_DEFAULT), f"open {path}")
    try:
        return read_group(f, ps)
    finally:
        h.H5Fclose(f)


def snapshot_path(prefix, step, width):
    """"$(prefix)_$(lpad(step, width, '0')).h5" (src/TimeIntegration.jl:528)."""
    return f"{prefix}_{str(step).rjust(width, '0')}.h5"


def save_snapshot(path, systems, ghosts, step, dt, virtual_systems=(), probes=()):
    """The body of _maybe_save! (src/TimeIntegration.jl:526-547): file attrs
    step and sim_time, one group per system, then one per ghost system, one
    per virtual system (named after the virtual system), one per probe."""
    d = os.path.dirname(str(path))
    if d:
        Path(d).mkdir(parents=True, exist_ok=True)
    h = _lib()
    f = _ok(h.H5Fcreate(_bytes(path), _F_ACC_TRUNC, _P_DEFAULT, _P_DEFAULT), f"create {path}")
    try:
        _attr_write(f, "step", int(step))
        _attr_write(f, "sim_time", float(step * dt))
        for ps in (list(systems) + [ge.ghost for ge in ghosts] + list(virtual_systems)
                   + list(probes)):
            name = ps.name
            ps = _inner(ps)
            g = _ok(h.H5Gcreate2(f, _bytes(name), _P_DEFAULT, _P_DEFAULT, _P_DEFAULT),
                    f"group {name}")
            try:
                write_group(g, ps)
            finally:
                h.H5Gclose(g)
    finally:
        h.H5Fclose(f)


def load_restart(path, systems):
    """_load_restart! (src/Driver.jl:79-96): returns the saved step (0 if the
    file has none) and reads every system whose group is present."""
    h = _lib()
    f = _ok(h.H5Fopen(_bytes(path), _F_ACC_RDONLY, _P_DEFAULT), f"open {path}")
    try:
        step = _attr_read(f, "step", int)
        for ps in systems:
            name = ps.name
            ps = _inner(ps)
            if _exists(f, name):
                g = _ok(h.H5Gopen2(f, _bytes(name), _P_DEFAULT), f"group {name}")
                try:
                    read_group(g, ps)
                finally:
                    h.H5Gclose(g)
        return int(step) if step is not None else 0
    finally:
        h.H5Fclose(f)
