"""Device-backed counterparts of the reference's entry points.

`time_integrate` is `time_integrate!(::LeapFrogTimeIntegrator, ...)`
(src/TimeIntegration.jl:580-697) with the per-step body executed by the HIP
library: the host uploads the mirror arrays, the device runs the steps, and
the host arrays are refreshed at print/save cadence and on return.  The
fine-grained functions mirror one reference call each and run it on the
device against the same host arrays.
"""
from __future__ import annotations

import time
import warnings

import numpy as np

from . import _abi as A
from . import model as M
from . import snapshot as SN
from .scene import DeviceScene

_PREC = {"f32": A.GR_PREC_F32, "f64": A.GR_PREC_F64}


class Session:
    """One device context (HIP stream + particle pool) and its scene."""

    def __init__(self, ndims, precision="f32", device=0):
        if precision not in _PREC:
            raise ValueError("precision must be 'f32' or 'f64'")
        self.backend = A.hip_backend(ndims, _PREC[precision], device)
        self.scene = DeviceScene(self.backend, ndims)
        self.precision = precision

    def close(self):
        self.backend.close()


def _session_for(obj, ndims, session, precision="f32"):
    if session is not None:
        return session
    s = getattr(obj, "_grasph_session", None)
    if s is None:
        s = Session(ndims, precision)
        obj._grasph_session = s
    return s


def _print_summary(ps, step, t):
    """Min/mean/max table of a few fields (src/Particles.jl:581-693, reduced)."""
    rows = []
    for f in ("rho", "p"):
        if hasattr(ps, f):
            a = getattr(ps, f)
            rows.append(f"  {f:>6}: min {a.min():.6g}  mean {a.mean():.6g}  max {a.max():.6g}")
    sp = np.linalg.norm(ps.v, axis=1)
    rows.append(f"  {'|v|':>6}: min {sp.min():.6g}  mean {sp.mean():.6g}  max {sp.max():.6g}")
    print(f"[{ps.name}] step {step} (t = {t:.6g})\n" + "\n".join(rows))


def time_integrate(integrator, num_timesteps, print_interval_step, save_interval_step, CFL,
                   output_prefix, *, step_offset=0, output_width=None, print_timer=True,
                   session=None, precision="f32"):
    """Run `num_timesteps` leapfrog steps on the device.

    dt = CFL·h/c (src/TimeIntegration.jl:598).  Every `save_interval_step`
    global steps an HDF5 snapshot in the reference's layout is written to
    "<output_prefix>_<padded step>.h5" (src/TimeIntegration.jl:520-549).
    Returns a dict of timings (seconds of wall time around device work, and
    steps/s).
    """
    lf = integrator
    nst = len(lf.interactions[0].pfns)
    if any(len(si.pfns) != nst for si in lf.interactions):
        raise AssertionError(
            "All interactions must have the same number of stages (pfns length), got: "
            f"{[len(si.pfns) for si in lf.interactions]}")
    for ps in lf.systems:
        n_upd = len(ps.state_updater)
        if n_upd != nst:
            warnings.warn(f'ParticleSystem "{ps.name}" has {n_upd} state updater(s) but '
                          f"num_stages={nst}; stages {n_upd + 1} and later will skip the "
                          "state update")
    s = _session_for(lf, lf.ndims, session, precision)
    sc, be = s.scene, s.backend
    h = sc.integrator(lf)
    sc.upload_all(lf)
    dt = be.integrator_dt(h, CFL)
    width = len(str(step_offset + num_timesteps)) if output_width is None else output_width
    t0 = time.perf_counter()
    done = 0
    while done < num_timesteps:
        g = step_offset + done
        nxt = num_timesteps - done
        for interval in (print_interval_step, save_interval_step if output_prefix else 0):
            if interval and interval > 0:
                nxt = min(nxt, interval - (g % interval))
        be.step(h, nxt, dt)
        done += nxt
        g = step_offset + done
        printing = print_interval_step > 0 and g % print_interval_step == 0
        saving = output_prefix is not None and save_interval_step > 0 and g % save_interval_step == 0
        if saving and getattr(lf, "probes", ()):
            sc.measure_probes(lf)  # _maybe_save! (src/TimeIntegration.jl:525-526)
        if printing or saving:
            sc.download_all(lf)
            if printing:
                for ps in lf.systems:
                    _print_summary(ps, g, g * dt)
            if saving:
                SN.save_snapshot(SN.snapshot_path(output_prefix, g, width), lf.systems,
                                 lf.ghosts, g, dt, lf.virtual_systems,
                                 getattr(lf, "probes", ()))
    sc.download_all(lf)
    el = time.perf_counter() - t0
    res = {"seconds": el, "steps": num_timesteps, "dt": dt,
           "steps_per_s": num_timesteps / el if el > 0 else float("inf")}
    if print_timer:
        print(f"time_integrate: {num_timesteps} steps in {el:.4g} s "
              f"({res['steps_per_s']:.4g} steps/s, dt = {dt:.6g})")
    return res


# ---------------------------------------------------------------------------
# Fine-grained drop-ins: each runs one reference call on the device against
# the host arrays of the objects involved.
# ---------------------------------------------------------------------------


def sort_particles(ps, cutoff, session=None):
    """sort_particles!(ps, cutoff, ...) (src/Sorting.jl:157)."""
    s = _session_for(ps, ps.ndims, session)
    s.scene.upload(ps)
    s.backend.sort(s.scene.system(ps), float(cutoff))
    s.scene.download(ps)


def create_grid(si, session=None):
    """create_grid!(si) (src/Interaction.jl:123); returns
    (mingridx, ngridx, cell_start, cell_start_a) with 1-based cell starts."""
    s = _session_for(si, si.ndims, session)
    for ps in (si.system_a, si.system_b):
        if ps is not None and not isinstance(ps, M.GhostParticleSystem):
            s.scene.upload(ps, ["x"])
    s.backend.create_grid(s.scene.interaction(si))
    return s.scene.grid(si)


def sweep(si, stage=1, session=None):
    """sweep!(si, stage) (src/Interaction.jl:292) on the device; the target
    fields of the host arrays are accumulated into, as in the reference."""
    s = _session_for(si, si.ndims, session)
    systems = [si.system_a] + ([si.system_b] if si.system_b is not None else [])
    for ps in systems:
        if not isinstance(ps, M.GhostParticleSystem):
            s.scene.upload(ps)
    h = s.scene.interaction(si)
    s.backend.create_grid(h)
    s.backend.sweep(h, int(stage))
    for ps in systems:
        if not isinstance(ps, M.GhostParticleSystem):
            s.scene.download(ps)


def adjust_v(si, session=None):
    """adjust_v!(si) (src/Interaction.jl:309) on the device: the interaction's
    velocity_adjust_pairwise_fn (XSPHPfn) accumulated into v_adjustment.  With
    a ghost partner the ghost's share lands in its source's array at the
    ghost's index, as the reference's forwarded field does."""
    s = _session_for(si, si.ndims, session)
    systems = [si.system_a] + ([si.system_b] if si.system_b is not None else [])
    real = [ps.source if isinstance(ps, M.GhostParticleSystem) else ps for ps in systems]
    for ps in real:
        s.scene.upload(ps)
    h = s.scene.interaction(si)
    s.backend.create_grid(h)
    s.backend.adjust_v(h)
    for ps in real:
        s.scene.download(ps)


def update_state(ps, stage=1, dt=0.0, session=None):
    """update_state!(ps, stage, dt) (src/Particles.jl:542)."""
    s = _session_for(ps, ps.ndims, session)
    s.scene.upload(ps)
    s.backend.update_state(s.scene.system(ps), int(stage), float(dt))
    s.scene.download(ps)
