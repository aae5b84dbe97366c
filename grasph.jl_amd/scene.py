"""Mirror host objects (model.py) onto one backend context.

`DeviceScene` assigns backend handles to particle systems, ghost entries,
interactions and integrators, and moves arrays between the host mirror and
the context.  The backend is the HIP library context returned by
`_abi.hip_backend`; any object exposing the same methods (the C ABI of
include/grasph_hip.h) can be passed, which is how the test-suite runs the
same scene on its CPU oracle for comparison.
"""
from __future__ import annotations

import numpy as np

from . import _abi as A
from . import model as M


def _field_comps(ps, name):
    D = ps.ndims
    if name in ("x", "v", "dvdt", "v_adjustment", "normals"):
        return D
    if name in ("rho", "drhodt", "p"):
        return 1
    if name in ("stress", "strain_rate", "strain", "strain_p"):
        return ps.ns if not isinstance(ps, M.GhostParticleSystem) else ps.source.ns
    if name == "vorticity":
        return 1 if D == 2 else 3
    raise KeyError(name)


class DeviceScene:
    def __init__(self, backend, ndims):
        self.be = backend
        self.ndims = int(ndims)
        self._sys = {}
        self._ghost = {}
        self._inter = {}
        self._integ = {}
        self._virt = {}
        self._objs = []  # keep the mirrored objects alive (ids stay valid)

    # ---- registration -------------------------------------------------
    def system(self, ps):
        """Handle of a real particle system (boundary wrappers → inner; a
        VirtualParticleSystem → its source, registered with the virtual
        system's updaters, zero fields and prescribed velocity)."""
        if isinstance(ps, (M.StaticBoundarySystem, M.DynamicBoundarySystem)):
            ps = ps.inner
        if isinstance(ps, M.VirtualParticleSystem):
            return self._virtual(ps)
        if isinstance(ps, M.ProbeParticleSystem):
            return self._probe(ps)
        k = id(ps)
        if k in self._sys:
            return self._sys[k]
        if isinstance(ps, M.GhostParticleSystem):
            raise ValueError(f"ghost system {ps.name} must be registered through its GhostEntry")
        if ps.ndims != self.ndims:
            raise ValueError("system dimension does not match the context")
        h = self.be.system_create(ps.kind, ps.n, ps.ns, ps.mass, ps.c, ps.source_v,
                                  ps.source_rho)
        for stage, entry in enumerate(ps.state_updater, start=1):
            items = entry if isinstance(entry, tuple) else (entry,)
            if not items:
                self.be.system_add_updater(h, stage, A.GR_UPD_NOTHING, [])
            for u in items:
                if u is None:
                    self.be.system_add_updater(h, stage, A.GR_UPD_NOTHING, [])
                else:
                    kind, prm = u.encode()
                    self.be.system_add_updater(h, stage, kind, prm)
        self._sys[k] = h
        self._objs.append(ps)
        return h

    def _virtual(self, vps):
        src = vps.source
        k = id(src)
        if k in self._sys:
            if self._virt.get(k) is not vps:
                raise ValueError(f"system {src.name} is already registered as a plain system")
            return self._sys[k]
        if src.ndims != self.ndims:
            raise ValueError("system dimension does not match the context")
        h = self.be.system_create(src.kind, src.n, src.ns, src.mass, src.c, src.source_v,
                                  src.source_rho)
        self.be.virtual_create(h, [A.FIELD_IDS[f] for f in vps.zero_fields], vps.prescribed_v)
        for stage, entry in enumerate(vps.state_updater, start=1):
            items = entry if isinstance(entry, tuple) else (entry,)
            if not items:
                self.be.system_add_updater(h, stage, A.GR_UPD_NOTHING, [])
            for u in items:
                if u is None:
                    self.be.system_add_updater(h, stage, A.GR_UPD_NOTHING, [])
                else:
                    kind, prm = u.encode()
                    self.be.system_add_updater(h, stage, kind, prm)
        self._sys[k] = h
        self._virt[k] = vps
        self._objs.extend([vps, src])
        return h

    def _probe(self, pr):
        k = id(pr)
        if k in self._sys:
            return self._sys[k]
        if pr.ndims != self.ndims:
            raise ValueError("probe dimension does not match the context")
        mt = pr.mirror_target
        mirror = self.system(mt) if mt is not None else -1
        mass = mt.mass if mt is not None else 0.0
        c = mt.c if mt is not None else 0.0
        h = self.be.system_create(pr.kind, pr.n, pr.ns, mass, c, np.zeros(pr.ndims), 0.0)
        self.be.probe_create(h, pr.prescribed_v, mirror, [f for _, f in pr.device_fields()])
        for stage, entry in enumerate(pr.state_updater, start=1):
            items = entry if isinstance(entry, tuple) else (entry,)
            if not items:
                self.be.system_add_updater(h, stage, A.GR_UPD_NOTHING, [])
            for u in items:
                if u is None:
                    self.be.system_add_updater(h, stage, A.GR_UPD_NOTHING, [])
                else:
                    kind, prm = u.encode()
                    self.be.system_add_updater(h, stage, kind, prm)
        self._sys[k] = h
        self._objs.append(pr)
        return h

    def ghost(self, entry):
        """Handle of the ghost system of a GhostEntry."""
        g = entry.ghost
        k = id(g)
        if k in self._ghost:
            return self._ghost[k]
        src = self.system(g.source)
        h = self.be.ghost_create(src, entry.cutoff, entry.normals, entry.points)
        for stage, cop in enumerate(g.updaters, start=1):
            if cop is None:
                continue
            fields = [A.FIELD_IDS[f] for f in cop.fields]
            self.be.ghost_set_copier(h, stage, fields, list(cop.modes))
        self._ghost[k] = h
        self._objs.extend([entry, g])
        return h

    def ghost_handle(self, g):
        try:
            return self._ghost[id(g)]
        except KeyError:
            raise ValueError(f"ghost system {g.name} has no registered GhostEntry") from None

    def interaction(self, si):
        k = id(si)
        if k in self._inter:
            return self._inter[k]
        a = self.system(si.system_a)
        b = si.system_b
        wrap, normal, point, beta, ljcut = A.GR_WRAP_NONE, None, None, 0.0, 0.0
        if b is None:
            partner = -1
        elif isinstance(b, M.GhostParticleSystem):
            partner = self.ghost_handle(b)
        else:
            if isinstance(b, M.StaticBoundarySystem):
                wrap, ljcut = A.GR_WRAP_STATIC, b.lj_cutoff
            elif isinstance(b, M.DynamicBoundarySystem):
                wrap = A.GR_WRAP_DYNAMIC
                normal, point, beta = b.boundary_normal, b.boundary_point, b.boundary_beta
            partner = self.system(b)
        pfns = [M.encode_pfn(p) for p in si.pfns]
        h = self.be.interaction_create(si.kernel.kind, si.kernel.h, a, partner, wrap, normal,
                                       point, beta, ljcut, pfns)
        if getattr(si, "vadjust_pfn", None) is not None:
            self.be.interaction_set_vadjust(h, si.vadjust_pfn.epsilon)
        self._inter[k] = h
        self._objs.append(si)
        return h

    def integrator(self, lf):
        k = id(lf)
        if k in self._integ:
            return self._integ[k]
        sys = [self.system(ps) for ps in lf.systems]
        sys += [self.system(v) for v in getattr(lf, "virtual_systems", ())]
        gh = [self.ghost(ge) for ge in lf.ghosts]
        its = [self.interaction(si) for si in lf.interactions]
        h = self.be.integrator_create(sys, its, gh, lf.gamma)
        if getattr(lf, "scheme", A.GR_SCHEME_LEAPFROG) != A.GR_SCHEME_LEAPFROG:
            self.be.integrator_set_scheme(h, lf.scheme)
        if getattr(lf, "probes", ()):
            self.be.integrator_set_probes(h, [self._probe(p) for p in lf.probes],
                                          [self.interaction(si) for si in lf.probe_interactions])
        self._integ[k] = h
        self._objs.append(lf)
        return h

    def release(self, lf):
        """Release an integrator and everything it owns on the context
        (gr_integrator_release); its pool segments are reused by systems
        registered later.  The host objects stay usable as plain arrays."""
        h = self._integ.pop(id(lf))
        self.be.integrator_release(h)
        for ps in lf.systems:
            inner = ps.inner if isinstance(ps, (M.StaticBoundarySystem,
                                                M.DynamicBoundarySystem)) else ps
            self._sys.pop(id(inner), None)
        for v in getattr(lf, "virtual_systems", ()):
            self._sys.pop(id(v.source), None)
            self._virt.pop(id(v.source), None)
        for ge in lf.ghosts:
            self._ghost.pop(id(ge.ghost), None)
        for si in lf.interactions:
            self._inter.pop(id(si), None)
        drop = {id(lf)} | {id(ps) for ps in lf.systems} | {id(si) for si in lf.interactions}
        drop |= {id(ge) for ge in lf.ghosts} | {id(ge.ghost) for ge in lf.ghosts}
        self._objs = [o for o in self._objs if id(o) not in drop]

    def measure_probes(self, lf):
        """_measure_probes! on the context, then the probes' arrays back to
        the host (the sources were re-sorted: download them with the rest)."""
        h = self.integrator(lf)
        self.be.measure_probes(h)
        for p in lf.probes:
            self.download(p)

    # ---- host <-> context -------------------------------------------------
    def upload(self, ps, fields=None):
        if isinstance(ps, (M.StaticBoundarySystem, M.DynamicBoundarySystem)):
            ps = ps.inner
        if isinstance(ps, M.VirtualParticleSystem):
            self._virtual(ps)
            ps = ps.source
        if isinstance(ps, M.ProbeParticleSystem):
            h = self._probe(ps)
            self.be.system_upload(h, A.GR_F_X, ps.x)
            for name, fid in ps.device_fields():
                self.be.system_upload(h, fid, np.asarray(ps.extras[name], dtype=np.float64))
            self.be.system_upload_index(h, A.GR_I_ID, ps.id)
            return
        if isinstance(ps, M.GhostParticleSystem):
            h = self.ghost_handle(ps)
            names = fields or ["x", "v", "rho"] + list(ps.extras)
            for f in names:
                self.be.system_upload(h, A.FIELD_IDS[f], getattr(ps, f))
            return
        h = self.system(ps)
        for f in fields or ps.field_names():
            self.be.system_upload(h, A.FIELD_IDS[f], getattr(ps, f))
        if fields is None:
            self.be.system_upload_index(h, A.GR_I_ID, ps.id)

    def download(self, ps, fields=None):
        if isinstance(ps, (M.StaticBoundarySystem, M.DynamicBoundarySystem)):
            ps = ps.inner
        if isinstance(ps, M.VirtualParticleSystem):
            h = self._virtual(ps)
            ps.w_sum[:] = self.be.system_download(h, A.GR_F_W_SUM, ps.n)
            ps = ps.source
        if isinstance(ps, M.ProbeParticleSystem):
            h = self._probe(ps)
            ps.x[...] = self.be.system_download(h, A.GR_F_X, ps.n * ps.ndims).reshape(ps.x.shape)
            ps.w_sum[:] = self.be.system_download(h, A.GR_F_W_SUM, ps.n)
            for name, fid in ps.device_fields():
                arr = ps.extras[name]
                comps = int(np.prod(arr.shape[1:])) if arr.ndim > 1 else 1
                arr[...] = self.be.system_download(h, fid, ps.n * comps).reshape(arr.shape)
            ps.id[:] = self.be.system_download_index(h, A.GR_I_ID, ps.n)
            return
        if isinstance(ps, M.GhostParticleSystem):
            h = self.ghost_handle(ps)
            n = self.be.system_get_n(h)
            ps._resize(n)
            if n == 0:
                return
            for f in fields or ["x", "v", "rho", "normals"] + list(ps.extras):
                arr = self.be.system_download(h, A.FIELD_IDS[f], n * _field_comps(ps, f))
                getattr(ps, f)[...] = arr.reshape(getattr(ps, f).shape)
            ps.idx_original[:] = self.be.system_download_index(h, A.GR_I_IDX_ORIGINAL, n)
            ps.idx_boundary[:] = self.be.system_download_index(h, A.GR_I_IDX_BOUNDARY, n)
            return
        h = self.system(ps)
        for f in fields or ps.field_names():
            arr = self.be.system_download(h, A.FIELD_IDS[f], ps.n * _field_comps(ps, f))
            getattr(ps, f)[...] = arr.reshape(getattr(ps, f).shape)
        if fields is None:
            ps.id[:] = self.be.system_download_index(h, A.GR_I_ID, ps.n)

    def upload_all(self, lf):
        self.integrator(lf)
        for ps in (list(lf.systems) + list(getattr(lf, "virtual_systems", ()))
                   + list(getattr(lf, "probes", ()))):
            self.upload(ps)

    def download_all(self, lf):
        for ps in (list(lf.systems) + list(getattr(lf, "virtual_systems", ()))
                   + list(getattr(lf, "probes", ()))):
            self.download(ps)
        for ge in lf.ghosts:
            self.download(ge.ghost)

    # ---- grid introspection -------------------------------------------------
    def grid(self, si):
        h = self.interaction(si)
        ming, ng, ncells = self.be.get_grid(h, self.ndims)
        cs = self.be.get_cell_start(h, 0, ncells + 1)
        csa = self.be.get_cell_start(h, 1, ncells + 1) if si.system_b is not None else None
        return ming, ng, cs, csa
