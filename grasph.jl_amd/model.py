"""Host-side mirror of Grasph.jl's data model and operator types.

Same names, argument meaning and errors as the reference constructors; the
arrays are numpy arrays in the reference's per-particle layout
(`Vector{SVector{ND,T}}` → shape (n, ND) float64, `Vector{T}` → (n,)).
Indices (`id`, `idx_original`, `idx_boundary`) are 1-based int64 as in Julia.

These objects carry no device state; `scene.DeviceScene` mirrors them onto a
backend context and moves arrays between host and device.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi as A

# ---------------------------------------------------------------------------
# Kernels (src/Kernels.jl)
# ---------------------------------------------------------------------------


class _Kernel:
    kind = None

    def __init__(self, h, ndims=3):
        self.h = float(h)
        self.ndims = int(ndims)

    @property
    def interaction_length(self):  # src/Kernels.jl:47-48
        return 2.0 * self.h


class CubicSplineKernel(_Kernel):
    """CubicSplineKernel(h; ndims=3) (src/Kernels.jl:34)."""

    kind = A.GR_KERNEL_CUBIC

    @property
    def norm_coeff(self):  # src/Kernels.jl:49-56
        H = self.h
        if self.ndims == 1:
            return 2.0 / (3.0 * H)
        if self.ndims == 2:
            return 10.0 / ((7.0 * math.pi) * (H * H))
        return 1.0 / (math.pi * (H * H * H))


class WenlandC2Kernel(_Kernel):
    """WenlandC2Kernel(h; ndims=3) (src/Kernels.jl:108)."""

    kind = A.GR_KERNEL_WENDLAND_C2

    @property
    def norm_coeff(self):  # src/Kernels.jl:123-128
        H = self.h
        if self.ndims == 2:
            return 7.0 / ((64.0 * math.pi) * (H * H))
        return 21.0 / ((256.0 * math.pi) * (H * H * H))


# ---------------------------------------------------------------------------
# State updaters (src/StateUpdaters.jl)
# ---------------------------------------------------------------------------


class ZeroFieldUpdater:
    """ZeroFieldUpdater(:f1, :f2, ...) (src/StateUpdaters.jl:30-42)."""

    def __init__(self, *fields):
        for f in fields:
            if f not in A.FIELD_IDS:
                raise ValueError(f"unknown field :{f}")
        self.fields = tuple(fields)

    def encode(self):
        mask = 0
        for f in self.fields:
            mask |= 1 << A.FIELD_IDS[f]
        return A.GR_UPD_ZERO, [float(mask)]


class TaitEOSUpdater:
    """TaitEOSUpdater(rho0; gamma=7) (src/StateUpdaters.jl:108-118)."""

    def __init__(self, rho0, gamma=7.0):
        self.rho0 = float(rho0)
        self.gamma = float(gamma)

    def encode(self):
        return A.GR_UPD_TAIT, [self.rho0, self.gamma]


class LinearEOSUpdater:
    """LinearEOSUpdater(rho0) (src/StateUpdaters.jl:131-140)."""

    def __init__(self, rho0):
        self.rho0 = float(rho0)

    def encode(self):
        return A.GR_UPD_LINEAR, [self.rho0]


class ElastoPlasticStressUpdater:
    """ElastoPlasticStressUpdater(E, nu, phi, psi, cohesion)
    (src/StateUpdaters.jl:202-335)."""

    def __init__(self, E, nu, phi, psi, cohesion):
        self.E, self.nu, self.phi, self.psi, self.cohesion = map(float, (E, nu, phi, psi, cohesion))

    def encode(self):
        return A.GR_UPD_EP, [self.E, self.nu, self.phi, self.psi, self.cohesion]


class HookeLawStressUpdater:
    """HookeLawStressUpdater(E, nu) (src/StateUpdaters.jl:338-405): linear
    elastic plane-strain stress update (NS = 4) with the Jaumann rate."""

    def __init__(self, E, nu):
        self.E, self.nu = float(E), float(nu)

    def encode(self):
        return A.GR_UPD_HOOKE, [self.E, self.nu]


class ViscoPlasticMCStressUpdater:
    """ViscoPlasticMCStressUpdater(eos, friction_angle, cohesion)
    (src/StateUpdaters.jl:144-189)."""

    def __init__(self, eos, friction_angle, cohesion):
        if not isinstance(eos, (TaitEOSUpdater, LinearEOSUpdater)):
            raise ValueError("ViscoPlasticMCStressUpdater needs a Tait or Linear EOS updater")
        self.eos = eos
        self.friction_angle = float(friction_angle)
        self.cohesion = float(cohesion)

    def encode(self):
        kind, prm = self.eos.encode()
        gamma = prm[1] if kind == A.GR_UPD_TAIT else 0.0
        return A.GR_UPD_VPMC, [float(kind), prm[0], gamma, self.friction_angle, self.cohesion]


class VirtualNormUpdater:
    """VirtualNormUpdater(v_mult, :f1, :f2, ...) (src/StateUpdaters.jl:58-84):
    divides each field by w_sum (zero where w_sum is 0), then v ← v·v_mult +
    prescribed_v.  Fields: v, rho, p, stress."""

    def __init__(self, v_mult, *fields):
        for f in fields:
            if f not in ("v", "rho", "p", "stress"):
                raise ValueError(f"VirtualNormUpdater field :{f} has no device implementation")
        self.v_mult = np.asarray(v_mult, dtype=np.float64).ravel()
        self.fields = tuple(fields)

    def encode(self):
        mask = 0
        for f in self.fields:
            mask |= 1 << A.FIELD_IDS[f]
        vm = list(self.v_mult) + [0.0] * (3 - len(self.v_mult))
        return A.GR_UPD_VNORM, vm[:3] + [float(mask)]


class PrescribedVelocityUpdater:
    """PrescribedVelocityUpdater() (src/StateUpdaters.jl:86-92): v ← v +
    prescribed_v of the virtual system."""

    def encode(self):
        return A.GR_UPD_PRESCRIBED_V, []


_UPDATER_TYPES = (ZeroFieldUpdater, TaitEOSUpdater, LinearEOSUpdater,
                  ElastoPlasticStressUpdater, ViscoPlasticMCStressUpdater, VirtualNormUpdater,
                  PrescribedVelocityUpdater, HookeLawStressUpdater)


def _normalise_updaters(state_updater):
    """`state_updater isa Tuple ? state_updater : (state_updater,)`
    (src/Particles.jl:87); each stage entry is None, one updater, or a tuple."""
    ups = state_updater if isinstance(state_updater, tuple) else (state_updater,)
    for u in ups:
        items = u if isinstance(u, tuple) else (u,)
        for it in items:
            if it is not None and not isinstance(it, _UPDATER_TYPES):
                raise ValueError(
                    f"state updater {type(it).__name__} has no device implementation")
    return ups


# ---------------------------------------------------------------------------
# Particle systems (src/Particles.jl)
# ---------------------------------------------------------------------------


class _ParticleSystem:
    kind = None

    def __init__(self, name, n, ndims, mass, c, *, source_v=None, source_rho=0.0,
                 state_updater=(), ns=0):
        n = int(n)
        ndims = int(ndims)
        if n <= 0:
            raise ValueError(f"n must be positive, got {n}")
        if ndims <= 0:
            raise ValueError(f"ndims must be positive, got {ndims}")
        sv = np.zeros(ndims) if source_v is None else np.asarray(source_v, dtype=np.float64)
        if sv.shape != (ndims,):
            raise ValueError(f"length(source_v) = {sv.size} must equal ndims = {ndims}")
        self.name = str(name)
        self.n = n
        self.ndims = ndims
        self.mass = float(mass)
        self.c = float(c)
        self.source_v = sv
        self.source_rho = float(source_rho)
        self.state_updater = _normalise_updaters(state_updater)
        self.ns = int(ns)
        self.id = np.arange(1, n + 1, dtype=np.int64)
        self.x = np.zeros((n, ndims))
        self.v = np.zeros((n, ndims))
        self.v_adjustment = np.zeros((n, ndims))
        self.rho = np.zeros(n)
        self.dvdt = np.zeros((n, ndims))
        self.drhodt = np.zeros(n)

    # fields moved between host and device, in the order of _particle_arrays
    # (src/Sorting.jl:15-26)
    def field_names(self):
        return ["x", "v", "rho", "dvdt", "drhodt", "v_adjustment"]


class BasicParticleSystem(_ParticleSystem):
    """BasicParticleSystem(name, n, ndims, mass, c; source_v, source_rho,
    state_updater) (src/Particles.jl:61)."""

    kind = A.GR_SYS_BASIC

    def __init__(self, name, n, ndims, mass, c, **kw):
        super().__init__(name, n, ndims, mass, c, **kw)


class FluidParticleSystem(_ParticleSystem):
    """FluidParticleSystem (src/Particles.jl:173): Basic + pressure p."""

    kind = A.GR_SYS_FLUID

    def __init__(self, name, n, ndims, mass, c, **kw):
        super().__init__(name, n, ndims, mass, c, **kw)
        self.p = np.zeros(self.n)

    def field_names(self):
        return super().field_names() + ["p"]


class StressParticleSystem(_ParticleSystem):
    """StressParticleSystem(name, n, ndims, ns, mass, c; ...)
    (src/Particles.jl:272): Fluid + stress, strain_rate (Voigt, ns ∈ {3,4,6})."""

    kind = A.GR_SYS_STRESS

    def __init__(self, name, n, ndims, ns, mass, c, **kw):
        ns = int(ns)
        if ns not in (3, 4, 6):
            raise ValueError(f"ns must be 3, 4, or 6, got {ns}")
        super().__init__(name, n, ndims, mass, c, ns=ns, **kw)
        self.p = np.zeros(self.n)
        self.stress = np.zeros((self.n, ns))
        self.strain_rate = np.zeros((self.n, ns))

    def field_names(self):
        return super().field_names() + ["p", "stress", "strain_rate"]


class ElastoPlasticParticleSystem(StressParticleSystem):
    """ElastoPlasticParticleSystem (src/Particles.jl:381): Stress + vorticity
    (scalar in 2D, 3-vector in 3D), strain, strain_p."""

    kind = A.GR_SYS_EP

    def __init__(self, name, n, ndims, ns, mass, c, **kw):
        super().__init__(name, n, ndims, ns, mass, c, **kw)
        self.vorticity = np.zeros(self.n) if self.ndims == 2 else np.zeros((self.n, 3))
        self.strain = np.zeros((self.n, self.ns))
        self.strain_p = np.zeros((self.n, self.ns))

    def field_names(self):
        return super().field_names() + ["vorticity", "strain", "strain_p"]


class VirtualParticleSystem:
    """VirtualParticleSystem(source, name, n, ndims, mass, c; state_updater,
    zero_fields, prescribed_v) (src/Particles.jl:476-511).

    Every field is the source system's (getproperty forwarding, :513-518);
    the virtual system adds w_sum, its own state updaters, the fields zeroed
    before each step's stage loop and a prescribed velocity."""

    def __init__(self, source, name, n, ndims, mass, c, state_updater=(), zero_fields=(),
                 prescribed_v=None):
        if not isinstance(source, _ParticleSystem):
            raise ValueError("source must be a particle system")
        if int(ndims) != source.ndims:
            raise ValueError(f"ndims={ndims} does not match source ndims={source.ndims}")
        if int(n) != source.n:
            raise ValueError(f"n={n} does not match source n={source.n}")
        for f in zero_fields:
            if f not in A.FIELD_IDS or f == "x" or not hasattr(source, f):
                raise ValueError(f"zero_fields: :{f} is not a field of the source")
        self.__dict__.update(
            source=source, name=str(name), state_updater=_normalise_updaters(state_updater),
            zero_fields=tuple(zero_fields), w_sum=np.zeros(source.n),
            prescribed_v=(np.zeros(source.ndims) if prescribed_v is None
                          else np.asarray(prescribed_v, dtype=np.float64).reshape(source.ndims)))

    def __getattr__(self, k):
        return getattr(self.__dict__["source"], k)

    def __setattr__(self, k, v):
        if k in self.__dict__:
            self.__dict__[k] = v
        else:
            setattr(self.__dict__["source"], k, v)

    @property
    def kind(self):
        return self.source.kind

    def field_names(self):
        return self.source.field_names()


# ---------------------------------------------------------------------------
# Boundary wrappers (src/BoundaryParticles.jl)
# ---------------------------------------------------------------------------


class StaticBoundarySystem:
    """StaticBoundarySystem(inner, lj_cutoff) (src/BoundaryParticles.jl:47)."""

    wrap = A.GR_WRAP_STATIC

    def __init__(self, inner, lj_cutoff):
        if not isinstance(inner, _ParticleSystem):
            raise ValueError("inner must be a particle system")
        self.inner = inner
        self.lj_cutoff = float(lj_cutoff)
        self.ndims = inner.ndims

    def __getattr__(self, k):  # forwarded getproperty (:13-20)
        return getattr(self.__dict__["inner"], k)


class DynamicBoundarySystem:
    """DynamicBoundarySystem(inner, normal, point, beta)
    (src/BoundaryParticles.jl:79)."""

    wrap = A.GR_WRAP_DYNAMIC

    def __init__(self, inner, normal, point, beta):
        if not isinstance(inner, _ParticleSystem):
            raise ValueError("inner must be a particle system")
        self.inner = inner
        self.boundary_normal = np.asarray(normal, dtype=np.float64).reshape(inner.ndims)
        self.boundary_point = np.asarray(point, dtype=np.float64).reshape(inner.ndims)
        self.boundary_beta = float(beta)
        self.ndims = inner.ndims

    def __getattr__(self, k):
        return getattr(self.__dict__["inner"], k)


# ---------------------------------------------------------------------------
# Ghosts (src/GhostParticles.jl)
# ---------------------------------------------------------------------------


class HouseholderReflect:
    """Marker selecting σ' = HσH, H = I − 2n̂n̂ᵀ (src/GhostParticles.jl:26)."""


class GhostCopier:
    """GhostCopier(:f, :g => HouseholderReflect(), ...) (src/GhostParticles.jl:55).

    Entries are field names ('p', 'stress') or (name, HouseholderReflect())."""

    def __init__(self, *entries):
        self.fields = []
        self.modes = []
        for e in entries:
            if isinstance(e, tuple):
                name, mode = e
                if not isinstance(mode, HouseholderReflect):
                    raise ValueError("GhostCopier mode must be HouseholderReflect()")
                self.fields.append(name)
                self.modes.append(A.GR_COPY_HOUSEHOLDER)
            else:
                self.fields.append(e)
                self.modes.append(A.GR_COPY_PLAIN)
        for f in self.fields:
            if f not in ("p", "stress"):
                raise ValueError(f"GhostCopier field :{f} has no device implementation")


class GhostParticleSystem:
    """GhostParticleSystem(ps, copiers...; name) (src/GhostParticles.jl:199).

    Host arrays are sized to the live ghost count after each download."""

    kind = A.GR_SYS_GHOST

    def __init__(self, ps, *updaters, name=None):
        if not isinstance(ps, _ParticleSystem):
            raise ValueError("ghost source must be a particle system")
        self.source = ps
        self.updaters = tuple(updaters)
        self.name = f"ghost({ps.name})" if name is None else str(name)
        self.ndims = ps.ndims
        extras = []
        for u in self.updaters:
            if u is None:
                continue
            for f in u.fields:
                if f not in extras:
                    extras.append(f)
        self.extras = tuple(extras)
        self._resize(0)

    @property
    def mass(self):
        return self.source.mass

    @property
    def c(self):
        return self.source.c

    def _resize(self, n):
        D = self.ndims
        self.n = int(n)
        self.x = np.zeros((n, D))
        self.v = np.zeros((n, D))
        self.rho = np.zeros(n)
        self.idx_original = np.zeros(n, dtype=np.int64)
        self.idx_boundary = np.zeros(n, dtype=np.int64)
        self.normals = np.zeros((n, D))
        if "p" in self.extras:
            self.p = np.zeros(n)
        if "stress" in self.extras:
            self.stress = np.zeros((n, self.source.ns))


class GhostEntry:
    """GhostEntry(ghost, cutoff, (normal, point), ...) (src/GhostParticles.jl:317)."""

    def __init__(self, ghost, cutoff, *boundary_pairs):
        if not boundary_pairs:
            raise ValueError("at least one boundary (normal, point) pair is required")
        self.ghost = ghost
        self.cutoff = float(cutoff)
        D = ghost.ndims
        self.normals = np.array([np.asarray(nm, dtype=np.float64).reshape(D) for nm, _ in boundary_pairs])
        self.points = np.array([np.asarray(pt, dtype=np.float64).reshape(D) for _, pt in boundary_pairs])


# ---------------------------------------------------------------------------
# Pair functors (src/PairwiseFunctors.jl)
# ---------------------------------------------------------------------------


def _nan_or(v):
    return math.nan if v is None else float(v)


class FluidPfn:
    """FluidPfn(alpha, beta, h; sigma=2, delta=nothing, epsilon=nothing)
    (src/PairwiseFunctors.jl:143)."""

    def __init__(self, alpha, beta, h, sigma=2, delta=None, epsilon=None):
        if sigma not in (1, 2):
            raise ValueError("FluidPfn sigma must be 1 or 2 on the device path")
        self.art_visc_alpha, self.art_visc_beta, self.h = float(alpha), float(beta), float(h)
        self.sigma, self.delta, self.epsilon = int(sigma), delta, epsilon

    def encode(self):
        return (A.GR_PFN_FLUID, self.sigma, self.art_visc_alpha, self.art_visc_beta, self.h,
                _nan_or(self.delta), _nan_or(self.epsilon))


class CauchyFluidPfn:
    """CauchyFluidPfn(alpha, beta, h; delta=nothing) (src/PairwiseFunctors.jl:258)."""

    def __init__(self, alpha, beta, h, delta=None):
        self.art_visc_alpha, self.art_visc_beta, self.h = float(alpha), float(beta), float(h)
        self.delta = delta

    def encode(self):
        return (A.GR_PFN_CAUCHY, 2, self.art_visc_alpha, self.art_visc_beta, self.h,
                _nan_or(self.delta), math.nan)


class FluidSolidPfn:
    """FluidSolidPfn(alpha, beta, h; sigma=2, delta=nothing)
    (src/PairwiseFunctors.jl:463-516): FluidPfn between a fluid system_a and
    a solid system_b with the fluid's pressure on both sides."""

    def __init__(self, alpha, beta, h, sigma=2, delta=None):
        if sigma not in (1, 2):
            raise ValueError("FluidSolidPfn sigma must be 1 or 2 on the device path")
        self.art_visc_alpha, self.art_visc_beta, self.h = float(alpha), float(beta), float(h)
        self.sigma, self.delta = int(sigma), delta

    def encode(self):
        return (A.GR_PFN_FLUIDSOLID, self.sigma, self.art_visc_alpha, self.art_visc_beta, self.h,
                _nan_or(self.delta), math.nan)


class StrainRatePfn:
    """StrainRatePfn() (src/PairwiseFunctors.jl:18)."""

    def encode(self):
        return (A.GR_PFN_SR, 2, 0.0, 0.0, 0.0, math.nan, math.nan)


class StrainRateVorticityPfn:
    """StrainRateVorticityPfn() (src/PairwiseFunctors.jl:78)."""

    def encode(self):
        return (A.GR_PFN_SRV, 2, 0.0, 0.0, 0.0, math.nan, math.nan)


class NeighborCounterPfn:
    """Adds 1 to dvdt[i][1] and drhodt[i] of both particles of every accepted
    pair — the counting functors of test/test_interaction_sweep.jl:14-30,
    expressed as an enumerable functor so the device can run it."""

    def encode(self):
        return (A.GR_PFN_COUNT, 2, 0.0, 0.0, 0.0, math.nan, math.nan)


class InterpolateFieldFn:
    """InterpolateFieldFn(:f1, :f2, ...; accumulate_wsum=true)
    (src/PairwiseFunctors.jl:405-434): the virtual side of a (real, virtual)
    interaction accumulates f += (m/ρ)_real · W · f_real (and w_sum).
    Fields: v, rho, p, stress."""

    def __init__(self, *fields, accumulate_wsum=True):
        for f in fields:
            if f not in ("v", "rho", "p", "stress"):
                raise ValueError(f"InterpolateFieldFn field :{f} has no device implementation")
        self.fields = tuple(fields)
        self.accumulate_wsum = bool(accumulate_wsum)

    def encode(self):
        mask = 0
        for f in self.fields:
            mask |= 1 << A.FIELD_IDS[f]
        return (A.GR_PFN_INTERP, mask, 1.0 if self.accumulate_wsum else 0.0, 0.0, 0.0,
                math.nan, math.nan)


class NeighborCountFn:
    """NeighborCountFn(:field) (src/PairwiseFunctors.jl:443-460): with a
    (real, ProbeParticleSystem) interaction, probe.field[j] += 1 for every
    source particle within the kernel support of probe j."""

    def __init__(self, field):
        self.field = str(field)

    def encode(self):
        return (A.GR_PFN_NCOUNT, 1 << A.GR_F_COUNT, 0.0, 0.0, 0.0, math.nan, math.nan)


_PFN_TYPES = (FluidPfn, CauchyFluidPfn, StrainRatePfn, StrainRateVorticityPfn, NeighborCounterPfn,
              InterpolateFieldFn, NeighborCountFn, FluidSolidPfn)


def encode_pfn(p):
    if p is None:
        return (A.GR_PFN_NONE, 2, 0.0, 0.0, 0.0, math.nan, math.nan)
    if not isinstance(p, _PFN_TYPES):
        raise ValueError(f"pairwise functor {type(p).__name__} has no device implementation")
    return p.encode()


# ---------------------------------------------------------------------------
# Interaction (src/Interaction.jl:61)
# ---------------------------------------------------------------------------


class XSPHPfn:
    """XSPHPfn(epsilon) (src/PairwiseFunctors.jl:360-390): velocity-adjust
    functor, v_adjustment += −ε·(v_i − v_j)·W/(ρ_i + ρ_j)·m (both sides)."""

    def __init__(self, epsilon):
        self.epsilon = float(epsilon)


# Fields a probe's extras can hold on the device: the interpolation targets
# map to the system's own arrays, any other name is a NeighborCountFn count.
PROBE_INTERP_FIELDS = ("v", "rho", "p", "stress")


class ProbeParticleSystem:
    """ProbeParticleSystem(name, positions | mirror_target; extras,
    state_updater, prescribed_v, dtype) (src/ProbeParticles.jl:22-89).

    A passive observer measured at save cadence: it owns x, id (1..n), w_sum
    and the `extras` arrays (a dict name → array of length n); any other
    attribute is forwarded to `mirror_target` when there is one.  Its
    positions advance by prescribed_v·dt every step."""

    def __init__(self, name, target, extras=None, state_updater=(), prescribed_v=None,
                 dtype=np.float64):
        extras = dict(extras or {})
        if isinstance(target, _ParticleSystem):
            mirror = target
            x = np.array(target.x, dtype=np.float64, copy=True)
        else:
            mirror = None
            pos = [np.asarray(p, dtype=dtype) for p in target]
            if not pos:
                raise ValueError("positions must not be empty")
            x = np.stack(pos).astype(dtype)
        n, nd = x.shape
        ext = {}
        for k, a in extras.items():
            a = np.asarray(a)
            if len(a) != n:
                raise ValueError(f"extras.{k} has length {len(a)} but n={n}")
            ext[k] = a
        pv = np.zeros(nd, dtype=dtype) if prescribed_v is None else np.asarray(
            prescribed_v, dtype=dtype).reshape(nd)
        self.__dict__.update(
            name=str(name), n=int(n), ndims=int(nd), x=x, id=np.arange(1, n + 1, dtype=np.int64),
            w_sum=np.zeros(n, dtype=dtype), extras=ext,
            state_updater=_normalise_updaters(state_updater), prescribed_v=pv,
            mirror_target=mirror)

    def __getattr__(self, k):  # getproperty (src/ProbeParticles.jl:95-104)
        d = self.__dict__
        if k in d.get("extras", {}):
            return d["extras"][k]
        if d.get("mirror_target") is not None:
            return getattr(d["mirror_target"], k)
        raise AttributeError(f"ProbeParticleSystem has no field {k}")

    @property
    def kind(self):
        if "stress" in self.extras:
            return A.GR_SYS_STRESS
        if "p" in self.extras:
            return A.GR_SYS_FLUID
        return A.GR_SYS_BASIC

    @property
    def ns(self):
        return int(self.extras["stress"].shape[1]) if "stress" in self.extras else 0

    def count_field(self):
        """The extras name stored in the device's count slot (None if none)."""
        other = [k for k in self.extras if k not in PROBE_INTERP_FIELDS]
        if len(other) > 1:
            raise ValueError(f"probe {self.name}: the device holds one NeighborCountFn field, "
                             f"got extras {other}")
        return other[0] if other else None

    def device_fields(self):
        """(extras name, GR_F_* id) of every extras array."""
        cf = self.count_field()
        return [(k, A.GR_F_COUNT if k == cf else A.FIELD_IDS[k]) for k in self.extras]


class SystemInteraction:
    """SystemInteraction(kernel, pairwise_fn, system_a[, system_b];
    velocity_adjust_pairwise_fn=nothing)."""

    def __init__(self, kernel, pairwise_fn, system_a, system_b=None,
                 velocity_adjust_pairwise_fn=None):
        if velocity_adjust_pairwise_fn is not None and not isinstance(
                velocity_adjust_pairwise_fn, XSPHPfn):
            raise ValueError("velocity_adjust_pairwise_fn must be an XSPHPfn on the device path")
        self.vadjust_pfn = velocity_adjust_pairwise_fn
        nd = system_a.ndims
        if kernel.ndims != nd:
            raise ValueError(f"kernel.ndims ({kernel.ndims}) must match system_a.ndims ({nd})")
        if system_b is not None and system_b.ndims != nd:
            raise ValueError(f"system_b.ndims ({system_b.ndims}) must match system_a.ndims ({nd})")
        self.kernel = kernel
        self.pfns = pairwise_fn if isinstance(pairwise_fn, tuple) else (pairwise_fn,)
        for p in self.pfns:
            encode_pfn(p)
            if isinstance(p, NeighborCountFn):
                if not isinstance(system_b, ProbeParticleSystem):
                    raise ValueError("NeighborCountFn needs a ProbeParticleSystem system_b")
                if p.field not in system_b.extras or system_b.count_field() != p.field:
                    raise ValueError(f"probe {system_b.name} has no count field :{p.field}")
            if isinstance(p, InterpolateFieldFn) and isinstance(system_b, ProbeParticleSystem):
                for f in p.fields:
                    if f not in system_b.extras:
                        raise ValueError(f"probe {system_b.name} has no extras field :{f}")
        self.system_a = system_a
        self.system_b = system_b
        self.ndims = nd

    def is_coupled(self):
        return self.system_b is not None


# ---------------------------------------------------------------------------
# Leapfrog integrator (src/TimeIntegration.jl:29-67)
# ---------------------------------------------------------------------------


class LeapFrogTimeIntegrator:
    """LeapFrogTimeIntegrator(systems, interactions; ghosts=(),
    virtual_systems=(), Γ=0)."""

    scheme = A.GR_SCHEME_LEAPFROG

    def __init__(self, systems, interactions, ghosts=(), gamma=0.0, virtual_systems=(),
                 probes=(), probe_interactions=()):
        self.probes = ((probes,) if isinstance(probes, ProbeParticleSystem) else tuple(probes))
        self.probe_interactions = ((probe_interactions,)
                                   if isinstance(probe_interactions, SystemInteraction)
                                   else tuple(probe_interactions))
        sys = (systems,) if isinstance(systems, _ParticleSystem) else tuple(systems)
        ints = (interactions,) if isinstance(interactions, SystemInteraction) else tuple(interactions)
        gsts = (ghosts,) if isinstance(ghosts, GhostEntry) else tuple(ghosts)
        vsys = ((virtual_systems,) if isinstance(virtual_systems, VirtualParticleSystem)
                else tuple(virtual_systems))
        self.virtual_systems = vsys
        if not sys:
            raise ValueError("systems must not be empty")
        if not ints:
            raise ValueError("interactions must not be empty")
        self.systems = sys
        self.interactions = ints
        self.ghosts = gsts
        self.gamma = float(gamma)
        self.c = max(ps.c for ps in sys)
        self.h = min(si.kernel.h for si in ints)
        self.ndims = sys[0].ndims


class RK4TimeIntegrator(LeapFrogTimeIntegrator):
    """RK4TimeIntegrator(systems, interactions; ghosts=(), virtual_systems=(),
    Γ=0) (src/TimeIntegration.jl:84-124): same construction as the leapfrog;
    time_integrate! runs four stage passes per step over one grid
    (:712-838)."""

    scheme = A.GR_SCHEME_RK4
