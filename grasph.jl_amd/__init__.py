"""grasph_amd — MI355X (gfx950) particle step for Grasph.jl's SPH model.

Host mirror of the reference's Julia API (particle systems, kernels, pair
functors, state updaters, ghosts, boundary wrappers, SystemInteraction,
LeapFrogTimeIntegrator, time_integrate!) over the C ABI of
include/grasph_hip.h, whose HIP kernels run the step on the GPU.
"""
from . import _abi  # noqa: F401
from .model import (  # noqa: F401
    BasicParticleSystem, CauchyFluidPfn, CubicSplineKernel, DynamicBoundarySystem,
    ElastoPlasticParticleSystem, ElastoPlasticStressUpdater, FluidParticleSystem, FluidPfn,
    GhostCopier, GhostEntry, GhostParticleSystem, HouseholderReflect, InterpolateFieldFn,
    LeapFrogTimeIntegrator, LinearEOSUpdater, RK4TimeIntegrator, NeighborCounterPfn,
    NeighborCountFn, PrescribedVelocityUpdater, ProbeParticleSystem, FluidSolidPfn,
    HookeLawStressUpdater,
    StaticBoundarySystem, StrainRatePfn, StrainRateVorticityPfn, StressParticleSystem,
    SystemInteraction, TaitEOSUpdater, ViscoPlasticMCStressUpdater, VirtualNormUpdater,
    VirtualParticleSystem, WenlandC2Kernel, XSPHPfn, ZeroFieldUpdater)
from .scene import DeviceScene  # noqa: F401
from .integration import (  # noqa: F401
    Session, adjust_v, create_grid, sort_particles, sweep, time_integrate, update_state)
from . import configs  # noqa: F401
from . import serving  # noqa: F401
from . import snapshot  # noqa: F401
from . import autotune  # noqa: F401
