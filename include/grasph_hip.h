/*
 * grasph_hip.h — C ABI of the MI355X (gfx950) particle-step library.
 *
 * The library replaces the execution of Grasph.jl's particle step
 * (sort → ghosts → grid → half-step → staged updaters/sweeps → full step →
 * positions) with HIP kernels.  The caller keeps its particle systems on the
 * host; the library keeps a device copy in a packed structure-of-arrays pool
 * and advances it.  Every entry point returns an int status (GR_OK on
 * success); gr_last_error() returns a message.  No C++ exception crosses the
 * ABI.  Host buffers passed in are borrowed for the duration of the call.
 *
 * Host array layout: the reference's `reinterpret(reshape, T, Vector{SVector})`
 * layout, i.e. interleaved AoS (x1,y1,z1,x2,...), doubles; index arrays are
 * int64 and 1-based, as in Julia.
 *
 * A context is not thread-safe.  One context = one HIP stream = one pool of a
 * single spatial dimension; it can hold many independent simulations
 * ("integrators"), which gr_step_many advances together in one launch per
 * phase (packed, offset-indexed pool).
 *
 * Reference interfaces each entry point replaces are cited per function
 * (paths relative to the reference repository root).
 */
#ifndef GRASPH_HIP_H
#define GRASPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
  GR_OK = 0,
  GR_ERR_ARGUMENT = 1,   /* reference: ArgumentError                     */
  GR_ERR_ASSERTION = 2,  /* reference: AssertionError (@assert)          */
  GR_ERR_HIP = 3,        /* HIP runtime failure                           */
  GR_ERR_CAPACITY = 4,   /* a grid/key range outgrew its device capacity  */
  GR_ERR_STATE = 5,      /* call made in the wrong state                  */
  GR_ERR_FAULT = 6       /* gr_step / gr_step_many: some integrators
                            faulted (gr_integrator_status); the others
                            advanced all their steps normally             */
};

/* ---- per-integrator fault bits (gr_integrator_status) ------------------ */
enum {
  GR_FAULT_NONFINITE = 1,  /* a position became NaN or infinite            */
  GR_FAULT_DOMAIN = 2      /* particles spread beyond the cell-key range or
                              the planned cell-grid capacity               */
};

/* ---- arithmetic precision of the pair sums (gr_ctx_create) ------------ */
enum {
  GR_PREC_F32 = 0,  /* FP32 pair math, FP64 sums of the FP32 pair terms;  */
                    /* FP64 x, v, rho, stress                             */
  GR_PREC_F64 = 1   /* FP64 everywhere                                    */
};

/* ---- particle system kinds (src/Particles.jl:38,149,245,351) ---------- */
enum {
  GR_SYS_BASIC = 0,   /* BasicParticleSystem                              */
  GR_SYS_FLUID = 1,   /* FluidParticleSystem (+ p)                        */
  GR_SYS_STRESS = 2,  /* StressParticleSystem (+ p, stress, strain_rate)  */
  GR_SYS_EP = 3,      /* ElastoPlasticParticleSystem (+ vorticity, strain, strain_p) */
  GR_SYS_GHOST = 4    /* GhostParticleSystem (src/GhostParticles.jl:174)  */
};

/* ---- per-particle fields (double arrays) ------------------------------ */
enum {
  GR_F_X = 0, GR_F_V = 1, GR_F_RHO = 2, GR_F_DVDT = 3, GR_F_DRHODT = 4,
  GR_F_P = 5, GR_F_STRESS = 6, GR_F_STRAIN_RATE = 7, GR_F_VORTICITY = 8,
  GR_F_STRAIN = 9, GR_F_STRAIN_P = 10, GR_F_V_ADJUSTMENT = 11,
  GR_F_NORMALS = 12,
  GR_F_W_SUM = 13,  /* VirtualParticleSystem / probe w_sum (download only) */
  GR_F_COUNT = 14,  /* probe neighbour count (NeighborCountFn's target field) */
  GR_NUM_FIELDS = 15
};
/* ---- per-particle index fields (int64, 1-based) ----------------------- */
enum { GR_I_ID = 0, GR_I_IDX_ORIGINAL = 1, GR_I_IDX_BOUNDARY = 2 };

/* ---- SPH kernels (src/Kernels.jl:18,92) -------------------------------- */
enum { GR_KERNEL_CUBIC = 0, GR_KERNEL_WENDLAND_C2 = 1 };

/* ---- pair functors (src/PairwiseFunctors.jl) -------------------------- */
enum {
  GR_PFN_NONE = 0,    /* `nothing` stage                                   */
  GR_PFN_FLUID = 1,   /* FluidPfn :136                                     */
  GR_PFN_CAUCHY = 2,  /* CauchyFluidPfn :252                               */
  GR_PFN_SR = 3,      /* StrainRatePfn :18                                 */
  GR_PFN_SRV = 4,     /* StrainRateVorticityPfn :78                        */
  GR_PFN_COUNT = 5,   /* neighbour counter: +1 to dvdt[1] and drhodt per pair,
                         both sides (the counting functors of
                         test/test_interaction_sweep.jl:14-30)              */
  GR_PFN_INTERP = 6,  /* InterpolateFieldFn :405-434 between a real system
                         and a virtual one: the virtual side accumulates
                         f += (m/ρ)·W·f_real; `sigma` holds the field bit
                         mask (1<<GR_F_V | 1<<GR_F_RHO | 1<<GR_F_P |
                         1<<GR_F_STRESS), `alpha` != 0 accumulates w_sum  */
  GR_PFN_XSPH = 7,    /* XSPHPfn :360-390, only as an interaction's
                         velocity-adjust functor (gr_interaction_set_vadjust) */
  GR_PFN_NCOUNT = 8,  /* NeighborCountFn :443-460: system_b (a probe) counts
                         its system_a partners into GR_F_COUNT              */
  GR_PFN_FLUIDSOLID = 9 /* FluidSolidPfn :463-516: FluidPfn between a fluid
                         system_a and a plain solid system_b with the fluid's
                         pressure on both sides (sigma, alpha, beta, h,
                         delta as FluidPfn; no epsilon)                    */
};

/* ---- partner wrappers (src/BoundaryParticles.jl:38,68) ---------------- */
enum { GR_WRAP_NONE = 0, GR_WRAP_STATIC = 1, GR_WRAP_DYNAMIC = 2 };

/* ---- state updaters (src/StateUpdaters.jl) ----------------------------- */
enum {
  GR_UPD_NOTHING = 0,  /* `nothing` stage entry                            */
  GR_UPD_ZERO = 1,     /* ZeroFieldUpdater :30; params[0] = field bit mask
                          (1<<GR_F_*)                                       */
  GR_UPD_TAIT = 2,     /* TaitEOSUpdater :108; params = {rho0, gamma}      */
  GR_UPD_LINEAR = 3,   /* LinearEOSUpdater :131; params = {rho0}           */
  GR_UPD_EP = 4,       /* ElastoPlasticStressUpdater :202;
                          params = {E, nu, phi, psi, cohesion}             */
  GR_UPD_VPMC = 5,     /* ViscoPlasticMCStressUpdater :144;
                          params = {eos_kind(GR_UPD_TAIT|LINEAR), rho0,
                                    gamma, phi, cohesion}                  */
  GR_UPD_VNORM = 6,    /* VirtualNormUpdater :58-77 (virtual systems);
                          params = {v_mult[0], v_mult[1], v_mult[2],
                                    field bit mask}                        */
  GR_UPD_PRESCRIBED_V = 7, /* PrescribedVelocityUpdater :88-92; no params */
  GR_UPD_HOOKE = 8     /* HookeLawStressUpdater :338-405 (ElastoPlastic,
                          NS = 4); params = {E, nu}: elastic increment with
                          the Jaumann rate, p = −tr(σ)/3                    */
};

/* ---- ghost copier modes (src/GhostParticles.jl:26,53) ------------------ */
enum { GR_COPY_PLAIN = 0, GR_COPY_HOUSEHOLDER = 1 };

/* Pair-functor descriptor: one per interaction stage.
 * delta / epsilon: NaN means `nothing` (FluidPfn :143, CauchyFluidPfn :258). */
typedef struct gr_pfn_desc {
  int32_t kind;     /* GR_PFN_*                         */
  int32_t sigma;    /* FluidPfn S (1 or 2)              */
  double alpha;     /* art_visc_alpha                   */
  double beta;      /* art_visc_beta                    */
  double h;         /* functor h (artificial viscosity) */
  double delta;     /* δ-SPH coefficient or NaN         */
  double epsilon;   /* surface-tension ε or NaN         */
} gr_pfn_desc;

typedef struct gr_ctx gr_ctx;

/* ---- context -----------------------------------------------------------
 * One HIP stream and one packed particle pool of spatial dimension `ndims`
 * (2 or 3) on HIP device `device`. */
int gr_ctx_create(gr_ctx **out, int device, int ndims, int precision);
int gr_ctx_destroy(gr_ctx *ctx);
const char *gr_last_error(const gr_ctx *ctx);
/* Device and build description (for logs). */
int gr_ctx_info(gr_ctx *ctx, char *buf, int64_t buflen);

/* ---- particle systems --------------------------------------------------
 * Replaces the constructors BasicParticleSystem / FluidParticleSystem /
 * StressParticleSystem / ElastoPlasticParticleSystem
 * (src/Particles.jl:61,173,272,381): n > 0, ns ∈ {3,4,6} for stress kinds,
 * source_v has ndims entries.  Fields start as the reference's do: x, v, rho
 * must be uploaded; everything else is zero; id = 1..n. */
int gr_system_create(gr_ctx *ctx, int kind, int64_t n, int ns, double mass,
                     double c, const double *source_v, double source_rho,
                     int *sys_out);
/* Register the updater at `stage` (1-based).  Several calls for one stage
 * form the tuple applied in order (src/Particles.jl:562-565).  The number of
 * updater stages is the largest stage registered (GR_UPD_NOTHING registers
 * a `nothing` stage), matching `length(ps.state_updater)`. */
int gr_system_add_updater(gr_ctx *ctx, int sys, int stage, int kind,
                          const double *params, int nparams);
/* Host → device copy of one field (count = n × components).  For ghost
 * systems only GR_F_X/V/RHO/P/STRESS of live ghosts are accepted. */
int gr_system_upload(gr_ctx *ctx, int sys, int field, const double *data,
                     int64_t count);
int gr_system_download(gr_ctx *ctx, int sys, int field, double *data,
                       int64_t count);
int gr_system_upload_index(gr_ctx *ctx, int sys, int field,
                           const int64_t *data, int64_t count);
int gr_system_download_index(gr_ctx *ctx, int sys, int field, int64_t *data,
                             int64_t count);
/* Current particle count (ghost systems: the count after the last
 * generate_ghosts; reference `ghost.n`, src/GhostParticles.jl:228). */
int gr_system_get_n(gr_ctx *ctx, int sys, int64_t *n_out);

/* Turn system `sys` into the source of a VirtualParticleSystem
 * (src/Particles.jl:488-511): it is sorted every step but not integrated,
 * `zero_fields` (GR_F_* ids) and its w_sum are zeroed before each step's
 * stage loop, updaters registered on it are the virtual system's (with
 * GR_UPD_VNORM / GR_UPD_PRESCRIBED_V), its positions advance by
 * prescribed_v·dt (ndims doubles, may be NULL for zero), and as an
 * interaction partner it is one-sided (src/PairwiseFunctors.jl:286,97,193).
 * Pass it in gr_integrator_create's `systems`; it is excluded from dt's c. */
/* Turn system `sys` into a ProbeParticleSystem (src/ProbeParticles.jl:
 * 22-89): a target of InterpolateFieldFn / NeighborCountFn interactions
 * that is measured only by gr_measure_probes.  `zero_fields` are its
 * extras (zeroed with w_sum before each measurement, auto_zero_probe!
 * :141-148); GR_F_COUNT holds a NeighborCountFn field.  `mirror` >= 0: the
 * real system whose positions the probe takes at each measurement
 * (mirror_target; same n).  Positions advance by prescribed_v·dt every step
 * (_advance_probe_positions!, src/TimeIntegration.jl:439-450). */
int gr_probe_create(gr_ctx *ctx, int sys, const double *prescribed_v, int mirror,
                    int nzero, const int *zero_fields);
int gr_virtual_create(gr_ctx *ctx, int sys, int nzero, const int *zero_fields,
                      const double *prescribed_v);

/* ---- ghosts --------------------------------------------------------------
 * GhostParticleSystem(ps, copiers...) + GhostEntry(ghost, cutoff,
 * (normal, point)...) (src/GhostParticles.jl:199,317).  normals/points hold
 * nb × ndims doubles. */
int gr_ghost_create(gr_ctx *ctx, int source_sys, double cutoff, int nb,
                    const double *normals, const double *points,
                    int *ghost_out);
/* GhostCopier(fields... [=> HouseholderReflect()]) registered at `stage`
 * (src/GhostParticles.jl:55).  fields ∈ {GR_F_P, GR_F_STRESS}. */
int gr_ghost_set_copier(gr_ctx *ctx, int ghost, int stage, int nfields,
                        const int *fields, const int *modes);

/* ---- interactions --------------------------------------------------------
 * SystemInteraction(kernel, pfns, system_a[, system_b]) (src/Interaction.jl:61).
 * partner < 0: self interaction.  Otherwise partner is a system or ghost id;
 * wrap selects StaticBoundarySystem(inner, lj_cutoff) or
 * DynamicBoundarySystem(inner, normal, point, beta) around a real partner
 * (src/BoundaryParticles.jl:47,79).  normal/point hold ndims doubles (may be
 * NULL for GR_WRAP_NONE/STATIC). */
int gr_interaction_create(gr_ctx *ctx, int kernel_kind, double h, int sys_a,
                          int partner, int wrap, const double *normal,
                          const double *point, double beta, double lj_cutoff,
                          int n_stages, const gr_pfn_desc *pfns,
                          int *inter_out);

/* Velocity-adjust functor of an interaction: XSPHPfn(epsilon)
 * (SystemInteraction's velocity_adjust_pairwise_fn, src/Interaction.jl:61-66,
 * applied by _xsph_correction!, src/TimeIntegration.jl:423-436, after the
 * full step and before positions move).  NaN removes it.  With a ghost
 * partner the reference writes through the ghost's forwarded v_adjustment,
 * i.e. into the source system at the ghost's own index
 * (src/GhostParticles.jl:226-236); this is reproduced, and a step in which
 * the ghosts outnumber the source particles (an out-of-bounds write in the
 * reference) fails with GR_ERR_ARGUMENT. */
int gr_interaction_set_vadjust(gr_ctx *ctx, int inter, double epsilon);

/* ---- integrator ----------------------------------------------------------
 * LeapFrogTimeIntegrator(systems, interactions; ghosts, Γ)
 * (src/TimeIntegration.jl:53).  Validates as the reference does and fixes
 * the per-step plan.  A context may hold many integrators (an ensemble). */
int gr_integrator_create(gr_ctx *ctx, int nsys, const int *systems,
                         int ninter, const int *inters, int nghost,
                         const int *ghosts, double gamma, int *integ_out);
/* Release an integrator with its systems, ghosts and interactions: their
 * pool segments and cell arrays are reused by systems created later (no
 * counterpart in the reference, where Julia's GC frees a finished
 * simulation; a serving loop that admits and retires simulations needs it).
 * Handles of released objects are rejected with GR_ERR_STATE. */
int gr_integrator_release(gr_ctx *ctx, int integ);
/* Time integration scheme of an integrator: GR_SCHEME_LEAPFROG (default,
 * LeapFrogTimeIntegrator) or GR_SCHEME_RK4 (RK4TimeIntegrator,
 * src/TimeIntegration.jl:84-124,712-838: one grid per step, four stage
 * passes). */
enum { GR_SCHEME_LEAPFROG = 0, GR_SCHEME_RK4 = 1 };
int gr_integrator_set_scheme(gr_ctx *ctx, int integ, int scheme);
/* dt = CFL·h/c with h = min kernel h, c = max system c
 * (src/TimeIntegration.jl:63-64,598). */
int gr_integrator_dt(gr_ctx *ctx, int integ, double cfl, double *dt_out);
/* probes / probe_interactions of an integrator (src/TimeIntegration.jl:53,
 * 98): each probe interaction's system_b is one of the probes. */
int gr_integrator_set_probes(gr_ctx *ctx, int integ, int nprobes, const int *probes,
                             int ninters, const int *inters);
/* _measure_probes! (src/TimeIntegration.jl:454-502): re-sort the probe
 * interactions' sources, mirror probe positions, sort the probes by cell,
 * build the grids, zero w_sum and the extras, sweep stage 1 of each probe
 * interaction, run the probes' stage-1 updaters, sort the probes by id. */
int gr_measure_probes(gr_ctx *ctx, int integ);
/* Override the step size of one integrator (ensembles: per-sim dt). */
int gr_integrator_set_dt(gr_ctx *ctx, int integ, double dt);
/* Advance one integrator `nsteps` leapfrog steps
 * (the per-step body of time_integrate!, src/TimeIntegration.jl:645-685;
 * print/save cadence stays with the caller). */
int gr_step(gr_ctx *ctx, int integ, int nsteps, double dt);
/* Advance every listed integrator `nsteps` steps with its own dt, all of
 * them in the same kernel launches (packed pool).
 * Fault isolation (no counterpart in the reference, whose only recovery is a
 * restart from a snapshot, src/Driver.jl:79-96): a simulation whose
 * positions become non-finite, or spread beyond its cell-key range or its
 * planned grid capacity, is faulted: from that step on its particles keep
 * their order and find no partners, and every other integrator of the call
 * advances exactly as it would alone.  The call then completes all nsteps
 * and returns GR_ERR_FAULT naming the faulted integrators; their fault bits
 * stay readable through gr_integrator_status until they are released.  The
 * same holds for gr_step. */
int gr_step_many(gr_ctx *ctx, int ninteg, const int *integs, int nsteps);
/* Fault bits (GR_FAULT_*) an integrator has collected since it was created;
 * 0 for a healthy one. */
int gr_integrator_status(gr_ctx *ctx, int integ, int *fault_out);
/* Capture one step of the listed integrators in a hipGraph and replay it
 * for subsequent gr_step_many calls (0 = off). */
int gr_set_graph_mode(gr_ctx *ctx, int on);

/* ---- launch-shape tuning (no counterpart in the reference: the knobs of
 * the neighbour search and functor gather that replace the hot loop of
 * src/Interaction.jl:477-541; grasph.jl_amd/autotune.py picks them from
 * timed trials).  No value changes any result: the pair lists, summation
 * order and every output are bit-identical for all of them. */
enum {
  GR_TUNE_NL_WIN = 0,       /* candidates staged in LDS per window of the
                               neighbour search: 64, 128, 256 (default),
                               512                                          */
  GR_TUNE_NL_SPAN = 1,      /* lanes whose candidate ranges start within
                               span × the widest range share one staging
                               pass: 1..16 (default 4)                      */
  GR_TUNE_SWEEP_BLOCK = 2,  /* target particles per workgroup of the
                               functor gather: 64, 128, 256 (default)       */
  GR_TUNE_MAXNB = 3,        /* list entries kept per target before it walks
                               its cells directly: 0 = default (48 in 3D,
                               24 in 2D) or 4..1024                         */
  GR_TUNE_INV_RECORDS = 4,  /* 1 (default): in FP32 stages whose sweeps are
                               all FluidPfn S = 2 without δ-SPH or surface
                               tension, the partner records hold 1/ρ and
                               p/ρ² (the same FP32 operations, done once per
                               particle); 0: ρ and p                        */
  GR_TUNE_FUSE_SORT = 5     /* 1 (default): a leapfrog step takes its half
                               step, and stage 0's partner records when no
                               updater or virtual system intervenes, in the
                               sort's permutation; 0: separate kernels      */
};
int gr_tune_set(gr_ctx *ctx, int param, int value);
int gr_tune_get(gr_ctx *ctx, int param, int *value);

/* ---- fine-grained drop-ins (one reference call each) --------------------- */
/* sort_particles!(ps, cutoff, ...) (src/Sorting.jl:157). */
int gr_sort(gr_ctx *ctx, int sys, double cutoff);
/* gr_sort for a real system without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same result as gr_sort (bit for bit).  Its scratch (keys, ranks, radix
 * histogram, block tables) is a static device array of the library, about
 * 34 MB per GPU, held until the process ends; limits: the system holds at
 * most 2,097,152 particles, the context at most 4096 systems and its real
 * systems at most 65536 blocks of 256 particles (beyond: GR_ERR_CAPACITY).
 * The system must be laid out (uploaded) already (else GR_ERR_STATE).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other threads, other streams) race. */
int gr_sort_static(gr_ctx *ctx, int sys, double cutoff);
/* create_grid!(si) (src/Interaction.jl:123). */
int gr_create_grid(gr_ctx *ctx, int inter);
/* sweep!(si, stage) (src/Interaction.jl:292): accumulates into the target
 * fields (+=), as the reference does. */
int gr_sweep(gr_ctx *ctx, int inter, int stage);
/* adjust_v!(si) (src/Interaction.jl:309): the velocity-adjust functor
 * (XSPHPfn) over the interaction's pairs, accumulating into v_adjustment;
 * a no-op when the interaction has none. */
int gr_adjust_v(gr_ctx *ctx, int inter);
/* update_state!(ps, stage, dt) (src/Particles.jl:542). */
int gr_update_state(gr_ctx *ctx, int sys, int stage, double dt);
/* generate_ghosts!(ge) / update_ghost_kinematics!(ge) / update_ghost!(ge, stage)
 * (src/GhostParticles.jl:344,404,423). */
int gr_generate_ghosts(gr_ctx *ctx, int ghost);
int gr_update_ghost_kinematics(gr_ctx *ctx, int ghost);
int gr_update_ghost(gr_ctx *ctx, int ghost, int stage);
/* Grid introspection (src/Interaction.jl:29-34): mingridx[ndims],
 * ngridx[ndims], cell_start (which = 0: _cell_start, 1: _cell_start_a),
 * length ncells+1, 1-based values.  A step builds only the cell arrays its
 * functors walk; asking for one the last grid build skipped (system_a's of
 * a coupled interaction whose functor writes only system_a) returns
 * GR_ERR_STATE, and gr_create_grid builds both. */
int gr_get_grid(gr_ctx *ctx, int inter, double *mingridx, int64_t *ngridx,
                int64_t *ncells);
int gr_get_cell_start(gr_ctx *ctx, int inter, int which, int64_t *out,
                      int64_t count);

/* ---- timing (replaces the TimerOutputs sections of src/TimeIntegration.jl) --
 * mode bit 0: record HIP events on the context stream at every phase of
 * every step of the following gr_step/gr_step_many calls (no extra
 * synchronisation inside a call); gr_get_timing returns the mean ms per step
 * of phases 0 sort, 1 ghosts, 2 grid, 3 half-step, 4 updaters + ghost
 * copies, 5 sweeps, 6 full step, 7 whole step, 8 neighbour search.
 * mode bit 1: count pairs with r < 2h in the sweeps; gr_get_pair_count
 * returns unique pairs per step (self pairs once, coupled pairs once),
 * gr_get_pair_stats the ordered self pairs (each unordered pair from both
 * sides) and the coupled pairs seen from the gathering side. */
int gr_set_timing(gr_ctx *ctx, int mode);
int gr_get_timing(gr_ctx *ctx, double *ms_out, int nphases);
int gr_get_pair_count(gr_ctx *ctx, int integ, int64_t *pairs_out);
int gr_get_pair_stats(gr_ctx *ctx, int64_t *self_ordered, int64_t *coupled);
/* Device-to-device copy of `bytes` bytes by a plain streaming kernel, `reps`
 * times (calibration of HBM traffic counters; no counterpart in the
 * reference). */
int gr_bench_copy(gr_ctx *ctx, int64_t bytes, int reps);

#ifdef __cplusplus
}
#endif
#endif /* GRASPH_HIP_H */
