// gr_device.h — device-side data layout and descriptors of the particle pool.
//
// One context owns one packed structure-of-arrays pool.  Every particle
// system (real or ghost) is a segment [off, off + cap) of the pool; a real
// system's count is fixed, a ghost system's count is written on the device
// each step.  Kernels are launched once per phase over a block table that
// maps each workgroup to (segment, first local index), so one launch serves
// every system of every simulation in the pool.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GR_MAXD 3
#define GR_MAXNS 6
#define GR_BLOCK 256

namespace gr {

// Partner relation of one gather source (functor dispatch of the reference
// on (system_a, system_b) types, src/PairwiseFunctors.jl).
enum PartnerKind : int32_t {
  PK_SELF = 0,       // same system, j != i
  PK_TWO_SIDED = 1,  // real partner, both sides written (each side gathers)
  PK_ONE_SIDED = 2,  // generic one-sided (ghosts, other real systems)
  PK_STATIC = 3,     // StaticBoundarySystem: LJ + artificial viscosity
  PK_DYNAMIC = 4     // DynamicBoundarySystem: mirrored velocity
};

// Functor families with a kernel each.
enum Family : int32_t {
  FAM_FORCE = 0,
  FAM_STRAIN = 1,
  FAM_COUNT = 2,
  FAM_INTERP = 3,
  FAM_XSPH = 4,  // velocity adjustment (XSPHPfn), once per step after the full step
  NFAM = 5
};

struct SysDesc {
  int64_t off;     // pool offset of local index 0
  int32_t cap;     // capacity (real: n; ghost: nb * n_source)
  int32_t kind;    // GR_SYS_*
  int32_t ns, nv;  // stress / vorticity components
  int32_t src;     // ghost: source system; else -1
  int32_t nb;      // ghost: boundaries
  int32_t bnd;     // ghost: first boundary in the boundary table
  int32_t virt;    // source of a VirtualParticleSystem
  double gcut;     // ghost cutoff
  double mass, c;  // ghosts carry their source's
  double sv[3], srho;
  double pv[3];            // virtual: prescribed velocity
  long long zero_mask;     // virtual: fields zeroed before the stage loop
};

// Per-(system) parameters of one step of a batch.
struct StepSys {
  double dt, half, gdt;   // dt, dt/2, Γ/dt
  double sort_cutoff;     // 2·h of the integrator
  int32_t reset_force;    // half-step writes dvdt/drhodt ← source
  int32_t fuse_eos;       // 0 none, GR_UPD_TAIT, GR_UPD_LINEAR
  double eos_rho0, eos_gamma;
  int32_t damping;        // Γ != 0
  // fields the step overwrites before reading them, which the sort's
  // permutation with the half step need not move: 1 dvdt and drhodt (the
  // first force stage assigns them), 2 strain rate, 4 vorticity (the first
  // strain-rate stage assigns them and no updater before it reads them)
  int32_t perm_skip;
};

struct Bnd {  // ghost boundary plane
  double n[3], p[3];
};

// Cell grid of one interaction (src/Interaction.jl:29-35), written on the
// device each step.
struct Grid {
  double ming[3];
  int32_t ng[3];
  int32_t ncells;
  int32_t ok;
  int32_t pad;
};

// One CSR array of one interaction (role a or b).
struct CellArr {
  int64_t off;    // offset in the cell pool
  int32_t cap;    // capacity in entries (ncells + 1 must fit)
  int32_t grid;   // interaction index
  int32_t sys;    // system whose particles it indexes
  int32_t pad;
};

// Position component arrays, passed by value so that the loads of all
// components issue together.
struct XP {
  const double *p[3];
  __device__ __forceinline__ const double *operator[](int d) const { return p[d]; }
};
struct XPW {
  double *p[3];
  __device__ __forceinline__ double *operator[](int d) const { return p[d]; }
};

// One gather source of a target system (one interaction, one side).
struct SrcDesc {
  int32_t pk, pfn, partner, grid;
  int32_t carr;      // CellArr index of the partner particles
  int32_t kk, sigma, has_delta;
  int32_t has_eps;
  int32_t side;      // 0: gathers for system_a, 1: for system_b (two-sided)
  int32_t lsrc;      // index of this source in the target's neighbour list
  int32_t psel;      // FluidSolidPfn: 1 the target's pressure on both sides, 2 the partner's
  double h, norm, cutoff2, cell;  // kernel
  double alpha, beta, hf, delta, eps;
  double mass_q;                  // partner mass used by the functor
  double c_t_piv, c_q_piv, c_t_psi, c_q_psi;
  double wn[3], wp[3], wbeta, ljcut;
};

// One target system of one (stage, family) sweep launch.
struct Task {
  int32_t sys;
  int32_t src_begin, src_end;
  int32_t assign;   // 1: write source(+)acc, 0: accumulate (+=)
  int32_t ns, nv;
  int32_t has_stress;
  int32_t write_vor;  // an SRV source writes vorticity
  int32_t ltask;      // neighbour-list task of the target (-1: walk cells)
  int32_t imask;      // FAM_INTERP: union of the sources' field masks; bit 30: w_sum
  // FAM_XSPH with a ghost target: ghost.v_adjustment is the source system's
  // array (src/GhostParticles.jl:226-236), so the sum of ghost j goes to the
  // source's entry j (pool index alias_off + j, j < alias_cap); -1: own entry
  int32_t alias;
  int32_t alias_cap;
  int64_t alias_off;
};

// Neighbour list of one target system: sources [src_begin, src_end) of the
// list source table; mask entry e of target i is the mask int2 {lo, hi} at
// lidx[list_off + 2·(e·cap + i)] and its base at
// lidx[list_off + 2·maxnb·cap + e·cap + i] (12 bytes an entry, list_off a
// multiple of 4), entry counts per source at lcnt[cnt_off + q·cap + i],
// overflow flag at lcnt[cnt_off + nsrc·cap + i]; maxnb = entries kept per
// target.
struct LTask {
  int32_t sys;
  int32_t src_begin, src_end;
  int32_t cap;
  int32_t maxnb;
  int32_t pad;
  int64_t list_off, cnt_off;
};

// One state updater of one (system, stage) (src/StateUpdaters.jl).
struct UpdDesc {
  int32_t kind;
  int32_t pad;
  long long mask;   // ZeroFieldUpdater field mask
  double prm[5];
};

struct Blk {  // block table entry
  int32_t item;   // system / task / cell-array index
  int32_t first;  // first local index
};

// Device view of the pool for one precision.
template <typename R>
struct Pool;

// Partner records: what a pair functor reads of a partner particle, packed
// per particle so that one pair costs three 16-byte loads (RecA: position and
// the scalar fields, RecB: half-step velocity) instead of eight scattered
// 4/8-byte gathers from separate arrays (two loads with RecC below).  Written
// by k_pack before the sweeps of each stage, from the arrays below.
template <typename R>
struct RecA;
template <>
struct __align__(16) RecA<float> {
  double x[3];
  float rho, p;
};
template <>
struct __align__(16) RecA<double> {
  double x[3];
  double rho;
};
template <typename R>
struct RecB;
template <>
struct __align__(16) RecB<float> {
  float v[3];
  float pad;
};
template <>
struct __align__(16) RecB<double> {
  double v[3];
  double p;
};
// Compact partner record of the FP32 fluid-only force sweep: one pair costs
// two 16-byte loads.  Positions are fixed-point multiples of the pool's
// quantum qs (a power of two, 2^-30 of the smallest power of two at or above
// every interaction's cutoff), kept modulo 2^32: the difference of two
// particles within a cutoff is exact in 32-bit arithmetic whatever their
// absolute position.
struct __align__(16) RecC {
  int32_t xq[3];
  float v[3];
  float rho, p;
};
// Partner stress of compact Cauchy sweeps (NS <= 4 in the first half).
struct __align__(16) RecS {
  float4 a, b;
};

// x -> its fixed-point coordinate modulo 2^32 (as k_pack and the target side
// both compute it)
__device__ inline uint32_t quant_of(double x, double iqs) {
  return (uint32_t)(unsigned long long)(long long)rint(x * iqs);
}

__device__ inline float rec_pres(const RecA<float> &a, const RecB<float> &) { return a.p; }
__device__ inline double rec_pres(const RecA<double> &, const RecB<double> &b) { return b.p; }

template <typename R>
struct Pool {
  RecA<R> *ra;
  RecB<R> *rb;
  RecC *rc;        // FP32 pools: compact records (force and strain-rate sweeps)
  RecS *rs;        // FP32 pools: partner stress of compact Cauchy sweeps
  double qs, iqs;  // fixed-point quantum of RecC positions and its inverse
  double *wsum;    // virtual systems: interpolation weight sums
  double *acc[4];  // RK4 accumulators of dvdt (D components) and drhodt
  double *vadj[3];  // v_adjustment (XSPH), swapped with the sort
  double *pcnt;     // probe neighbour counts (NeighborCountFn)
  double *x[3], *v[3], *rho, *rho_h;
  R *vs[3], *rhos, *dvdt[3], *drhodt, *p;
  double *stress[GR_MAXNS];
  R *stress_s[GR_MAXNS], *sr[GR_MAXNS], *vor[3];
  double *strain[GR_MAXNS], *strainp[GR_MAXNS];
  int32_t *id, *gidx, *gbnd, *gnrm;
};

// Double → ordered int64 (monotone, -0.0 < +0.0) for atomic min/max.
__device__ inline long long ord_of(double d) {
  long long i = __double_as_longlong(d);
  return i < 0 ? (i ^ 0x7FFFFFFFFFFFFFFFLL) : i;
}
__device__ inline double double_of(long long i) {
  return __longlong_as_double(i < 0 ? (i ^ 0x7FFFFFFFFFFFFFFFLL) : i);
}

}  // namespace gr
