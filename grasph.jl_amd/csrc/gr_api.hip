// gr_api.hip — C ABI (include/grasph_hip.h), particle pool and step planner.
//
// Host side of the library: registry of systems / ghosts / interactions /
// integrators, the packed device pool, per-batch launch plans (block tables,
// gather-source tables, updater tables) and the step sequence of
// time_integrate! (src/TimeIntegration.jl:645-685) expressed as kernel
// launches on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grasph_hip.h"
#include "gr_device.h"
#include "gr_launch.h"

using namespace gr;

// ---------------------------------------------------------------------------
// Scratch of gr_sort_static: a file-scope device array, one copy per GPU,
// held until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host it only stands for the symbol): it uses the
// device address hipGetSymbolAddress returns, taken once per device in
// gr_ctx_create with that device current.
//   keys, vals, keys2, vals2  4 × GR_STATIC_SORT_MAXN × 4 B
//   radix scratch             radix_scratch_bytes(GR_STATIC_SORT_MAXN)
//   per-system key bases      GR_STATIC_MAXSYS × 4 B
//   the sorted system's list and block table, every real system's blocks
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two threads) race.
// ---------------------------------------------------------------------------
#define GR_STATIC_SORT_MAXN (1 << 21)  // particles of the sorted system
#define GR_STATIC_MAXSYS 4096          // systems of the context
#define GR_STATIC_MAXBLK (1 << 16)     // real-system blocks of the pool (16.7M particles)
namespace {
constexpr size_t kScrKeys = 0;
constexpr size_t kScrVals = kScrKeys + (size_t)GR_STATIC_SORT_MAXN * 4;
constexpr size_t kScrKeys2 = kScrVals + (size_t)GR_STATIC_SORT_MAXN * 4;
constexpr size_t kScrVals2 = kScrKeys2 + (size_t)GR_STATIC_SORT_MAXN * 4;
constexpr size_t kScrHist = kScrVals2 + (size_t)GR_STATIC_SORT_MAXN * 4;
// radix scratch: 4 KB of digit totals, then two arrays of 256 tile counts of
// 4 B per tile of 4096 keys (radix_scratch_bytes)
constexpr size_t kScrHistBytes = 4096 + 2 * ((size_t)GR_STATIC_SORT_MAXN / 4096 + 1) * 256 * 4;
constexpr size_t kScrBase = kScrHist + kScrHistBytes;
constexpr size_t kScrList = kScrBase + (size_t)GR_STATIC_MAXSYS * 4;
constexpr size_t kScrBlkSys = kScrList + 256;
constexpr size_t kScrBlkAll = kScrBlkSys + ((size_t)GR_STATIC_SORT_MAXN / GR_BLOCK + 1) * 8;
constexpr size_t kScrBytes = kScrBlkAll + (size_t)GR_STATIC_MAXBLK * 8;
}  // namespace
static __device__ __align__(256) unsigned char scratch[kScrBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per GPU

// hipGetSymbolAddress once per device, with that device current.
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = ptr;
  return hipSuccess;
}
void *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace

namespace {

struct GrError : std::runtime_error {
  int code;
  GrError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      throw GrError(GR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Device blocks of one context kept for reuse: plans are rebuilt whenever a
// simulation is admitted or released, and their tables come and go.  Blocks
// up to kCacheBlock bytes are kept in power-of-two size classes, at most
// kCacheTotal bytes in all; a larger block is freed.  All work of a context
// runs on its one stream, so a block returned while that stream may still
// read it is only handed out again to work queued behind it.
//
// Blocks come from a memory pool of the context and go back to it in stream
// order (hipMallocFromPoolAsync / hipFreeAsync on the context's stream).
// hipFree waits for every stream of the device: a context that grows its
// particle pool (dozens of arrays) while other contexts step would wait for
// their queued steps once per array.  Without a pool (creation failed) the
// blocks come from hipMalloc and go back with hipFree.
struct DevCache {
  static constexpr size_t kCacheBlock = 32u << 20;
  static constexpr size_t kCacheTotal = 1u << 30;
  std::map<size_t, std::vector<void *>> bins;
  size_t cached = 0;
  hipMemPool_t pool = nullptr;
  hipStream_t stream = nullptr;
  DevCache() = default;
  DevCache(const DevCache &) = delete;
  DevCache &operator=(const DevCache &) = delete;
  ~DevCache() {
    for (auto &kv : bins)
      for (void *q : kv.second) raw_free(q);
    if (pool) {
      (void)hipStreamSynchronize(stream);
      (void)hipMemPoolDestroy(pool);
    }
  }
  void *raw_alloc(size_t b) {
    void *q = nullptr;
    if (pool)
      HIPCHK(hipMallocFromPoolAsync(&q, b, pool, stream));
    else
      HIPCHK(hipMalloc(&q, b));
    return q;
  }
  void raw_free(void *q) {
    if (pool)
      (void)hipFreeAsync(q, stream);
    else
      (void)hipFree(q);
  }
  static size_t size_class(size_t b) {
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
  }
  void *take(size_t b, size_t &got) {
    if (b <= kCacheBlock) {
      got = size_class(b);
      auto it = bins.find(got);
      if (it != bins.end() && !it->second.empty()) {
        void *q = it->second.back();
        it->second.pop_back();
        cached -= got;
        return q;
      }
    } else {
      got = b;
    }
    return raw_alloc(got);
  }
  void give(void *q, size_t b) {
    if (b <= kCacheBlock && cached + b <= kCacheTotal) {
      bins[b].push_back(q);
      cached += b;
    } else {
      raw_free(q);
    }
  }
};
// The cache of the context whose C-ABI call is running on this thread
// (set by guard()); null outside a call.
thread_local DevCache *tl_cache = nullptr;

// Owned device buffer.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevCache *cache = nullptr;  // where the block goes back (null: hipFree)
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      if (cache)
        cache->give(p, bytes);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    cache = nullptr;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    if (tl_cache) {
      p = tl_cache->take(b, bytes);
      cache = tl_cache;
    } else {
      HIPCHK(hipMalloc(&p, b));
      bytes = b;
    }
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
  // Copy a host vector in, reusing the allocation when it is large enough
  // (pointers baked into captured graphs stay valid), then wait for it.
  template <typename T>
  void upload(const std::vector<T> &v, hipStream_t s) {
    const size_t need = std::max<size_t>(v.size() * sizeof(T), 16);
    if (!p || bytes < need) alloc(need);
    if (!v.empty()) {
      HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
    }
  }
};

struct SysH {
  int kind = GR_SYS_BASIC, ns = 0, nv = 0;
  int64_t n = 0, cap = 0, off = -1;
  double mass = 0, c = 0, sv[3] = {0, 0, 0}, srho = 0;
  std::vector<std::vector<UpdDesc>> upd;  // per stage
  // ghost
  int src = -1, nb = 0, bnd = -1;
  double gcut = 0;
  bool ex_p = false, ex_stress = false;
  std::vector<int2> copy;  // per stage: {copy_p, stress_mode}
  bool dead = false;       // released with its integrator; segment reusable
  // source of a VirtualParticleSystem (src/Particles.jl:476-511)
  bool virt = false;
  long long zero_mask = 0;
  double pv[3] = {0, 0, 0};
  // ProbeParticleSystem (src/ProbeParticles.jl:22-89): virtual-like target
  // measured by gr_measure_probes; mirror: mirror_target system or -1
  bool probe = false;
  int mirror = -1;
};

struct InterH {
  int kk = 0;
  double h = 0, norm = 0, cell = 0;
  int a = -1, b = -1, wrap = GR_WRAP_NONE;
  double wn[3] = {0, 0, 0}, wp[3] = {0, 0, 0}, beta = 0, ljcut = 0;
  std::vector<gr_pfn_desc> pfns;
  int carr_b = -1, carr_a = -1;  // CellArr indices (b or self; a when coupled)
  bool has_xsph = false;  // velocity_adjust_pairwise_fn = XSPHPfn(xsph_eps)
  double xsph_eps = 0;
  bool grid_built = false;
  bool dead = false;
};

struct IntegH {
  std::vector<int> sys, inter, ghost, virt, probes, probe_inter;
  double gamma = 0, h = 0, c = 0, dt = 0;
  int nst = 0;
  bool dead = false;
  int scheme = GR_SCHEME_LEAPFROG;
  int fault = 0;  // GR_FAULT_* bits collected by its steps
};

// Free ranges [off, off + len) of a pool, kept sorted and coalesced.
struct FreeList {
  std::vector<std::pair<int64_t, int64_t>> seg;
  void put(int64_t off, int64_t len) {
    if (len <= 0) return;
    seg.push_back({off, len});
    std::sort(seg.begin(), seg.end());
    std::vector<std::pair<int64_t, int64_t>> m;
    for (auto &q : seg) {
      if (!m.empty() && m.back().first + m.back().second == q.first)
        m.back().second += q.second;
      else
        m.push_back(q);
    }
    seg.swap(m);
  }
  // First fit; -1 if no range is long enough.
  int64_t take(int64_t len) {
    for (size_t k = 0; k < seg.size(); ++k)
      if (seg[k].second >= len) {
        const int64_t off = seg[k].first;
        seg[k].first += len;
        seg[k].second -= len;
        if (seg[k].second == 0) seg.erase(seg.begin() + (long)k);
        return off;
      }
    return -1;
  }
};

enum PK { P_SELF = PK_SELF, P_TWO = PK_TWO_SIDED, P_ONE = PK_ONE_SIDED, P_STATIC = PK_STATIC, P_DYN = PK_DYNAMIC };

struct FamPlan {
  DBuf tasks, srcs, blk;
  int nblk = 0;
  bool fluid_only = false;  // every source's functor is FluidPfn / FluidSolidPfn
  bool cauchy = false;      // some source's functor is CauchyFluidPfn
  bool inv = false;         // every source: FluidPfn S = 2, no δ-SPH, no surface tension
};
struct StagePlan {
  DBuf blk_upd, urange, upd;
  int nblk_upd = 0;
  int upd_kinds = 0;  // the updater kinds of the stage (bit GR_UPD_x)
  DBuf blk_gcopy, gcopy;
  int nblk_gcopy = 0;
  FamPlan fam[NFAM];
};

struct Plan {
  std::vector<int> integs;
  // real sort
  std::vector<int> real_list;
  DBuf d_real_list, d_bbase_real, blk_real;
  int nblk_real = 0, n_real = 0;
  // ghosts
  std::vector<int> ghost_list;
  // ghost generation: blocks of source particles per ghost system, and the
  // first (boundary, source block) count slot of each ghost system
  DBuf d_ghost_list, d_bbase_ghost, d_gslot, blk_ghost, blk_gsrc;
  int nblk_ghost = 0, nblk_gsrc = 0, n_ghost = 0, n_gslot = 0;
  // grids
  std::vector<int> inters;
  std::vector<std::pair<int, bool>> carr_state;  // (cell array, built by this plan)
  DBuf d_inter_list, blk_cells, blk_mark;
  int nblk_cells = 0, nblk_mark = 0;
  // stages
  std::vector<std::unique_ptr<StagePlan>> stages;
  std::vector<std::pair<int, StepSys>> stepsys;  // per-system step parameters
  // neighbour lists
  DBuf ltasks, lsrcs, blk_list;
  int nblk_list = 0;
  int scheme = GR_SCHEME_LEAPFROG;
  // velocity adjustment after the full step (XSPH sweep tasks; vadj_live:
  // the correction runs, also with no XSPH sweep when v_adjustment was set)
  FamPlan xsph;
  // ghost-target XSPH tasks, one launch each after the others: they add into
  // their source's v_adjustment, which the source's own task also writes
  std::vector<std::unique_ptr<FamPlan>> xsph_al;
  std::vector<int2> xsph_alias;  // (ghost, source n) of those tasks
  int vadj_live = 0;
  // targets per workgroup of the plan's sweep block tables
  int sweep_block = GR_BLOCK;
  // integrated (non-virtual) real systems; virtual systems
  DBuf blk_int, blk_virt, blk_probe;
  DBuf d_halfsys;  // per system: 1 if integrated by this plan (its half step can ride the sort)
  int nblk_int = 0, nblk_virt = 0, nblk_probe = 0;
  bool do_sort = true, do_ghosts = true, do_grid = true, do_step = true;
  // graph capture per A/B parity
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  ~Plan() {
    for (auto &g : graph)
      if (g) (void)hipGraphExecDestroy(g);
  }
};

}  // namespace

struct gr_ctx {
  // first member: destroyed last, after every buffer has returned its block
  DevCache cache;
  int device = 0, D = 2, prec = GR_PREC_F32;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<SysH> sys;
  std::vector<InterH> inter;
  std::vector<IntegH> integ;
  std::vector<Bnd> bnds;
  // pool
  int64_t pool_cap = 0;
  int64_t pool_end = 0;  // end of the laid-out segments
  FreeList free_pool, free_cells;
  bool dirty_layout = true;
  int nsmax = 4;
  int rsize = 4;  // sizeof(R)
  // swapped families: index 0/1 = A/B
  DBuf x[2], v[2], rho[2], dvdt[2], drhodt[2], p[2], id[2], stress[2], stress_s[2], sr[2],
      vor[2], strain[2], strainp[2];
  int cur = 0;
  DBuf vs, rhos, rho_h, gidx, gbnd, gnrm, gxs, gidx_s, gbnd_s;
  std::vector<char> *carr_used = nullptr;  // set while a step plan collects its sources
  std::vector<char> carr_valid;  // cell array holds the last grid build's starts
  DBuf ra, rb;  // partner records (RecA / RecB), rebuilt before every stage's sweeps
  DBuf rc;      // compact records (RecC) of FP32 pools
  DBuf rs;      // partner stress records (RecS) of FP32 pools
  double qscale = 1.0;  // fixed-point quantum of RecC positions
  DBuf wsum;    // w_sum of virtual systems
  DBuf pcnt;    // neighbour counts of probe systems
  DBuf probe_scr;  // scratch of the probes' sort by id
  DBuf acc;     // RK4 accumulators (4 components)
  DBuf vadj[2];  // v_adjustment (XSPH), swapped with the sort
  bool vadj_any = false;  // some v_adjustment may be nonzero: permute and correct it
  // tables
  DBuf d_sd, d_n, d_st, d_bnd, d_kmin, d_kmax, d_bbmin, d_bbmax, d_kbase, d_krange, d_flags;
  DBuf d_sysfault;  // per system: GR_FAULT_* bits raised by the step's kernels
  int bb_rs = 0;  // entries per replica of d_bbmin / d_bbmax (GR_BB_REP replicas)
  bool bb_valid = false;  // d_bbmin / d_bbmax hold the bounds of a step's key pass
  DBuf d_grids, d_carr, d_cellsz, d_inter_sys, d_inter_carr, cells;
  std::vector<CellArr> carr;
  int64_t cells_cap = 0;
  std::vector<StepSys> stepsys;
  // sort / scan scratch
  DBuf keys, vals, keys2, vals2, hist, cand, cand_scan, scan_bsum, bagg;
  int64_t sort_cap = 0, cand_cap = 0, cellblk_cap = 0;
  DBuf d_pairs;
  DBuf lidx, lcnt;
  int64_t lidx_cap = 0, lcnt_cap = 0;
  int use_lists = 1;
  int list_maxnb = 0;  // mask entries kept per target (0: 48 in 3D, 24 in 2D)
  // launch shapes chosen by gr_tune_set (results do not depend on them)
  // default shapes: nl_win 256 and span 4, the autotuner's choice for the dam
  // break and the 2D column ensemble, 2–3 % faster than 128 and 2 on batches
  // of 2D and 3D columns
  int tune_win = 256, tune_span = 4, tune_block = GR_BLOCK, tune_inv = 1, tune_fuse = 1;
  std::map<std::vector<int>, std::unique_ptr<Plan>> plans;
  // cached device tables derived from the layout
  DBuf allreal_blk;
  int allreal_nblk = 0;
  int64_t allreal_cap = -1;
  size_t allreal_nsys = 0;
  std::vector<char> id_written;
  bool graph_mode = false, timing = false, count_pairs = false;
  hipEvent_t ev[9] = {};
  std::vector<std::vector<hipEvent_t>> tev;  // per-step phase events (timing mode)
  double phase_ms[9] = {0};
  long long last_pairs[2] = {0, 0};
  int last_pairs_integ_key = -1;
};

namespace {
// Position component pointers of pool family w and of the ghost scratch.
XP xp_pool(gr_ctx *c, int w) {
  XP r;
  for (int d = 0; d < 3; ++d) r.p[d] = c->x[w].as<double>() + d * c->pool_cap;
  return r;
}
XPW xpw_pool(gr_ctx *c, int w) {
  XPW r;
  for (int d = 0; d < 3; ++d) r.p[d] = c->x[w].as<double>() + d * c->pool_cap;
  return r;
}
XP xp_gxs(gr_ctx *c) {
  XP r;
  for (int d = 0; d < 3; ++d) r.p[d] = c->gxs.as<double>() + d * c->pool_cap;
  return r;
}
XPW xpw_gxs(gr_ctx *c) {
  XPW r;
  for (int d = 0; d < 3; ++d) r.p[d] = c->gxs.as<double>() + d * c->pool_cap;
  return r;
}

}  // namespace

namespace {

int fail(gr_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

template <typename F>
int guard(gr_ctx *c, F &&f) {
  if (!c) return GR_ERR_ARGUMENT;
  struct CacheScope {  // device blocks of this call come from c's cache
    DevCache *prev;
    explicit CacheScope(DevCache *cc) : prev(tl_cache) { tl_cache = cc; }
    ~CacheScope() { tl_cache = prev; }
  } scope(&c->cache);
  try {
    f();
    return GR_OK;
  } catch (const GrError &e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    c->err = e.what();
    return GR_ERR_STATE;
  } catch (...) {  // nothing may unwind across the C ABI
    c->err = "unknown exception";
    return GR_ERR_STATE;
  }
}

void need(bool cond, int code, const std::string &m) {
  if (!cond) throw GrError(code, m);
}

// ---------------------------------------------------------------------------
// Pool
// ---------------------------------------------------------------------------

template <typename R>
Pool<R> pool_view(gr_ctx *c, int w) {
  Pool<R> P;
  const int64_t cap = c->pool_cap;
  P.ra = c->ra.as<RecA<R>>();
  P.rb = c->rb.as<RecB<R>>();
  P.rc = c->rc.as<RecC>();
  P.rs = c->rs.as<RecS>();
  P.qs = c->qscale;
  P.iqs = 1.0 / c->qscale;
  P.wsum = c->wsum.as<double>();
  P.pcnt = c->pcnt.as<double>();
  for (int k = 0; k < 4; ++k) P.acc[k] = c->acc.as<double>() + k * cap;
  for (int d = 0; d < 3; ++d) P.vadj[d] = c->vadj[w].as<double>() + d * cap;
  for (int d = 0; d < 3; ++d) {
    P.x[d] = c->x[w].as<double>() + d * cap;
    P.v[d] = c->v[w].as<double>() + d * cap;
    P.vs[d] = c->vs.as<R>() + d * cap;
    P.dvdt[d] = c->dvdt[w].as<R>() + d * cap;
    P.vor[d] = c->vor[w].as<R>() + d * cap;
  }
  P.rho = c->rho[w].as<double>();
  P.rho_h = c->rho_h.as<double>();
  P.rhos = c->rhos.as<R>();
  P.drhodt = c->drhodt[w].as<R>();
  P.p = c->p[w].as<R>();
  for (int k = 0; k < GR_MAXNS; ++k) {
    const int kk = k < c->nsmax ? k : 0;
    P.stress[k] = c->stress[w].as<double>() + kk * cap;
    P.stress_s[k] = c->stress_s[w].as<R>() + kk * cap;
    P.sr[k] = c->sr[w].as<R>() + kk * cap;
    P.strain[k] = c->strain[w].as<double>() + kk * cap;
    P.strainp[k] = c->strainp[w].as<double>() + kk * cap;
  }
  P.id = c->id[w].as<int>();
  P.gidx = c->gidx.as<int>();
  P.gbnd = c->gbnd.as<int>();
  P.gnrm = c->gnrm.as<int>();
  return P;
}

// Grow (or create) one pool array family, keeping the first old_cap entries
// of each of its `comps` components.
void grow(gr_ctx *c, DBuf &b, int comps, size_t esz, int64_t old_cap, int64_t new_cap) {
  DBuf nb;
  nb.alloc((size_t)comps * new_cap * esz);
  HIPCHK(hipMemsetAsync(nb.p, 0, nb.bytes, c->stream));
  if (b.p && old_cap > 0)
    for (int k = 0; k < comps; ++k)
      HIPCHK(hipMemcpyAsync((char *)nb.p + (size_t)k * new_cap * esz,
                            (char *)b.p + (size_t)k * old_cap * esz, old_cap * esz,
                            hipMemcpyDeviceToDevice, c->stream));
  std::swap(b.p, nb.p);
  std::swap(b.bytes, nb.bytes);
  std::swap(b.cache, nb.cache);
}

void commit(gr_ctx *c) {
  if (!c->dirty_layout) return;
  // new systems take a released segment (first fit) or go at the end
  std::vector<std::pair<int64_t, int64_t>> fresh;
  for (auto &s : c->sys) {
    if (s.dead || s.off >= 0) continue;
    s.off = c->free_pool.take(s.cap);
    if (s.off < 0) {
      s.off = c->pool_end;
      c->pool_end += s.cap;
    } else {
      fresh.push_back({s.off, s.cap});  // a reused segment keeps old contents
    }
  }
  const int64_t need_cap = std::max<int64_t>(c->pool_end, 1);
  if (need_cap > c->pool_cap) {
    const int64_t old = c->pool_cap;
    const int64_t nc = std::max(need_cap, old + old / 2);
    const size_t R = c->rsize;
    for (int w = 0; w < 2; ++w) {
      grow(c, c->x[w], 3, 8, old, nc);
      grow(c, c->v[w], 3, 8, old, nc);
      grow(c, c->rho[w], 1, 8, old, nc);
      grow(c, c->dvdt[w], 3, R, old, nc);
      grow(c, c->drhodt[w], 1, R, old, nc);
      grow(c, c->p[w], 1, R, old, nc);
      grow(c, c->id[w], 1, 4, old, nc);
      grow(c, c->stress[w], c->nsmax, 8, old, nc);
      grow(c, c->stress_s[w], c->nsmax, R, old, nc);
      grow(c, c->sr[w], c->nsmax, R, old, nc);
      grow(c, c->vor[w], 3, R, old, nc);
      grow(c, c->strain[w], c->nsmax, 8, old, nc);
      grow(c, c->strainp[w], c->nsmax, 8, old, nc);
      grow(c, c->vadj[w], 3, 8, old, nc);
    }
    grow(c, c->vs, 3, R, old, nc);
    grow(c, c->rhos, 1, R, old, nc);
    grow(c, c->rho_h, 1, 8, old, nc);
    grow(c, c->gidx, 1, 4, old, nc);
    grow(c, c->gbnd, 1, 4, old, nc);
    grow(c, c->gnrm, 1, 4, old, nc);
    grow(c, c->gxs, 3, 8, old, nc);
    grow(c, c->gidx_s, 1, 4, old, nc);
    grow(c, c->gbnd_s, 1, 4, old, nc);
    grow(c, c->ra, 1, 32, old, nc);
    grow(c, c->wsum, 1, 8, old, nc);
    grow(c, c->pcnt, 1, 8, old, nc);
    grow(c, c->acc, 4, 8, old, nc);
    grow(c, c->rb, 1, R == 4 ? 16 : 32, old, nc);
    if (R == 4) grow(c, c->rc, 1, 32, old, nc);
    if (R == 4) grow(c, c->rs, 1, 32, old, nc);
    c->pool_cap = nc;
  }
  // v_adjustment starts at zero (src/Particles.jl:82) in both families
  for (auto &f : fresh)
    for (int w = 0; w < 2; ++w)
      for (int d = 0; d < 3; ++d)
        HIPCHK(hipMemsetAsync(c->vadj[w].as<double>() + d * c->pool_cap + f.first, 0,
                              (size_t)f.second * 8, c->stream));
  c->plans.clear();  // block tables and captured graphs refer to the layout
  // id = 1..n for systems laid out for the first time (src/Particles.jl:92)
  c->id_written.resize(c->sys.size(), 0);
  for (size_t i = 0; i < c->sys.size(); ++i) {
    if (c->id_written[i] || c->sys[i].kind == GR_SYS_GHOST || c->sys[i].dead) continue;
    std::vector<int> ids(c->sys[i].n);
    for (int64_t k = 0; k < c->sys[i].n; ++k) ids[k] = (int)(k + 1);
    HIPCHK(hipMemcpyAsync(c->id[c->cur].as<int>() + c->sys[i].off, ids.data(), ids.size() * 4,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->id_written[i] = 1;
  }
  // descriptor tables
  const int ns = (int)c->sys.size();
  std::vector<SysDesc> sd(ns);
  std::vector<int> nn(ns);
  for (int i = 0; i < ns; ++i) {
    const SysH &s = c->sys[i];
    SysDesc &d = sd[i];
    memset(&d, 0, sizeof(d));
    d.off = s.off;
    d.cap = (int32_t)s.cap;
    d.kind = s.kind;
    d.ns = s.ns;
    d.nv = s.nv;
    d.src = s.src;
    d.nb = s.nb;
    d.bnd = s.bnd;
    d.gcut = s.gcut;
    d.mass = s.mass;
    d.c = s.c;
    for (int k = 0; k < 3; ++k) {
      d.sv[k] = s.sv[k];
      d.pv[k] = s.pv[k];
    }
    d.srho = s.srho;
    d.virt = s.virt ? 1 : 0;
    d.zero_mask = s.zero_mask;
    nn[i] = s.kind == GR_SYS_GHOST ? 0 : (int)s.n;
  }
  // keep device ghost counts across commits
  std::vector<int> oldn;
  if (c->d_n.p && c->d_n.bytes >= sizeof(int)) {
    oldn.resize(c->d_n.bytes / sizeof(int));
    HIPCHK(hipMemcpyAsync(oldn.data(), c->d_n.p, oldn.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < ns && i < (int)oldn.size(); ++i)
      if (c->sys[i].kind == GR_SYS_GHOST) nn[i] = oldn[i];
  }
  c->d_sd.upload(sd, c->stream);
  c->d_n.upload(nn, c->stream);
  c->stepsys.resize(ns);
  c->d_st.upload(c->stepsys, c->stream);
  std::vector<Bnd> bnds = c->bnds;
  if (bnds.empty()) bnds.resize(1);
  c->d_bnd.upload(bnds, c->stream);
  std::vector<int> z(ns * 3, 0);
  std::vector<long long> zl((size_t)ns * 3 * GR_BB_REP, 0);
  c->d_kmin.upload(z, c->stream);
  c->d_kmax.upload(z, c->stream);
  c->d_bbmin.upload(zl, c->stream);
  c->d_bbmax.upload(zl, c->stream);
  c->bb_valid = false;
  c->bb_rs = ns * 3;
  std::vector<unsigned> zu(ns * 3, 0);
  c->d_kbase.upload(zu, c->stream);
  c->d_krange.upload(zu, c->stream);
  c->d_sysfault.upload(std::vector<int>(std::max(ns, 1), 0), c->stream);
  if (!c->d_flags.p) {
    std::vector<int> f(4, 0);
    c->d_flags.upload(f, c->stream);
  }
  if (!c->d_pairs.p) {
    std::vector<unsigned long long> z2(2, 0);
    c->d_pairs.upload(z2, c->stream);
  }
  // sort / ghost generation scratch (one count per boundary and block of
  // GR_BLOCK source particles)
  int64_t nreal = 0, nghost = 0, ncand = 0;
  for (auto &s : c->sys) {
    if (s.kind == GR_SYS_GHOST) {
      nghost += s.cap;
      ncand += (int64_t)s.nb * ((c->sys[s.src].cap + GR_BLOCK - 1) / GR_BLOCK);
    } else {
      nreal += s.cap;
    }
  }
  const int64_t sc = std::max<int64_t>(std::max(nreal, nghost), 1);
  if (sc > c->sort_cap) {
    c->keys.alloc(sc * 4);
    c->vals.alloc(sc * 4);
    c->keys2.alloc(sc * 4);
    c->vals2.alloc(sc * 4);
    // the radix sort's tile counts start at zero (each sort leaves them so)
    c->hist.alloc(radix_scratch_bytes((int)sc));
    HIPCHK(hipMemsetAsync(c->hist.p, 0, c->hist.bytes, c->stream));
    c->sort_cap = sc;
  }
  if (ncand + 1 > c->cand_cap) {
    c->cand.alloc((ncand + 1) * 4);
    c->cand_scan.alloc((ncand + 2) * 4);
    c->scan_bsum.alloc((size_t)(scan_block_count((int)ncand + 1) + 1) * 4);
    c->cand_cap = ncand + 1;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->dirty_layout = false;
}

void upload_stepsys(gr_ctx *c) { c->d_st.upload(c->stepsys, c->stream); }

// ---------------------------------------------------------------------------
// Field transfer (host AoS doubles <-> device SoA)
// ---------------------------------------------------------------------------

int ncomp(gr_ctx *c, const SysH &s, int field) {
  const int D = c->D;
  const int NS = s.kind == GR_SYS_GHOST ? c->sys[s.src].ns : s.ns;
  switch (field) {
    case GR_F_X:
    case GR_F_V:
    case GR_F_DVDT:
    case GR_F_V_ADJUSTMENT:
    case GR_F_NORMALS:
      return D;
    case GR_F_RHO:
    case GR_F_DRHODT:
    case GR_F_P:
    case GR_F_W_SUM:
    case GR_F_COUNT:
      return 1;
    case GR_F_STRESS:
    case GR_F_STRAIN_RATE:
    case GR_F_STRAIN:
    case GR_F_STRAIN_P:
      return NS;
    case GR_F_VORTICITY:
      return s.nv;
  }
  return 0;
}

bool has_field(const SysH &s, int field) {
  if (s.kind == GR_SYS_GHOST) {
    switch (field) {
      case GR_F_X:
      case GR_F_V:
      case GR_F_RHO:
      case GR_F_NORMALS:
        return true;
      case GR_F_P:
        return s.ex_p;
      case GR_F_STRESS:
        return s.ex_stress;
    }
    return false;
  }
  switch (field) {
    case GR_F_X:
    case GR_F_V:
    case GR_F_RHO:
    case GR_F_DVDT:
    case GR_F_DRHODT:
    case GR_F_V_ADJUSTMENT:
      return true;
    case GR_F_P:
      return s.kind >= GR_SYS_FLUID;
    case GR_F_STRESS:
    case GR_F_STRAIN_RATE:
      return s.kind >= GR_SYS_STRESS;
    case GR_F_VORTICITY:
    case GR_F_STRAIN:
    case GR_F_STRAIN_P:
      return s.kind == GR_SYS_EP;
    case GR_F_W_SUM:
      return s.virt;
    case GR_F_COUNT:
      return s.probe;
  }
  return false;
}

// Component pointers (element size) of a field in the current pool.
struct FieldRef {
  std::vector<char *> comp;
  size_t esz;
};

FieldRef field_ref(gr_ctx *c, int field, bool shadow_only) {
  const int w = c->cur;
  const int64_t cap = c->pool_cap;
  const size_t R = c->rsize;
  FieldRef f;
  auto add = [&](DBuf &b, int comps, size_t esz) {
    f.esz = esz;
    for (int k = 0; k < comps; ++k) f.comp.push_back((char *)b.p + (size_t)k * cap * esz);
  };
  (void)shadow_only;
  switch (field) {
    case GR_F_X: add(c->x[w], c->D, 8); break;
    case GR_F_V: add(c->v[w], c->D, 8); break;
    case GR_F_RHO: add(c->rho[w], 1, 8); break;
    case GR_F_DVDT: add(c->dvdt[w], c->D, R); break;
    case GR_F_DRHODT: add(c->drhodt[w], 1, R); break;
    case GR_F_P: add(c->p[w], 1, R); break;
    case GR_F_STRESS: add(c->stress[w], c->nsmax, 8); break;
    case GR_F_STRAIN_RATE: add(c->sr[w], c->nsmax, R); break;
    case GR_F_VORTICITY: add(c->vor[w], 3, R); break;
    case GR_F_STRAIN: add(c->strain[w], c->nsmax, 8); break;
    case GR_F_STRAIN_P: add(c->strainp[w], c->nsmax, 8); break;
    case GR_F_W_SUM: add(c->wsum, 1, 8); break;
    case GR_F_COUNT: add(c->pcnt, 1, 8); break;
    case GR_F_V_ADJUSTMENT: add(c->vadj[w], c->D, 8); break;
  }
  return f;
}

void copy_comp_to_dev(gr_ctx *c, char *dst, size_t esz, int64_t off, const double *src,
                      int64_t n, int stride, int k) {
  if (n <= 0) return;
  if (esz == 8) {
    std::vector<double> t(n);
    for (int64_t i = 0; i < n; ++i) t[i] = src[i * stride + k];
    HIPCHK(hipMemcpyAsync(dst + off * 8, t.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> t(n);
    for (int64_t i = 0; i < n; ++i) t[i] = (float)src[i * stride + k];
    HIPCHK(hipMemcpyAsync(dst + off * 4, t.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
}

void copy_comp_from_dev(gr_ctx *c, const char *src, size_t esz, int64_t off, double *dst,
                        int64_t n, int stride, int k) {
  if (n <= 0) return;
  if (esz == 8) {
    std::vector<double> t(n);
    HIPCHK(hipMemcpyAsync(t.data(), src + off * 8, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) dst[i * stride + k] = t[i];
  } else if (esz == 4) {
    std::vector<float> t(n);
    HIPCHK(hipMemcpyAsync(t.data(), src + off * 4, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) dst[i * stride + k] = (double)t[i];
  }
}

std::vector<int> read_ints(gr_ctx *c, const DBuf &b, int64_t off, int64_t n) {
  std::vector<int> t(std::max<int64_t>(n, 0));
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(t.data(), (const char *)b.p + off * 4, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return t;
}

int64_t live_n(gr_ctx *c, int sysid) {
  const SysH &s = c->sys[sysid];
  if (s.kind != GR_SYS_GHOST) return s.n;
  commit(c);
  auto v = read_ints(c, c->d_n, sysid, 1);
  return v[0];
}

// ---------------------------------------------------------------------------
// Planning helpers
// ---------------------------------------------------------------------------

void add_blocks(std::vector<Blk> &v, int item, int64_t count, int per) {
  for (int64_t f = 0; f < count; f += per) v.push_back(Blk{item, (int32_t)f});
}

// Functor dispatch on the partner's type, as the oracle's classify().
int classify(gr_ctx *c, const InterH &in, int pfn) {
  if (in.b < 0) return P_SELF;
  if (in.wrap == GR_WRAP_STATIC) {
    if (pfn == GR_PFN_SR || pfn == GR_PFN_SRV || pfn == GR_PFN_COUNT) return P_ONE;
    return P_STATIC;
  }
  if (in.wrap == GR_WRAP_DYNAMIC) return pfn == GR_PFN_COUNT ? P_ONE : P_DYN;
  const SysH &b = c->sys[in.b];
  const SysH &a = c->sys[in.a];
  if (b.kind == GR_SYS_GHOST || b.virt || a.virt) return P_ONE;
  switch (pfn) {
    case GR_PFN_FLUID:
      return (a.kind == GR_SYS_FLUID && b.kind == GR_SYS_FLUID) ? P_TWO : P_ONE;
    case GR_PFN_CAUCHY:
    case GR_PFN_COUNT:
    case GR_PFN_FLUIDSOLID:
      return P_TWO;
    default:
      return P_ONE;
  }
}

int family_of(int pfn) {
  if (pfn == GR_PFN_FLUID || pfn == GR_PFN_CAUCHY || pfn == GR_PFN_FLUIDSOLID) return FAM_FORCE;
  if (pfn == GR_PFN_SR || pfn == GR_PFN_SRV) return FAM_STRAIN;
  if (pfn == GR_PFN_COUNT) return FAM_COUNT;
  if (pfn == GR_PFN_INTERP) return FAM_INTERP;
  if (pfn == GR_PFN_XSPH) return FAM_XSPH;
  if (pfn == GR_PFN_NCOUNT) return FAM_INTERP;
  return -1;
}

double src_mass(gr_ctx *c, int sysid) {
  const SysH &s = c->sys[sysid];
  return s.kind == GR_SYS_GHOST ? c->sys[s.src].mass : s.mass;
}
double src_c(gr_ctx *c, int sysid) {
  const SysH &s = c->sys[sysid];
  return s.kind == GR_SYS_GHOST ? c->sys[s.src].c : s.c;
}

SrcDesc make_src(gr_ctx *c, int it, int target, bool b_side, int pk) {
  const InterH &in = c->inter[it];
  (void)target;
  SrcDesc q;
  memset(&q, 0, sizeof(q));
  q.pk = pk;
  q.grid = it;
  const int a = in.a;
  const int partner = b_side ? a : (in.b < 0 ? a : in.b);
  q.partner = partner;
  q.carr = b_side ? in.carr_a : in.carr_b;
  if (c->carr_used && q.carr >= 0) {
    if ((int)c->carr_used->size() <= q.carr) c->carr_used->resize(q.carr + 1, 0);
    (*c->carr_used)[q.carr] = 1;
  }
  q.side = b_side ? 1 : 0;
  q.lsrc = 0;
  q.kk = in.kk;
  q.h = in.h;
  q.norm = in.norm;
  q.cell = in.cell;
  q.cutoff2 = in.cell * in.cell;
  q.wbeta = in.beta;
  q.ljcut = in.ljcut;
  for (int d = 0; d < 3; ++d) {
    q.wn[d] = in.wn[d];
    q.wp[d] = in.wp[d];
  }
  const int t = b_side ? in.b : a;
  const double ct = src_c(c, t), cq = src_c(c, partner), ca = c->sys[a].c;
  switch (pk) {
    case P_SELF:
      q.mass_q = c->sys[a].mass;
      break;
    case P_TWO:
      q.mass_q = src_mass(c, partner);
      break;
    case P_ONE:
      q.mass_q = src_mass(c, partner);
      break;
    case P_STATIC:
      q.mass_q = src_mass(c, partner);
      break;
    case P_DYN:
      q.mass_q = c->sys[a].mass;
      break;
  }
  q.c_t_piv = q.c_q_piv = q.c_t_psi = q.c_q_psi = ca;
  if (pk == P_TWO) {
    q.c_t_piv = q.c_t_psi = ct;
    q.c_q_piv = q.c_q_psi = cq;
  }
  return q;
}

void set_pfn(SrcDesc &q, const gr_pfn_desc &f, int pk, gr_ctx *c, const InterH &in) {
  q.pfn = f.kind;
  q.psel = f.kind == GR_PFN_FLUIDSOLID ? (q.side == 0 ? 1 : 2) : 0;
  q.sigma = f.sigma;
  q.alpha = f.alpha;
  q.beta = f.beta;
  q.hf = f.h;
  q.has_delta = std::isnan(f.delta) ? 0 : 1;
  q.delta = q.has_delta ? f.delta : 0.0;
  q.has_eps = (!std::isnan(f.epsilon) && pk == P_TWO && f.kind == GR_PFN_FLUID) ? 1 : 0;
  q.eps = q.has_eps ? f.epsilon : 0.0;
  // CauchyFluidPfn with a ghost / virtual partner uses the partner's c in
  // the artificial viscosity (src/PairwiseFunctors.jl:293)
  if (f.kind == GR_PFN_CAUCHY && pk == P_ONE && in.b >= 0 &&
      (c->sys[in.b].kind == GR_SYS_GHOST || c->sys[in.b].virt))
    q.c_q_piv = src_c(c, in.b);
}

// Cell capacity of an interaction: grid of the union bounding box of the
// given systems, padded by the ghost cutoffs and doubled per dimension.
int64_t cell_capacity(gr_ctx *c, const InterH &in, const double *mn, const double *mx) {
  double pad = 0.0;
  for (auto &s : c->sys)
    if (s.kind == GR_SYS_GHOST) pad = std::max(pad, s.gcut);
  int64_t cells = 1;
  for (int d = 0; d < c->D; ++d) {
    double ext = (mx[d] - mn[d]) + 2 * pad + 6 * in.cell;
    double n = std::floor(ext / in.cell) + 3;
    n = n * 1.6 + 4;
    if (!(n < 4e6)) n = 4e6;
    cells *= (int64_t)n;
  }
  return std::min<int64_t>(cells + 1, (int64_t)1 << 30);
}

// Bounding boxes of the given real systems (device reduction, host readback).
void host_bbox(gr_ctx *c, const std::vector<int> &systems, std::vector<double> &mn,
               std::vector<double> &mx) {
  commit(c);
  std::vector<int> list;
  std::vector<Blk> blk;
  for (int s : systems) {
    list.push_back(s);
    add_blocks(blk, s, c->sys[s].cap, GR_BLOCK);
  }
  DBuf dl, db;
  dl.upload(list, c->stream);
  db.upload(blk, c->stream);
  launch_minmax_reset(c->stream, dl.as<int>(), (int)list.size(), c->d_bbmin.as<long long>(),
                      c->d_bbmax.as<long long>(), c->bb_rs);
  launch_key_minmax(c->stream, xp_pool(c, c->cur), c->d_sd.as<SysDesc>(), c->d_n.as<int>(),
                    db.as<Blk>(), (int)blk.size(), c->D, c->d_bbmin.as<long long>(),
                    c->d_bbmax.as<long long>(), c->bb_rs, dl.as<int>(), (int)list.size(), true);
  const int ns = (int)c->sys.size();
  std::vector<long long> bmn(ns * 3), bmx(ns * 3);
  HIPCHK(hipMemcpyAsync(bmn.data(), c->d_bbmin.p, ns * 3 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(bmx.data(), c->d_bbmax.p, ns * 3 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  auto dof = [](long long i) {
    long long u = i < 0 ? (i ^ 0x7FFFFFFFFFFFFFFFLL) : i;
    double d;
    memcpy(&d, &u, 8);
    return d;
  };
  mn.assign(3, INFINITY);
  mx.assign(3, -INFINITY);
  for (int s : systems)
    for (int d = 0; d < c->D; ++d) {
      if (c->sys[s].n == 0) continue;
      mn[d] = std::min(mn[d], dof(bmn[s * 3 + d]));
      mx[d] = std::max(mx[d], dof(bmx[s * 3 + d]));
    }
}

// Destroy every captured step graph (device tables or pools they point to
// are about to move).
void drop_graphs(gr_ctx *c) {
  for (auto &pk : c->plans)
    for (auto &gg : pk.second->graph)
      if (gg) {
        (void)hipGraphExecDestroy(gg);
        gg = nullptr;
      }
}

// Allocate cell arrays of the interactions that have none yet.
void ensure_cells(gr_ctx *c, const std::vector<int> &inters, const std::vector<int> &union_sys) {
  bool any = false;
  for (int it : inters)
    if (c->inter[it].carr_b < 0) any = true;
  if (!any) return;
  std::vector<int> reals;
  for (int s : union_sys)
    if (c->sys[s].kind != GR_SYS_GHOST) reals.push_back(s);
  for (int it : inters) {
    InterH &in = c->inter[it];
    if (in.carr_b >= 0) continue;
    std::vector<int> own;
    for (int s : reals) own.push_back(s);
    if (own.empty()) {
      own.push_back(in.a);
      if (in.b >= 0 && c->sys[in.b].kind != GR_SYS_GHOST) own.push_back(in.b);
    }
    std::vector<double> mn, mx;
    host_bbox(c, own, mn, mx);
    const int64_t cap = cell_capacity(c, in, mn.data(), mx.data());
    auto mk = [&](int sysid) {
      CellArr a;
      a.off = c->free_cells.take(cap);
      if (a.off < 0) {
        a.off = c->cells_cap;
        c->cells_cap += cap;
      }
      a.cap = (int32_t)cap;
      a.grid = it;
      a.sys = sysid;
      a.pad = 0;
      c->carr.push_back(a);
      return (int)c->carr.size() - 1;
    };
    if (in.b < 0) {
      in.carr_b = mk(in.a);
    } else {
      in.carr_b = mk(in.b);
      in.carr_a = mk(in.a);
    }
  }
  // grow the cell pool (grids are rebuilt every step, so no copy); captured
  // graphs of other plans hold the old table and pool pointers
  drop_graphs(c);
  if ((size_t)std::max<int64_t>(c->cells_cap, 1) * 4 > c->cells.bytes)
    c->cells.alloc(std::max<int64_t>(c->cells_cap, 1) * 4);
  c->d_carr.upload(c->carr, c->stream);
  const int ni = (int)c->inter.size();
  std::vector<int2> isys(ni), icarr(ni);
  std::vector<double> cellsz(ni);
  double cmax = 0.0;
  for (int i = 0; i < ni; ++i) {
    isys[i] = make_int2(c->inter[i].a, c->inter[i].b);
    icarr[i] = make_int2(c->inter[i].carr_b, c->inter[i].carr_a);
    cellsz[i] = c->inter[i].cell;
    cmax = std::max(cmax, c->inter[i].cell);
  }
  // RecC quantum: 2^-30 of the power of two at or above the largest cutoff,
  // so a partner's offset fits 31 bits
  {
    int e = 0;
    if (cmax > 0.0 && std::isfinite(cmax)) frexp(cmax, &e);
    const double q = cmax > 0.0 && std::isfinite(cmax) ? ldexp(1.0, e - 30) : 1.0;
    if (q != c->qscale) drop_graphs(c);
    c->qscale = q;
  }
  c->d_inter_sys.upload(isys, c->stream);
  c->d_inter_carr.upload(icarr, c->stream);
  c->d_cellsz.upload(cellsz, c->stream);
  std::vector<Grid> g(std::max(ni, 1));
  memset(g.data(), 0, g.size() * sizeof(Grid));
  if (!c->d_grids.p || c->d_grids.bytes < g.size() * sizeof(Grid)) c->d_grids.upload(g, c->stream);
  for (auto &kv : c->plans)
    for (auto &gg : kv.second->graph)
      if (gg) {
        (void)hipGraphExecDestroy(gg);
        gg = nullptr;
      }
}

// Grid plan (block tables over cell arrays and marked particles).  With
// `used`, only the cell arrays some source of the plan reads are built (a
// coupled interaction whose functor writes only system_a never walks
// system_a's cells).
void plan_grids(gr_ctx *c, Plan &P, const std::vector<int> &inters,
                const std::vector<char> *used = nullptr) {
  P.inters = inters;
  P.d_inter_list.upload(inters, c->stream);
  std::vector<Blk> cb, mb;
  for (int it : inters) {
    const InterH &in = c->inter[it];
    for (int ca : {in.carr_b, in.carr_a}) {
      if (ca < 0) continue;
      const bool skip = used && (ca >= (int)used->size() || !(*used)[ca]);
      P.carr_state.emplace_back(ca, !skip);
      if (skip) continue;
      add_blocks(cb, ca, c->carr[ca].cap, GR_BLOCK * 4);
      add_blocks(mb, ca, c->sys[c->carr[ca].sys].cap, GR_BLOCK);
    }
  }
  P.blk_cells.upload(cb, c->stream);
  P.blk_mark.upload(mb, c->stream);
  P.nblk_cells = (int)cb.size();
  P.nblk_mark = (int)mb.size();
  if ((int64_t)cb.size() + 1 > c->cellblk_cap) {
    c->bagg.alloc((cb.size() + 1) * 4);
    c->cellblk_cap = cb.size() + 1;
  }
}

void plan_sort_real(gr_ctx *c, Plan &P, const std::vector<int> &systems) {
  P.real_list = systems;
  P.d_real_list.upload(systems, c->stream);
  std::vector<int> bbase(c->sys.size(), -1);
  std::vector<Blk> blk;
  int64_t acc = 0;
  for (int s : systems) {
    bbase[s] = (int)acc;
    acc += c->sys[s].cap;
    add_blocks(blk, s, c->sys[s].cap, GR_BLOCK);
  }
  P.n_real = (int)acc;
  P.d_bbase_real.upload(bbase, c->stream);
  P.blk_real.upload(blk, c->stream);
  P.nblk_real = (int)blk.size();
}

// Count slots of ghost generation: one per (boundary, block of GR_BLOCK
// source particles) of each ghost system.
int64_t ghost_slots(const gr_ctx *c, const SysH &g) {
  return (int64_t)g.nb * ((c->sys[g.src].cap + GR_BLOCK - 1) / GR_BLOCK);
}

void plan_ghosts(gr_ctx *c, Plan &P, const std::vector<int> &ghosts) {
  P.ghost_list = ghosts;
  P.d_ghost_list.upload(ghosts, c->stream);
  std::vector<int> bbase(c->sys.size(), -1), slot(c->sys.size(), 0);
  std::vector<Blk> blk, sblk;
  int64_t acc = 0, sacc = 0;
  for (int g : ghosts) {
    bbase[g] = (int)acc;
    acc += c->sys[g].cap;
    slot[g] = (int)sacc;
    sacc += ghost_slots(c, c->sys[g]);
    add_blocks(blk, g, c->sys[g].cap, GR_BLOCK);
    add_blocks(sblk, g, c->sys[c->sys[g].src].cap, GR_BLOCK);
  }
  P.n_ghost = (int)acc;
  P.n_gslot = (int)sacc;
  P.d_bbase_ghost.upload(bbase, c->stream);
  P.d_gslot.upload(slot, c->stream);
  P.blk_ghost.upload(blk, c->stream);
  P.blk_gsrc.upload(sblk, c->stream);
  P.nblk_ghost = (int)blk.size();
  P.nblk_gsrc = (int)sblk.size();
}

// Stage plans of a batch of integrators: updater tables (with the fused EOS
// and the zeroing absorbed by assigning sweeps), ghost copiers, and the
// gather tasks of each functor family.
void plan_stages(gr_ctx *c, Plan &P, bool count_pairs_unused) {
  (void)count_pairs_unused;
  int nst = 0;
  for (int ig : P.integs) nst = std::max(nst, c->integ[ig].nst);
  P.stages.clear();
  // per-system effective updater lists
  const int ns = (int)c->sys.size();
  std::vector<std::vector<std::vector<UpdDesc>>> upd(ns);
  std::vector<int> sys_integ(ns, -1);
  for (int ig : P.integs) {
    for (int s : c->integ[ig].sys) sys_integ[s] = ig;
    for (int v : c->integ[ig].virt) sys_integ[v] = ig;
  }
  for (int ig : P.integs) {
    const IntegH &I = c->integ[ig];
    for (int s : I.sys) {
      SysH &S = c->sys[s];
      StepSys &T = c->stepsys[s];
      memset(&T, 0, sizeof(T));
      T.dt = I.dt;
      T.half = I.dt / 2;
      T.gdt = I.gamma / I.dt;
      T.damping = I.gamma != 0.0;
      T.sort_cutoff = 2.0 * I.h;
      T.reset_force = 1;
      upd[s].assign(I.nst, {});
      if ((int)S.upd.size() == I.nst) {  // _sweep_all_stages! :405
        for (int k = 0; k < I.nst; ++k)
          for (const UpdDesc &u : S.upd[k])
            if (u.kind != GR_UPD_NOTHING) upd[s][k].push_back(u);
        if (I.nst >= 1 && upd[s][0].size() == 1 &&
            (upd[s][0][0].kind == GR_UPD_TAIT || upd[s][0][0].kind == GR_UPD_LINEAR)) {
          T.fuse_eos = upd[s][0][0].kind;
          T.eos_rho0 = upd[s][0][0].prm[0];
          T.eos_gamma = upd[s][0][0].prm[1];
          upd[s][0].clear();
        }
      }
    }
    for (int v : I.virt) {  // sorted, updated per stage, never integrated
      SysH &S = c->sys[v];
      StepSys &T = c->stepsys[v];
      memset(&T, 0, sizeof(T));
      T.dt = I.dt;
      T.half = I.dt / 2;
      T.sort_cutoff = 2.0 * I.h;
      upd[v].assign(I.nst, {});
      if ((int)S.upd.size() == I.nst)  // _sweep_all_stages! :411-414
        for (int k = 0; k < I.nst; ++k)
          for (const UpdDesc &u : S.upd[k])
            if (u.kind != GR_UPD_NOTHING) upd[v][k].push_back(u);
    }
    for (int g : I.ghost) {
      StepSys &T = c->stepsys[g];
      memset(&T, 0, sizeof(T));
      T.sort_cutoff = 2.0 * I.h;
      T.dt = I.dt;
    }
    for (int q : I.probes) {  // advanced by prescribed_v·dt every step
      StepSys &T = c->stepsys[q];
      memset(&T, 0, sizeof(T));
      T.sort_cutoff = 2.0 * I.h;
      T.dt = I.dt;
    }
  }
  // gather tasks: (stage, family, target) → sources, in interaction order
  struct TaskB {
    Task t;
    std::vector<SrcDesc> src;
  };
  // [stage·NFAM + fam] target → task; stage slot nst holds the XSPH sweep
  std::vector<std::map<int, TaskB>> tasks((nst + 1) * NFAM);
  for (int k = 0; k <= nst; ++k)
    for (int ig : P.integs)
      for (int it : c->integ[ig].inter) {
        const InterH &in = c->inter[it];
        gr_pfn_desc xf;
        const gr_pfn_desc *fp = nullptr;
        if (k == nst) {  // adjust_v! (src/Interaction.jl:309): the generic two-sided method
          if (!in.has_xsph) continue;
          memset(&xf, 0, sizeof(xf));
          xf.kind = GR_PFN_XSPH;
          xf.alpha = in.xsph_eps;
          fp = &xf;
        } else {
          if (k >= (int)in.pfns.size()) continue;
          fp = &in.pfns[k];
        }
        const gr_pfn_desc &f = *fp;
        const int fam = family_of(f.kind);
        if (fam < 0) continue;
        const int pk = f.kind == GR_PFN_XSPH ? (in.b < 0 ? P_SELF : P_TWO) : classify(c, in, f.kind);
        auto add = [&](int target, bool b_side) {
          auto &m = tasks[k * NFAM + fam];
          TaskB &tb = m[target];
          if (tb.src.empty()) {
            memset(&tb.t, 0, sizeof(Task));
            tb.t.sys = target;
            tb.t.ns = c->sys[target].ns;
            tb.t.nv = c->sys[target].nv;
            tb.t.alias = -1;
            if (f.kind == GR_PFN_XSPH && c->sys[target].kind == GR_SYS_GHOST) {
              const SysH &src = c->sys[c->sys[target].src];
              tb.t.alias = c->sys[target].src;
              tb.t.alias_cap = (int32_t)src.n;
              tb.t.alias_off = src.off;
            }
          }
          SrcDesc q = make_src(c, it, target, b_side, pk);
          set_pfn(q, f, pk, c, in);
          if (f.kind == GR_PFN_CAUCHY) tb.t.has_stress = 1;
          if (f.kind == GR_PFN_SRV) tb.t.write_vor = 1;
          if (f.kind == GR_PFN_INTERP) tb.t.imask |= f.sigma | (f.alpha != 0.0 ? 1 << 30 : 0);
          if (f.kind == GR_PFN_NCOUNT) tb.t.imask |= 1 << GR_F_COUNT;
          tb.src.push_back(q);
        };
        if (f.kind == GR_PFN_NCOUNT) {  // the probe side counts (:443-460)
          need(in.b >= 0 && c->sys[in.b].probe && !c->sys[in.a].virt && in.wrap == GR_WRAP_NONE,
               GR_ERR_ARGUMENT,
               "NeighborCountFn needs a real system_a and a ProbeParticleSystem system_b");
          add(in.b, true);
          continue;
        }
        if (f.kind == GR_PFN_INTERP) {  // the virtual side gathers (:424-434)
          need(in.b >= 0 && in.wrap == GR_WRAP_NONE &&
                   (c->sys[in.b].virt != c->sys[in.a].virt) &&
                   c->sys[in.b].kind != GR_SYS_GHOST,
               GR_ERR_ARGUMENT, "InterpolateFieldFn needs a real and a virtual particle system");
          if (c->sys[in.b].virt)
            add(in.b, true);
          else
            add(in.a, false);
          continue;
        }
        add(in.a, false);
        if (pk == P_TWO && in.b >= 0) add(in.b, true);
      }
  // assignment: FORCE group
  for (int s = 0; s < ns; ++s) {
    if (sys_integ[s] < 0 || c->sys[s].kind == GR_SYS_GHOST) continue;
    int first = -1, firstfam = -1;
    for (int k = 0; k < nst && first < 0; ++k)
      for (int fam : {FAM_FORCE, FAM_COUNT})
        if (tasks[k * NFAM + fam].count(s)) {
          first = k;
          firstfam = fam;
          break;
        }
    if (first < 0) continue;
    bool zeroed = false;
    for (int k = 0; k <= first && k < (int)upd[s].size(); ++k)
      for (auto &u : upd[s][k])
        if (u.kind == GR_UPD_ZERO && (u.mask & ((1 << GR_F_DVDT) | (1 << GR_F_DRHODT))))
          zeroed = true;
    if (!zeroed) {
      c->stepsys[s].reset_force = 0;
      c->stepsys[s].perm_skip |= 1;
      tasks[first * NFAM + firstfam][s].t.assign = 1;
    }
  }
  // assignment: STRAIN group absorbs a stage's ZeroFieldUpdater
  for (int k = 0; k < nst; ++k)
    for (auto &kv : tasks[k * NFAM + FAM_STRAIN]) {
      const int s = kv.first;
      if (k >= (int)upd[s].size()) continue;
      auto &L = upd[s][k];
      bool only_zero = !L.empty();
      long long m = 0;
      for (auto &u : L) {
        if (u.kind != GR_UPD_ZERO) only_zero = false;
        m |= u.mask;
      }
      long long need = 1 << GR_F_STRAIN_RATE;
      if (kv.second.t.write_vor) need |= 1 << GR_F_VORTICITY;
      if (only_zero && (m & need) == need) {
        kv.second.t.assign = 1;
        for (auto &u : L) u.mask &= ~need;
        L.erase(std::remove_if(L.begin(), L.end(), [](const UpdDesc &u) { return u.mask == 0; }),
                L.end());
      }
    }
  // strain rate and vorticity are dead from the sort to the first strain-rate
  // stage of a system when that stage assigns them and no updater up to it
  // reads them (the stress updaters run before the stage's sweeps)
  for (int s = 0; s < ns; ++s) {
    if (sys_integ[s] < 0 || c->sys[s].kind == GR_SYS_GHOST) continue;
    for (int k = 0; k < nst; ++k) {
      auto it = tasks[k * NFAM + FAM_STRAIN].find(s);
      bool reads = false;
      if (k < (int)upd[s].size())
        for (auto &u : upd[s][k])
          reads = reads || u.kind == GR_UPD_EP || u.kind == GR_UPD_HOOKE || u.kind == GR_UPD_VPMC;
      if (reads) break;
      if (it == tasks[k * NFAM + FAM_STRAIN].end()) continue;
      if (it->second.t.assign) {
        c->stepsys[s].perm_skip |= 2;
        if (it->second.t.write_vor) c->stepsys[s].perm_skip |= 4;
      }
      break;
    }
  }
  // neighbour lists: one per target, over the union of its sources
  std::map<int, std::vector<SrcDesc>> tl;  // target → list sources
  for (int k = 0; k <= nst; ++k)
    for (int fam = 0; fam < NFAM; ++fam)
      for (auto &kv : tasks[k * NFAM + fam]) {
        if (kv.second.t.alias >= 0) continue;  // ghost targets walk the cells
        auto &L = tl[kv.first];
        for (auto &q : kv.second.src) {
          int idx = -1;
          for (int l = 0; l < (int)L.size(); ++l)
            if (L[l].grid == q.grid && L[l].side == q.side) idx = l;
          if (idx < 0) {
            idx = (int)L.size();
            L.push_back(q);
          }
          q.lsrc = idx;
        }
      }
  std::vector<LTask> ltv;
  std::vector<SrcDesc> lsv;
  std::vector<Blk> lbv;
  std::map<int, int> ltask_of;
  int64_t lacc = 0, cacc = 0;
  // mask entries kept per target (12 B each): a strip of <= 64 candidates
  // takes one entry, so 9 (3D) or 3 (2D) per source for interior targets
  const int maxnb = c->list_maxnb > 0 ? c->list_maxnb : (c->D == 3 ? 48 : 24);
  for (auto &kv : tl) {
    LTask L;
    memset(&L, 0, sizeof(L));
    L.sys = kv.first;
    L.cap = (int32_t)c->sys[kv.first].cap;
    L.maxnb = maxnb;
    L.src_begin = (int)lsv.size();
    for (auto &q : kv.second) lsv.push_back(q);
    L.src_end = (int)lsv.size();
    L.list_off = lacc;
    L.cnt_off = cacc;
    lacc += ((int64_t)L.cap * maxnb * 3 + 3) & ~(int64_t)3;  // 16-byte aligned lists
    cacc += (int64_t)L.cap * (L.src_end - L.src_begin + 1);
    add_blocks(lbv, (int)ltv.size(), L.cap, 64);  // one wavefront per workgroup
    ltask_of[kv.first] = (int)ltv.size();
    ltv.push_back(L);
  }
  if (c->use_lists) {
    P.ltasks.upload(ltv, c->stream);
    P.lsrcs.upload(lsv, c->stream);
    P.blk_list.upload(lbv, c->stream);
    P.nblk_list = (int)lbv.size();
    if (lacc > c->lidx_cap || cacc > c->lcnt_cap) {
      c->lidx.alloc(std::max<int64_t>(lacc, 1) * 4);
      c->lcnt.alloc(std::max<int64_t>(cacc, 1) * 4);
      c->lidx_cap = lacc;
      c->lcnt_cap = cacc;
      drop_graphs(c);
    }
  }
  // materialise stage plans
  for (int k = 0; k < nst; ++k) {
    auto sp = std::make_unique<StagePlan>();
    std::vector<UpdDesc> ud;
    std::vector<int2> ur(ns, make_int2(0, 0));
    std::vector<Blk> ub;
    for (int s = 0; s < ns; ++s) {
      if (sys_integ[s] < 0 || k >= (int)upd[s].size() || upd[s][k].empty()) continue;
      ur[s] = make_int2((int)ud.size(), (int)(ud.size() + upd[s][k].size()));
      for (auto &u : upd[s][k]) {
        ud.push_back(u);
        sp->upd_kinds |= 1 << u.kind;
      }
      add_blocks(ub, s, c->sys[s].cap, GR_BLOCK);
    }
    sp->upd.upload(ud, c->stream);
    sp->urange.upload(ur, c->stream);
    sp->blk_upd.upload(ub, c->stream);
    sp->nblk_upd = (int)ub.size();
    std::vector<int2> gc(ns, make_int2(0, -1));
    std::vector<Blk> gb;
    for (int ig : P.integs)
      for (int g : c->integ[ig].ghost) {
        const SysH &G = c->sys[g];
        if (k < (int)G.copy.size() && (G.copy[k].x || G.copy[k].y >= 0)) {
          gc[g] = G.copy[k];
          add_blocks(gb, g, G.cap, GR_BLOCK);
        }
      }
    sp->gcopy.upload(gc, c->stream);
    sp->blk_gcopy.upload(gb, c->stream);
    sp->nblk_gcopy = (int)gb.size();
    for (int fam = 0; fam < NFAM; ++fam) {
      std::vector<Task> tv;
      std::vector<SrcDesc> sv;
      std::vector<Blk> bv;
      for (auto &kv : tasks[k * NFAM + fam]) {
        Task t = kv.second.t;
        t.ltask = c->use_lists ? ltask_of[kv.first] : -1;
        t.src_begin = (int)sv.size();
        for (auto &q : kv.second.src) sv.push_back(q);
        t.src_end = (int)sv.size();
        add_blocks(bv, (int)tv.size(), c->sys[t.sys].cap, P.sweep_block);
        tv.push_back(t);
      }
      sp->fam[fam].tasks.upload(tv, c->stream);
      sp->fam[fam].srcs.upload(sv, c->stream);
      sp->fam[fam].blk.upload(bv, c->stream);
      sp->fam[fam].nblk = (int)bv.size();
      bool fo = fam == FAM_FORCE;
      for (const SrcDesc &q : sv) fo = fo && (q.pfn == GR_PFN_FLUID || q.pfn == GR_PFN_FLUIDSOLID);
      sp->fam[fam].fluid_only = fo;
      bool cy = false;
      for (const SrcDesc &q : sv) cy = cy || q.pfn == GR_PFN_CAUCHY;
      sp->fam[fam].cauchy = cy;
      bool iv = fo && c->tune_inv && !sv.empty();
      for (const SrcDesc &q : sv)
        iv = iv && q.pfn == GR_PFN_FLUID && q.sigma == 2 && !q.has_delta && !q.has_eps &&
             q.psel == 0;
      sp->fam[fam].inv = iv;
    }
    P.stages.push_back(std::move(sp));
  }
  {
    auto build = [&](FamPlan &F, bool alias, int only) {
      std::vector<Task> tv;
      std::vector<SrcDesc> sv;
      std::vector<Blk> bv;
      for (auto &kv : tasks[nst * NFAM + FAM_XSPH]) {
        Task t = kv.second.t;
        if ((t.alias >= 0) != alias || (alias && kv.first != only)) continue;
        t.ltask = (c->use_lists && t.alias < 0) ? ltask_of[kv.first] : -1;
        t.src_begin = (int)sv.size();
        for (auto &q : kv.second.src) sv.push_back(q);
        t.src_end = (int)sv.size();
        add_blocks(bv, (int)tv.size(), c->sys[t.sys].cap, P.sweep_block);
        tv.push_back(t);
      }
      F.tasks.upload(tv, c->stream);
      F.srcs.upload(sv, c->stream);
      F.blk.upload(bv, c->stream);
      F.nblk = (int)bv.size();
    };
    build(P.xsph, false, -1);
    for (auto &kv : tasks[nst * NFAM + FAM_XSPH]) {
      if (kv.second.t.alias < 0) continue;
      P.xsph_al.push_back(std::make_unique<FamPlan>());
      build(*P.xsph_al.back(), true, kv.first);
      P.xsph_alias.push_back(make_int2(kv.second.t.sys, kv.second.t.alias_cap));
    }
    P.vadj_live = c->vadj_any ? 1 : 0;
  }
  P.stepsys.clear();
  for (int ig : P.integs) {
    for (int s : c->integ[ig].sys) P.stepsys.push_back({s, c->stepsys[s]});
    for (int v : c->integ[ig].virt) P.stepsys.push_back({v, c->stepsys[v]});
    for (int g : c->integ[ig].ghost) P.stepsys.push_back({g, c->stepsys[g]});
    for (int q : c->integ[ig].probes) P.stepsys.push_back({q, c->stepsys[q]});
  }
  upload_stepsys(c);
}

Plan &get_plan(gr_ctx *c, const std::vector<int> &integs) {
  commit(c);
  auto it = c->plans.find(integs);
  if (it != c->plans.end()) return *it->second;
  auto P = std::make_unique<Plan>();
  P->integs = integs;
  P->sweep_block = c->tune_block;
  std::vector<int> reals, ghosts, inters, all;
  for (int ig : integs) {
    const IntegH &I = c->integ[ig];
    need(!I.dead, GR_ERR_STATE, "integrator was released");
    for (int s : I.sys) reals.push_back(s);
    for (int v : I.virt) reals.push_back(v);  // sorted with the real systems
    for (int g : I.ghost) ghosts.push_back(g);
    for (int t : I.inter) inters.push_back(t);
  }
  all = reals;
  for (int g : ghosts) all.push_back(g);
  {
    std::vector<int> u = all;
    std::sort(u.begin(), u.end());
    need(std::unique(u.begin(), u.end()) == u.end(), GR_ERR_ARGUMENT,
         "a system belongs to more than one integrator of the batch");
  }
  ensure_cells(c, inters, reals);
  P->scheme = c->integ[integs[0]].scheme;
  for (int ig : integs)
    need(c->integ[ig].scheme == P->scheme, GR_ERR_ARGUMENT,
         "integrators advanced together must use the same time integration scheme");
  plan_sort_real(c, *P, reals);
  {
    std::vector<Blk> bi, bv, bp;
    for (int ig : integs) {
      for (int s : c->integ[ig].sys) add_blocks(bi, s, c->sys[s].cap, GR_BLOCK);
      for (int v : c->integ[ig].virt) add_blocks(bv, v, c->sys[v].cap, GR_BLOCK);
      for (int q : c->integ[ig].probes) {
        const SysH &S = c->sys[q];
        if (S.pv[0] != 0.0 || S.pv[1] != 0.0 || S.pv[2] != 0.0)  // iszero(pv) skips
          add_blocks(bp, q, S.cap, GR_BLOCK);
      }
    }
    P->blk_int.upload(bi, c->stream);
    std::vector<int> hs(std::max<size_t>(c->sys.size(), 1), 0);
    for (int ig : integs)
      for (int s : c->integ[ig].sys) hs[s] = 1;
    P->d_halfsys.upload(hs, c->stream);
    P->blk_virt.upload(bv, c->stream);
    P->blk_probe.upload(bp, c->stream);
    P->nblk_int = (int)bi.size();
    P->nblk_virt = (int)bv.size();
    P->nblk_probe = (int)bp.size();
  }
  plan_ghosts(c, *P, ghosts);
  std::vector<char> used;
  c->carr_used = &used;
  try {
    plan_stages(c, *P, false);
  } catch (...) {
    c->carr_used = nullptr;
    throw;
  }
  c->carr_used = nullptr;
  plan_grids(c, *P, inters, &used);
  HIPCHK(hipStreamSynchronize(c->stream));
  Plan &ref = *P;
  c->plans[integs] = std::move(P);
  return ref;
}

// Block table over every real system of the pool (the permutation writes all
// of them, so the A/B swap keeps systems outside the batch intact).
void all_real(gr_ctx *c, const Blk *&blk, int &nblk) {
  if (c->allreal_cap != c->pool_cap || c->allreal_nsys != c->sys.size()) {
    std::vector<Blk> v;
    for (int s = 0; s < (int)c->sys.size(); ++s)
      if (c->sys[s].kind != GR_SYS_GHOST && !c->sys[s].dead)
        add_blocks(v, s, c->sys[s].cap, GR_BLOCK);
    c->allreal_blk.release();
    c->allreal_blk.upload(v, c->stream);
    c->allreal_nblk = (int)v.size();
    c->allreal_cap = c->pool_cap;
    c->allreal_nsys = c->sys.size();
  }
  blk = c->allreal_blk.as<Blk>();
  nblk = c->allreal_nblk;
}

// ---------------------------------------------------------------------------
// Step phases
// ---------------------------------------------------------------------------

// half: also take the leapfrog half step of the plan's integrated systems
// in the permutation (run_step, when it follows the sort directly in data).
template <typename R>
void phase_sort_real(gr_ctx *c, Plan &P, bool half = false, int pack = 0) {
  if (P.real_list.empty()) return;
  hipStream_t s = c->stream;
  const XP xs = xp_pool(c, c->cur);
  launch_minmax_reset(s, P.d_real_list.as<int>(), (int)P.real_list.size(),
                      c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(), c->bb_rs);
  launch_key_minmax(s, xs, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), P.blk_real.as<Blk>(),
                    P.nblk_real, c->D, c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(),
                    c->bb_rs, P.d_real_list.as<int>(), (int)P.real_list.size(), false);
  launch_key_base(s, P.d_real_list.as<int>(), (int)P.real_list.size(), c->D, 0, c->d_n.as<int>(),
                  c->d_st.as<StepSys>(), c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(),
                  c->bb_rs, c->d_kmin.as<int>(), c->d_kmax.as<int>(), c->d_kbase.as<unsigned>(),
                  c->d_krange.as<unsigned>(), c->d_flags.as<int>(), c->d_sysfault.as<int>());
  launch_key_write(s, xs, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), c->d_st.as<StepSys>(),
                   P.blk_real.as<Blk>(), P.nblk_real, P.d_bbase_real.as<int>(), c->D,
                   c->d_kmin.as<int>(), c->d_krange.as<unsigned>(), c->d_kbase.as<unsigned>(),
                   c->keys.as<unsigned>(), c->vals.as<int>(), c->d_flags.as<int>(),
                   c->hist.as<unsigned>());
  launch_radix_sort(s, c->keys.as<unsigned>(), c->vals.as<int>(), c->keys2.as<unsigned>(),
                    c->vals2.as<int>(), c->hist.as<unsigned>(), P.n_real, c->d_flags.as<int>());
  const Blk *arb;
  int arn;
  all_real(c, arb, arn);
  Pool<R> A = pool_view<R>(c, c->cur), B = pool_view<R>(c, 1 - c->cur);
  launch_permute_real<R>(s, A, B, c->d_sd.as<SysDesc>(), arb, arn,
                         P.d_bbase_real.as<int>(), c->vals.as<int>(), c->vals2.as<int>(),
                         c->d_flags.as<int>(), c->vadj_any ? 1 : 0, c->D,
                         c->d_st.as<StepSys>(), half ? P.d_halfsys.as<int>() : nullptr, pack);
  c->cur = 1 - c->cur;
}

// Generation (flags, scan, scatter to scratch) + sort (scratch → pool).
template <typename R>
void phase_ghosts(gr_ctx *c, Plan &P, bool generate, bool sort) {
  if (P.ghost_list.empty()) return;
  hipStream_t s = c->stream;
  const XP x = xp_pool(c, c->cur);
  Pool<R> A = pool_view<R>(c, c->cur);
  if (generate) {
    launch_ghost_count(s, x, c->d_sd.as<SysDesc>(), c->d_bnd.as<Bnd>(), P.blk_gsrc.as<Blk>(),
                       P.nblk_gsrc, P.d_gslot.as<int>(), c->D, c->cand.as<int>());
    launch_scan(s, c->cand.as<int>(), c->cand_scan.as<int>(), c->scan_bsum.as<int>(), P.n_gslot);
    launch_ghost_scatter(s, x, xpw_gxs(c), c->gidx_s.as<int>(), c->gbnd_s.as<int>(),
                         c->gnrm.as<int>(), c->d_sd.as<SysDesc>(), c->d_bnd.as<Bnd>(),
                         P.blk_gsrc.as<Blk>(), P.nblk_gsrc, P.d_gslot.as<int>(),
                         c->cand_scan.as<int>(), c->D, c->d_n.as<int>());
  }
  const XP gx = xp_gxs(c);
  launch_minmax_reset(s, P.d_ghost_list.as<int>(), (int)P.ghost_list.size(),
                      c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(), c->bb_rs);
  launch_key_minmax(s, gx, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), P.blk_ghost.as<Blk>(),
                    P.nblk_ghost, c->D, c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(),
                    c->bb_rs, P.d_ghost_list.as<int>(), (int)P.ghost_list.size(), !sort);
  if (sort) {
    launch_key_base(s, P.d_ghost_list.as<int>(), (int)P.ghost_list.size(), c->D, 1,
                    c->d_n.as<int>(), c->d_st.as<StepSys>(), c->d_bbmin.as<long long>(),
                    c->d_bbmax.as<long long>(), c->bb_rs, c->d_kmin.as<int>(), c->d_kmax.as<int>(),
                    c->d_kbase.as<unsigned>(), c->d_krange.as<unsigned>(), c->d_flags.as<int>(),
                    c->d_sysfault.as<int>());
    launch_key_write(s, gx, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), c->d_st.as<StepSys>(),
                     P.blk_ghost.as<Blk>(), P.nblk_ghost, P.d_bbase_ghost.as<int>(), c->D,
                     c->d_kmin.as<int>(), c->d_krange.as<unsigned>(), c->d_kbase.as<unsigned>(),
                     c->keys.as<unsigned>(), c->vals.as<int>(), c->d_flags.as<int>(),
                     c->hist.as<unsigned>());
    launch_radix_sort(s, c->keys.as<unsigned>(), c->vals.as<int>(), c->keys2.as<unsigned>(),
                      c->vals2.as<int>(), c->hist.as<unsigned>(), P.n_ghost, c->d_flags.as<int>());
  } else {
    launch_fill_i32(s, c->d_flags.as<int>(), 0, 1);
  }
  launch_permute_ghost(s, gx, c->gidx_s.as<int>(), c->gbnd_s.as<int>(),
                       xpw_pool(c, c->cur), A.gidx, A.gbnd, c->d_sd.as<SysDesc>(),
                       c->d_n.as<int>(), P.blk_ghost.as<Blk>(), P.nblk_ghost,
                       P.d_bbase_ghost.as<int>(), c->vals.as<int>(), c->vals2.as<int>(),
                       c->d_flags.as<int>(), c->D);
}

// Host record of which cell arrays the plan's grid build fills.
void mark_grids(gr_ctx *c, const Plan &P) {
  for (const auto &e : P.carr_state) {
    if ((int)c->carr_valid.size() <= e.first) c->carr_valid.resize(e.first + 1, 0);
    c->carr_valid[e.first] = e.second ? 1 : 0;
  }
}

void phase_grids(gr_ctx *c, Plan &P) {
  if (P.inters.empty()) return;
  hipStream_t s = c->stream;
  launch_grid_params(s, P.d_inter_list.as<int>(), (int)P.inters.size(), c->d_inter_sys.as<int2>(),
                     c->d_cellsz.as<double>(), c->d_bbmin.as<long long>(),
                     c->d_bbmax.as<long long>(), c->d_n.as<int>(), c->d_grids.as<Grid>(),
                     c->d_inter_carr.as<int2>(), c->d_carr.as<CellArr>(), c->D, c->d_flags.as<int>(),
                     c->d_sysfault.as<int>());
  launch_cells_build(s, xp_pool(c, c->cur), c->cells.as<int>(),
                     c->d_carr.as<CellArr>(), c->d_grids.as<Grid>(), c->d_cellsz.as<double>(),
                     c->d_sd.as<SysDesc>(), c->d_n.as<int>(), P.blk_cells.as<Blk>(), P.nblk_cells,
                     P.blk_mark.as<Blk>(), P.nblk_mark, c->bagg.as<int>(), c->D);
}

template <typename R>
void phase_stage(gr_ctx *c, Plan &P, int k, bool cnt, hipEvent_t *mid, bool real_packed = false) {
  hipStream_t s = c->stream;
  StagePlan &S = *P.stages[k];
  Pool<R> A = pool_view<R>(c, c->cur);
  launch_updaters<R>(s, A, c->d_sd.as<SysDesc>(), c->d_st.as<StepSys>(), S.blk_upd.as<Blk>(),
                     S.nblk_upd, S.urange.as<int2>(), S.upd.as<UpdDesc>(), c->D, S.upd_kinds);
  launch_ghost_copy<R>(s, A, c->d_sd.as<SysDesc>(), c->d_bnd.as<Bnd>(), c->d_n.as<int>(),
                       S.blk_gcopy.as<Blk>(), S.nblk_gcopy, S.gcopy.as<int2>(), c->D);
  if (mid) HIPCHK(hipEventRecord(*mid, s));
  // a stage whose only sweeps are FP32 force and strain-rate sweeps reads the
  // compact records alone
  const bool c_only = sizeof(R) == 4 && S.fam[FAM_INTERP].nblk == 0 &&
                      S.fam[FAM_COUNT].nblk == 0;
  // the partner stress records only for stages with Cauchy force sweeps
  const bool with_s = S.fam[FAM_FORCE].cauchy;
  // records of 1/ρ and p/ρ² when the stage's only sweeps are FluidPfn S = 2
  // force sweeps without δ-SPH or surface tension
  const bool inv = c_only && S.fam[FAM_FORCE].inv && S.fam[FAM_STRAIN].nblk == 0;
  // the compact records of the real systems packed for the previous stage
  // still hold when they are of the same kind and nothing in between wrote
  // x, v, ρ or p: the stage's only updaters are Drucker-Prager (stress and
  // strains) and no virtual system is interpolated; then only the stress
  // records are packed
  auto rec_kind = [](const StagePlan &T) {  // 0: no compact records
    const int nsw = T.fam[FAM_FORCE].nblk + T.fam[FAM_COUNT].nblk + T.fam[FAM_STRAIN].nblk +
                    T.fam[FAM_INTERP].nblk;
    if (sizeof(R) != 4 || nsw == 0 || T.fam[FAM_INTERP].nblk || T.fam[FAM_COUNT].nblk) return 0;
    return T.fam[FAM_FORCE].inv && T.fam[FAM_STRAIN].nblk == 0 ? 2 : 1;
  };
  const bool keep_c = k > 0 && P.nblk_virt == 0 && rec_kind(S) != 0 &&
                      rec_kind(S) == rec_kind(*P.stages[k - 1]) &&
                      (S.upd_kinds & ~(1 << GR_UPD_EP | 1 << GR_UPD_NOTHING)) == 0;
  auto pack = [&]() {
    if (!real_packed)
      launch_pack<R>(s, A, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), P.blk_real.as<Blk>(),
                     P.nblk_real, c->D, c_only, with_s, inv, keep_c);
    launch_pack<R>(s, A, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), P.blk_ghost.as<Blk>(),
                   P.nblk_ghost, c->D, c_only, with_s, inv);
  };
  const int nother = S.fam[FAM_FORCE].nblk + S.fam[FAM_COUNT].nblk + S.fam[FAM_STRAIN].nblk;
  if (nother + S.fam[FAM_INTERP].nblk > 0) pack();
  // interpolation into virtual systems runs first; the functors of the same
  // stage then see the interpolated fields
  for (int fam : {FAM_INTERP, FAM_FORCE, FAM_COUNT, FAM_STRAIN}) {
    FamPlan &F = S.fam[fam];
    if (F.nblk == 0) continue;
    if (fam == FAM_FORCE && S.fam[FAM_INTERP].nblk > 0) pack();
    if (fam == FAM_COUNT && S.fam[FAM_INTERP].nblk > 0 && S.fam[FAM_FORCE].nblk == 0) pack();
    if (fam == FAM_STRAIN && S.fam[FAM_INTERP].nblk > 0 && S.fam[FAM_FORCE].nblk == 0 &&
        S.fam[FAM_COUNT].nblk == 0)
      pack();
    SweepArgs a;
    a.sd = c->d_sd.as<SysDesc>();
    a.n = c->d_n.as<int>();
    a.tasks = F.tasks.as<Task>();
    a.srcs = F.srcs.as<SrcDesc>();
    a.grids = c->d_grids.as<Grid>();
    a.carr = c->d_carr.as<CellArr>();
    a.cells = c->cells.as<int>();
    a.blk = F.blk.as<Blk>();
    a.nblk = F.nblk;
    a.pair_count = c->d_pairs.as<unsigned long long>();
    if (P.nblk_list > 0) {
      a.lt = P.ltasks.as<LTask>();
      a.lcnt = c->lcnt.as<int>();
      a.lidx = c->lidx.as<int>();
    }
    a.fluid_only = F.fluid_only;
    a.inv = inv && fam == FAM_FORCE;
    a.block = P.sweep_block;
    launch_sweep<R>(s, A, a, c->D, fam, cnt);
  }
}

// _xsph_correction! (src/TimeIntegration.jl:423-436) between the full step
// and the position update, which it performs.
template <typename R>
void phase_xsph(gr_ctx *c, Plan &P) {
  hipStream_t s = c->stream;
  Pool<R> A = pool_view<R>(c, c->cur);
  const SysDesc *sd = c->d_sd.as<SysDesc>();
  const StepSys *st = c->d_st.as<StepSys>();
  launch_xsph<R>(s, A, sd, st, c->d_n.as<int>(), P.blk_int.as<Blk>(), P.nblk_int, c->D, 0);
  if (P.xsph.nblk > 0 || !P.xsph_al.empty()) {
    launch_xsph<R>(s, A, sd, st, c->d_n.as<int>(), P.blk_virt.as<Blk>(), P.nblk_virt, c->D, 1);
    launch_pack<R>(s, A, sd, c->d_n.as<int>(), P.blk_real.as<Blk>(), P.nblk_real, c->D);
    if (!P.xsph_alias.empty())
      launch_pack<R>(s, A, sd, c->d_n.as<int>(), P.blk_ghost.as<Blk>(), P.nblk_ghost, c->D);
    SweepArgs a;
    a.sd = sd;
    a.n = c->d_n.as<int>();
    a.tasks = P.xsph.tasks.as<Task>();
    a.srcs = P.xsph.srcs.as<SrcDesc>();
    a.grids = c->d_grids.as<Grid>();
    a.carr = c->d_carr.as<CellArr>();
    a.cells = c->cells.as<int>();
    a.blk = P.xsph.blk.as<Blk>();
    a.nblk = P.xsph.nblk;
    a.block = P.sweep_block;
    a.pair_count = c->d_pairs.as<unsigned long long>();
    if (P.nblk_list > 0) {
      a.lt = P.ltasks.as<LTask>();
      a.lcnt = c->lcnt.as<int>();
      a.lidx = c->lidx.as<int>();
    }
    launch_sweep<R>(s, A, a, c->D, FAM_XSPH, false);
    for (auto &F : P.xsph_al) {
      a.tasks = F->tasks.as<Task>();
      a.srcs = F->srcs.as<SrcDesc>();
      a.blk = F->blk.as<Blk>();
      a.nblk = F->nblk;
      launch_sweep<R>(s, A, a, c->D, FAM_XSPH, false);
    }
    for (const int2 &g : P.xsph_alias)
      launch_alias_check(s, c->d_n.as<int>(), g.x, g.y, c->d_flags.as<int>());
  }
  launch_xsph<R>(s, A, sd, st, c->d_n.as<int>(), P.blk_int.as<Blk>(), P.nblk_int, c->D, 2);
}

// Steps between two checks of the cell arrays' capacity, and the cells kept
// free on every side of the systems' current extent.  With dt = CFL·h/c a
// particle at speed v moves CFL·(v/c)·h per step: a chunk moves a subsonic
// particle (v < c, CFL <= 1) at most 16 h = 8 cells.
constexpr int kGrowChunk = 16;
constexpr int kGrowMargin = 8;

// The bounds of the last key pass (replica 0 of d_bbmin / d_bbmax, exact
// doubles) against each interaction's cell arrays: an interaction whose grid,
// kGrowMargin cells wider on every side, would not fit gets new arrays sized
// for its current extent (cell_capacity).  Cell arrays are rebuilt every
// step, so nothing is copied; plans and captured graphs refer to the tables
// and are dropped.  Returns whether anything was reallocated.
bool grow_cells(gr_ctx *c, const Plan &P) {
  if (P.inters.empty()) return false;
  const int ns = (int)c->sys.size();
  std::vector<long long> bmn((size_t)ns * 3), bmx((size_t)ns * 3);
  HIPCHK(hipMemcpyAsync(bmn.data(), c->d_bbmin.p, bmn.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(bmx.data(), c->d_bbmax.p, bmx.size() * 8, hipMemcpyDeviceToHost, c->stream));
  std::vector<int> nn(ns);
  HIPCHK(hipMemcpyAsync(nn.data(), c->d_n.p, ns * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  auto dbl = [](long long i) {  // ordered int64 -> double (ord_of's inverse)
    const long long b = i < 0 ? (i ^ 0x7FFFFFFFFFFFFFFFLL) : i;
    double d;
    memcpy(&d, &b, 8);
    return d;
  };
  bool grew = false;
  for (int it : P.inters) {
    InterH &in = c->inter[it];
    if (in.carr_b < 0) continue;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    bool any = false, finite = true;
    for (int sid : {in.a, in.b}) {
      if (sid < 0 || nn[sid] <= 0) continue;
      for (int d = 0; d < c->D; ++d) {
        const double lo = dbl(bmn[sid * 3 + d]), hi = dbl(bmx[sid * 3 + d]);
        finite = finite && std::isfinite(lo) && std::isfinite(hi);
        mn[d] = any ? std::min(mn[d], lo) : lo;
        mx[d] = any ? std::max(mx[d], hi) : hi;
      }
      any = true;
    }
    if (!any || !finite) continue;  // a faulted system is reported, not grown for
    const double cell = in.cell;
    double need = 1.0;
    for (int d = 0; d < c->D; ++d) {
      const double ming = std::floor((mn[d] - 2.0 * cell) / cell) * cell;
      need *= std::floor((mx[d] + 2.0 * cell - ming) / cell) + 1.0 + 2.0 * kGrowMargin;
    }
    int64_t cap = c->carr[in.carr_b].cap;
    if (in.carr_a >= 0) cap = std::min<int64_t>(cap, c->carr[in.carr_a].cap);
    if (need + 1.0 <= (double)cap) continue;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = d < c->D ? mn[d] - kGrowMargin * cell : 0.0;
      hi[d] = d < c->D ? mx[d] + kGrowMargin * cell : 0.0;
    }
    const int64_t ncap = cell_capacity(c, in, lo, hi);
    if (ncap <= cap) continue;
    for (int ca : {in.carr_b, in.carr_a}) {
      if (ca < 0) continue;
      CellArr &a = c->carr[ca];
      c->free_cells.put(a.off, a.cap);
      a.off = c->free_cells.take(ncap);
      if (a.off < 0) {
        a.off = c->cells_cap;
        c->cells_cap += ncap;
      }
      a.cap = (int32_t)ncap;
    }
    grew = true;
  }
  if (!grew) return false;
  drop_graphs(c);
  c->plans.clear();
  if ((size_t)std::max<int64_t>(c->cells_cap, 1) * 4 > c->cells.bytes)
    c->cells.alloc(std::max<int64_t>(c->cells_cap, 1) * 4);
  c->d_carr.upload(c->carr, c->stream);
  std::fill(c->carr_valid.begin(), c->carr_valid.end(), 0);
  return true;
}

// Phase events of one step (timing mode): 0 start, 1 sorted, 2 ghosts,
// 3 grids, 4 half-step, then per stage (updaters end, sweeps end), last
// full-step end.  Recorded without synchronisation; read after the call.
std::vector<hipEvent_t> &step_events(gr_ctx *c, int k, int nst) {
  if ((int)c->tev.size() <= k) c->tev.resize(k + 1);
  auto &v = c->tev[k];
  const size_t need = 7 + 2 * nst;
  while (v.size() < need) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    v.push_back(e);
  }
  return v;
}

template <typename R>
void run_step(gr_ctx *c, Plan &P, bool cnt, std::vector<hipEvent_t> *ev) {
  hipStream_t s = c->stream;
  auto rec = [&](size_t q) {
    if (ev) HIPCHK(hipEventRecord((*ev)[q], s));
  };
  rec(0);
  // the leapfrog half step changes v and ρ only, which nothing reads before
  // it (ghost generation, grids and lists read positions): it rides the
  // sort's permutation
  const bool half_in_sort =
      c->tune_fuse && P.scheme != GR_SCHEME_RK4 && !P.real_list.empty();
  // and so can stage 0's compact partner records of the real systems, when
  // no updater or virtual system changes their fields in between (the
  // choice of records follows phase_stage's)
  int pack_in_sort = 0;
  if (half_in_sort && sizeof(R) == 4 && !P.stages.empty() && P.nblk_virt == 0) {
    const StagePlan &S0 = *P.stages[0];
    const bool c_only = S0.fam[FAM_INTERP].nblk == 0 && S0.fam[FAM_COUNT].nblk == 0;
    const bool inv = c_only && S0.fam[FAM_FORCE].inv && S0.fam[FAM_STRAIN].nblk == 0;
    if (c_only && !S0.fam[FAM_FORCE].cauchy && S0.nblk_upd == 0 &&
        S0.fam[FAM_FORCE].nblk + S0.fam[FAM_STRAIN].nblk > 0)
      pack_in_sort = inv ? 2 : 1;
  }
  phase_sort_real<R>(c, P, half_in_sort, pack_in_sort);
  rec(1);
  phase_ghosts<R>(c, P, true, true);
  rec(2);
  phase_grids(c, P);
  rec(3);
  Pool<R> A = pool_view<R>(c, c->cur);
  if (P.nblk_list > 0) {
    NlistArgs a;
    a.sd = c->d_sd.as<SysDesc>();
    a.n = c->d_n.as<int>();
    a.lt = P.ltasks.as<LTask>();
    a.srcs = P.lsrcs.as<SrcDesc>();
    a.grids = c->d_grids.as<Grid>();
    a.carr = c->d_carr.as<CellArr>();
    a.cells = c->cells.as<int>();
    a.blk = P.blk_list.as<Blk>();
    a.nblk = P.nblk_list;
    a.lcnt = c->lcnt.as<int>();
    a.lidx = c->lidx.as<int>();
    a.bbmin = c->d_bbmin.as<long long>();
    a.bbmax = c->d_bbmax.as<long long>();
    a.win = c->tune_win;
    a.span = c->tune_span;
    launch_nlist<R>(s, A, a, c->D);
  }
  rec(4);
  const int nst = (int)P.stages.size();
  const SysDesc *sd = c->d_sd.as<SysDesc>();
  const StepSys *st = c->d_st.as<StepSys>();
  // the stage loop of one pass (_sweep_all_stages!) after the stage state
  auto pass = [&]() {
    launch_ghost_kin<R>(s, A, sd, c->d_bnd.as<Bnd>(), c->d_n.as<int>(), P.blk_ghost.as<Blk>(),
                        P.nblk_ghost, c->D);
    launch_virtual<R>(s, A, sd, st, c->d_n.as<int>(), P.blk_virt.as<Blk>(), P.nblk_virt, 0, c->D);
    rec(5);
    for (int k = 0; k < nst; ++k) {
      phase_stage<R>(c, P, k, cnt, ev ? &(*ev)[6 + 2 * k] : nullptr,
                     k == 0 && pack_in_sort != 0);
      rec(7 + 2 * k);
    }
  };
  if (P.scheme == GR_SCHEME_RK4) {  // src/TimeIntegration.jl:789-815
    const double adv[4] = {0.0, 0.5, 0.5, 1.0};
    const double wts[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    for (int r = 0; r < 4; ++r) {
      launch_halfstep<R>(s, A, sd, st, P.blk_int.as<Blk>(), P.nblk_int, c->D, adv[r], r == 0);
      pass();
      launch_rk4_acc<R>(s, A, sd, st, P.blk_int.as<Blk>(), P.nblk_int, c->D, wts[r], r == 0);
    }
    launch_rk4_apply<R>(s, A, sd, st, P.blk_int.as<Blk>(), P.nblk_int, c->D, !P.vadj_live);
  } else {
    if (!half_in_sort)
      launch_halfstep<R>(s, A, sd, st, P.blk_int.as<Blk>(), P.nblk_int, c->D, 0.5, 0);
    pass();
    launch_fullstep<R>(s, A, sd, st, P.blk_int.as<Blk>(), P.nblk_int, c->D, !P.vadj_live);
  }
  if (P.vadj_live) phase_xsph<R>(c, P);
  launch_virtual<R>(s, A, sd, st, c->d_n.as<int>(), P.blk_virt.as<Blk>(), P.nblk_virt, 1, c->D);
  launch_virtual<R>(s, A, sd, st, c->d_n.as<int>(), P.blk_probe.as<Blk>(), P.nblk_probe, 1, c->D);
  rec(6 + 2 * nst);
}

// Sum the phase times of `nsteps` recorded steps (ms per step).
void collect_timing(gr_ctx *c, int nsteps, int nst) {
  double acc[9] = {0};
  for (int k = 0; k < nsteps; ++k) {
    auto &v = c->tev[k];
    HIPCHK(hipEventSynchronize(v[6 + 2 * nst]));
    auto el = [&](size_t a, size_t b) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, v[a], v[b]));
      return (double)ms;
    };
    acc[0] += el(0, 1);
    acc[1] += el(1, 2);
    acc[2] += el(2, 3);
    acc[8] += el(3, 4);
    acc[3] += el(4, 5);
    size_t prev = 5;
    for (int q = 0; q < nst; ++q) {
      acc[4] += el(prev, 6 + 2 * q);
      acc[5] += el(6 + 2 * q, 7 + 2 * q);
      prev = 7 + 2 * q;
    }
    acc[6] += el(prev, 6 + 2 * nst);
    acc[7] += el(0, 6 + 2 * nst);
  }
  for (int q = 0; q < 9; ++q) c->phase_ms[q] = acc[q] / std::max(nsteps, 1);
}

// Graph-captured step: two executables, one per A/B parity of the pool.
template <typename R>
void step_graph(gr_ctx *c, Plan &P) {
  hipGraphExec_t &g = P.graph[c->cur];
  if (!g) {
    const int cur0 = c->cur;
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    run_step<R>(c, P, false, nullptr);
    HIPCHK(hipStreamEndCapture(c->stream, &graph));
    HIPCHK(hipGraphInstantiate(&g, graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
    c->cur = cur0;  // capture did not execute; restore the parity
  }
  HIPCHK(hipGraphLaunch(g, c->stream));
  c->cur = 1 - c->cur;  // one real-sort swap per step
}

template <typename R>
void do_steps(gr_ctx *c, const std::vector<int> &integs, int nsteps) {
  Plan *P = &get_plan(c, integs);
  // the plan's step parameters (fine-grained calls may have changed them)
  for (auto &e : P->stepsys) c->stepsys[e.first] = e.second;
  upload_stepsys(c);
  {
    const Blk *b;
    int nb;
    all_real(c, b, nb);  // build cached tables before any capture
  }
  if (c->count_pairs) launch_fill_i32(c->stream, c->d_pairs.as<int>(), 0, 4);
  if (nsteps > 0) mark_grids(c, *P);
  const int nst = (int)P->stages.size();
  // chunks of at most kGrowChunk steps; between chunks the cell arrays
  // follow the systems' spread (grow_cells), so a flow that runs out over
  // thousands of steps never outgrows its grids
  for (int k0 = 0; k0 < nsteps; k0 += kGrowChunk) {
    if ((k0 > 0 || c->bb_valid) && grow_cells(c, *P)) {
      P = &get_plan(c, integs);
      mark_grids(c, *P);
    }
    const int k1 = std::min(nsteps, k0 + kGrowChunk);
    for (int k = k0; k < k1; ++k) {
      if (c->timing)
        run_step<R>(c, *P, c->count_pairs, &step_events(c, k, nst));
      else if (c->graph_mode && !c->count_pairs && !P->real_list.empty())
        step_graph<R>(c, *P);
      else
        run_step<R>(c, *P, c->count_pairs, nullptr);
    }
    c->bb_valid = !P->real_list.empty();
  }
  int flags[4];
  HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->timing) collect_timing(c, nsteps, nst);
  if (c->count_pairs) {
    unsigned long long pc[2];
    HIPCHK(hipMemcpyAsync(pc, c->d_pairs.p, sizeof(pc), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last_pairs[0] = (long long)(pc[0] / std::max(nsteps, 1));
    c->last_pairs[1] = (long long)(pc[1] / std::max(nsteps, 1));
  }
  std::string faulted;
  if (flags[1] & (2 | 8)) {
    // per-system faults → the integrators they belong to (the kernels have
    // isolated those systems; every other system stepped normally)
    const size_t ns = c->sys.size();
    std::vector<int> sf(ns, 0);
    HIPCHK(hipMemcpyAsync(sf.data(), c->d_sysfault.p, ns * sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemsetAsync(c->d_sysfault.p, 0, ns * sizeof(int), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int ig : integs) {
      IntegH &I = c->integ[ig];
      int f = 0;
      for (const auto *lst : {&I.sys, &I.ghost, &I.virt, &I.probes})
        for (int q : *lst)
          if (q >= 0 && q < (int)ns) f |= sf[q];
      if (f) {
        I.fault |= f;
        faulted += (faulted.empty() ? "" : ", ") + std::to_string(ig);
      }
    }
    if (!faulted.empty()) flags[1] &= ~(2 | 8);
  }
  if (flags[1] || !faulted.empty()) {
    int z[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(c->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (flags[1]) {
    if (flags[1] & 16)
      throw GrError(GR_ERR_HIP, "device radix sort: the tile counts did not place every key");
    if (flags[1] & 4)
      throw GrError(GR_ERR_ARGUMENT,
                    "XSPHPfn with a ghost partner: more ghosts than source particles "
                    "(ps_b.v_adjustment[j] would index past the source's array)");
    throw GrError(GR_ERR_CAPACITY,
                  (flags[1] & 2) ? "cell grid outgrew its capacity (particles left the planned domain)"
                                 : "cell-key range exceeds 32 bits");
  }
  if (!faulted.empty())
    throw GrError(GR_ERR_FAULT,
                  "integrator(s) " + faulted +
                      " faulted (non-finite positions or particles beyond the cell range); "
                      "the others advanced normally");
}

void validate_sys(gr_ctx *c, int s) {
  need(s >= 0 && s < (int)c->sys.size(), GR_ERR_ARGUMENT, "bad system handle");
  need(!c->sys[s].dead, GR_ERR_STATE, "system was released");
}

template <typename R>
void refresh_shadows(gr_ctx *c, const std::vector<int> &systems);

}  // namespace

// ===========================================================================
// Shadow refresh kernel (fine-grained calls see the uploaded state)
// ===========================================================================

template <typename R>
__global__ void k_refresh(Pool<R> P, const SysDesc *sd, const Blk *blk, int D) {
  const Blk b = blk[blockIdx.x];
  const SysDesc S = sd[b.item];
  const int i = b.first + threadIdx.x;
  if (i >= S.cap) return;
  const int64_t g = S.off + i;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d < D) P.vs[d][g] = (R)P.v[d][g];
  P.rhos[g] = (R)P.rho[g];
  P.rho_h[g] = P.rho[g];
  if (S.kind >= GR_SYS_STRESS) {
#pragma unroll
    for (int k = 0; k < GR_MAXNS; ++k)
      if (k < S.ns) P.stress_s[k][g] = (R)P.stress[k][g];
  }
}

namespace {
template <typename R>
void refresh_shadows(gr_ctx *c, const std::vector<int> &systems) {
  std::vector<Blk> v;
  for (int s : systems)
    if (c->sys[s].kind != GR_SYS_GHOST) add_blocks(v, s, c->sys[s].cap, GR_BLOCK);
  if (v.empty()) return;
  DBuf b;
  b.upload(v, c->stream);
  k_refresh<R><<<(unsigned)v.size(), GR_BLOCK, 0, c->stream>>>(pool_view<R>(c, c->cur),
                                                              c->d_sd.as<SysDesc>(), b.as<Blk>(), c->D);
  HIPCHK(hipStreamSynchronize(c->stream));
}

// Partner records of the listed systems (real and ghost) before a
// fine-grained sweep.
template <typename R>
void pack_systems(gr_ctx *c, const std::vector<int> &systems) {
  std::vector<Blk> v;
  for (int s : systems) add_blocks(v, s, c->sys[s].cap, GR_BLOCK);
  if (v.empty()) return;
  DBuf b;
  b.upload(v, c->stream);
  launch_pack<R>(c->stream, pool_view<R>(c, c->cur), c->d_sd.as<SysDesc>(), c->d_n.as<int>(),
                 b.as<Blk>(), (int)v.size(), c->D);
  HIPCHK(hipStreamSynchronize(c->stream));
}

template <typename F>
void by_prec(gr_ctx *c, F &&f) {
  if (c->prec == GR_PREC_F64)
    f((double *)nullptr);
  else
    f((float *)nullptr);
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

static thread_local std::string g_create_err;

extern "C" {

int gr_ctx_create(gr_ctx **out, int device, int ndims, int precision) {
  if (!out) return GR_ERR_ARGUMENT;
  *out = nullptr;
  if (ndims != 2 && ndims != 3) {
    g_create_err = "ndims must be 2 or 3";
    return GR_ERR_ARGUMENT;
  }
  if (precision != GR_PREC_F32 && precision != GR_PREC_F64) {
    g_create_err = "precision must be GR_PREC_F32 or GR_PREC_F64";
    return GR_ERR_ARGUMENT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    g_create_err = "no HIP device " + std::to_string(device);
    return GR_ERR_HIP;
  }
  if (hipSetDevice(device) != hipSuccess) {
    g_create_err = "hipSetDevice failed";
    return GR_ERR_HIP;
  }
  if (hipError_t e = scratch_setup(device); e != hipSuccess) {
    g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return GR_ERR_HIP;
  }
  gr_ctx *c = new gr_ctx();
  c->device = device;
  c->D = ndims;
  c->prec = precision;
  c->rsize = precision == GR_PREC_F64 ? 8 : 4;
  c->nsmax = ndims == 2 ? 4 : 6;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    g_create_err = "hipStreamCreate failed";
    return GR_ERR_HIP;
  }
  {
    hipMemPoolProps pp;
    memset(&pp, 0, sizeof(pp));
    pp.allocType = hipMemAllocationTypePinned;
    pp.handleTypes = hipMemHandleTypeNone;
    pp.location.type = hipMemLocationTypeDevice;
    pp.location.id = device;
    hipMemPool_t pool = nullptr;
    if (hipMemPoolCreate(&pool, &pp) == hipSuccess) {
      c->cache.pool = pool;
      c->cache.stream = c->stream;
    } else {
      (void)hipGetLastError();  // fall back to hipMalloc / hipFree
    }
  }
  for (auto &e : c->ev) (void)hipEventCreate(&e);
  // GRASPH_NO_NLIST=1: walk the cells in every sweep; GRASPH_NLIST_MAXNB=k:
  // keep at most k partners per target (targets with more walk the cells).
  if (const char *e = getenv("GRASPH_NO_NLIST")) c->use_lists = atoi(e) ? 0 : 1;
  if (const char *e = getenv("GRASPH_NLIST_MAXNB")) c->list_maxnb = std::max(0, atoi(e));
  *out = c;
  return GR_OK;
}

int gr_ctx_destroy(gr_ctx *c) {
  if (!c) return GR_OK;
  (void)hipStreamSynchronize(c->stream);
  c->plans.clear();
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &v : c->tev)
    for (auto &e : v) (void)hipEventDestroy(e);
  hipStream_t s = c->stream;
  delete c;
  (void)hipStreamDestroy(s);
  return GR_OK;
}

const char *gr_last_error(const gr_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int gr_ctx_info(gr_ctx *c, char *buf, int64_t buflen) {
  return guard(c, [&] {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, c->device));
    std::string s = std::string(p.name) + " " + p.gcnArchName + " CUs=" +
                    std::to_string(p.multiProcessorCount) +
                    (c->prec == GR_PREC_F64 ? " prec=f64" : " prec=f32");
    snprintf(buf, (size_t)buflen, "%s", s.c_str());
  });
}

int gr_system_create(gr_ctx *c, int kind, int64_t n, int ns, double mass, double cc,
                     const double *source_v, double source_rho, int *sys_out) {
  return guard(c, [&] {
    need(n > 0, GR_ERR_ARGUMENT, "n must be positive, got " + std::to_string(n));
    need(n < (1LL << 30), GR_ERR_ARGUMENT, "n too large for one system");
    need(kind >= GR_SYS_BASIC && kind <= GR_SYS_EP, GR_ERR_ARGUMENT, "bad system kind");
    if (kind >= GR_SYS_STRESS)
      need(ns == 3 || ns == 4 || ns == 6, GR_ERR_ARGUMENT,
           "ns must be 3, 4, or 6, got " + std::to_string(ns));
    if (kind >= GR_SYS_STRESS)
      need((c->D == 2 && ns != 6) || (c->D == 3 && ns == 6), GR_ERR_ARGUMENT,
           "stress length must be 3/4 in 2D and 6 in 3D");
    SysH s;
    s.kind = kind;
    s.ns = kind >= GR_SYS_STRESS ? ns : 0;
    s.nv = kind == GR_SYS_EP ? (c->D == 2 ? 1 : 3) : 0;
    s.n = n;
    s.cap = n;
    s.mass = mass;
    s.c = cc;
    for (int d = 0; d < c->D; ++d) s.sv[d] = source_v ? source_v[d] : 0.0;
    s.srho = source_rho;
    c->sys.push_back(s);
    c->dirty_layout = true;
    *sys_out = (int)c->sys.size() - 1;
  });
}

int gr_system_add_updater(gr_ctx *c, int sys, int stage, int kind, const double *params,
                          int nparams) {
  return guard(c, [&] {
    validate_sys(c, sys);
    need(stage >= 1, GR_ERR_ARGUMENT, "stage must be >= 1");
    SysH &s = c->sys[sys];
    need(s.kind != GR_SYS_GHOST, GR_ERR_ARGUMENT, "ghost systems take copiers, not updaters");
    if (kind == GR_UPD_EP)
      need(s.ns == 4, GR_ERR_ASSERTION,
           "ElastoPlasticStressUpdater requires NS=4 (plane-strain Voigt). Got NS=" +
               std::to_string(s.ns) + ".");
    if (kind == GR_UPD_HOOKE)
      need(s.ns == 4, GR_ERR_ASSERTION,
           "HookeLawStressUpdater requires NS=4 (plane-strain Voigt). Got NS=" +
               std::to_string(s.ns) + ".");
    need(!(((kind == GR_UPD_EP || kind == GR_UPD_HOOKE) && s.kind != GR_SYS_EP) ||
           (kind == GR_UPD_VPMC && s.kind < GR_SYS_STRESS) ||
           ((kind == GR_UPD_TAIT || kind == GR_UPD_LINEAR) && s.kind < GR_SYS_FLUID) ||
           ((kind == GR_UPD_VNORM || kind == GR_UPD_PRESCRIBED_V) && !s.virt)),
         GR_ERR_ARGUMENT, "updater needs fields the system does not have");
    need(kind >= GR_UPD_NOTHING && kind <= GR_UPD_HOOKE, GR_ERR_ARGUMENT,
         "unknown updater kind");
    if ((int)s.upd.size() < stage) s.upd.resize(stage);
    UpdDesc u;
    memset(&u, 0, sizeof(u));
    u.kind = kind;
    double prm[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < nparams && k < 5; ++k) prm[k] = params[k];
    switch (kind) {
      case GR_UPD_ZERO:
        u.mask = (long long)prm[0];
        break;
      case GR_UPD_TAIT:
        u.prm[0] = prm[0];
        u.prm[1] = nparams >= 2 ? prm[1] : 7.0;
        break;
      case GR_UPD_LINEAR:
        u.prm[0] = prm[0];
        break;
      case GR_UPD_EP: {  // trig constants as src/StateUpdaters.jl:264-266,288
        const double phi = prm[2], psi = prm[3], coh = prm[4];
        const double sqrt3 = std::sqrt(3.0);
        u.prm[0] = prm[0];
        u.prm[1] = prm[1];
        u.prm[2] = 2 * std::sin(phi) / (sqrt3 * (3 - std::sin(phi)));
        u.prm[3] = 6 * coh * std::cos(phi) / (sqrt3 * (3 - std::sin(phi)));
        u.prm[4] = 2 * std::sin(psi) / (sqrt3 * (3 - std::sin(psi)));
        break;
      }
      case GR_UPD_VPMC:
        u.prm[0] = prm[0];
        u.prm[1] = prm[1];
        u.prm[2] = prm[2];
        u.prm[3] = std::tan(prm[3]);
        u.prm[4] = prm[4];
        need((int)prm[0] == GR_UPD_TAIT || (int)prm[0] == GR_UPD_LINEAR, GR_ERR_ARGUMENT,
             "ViscoPlasticMCStressUpdater needs a Tait or Linear EOS");
        break;
      case GR_UPD_HOOKE:
        u.prm[0] = prm[0];
        u.prm[1] = prm[1];
        break;
      case GR_UPD_VNORM:  // v_mult, field mask
        u.prm[0] = prm[0];
        u.prm[1] = prm[1];
        u.prm[2] = prm[2];
        u.mask = (long long)prm[3];
        break;
    }
    s.upd[stage - 1].push_back(u);
    c->plans.clear();
  });
}

int gr_system_upload(gr_ctx *c, int sys, int field, const double *data, int64_t count) {
  return guard(c, [&] {
    validate_sys(c, sys);
    commit(c);
    SysH &s = c->sys[sys];
    need(has_field(s, field) && field != GR_F_NORMALS, GR_ERR_ARGUMENT,
         "field not present on this system");
    const int nc = ncomp(c, s, field);
    const int64_t n = live_n(c, sys);
    need(count == n * nc, GR_ERR_ARGUMENT, "count mismatch");
    if (field == GR_F_V_ADJUSTMENT && !c->vadj_any) {
      bool nz = false;
      for (int64_t i = 0; i < count; ++i) nz = nz || data[i] != 0.0;
      if (nz) {  // from now on v_adjustment follows the sort and the correction runs
        c->vadj_any = true;
        c->plans.clear();
      }
    }
    const int64_t off = s.off;
    const size_t R = c->rsize;
    const int64_t cap = c->pool_cap;
    if (s.kind == GR_SYS_GHOST) {
      switch (field) {
        case GR_F_X:
          for (int k = 0; k < nc; ++k)
            copy_comp_to_dev(c, (char *)(c->x[c->cur].as<double>() + k * cap), 8, off, data, n, nc, k);
          break;
        case GR_F_V:
          for (int k = 0; k < nc; ++k)
            copy_comp_to_dev(c, (char *)c->vs.p + k * cap * R, R, off, data, n, nc, k);
          break;
        case GR_F_RHO:
          copy_comp_to_dev(c, (char *)c->rhos.p, R, off, data, n, 1, 0);
          break;
        case GR_F_P:
          copy_comp_to_dev(c, (char *)c->p[c->cur].p, R, off, data, n, 1, 0);
          break;
        case GR_F_STRESS:
          for (int k = 0; k < nc; ++k)
            copy_comp_to_dev(c, (char *)c->stress_s[c->cur].p + k * cap * R, R, off, data, n, nc, k);
          break;
      }
      return;
    }
    FieldRef f = field_ref(c, field, false);
    for (int k = 0; k < nc; ++k) copy_comp_to_dev(c, f.comp[k], f.esz, off, data, n, nc, k);
    // shadows follow their masters
    if (field == GR_F_V)
      for (int k = 0; k < nc; ++k) copy_comp_to_dev(c, (char *)c->vs.p + k * cap * R, R, off, data, n, nc, k);
    if (field == GR_F_RHO) {
      copy_comp_to_dev(c, (char *)c->rhos.p, R, off, data, n, 1, 0);
      copy_comp_to_dev(c, (char *)c->rho_h.p, 8, off, data, n, 1, 0);
    }
    if (field == GR_F_STRESS)
      for (int k = 0; k < nc; ++k)
        copy_comp_to_dev(c, (char *)c->stress_s[c->cur].p + k * cap * R, R, off, data, n, nc, k);
  });
}

int gr_system_download(gr_ctx *c, int sys, int field, double *data, int64_t count) {
  return guard(c, [&] {
    validate_sys(c, sys);
    commit(c);
    SysH &s = c->sys[sys];
    need(has_field(s, field), GR_ERR_ARGUMENT, "field not present on this system");
    const int nc = ncomp(c, s, field);
    const int64_t n = live_n(c, sys);
    need(count == n * nc, GR_ERR_ARGUMENT, "count mismatch");
    const int64_t off = s.off, cap = c->pool_cap;
    const size_t R = c->rsize;
    if (s.kind == GR_SYS_GHOST) {
      switch (field) {
        case GR_F_X:
          for (int k = 0; k < nc; ++k)
            copy_comp_from_dev(c, (char *)(c->x[c->cur].as<double>() + k * cap), 8, off, data, n, nc, k);
          break;
        case GR_F_V:
          for (int k = 0; k < nc; ++k)
            copy_comp_from_dev(c, (char *)c->vs.p + k * cap * R, R, off, data, n, nc, k);
          break;
        case GR_F_RHO:
          copy_comp_from_dev(c, (char *)c->rhos.p, R, off, data, n, 1, 0);
          break;
        case GR_F_P:
          copy_comp_from_dev(c, (char *)c->p[c->cur].p, R, off, data, n, 1, 0);
          break;
        case GR_F_STRESS:
          for (int k = 0; k < nc; ++k)
            copy_comp_from_dev(c, (char *)c->stress_s[c->cur].p + k * cap * R, R, off, data, n, nc, k);
          break;
        case GR_F_NORMALS: {
          auto nb = read_ints(c, c->gnrm, off, n);
          for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < nc; ++d) data[i * nc + d] = c->bnds[s.bnd + nb[i]].n[d];
          break;
        }
      }
      return;
    }
    FieldRef f = field_ref(c, field, false);
    for (int k = 0; k < nc; ++k) copy_comp_from_dev(c, f.comp[k], f.esz, off, data, n, nc, k);
  });
}

int gr_system_upload_index(gr_ctx *c, int sys, int field, const int64_t *data, int64_t count) {
  return guard(c, [&] {
    validate_sys(c, sys);
    commit(c);
    SysH &s = c->sys[sys];
    const int64_t n = live_n(c, sys);
    need(count == n, GR_ERR_ARGUMENT, "count mismatch");
    need((s.kind == GR_SYS_GHOST) == (field != GR_I_ID), GR_ERR_ARGUMENT, "index field not present");
    std::vector<int> t(n);
    for (int64_t i = 0; i < n; ++i) t[i] = (int)(field == GR_I_ID ? data[i] : data[i] - 1);
    int *dst = field == GR_I_ID ? c->id[c->cur].as<int>() : field == GR_I_IDX_ORIGINAL ? c->gidx.as<int>() : c->gbnd.as<int>();
    if (n) {
      HIPCHK(hipMemcpyAsync(dst + s.off, t.data(), n * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  });
}

int gr_system_download_index(gr_ctx *c, int sys, int field, int64_t *data, int64_t count) {
  return guard(c, [&] {
    validate_sys(c, sys);
    commit(c);
    SysH &s = c->sys[sys];
    const int64_t n = live_n(c, sys);
    need(count == n, GR_ERR_ARGUMENT, "count mismatch");
    need((s.kind == GR_SYS_GHOST) == (field != GR_I_ID), GR_ERR_ARGUMENT, "index field not present");
    const DBuf &b = field == GR_I_ID ? c->id[c->cur] : field == GR_I_IDX_ORIGINAL ? c->gidx : c->gbnd;
    auto t = read_ints(c, b, s.off, n);
    for (int64_t i = 0; i < n; ++i) data[i] = field == GR_I_ID ? t[i] : t[i] + 1;
  });
}

int gr_system_get_n(gr_ctx *c, int sys, int64_t *n_out) {
  return guard(c, [&] {
    validate_sys(c, sys);
    *n_out = live_n(c, sys);
  });
}

int gr_virtual_create(gr_ctx *c, int sys, int nzero, const int *zero_fields,
                      const double *prescribed_v) {
  return guard(c, [&] {
    validate_sys(c, sys);
    SysH &s = c->sys[sys];
    need(s.kind != GR_SYS_GHOST && !s.virt, GR_ERR_ARGUMENT, "not a plain particle system");
    long long m = 0;
    for (int k = 0; k < nzero; ++k) {
      const int f = zero_fields[k];
      need(f >= 0 && f < GR_NUM_FIELDS && f != GR_F_X && f != GR_F_W_SUM && has_field(s, f),
           GR_ERR_ARGUMENT, "zero_fields: field not present on the virtual system");
      m |= 1LL << f;
    }
    s.virt = true;
    s.zero_mask = m;
    for (int d = 0; d < 3; ++d) s.pv[d] = (prescribed_v && d < c->D) ? prescribed_v[d] : 0.0;
    s.upd.clear();  // the virtual system's own updaters follow
    c->dirty_layout = true;
    c->plans.clear();
  });
}

int gr_probe_create(gr_ctx *c, int sys, const double *prescribed_v, int mirror, int nzero,
                    const int *zero_fields) {
  return guard(c, [&] {
    validate_sys(c, sys);
    SysH &s = c->sys[sys];
    need(s.kind != GR_SYS_GHOST && !s.virt, GR_ERR_ARGUMENT, "not a plain particle system");
    if (mirror >= 0) {
      validate_sys(c, mirror);
      need(c->sys[mirror].kind != GR_SYS_GHOST && !c->sys[mirror].virt, GR_ERR_ARGUMENT,
           "mirror_target must be a real particle system");
      need(c->sys[mirror].n == s.n, GR_ERR_ARGUMENT,
           "a mirroring probe has one slot per target particle");
    }
    s.probe = true;  // has_field(GR_F_COUNT) for the zero_fields check
    const int rc = gr_virtual_create(c, sys, nzero, zero_fields, prescribed_v);
    if (rc) {
      s.probe = false;
      throw GrError(rc, c->err);
    }
    s.mirror = mirror;
  });
}

int gr_ghost_create(gr_ctx *c, int source_sys, double cutoff, int nb, const double *normals,
                    const double *points, int *ghost_out) {
  return guard(c, [&] {
    validate_sys(c, source_sys);
    need(c->sys[source_sys].kind != GR_SYS_GHOST, GR_ERR_ARGUMENT, "bad ghost source");
    need(nb >= 1, GR_ERR_ARGUMENT, "at least one boundary (normal, point) pair is required");
    SysH g;
    g.kind = GR_SYS_GHOST;
    g.src = source_sys;
    g.nb = nb;
    g.gcut = cutoff;
    g.cap = (int64_t)nb * c->sys[source_sys].cap;
    need(g.cap < (1LL << 30), GR_ERR_ARGUMENT, "ghost capacity too large");
    g.n = 0;
    g.ns = c->sys[source_sys].ns;
    g.mass = c->sys[source_sys].mass;
    g.c = c->sys[source_sys].c;
    g.bnd = (int)c->bnds.size();
    for (int b = 0; b < nb; ++b) {
      Bnd B;
      memset(&B, 0, sizeof(B));
      for (int d = 0; d < c->D; ++d) {
        B.n[d] = normals[b * c->D + d];
        B.p[d] = points[b * c->D + d];
      }
      c->bnds.push_back(B);
    }
    c->sys.push_back(g);
    c->dirty_layout = true;
    *ghost_out = (int)c->sys.size() - 1;
  });
}

int gr_ghost_set_copier(gr_ctx *c, int ghost, int stage, int nfields, const int *fields,
                        const int *modes) {
  return guard(c, [&] {
    validate_sys(c, ghost);
    SysH &g = c->sys[ghost];
    need(g.kind == GR_SYS_GHOST && stage >= 1, GR_ERR_ARGUMENT, "bad ghost/stage");
    const SysH &src = c->sys[g.src];
    if ((int)g.copy.size() < stage) g.copy.resize(stage, make_int2(0, -1));
    for (int k = 0; k < nfields; ++k) {
      const int f = fields[k];
      const int m = modes ? modes[k] : GR_COPY_PLAIN;
      if (f == GR_F_P) {
        need(src.kind >= GR_SYS_FLUID, GR_ERR_ARGUMENT, "source has no p");
        g.ex_p = true;
        g.copy[stage - 1].x = 1;
      } else if (f == GR_F_STRESS) {
        need(src.kind >= GR_SYS_STRESS, GR_ERR_ARGUMENT, "source has no stress");
        g.ex_stress = true;
        g.copy[stage - 1].y = m;
      } else {
        throw GrError(GR_ERR_ARGUMENT, "GhostCopier supports :p and :stress");
      }
    }
    c->plans.clear();
  });
}

int gr_interaction_create(gr_ctx *c, int kernel_kind, double h, int sys_a, int partner, int wrap,
                          const double *normal, const double *point, double beta,
                          double lj_cutoff, int n_stages, const gr_pfn_desc *pfns,
                          int *inter_out) {
  return guard(c, [&] {
    validate_sys(c, sys_a);
    need(partner < (int)c->sys.size(), GR_ERR_ARGUMENT, "bad system_b");
    need(kernel_kind == GR_KERNEL_CUBIC || kernel_kind == GR_KERNEL_WENDLAND_C2, GR_ERR_ARGUMENT,
         "bad kernel kind");
    need(c->sys[sys_a].kind != GR_SYS_GHOST, GR_ERR_ARGUMENT, "a ghost system cannot be system_a");
    need(wrap == GR_WRAP_NONE || (partner >= 0 && c->sys[partner].kind != GR_SYS_GHOST),
         GR_ERR_ARGUMENT, "boundary wrappers need a real partner system");
    InterH in;
    in.kk = kernel_kind;
    in.h = h;
    const int D = c->D;
    const double H = h, pi = 3.141592653589793;
    if (kernel_kind == GR_KERNEL_CUBIC)
      in.norm = D == 2 ? 10.0 / ((7.0 * pi) * (H * H)) : 1.0 / (pi * (H * H * H));
    else
      in.norm = D == 2 ? 7.0 / ((64.0 * pi) * (H * H)) : 21.0 / ((256.0 * pi) * (H * H * H));
    in.cell = 2.0 * h;
    in.a = sys_a;
    in.b = partner;
    in.wrap = wrap;
    for (int d = 0; d < D; ++d) {
      in.wn[d] = normal ? normal[d] : 0.0;
      in.wp[d] = point ? point[d] : 0.0;
    }
    in.beta = beta;
    in.ljcut = lj_cutoff;
    in.pfns.assign(pfns, pfns + n_stages);
    const SysH &A = c->sys[sys_a];
    for (const gr_pfn_desc &f : in.pfns) {
      if (f.kind == GR_PFN_NONE) continue;
      need(f.kind >= GR_PFN_FLUID && f.kind <= GR_PFN_FLUIDSOLID, GR_ERR_ARGUMENT,
           "unknown functor kind");
      need(f.kind != GR_PFN_FLUIDSOLID ||
               (partner >= 0 && wrap == GR_WRAP_NONE && c->sys[partner].kind != GR_SYS_GHOST &&
                !c->sys[partner].virt && A.kind >= GR_SYS_FLUID),
           GR_ERR_ARGUMENT, "FluidSolidPfn needs a fluid system_a and a plain solid system_b");
      need(f.kind != GR_PFN_XSPH, GR_ERR_ARGUMENT,
           "XSPHPfn runs as the interaction's velocity_adjust_pairwise_fn");
      if (f.kind == GR_PFN_FLUID)
        need(f.sigma == 1 || f.sigma == 2, GR_ERR_ARGUMENT, "FluidPfn sigma must be 1 or 2");
      need(!(f.kind == GR_PFN_FLUID && A.kind < GR_SYS_FLUID), GR_ERR_ARGUMENT, "FluidPfn needs p on system_a");
      need(!(f.kind == GR_PFN_CAUCHY && A.kind < GR_SYS_STRESS), GR_ERR_ARGUMENT, "CauchyFluidPfn needs stress on system_a");
      need(!(f.kind == GR_PFN_SR && A.kind < GR_SYS_STRESS), GR_ERR_ARGUMENT, "StrainRatePfn needs strain_rate on system_a");
      need(!(f.kind == GR_PFN_SRV && A.kind != GR_SYS_EP), GR_ERR_ARGUMENT, "StrainRateVorticityPfn needs vorticity on system_a");
      if (partner >= 0) {
        const SysH &B = c->sys[partner];
        const int pk = classify(c, in, f.kind);
        if (B.kind == GR_SYS_GHOST) {
          need(!(f.kind == GR_PFN_FLUID && !B.ex_p), GR_ERR_ARGUMENT,
               "ghost partner of FluidPfn must copy :p (GhostCopier)");
          need(!(f.kind == GR_PFN_CAUCHY && !B.ex_stress), GR_ERR_ARGUMENT,
               "ghost partner of CauchyFluidPfn must copy :stress (GhostCopier)");
        } else if (pk == P_TWO || pk == P_ONE) {
          need(!(f.kind == GR_PFN_FLUID && B.kind < GR_SYS_FLUID), GR_ERR_ARGUMENT, "FluidPfn partner needs p");
          need(!(f.kind == GR_PFN_CAUCHY && B.kind < GR_SYS_STRESS), GR_ERR_ARGUMENT, "CauchyFluidPfn partner needs stress");
          need(!(f.kind == GR_PFN_CAUCHY && B.ns != A.ns), GR_ERR_ARGUMENT, "stress length mismatch");
        }
      }
    }
    c->inter.push_back(in);
    *inter_out = (int)c->inter.size() - 1;
  });
}

int gr_interaction_set_vadjust(gr_ctx *c, int inter, double epsilon) {
  return guard(c, [&] {
    need(inter >= 0 && inter < (int)c->inter.size() && !c->inter[inter].dead, GR_ERR_ARGUMENT,
         "bad interaction");
    InterH &in = c->inter[inter];
    if (std::isnan(epsilon)) {  // velocity_adjust_pairwise_fn = nothing
      if (in.has_xsph) c->plans.clear();
      in.has_xsph = false;
      return;
    }
    // a ghost partner's share goes to its source's v_adjustment at the
    // ghost's index (Task.alias), as the reference's forwarded field does
    in.has_xsph = true;
    in.xsph_eps = epsilon;
    c->vadj_any = true;
    c->plans.clear();  // tasks, lists and captured graphs change
  });
}

int gr_integrator_create(gr_ctx *c, int nsys, const int *systems, int ninter, const int *inters,
                         int nghost, const int *ghosts, double gamma, int *integ_out) {
  return guard(c, [&] {
    need(nsys >= 1, GR_ERR_ARGUMENT, "systems must not be empty");
    need(ninter >= 1, GR_ERR_ARGUMENT, "interactions must not be empty");
    IntegH I;
    for (int k = 0; k < nsys; ++k) {
      validate_sys(c, systems[k]);
      need(c->sys[systems[k]].kind != GR_SYS_GHOST, GR_ERR_ARGUMENT,
           "ghost systems go in `ghosts`");
      (c->sys[systems[k]].virt ? I.virt : I.sys).push_back(systems[k]);
    }
    need(!I.sys.empty(), GR_ERR_ARGUMENT, "systems must not be empty");
    I.inter.assign(inters, inters + ninter);
    if (nghost > 0) I.ghost.assign(ghosts, ghosts + nghost);
    for (int g : I.ghost) {
      validate_sys(c, g);
      need(c->sys[g].kind == GR_SYS_GHOST, GR_ERR_ARGUMENT, "not a ghost system");
    }
    for (int t : I.inter) {
      need(t >= 0 && t < (int)c->inter.size(), GR_ERR_ARGUMENT, "bad interaction");
      const InterH &in = c->inter[t];
      auto owned = [&](int s) {
        return std::find(I.sys.begin(), I.sys.end(), s) != I.sys.end() ||
               std::find(I.virt.begin(), I.virt.end(), s) != I.virt.end() ||
               std::find(I.ghost.begin(), I.ghost.end(), s) != I.ghost.end();
      };
      need(owned(in.a) && (in.b < 0 || owned(in.b)), GR_ERR_ARGUMENT,
           "every system of an interaction must be in the integrator's systems or ghosts");
    }
    I.gamma = gamma;
    I.c = -INFINITY;
    for (int s : I.sys) I.c = std::max(I.c, c->sys[s].c);
    I.h = INFINITY;
    for (int t : I.inter) I.h = std::min(I.h, c->inter[t].h);
    I.nst = (int)c->inter[I.inter[0]].pfns.size();
    for (int t : I.inter)
      need((int)c->inter[t].pfns.size() == I.nst, GR_ERR_ASSERTION,
           "All interactions must have the same number of stages (pfns length)");
    c->integ.push_back(I);
    *integ_out = (int)c->integ.size() - 1;
  });
}

int gr_integrator_release(gr_ctx *c, int integ) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
    IntegH &I = c->integ[integ];
    need(!I.dead, GR_ERR_STATE, "integrator was released");
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    c->plans.clear();
    I.dead = true;
    auto free_sys = [&](int s) {
      SysH &S = c->sys[s];
      if (S.dead) return;
      S.dead = true;
      if (S.off >= 0) c->free_pool.put(S.off, S.cap);
    };
    for (int s : I.sys) free_sys(s);
    for (int v : I.virt) free_sys(v);
    for (int g : I.ghost) free_sys(g);
    for (int t : I.inter) {
      InterH &in = c->inter[t];
      if (in.dead) continue;
      in.dead = true;
      for (int ca : {in.carr_b, in.carr_a})
        if (ca >= 0) c->free_cells.put(c->carr[ca].off, c->carr[ca].cap);
    }
    c->allreal_cap = -1;  // rebuild the all-real-systems table
  });
}

int gr_integrator_set_scheme(gr_ctx *c, int integ, int scheme) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
    need(!c->integ[integ].dead, GR_ERR_STATE, "integrator was released");
    need(scheme == GR_SCHEME_LEAPFROG || scheme == GR_SCHEME_RK4, GR_ERR_ARGUMENT,
         "unknown time integration scheme");
    if (c->integ[integ].scheme != scheme) {
      c->integ[integ].scheme = scheme;
      drop_graphs(c);
      c->plans.clear();
    }
  });
}

int gr_integrator_dt(gr_ctx *c, int integ, double cfl, double *dt_out) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
    const IntegH &I = c->integ[integ];
    *dt_out = cfl * I.h / I.c;
  });
}

int gr_integrator_set_probes(gr_ctx *c, int integ, int nprobes, const int *probes, int ninters,
                             const int *inters) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size() && !c->integ[integ].dead, GR_ERR_ARGUMENT,
         "bad integrator");
    std::vector<int> pr(probes, probes + nprobes), it(inters, inters + ninters);
    for (int q : pr) {
      validate_sys(c, q);
      need(c->sys[q].probe, GR_ERR_ARGUMENT, "probes must be ProbeParticleSystems");
    }
    for (int t : it) {
      need(t >= 0 && t < (int)c->inter.size() && !c->inter[t].dead, GR_ERR_ARGUMENT,
           "bad interaction");
      const InterH &in = c->inter[t];
      need(in.b >= 0 && std::find(pr.begin(), pr.end(), in.b) != pr.end(), GR_ERR_ARGUMENT,
           "a probe interaction's system_b must be one of the probes");
    }
    c->integ[integ].probes = pr;
    c->integ[integ].probe_inter = it;
    c->plans.clear();
  });
}

int gr_integrator_set_dt(gr_ctx *c, int integ, double dt) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
    need(dt > 0, GR_ERR_ARGUMENT, "dt must be positive");
    if (c->integ[integ].dt != dt) {
      c->integ[integ].dt = dt;
      c->plans.clear();
    }
  });
}

int gr_step(gr_ctx *c, int integ, int nsteps, double dt) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
    need(nsteps >= 0, GR_ERR_ARGUMENT, "nsteps must be >= 0");
    need(dt > 0, GR_ERR_ARGUMENT, "dt must be positive");
    if (c->integ[integ].dt != dt) {
      c->integ[integ].dt = dt;
      c->plans.clear();
    }
    std::vector<int> b{integ};
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      do_steps<R>(c, b, nsteps);
    });
  });
}

int gr_step_many(gr_ctx *c, int ninteg, const int *integs, int nsteps) {
  return guard(c, [&] {
    std::vector<int> b(integs, integs + ninteg);
    for (int ig : b) {
      need(ig >= 0 && ig < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
      need(c->integ[ig].dt > 0, GR_ERR_STATE, "integrator has no dt (gr_integrator_set_dt)");
    }
    std::vector<int> sorted = b;
    std::sort(sorted.begin(), sorted.end());
    need(std::unique(sorted.begin(), sorted.end()) == sorted.end(), GR_ERR_ARGUMENT, "duplicate integrator");
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      do_steps<R>(c, b, nsteps);
    });
  });
}

int gr_integrator_status(gr_ctx *c, int integ, int *fault_out) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size(), GR_ERR_ARGUMENT, "bad integrator");
    need(fault_out != nullptr, GR_ERR_ARGUMENT, "fault_out is NULL");
    *fault_out = c->integ[integ].fault;
  });
}

int gr_tune_set(gr_ctx *c, int param, int value) {
  return guard(c, [&] {
    switch (param) {
      case GR_TUNE_NL_WIN:
        need(value == 64 || value == 128 || value == 256 || value == 512, GR_ERR_ARGUMENT,
             "GR_TUNE_NL_WIN must be 64, 128, 256 or 512");
        c->tune_win = value;
        break;
      case GR_TUNE_NL_SPAN:
        need(value >= 1 && value <= 16, GR_ERR_ARGUMENT, "GR_TUNE_NL_SPAN must be in 1..16");
        c->tune_span = value;
        break;
      case GR_TUNE_SWEEP_BLOCK:
        need(value == 64 || value == 128 || value == 256, GR_ERR_ARGUMENT,
             "GR_TUNE_SWEEP_BLOCK must be 64, 128 or 256");
        c->tune_block = value;
        break;
      case GR_TUNE_MAXNB:
        need(value == 0 || (value >= 4 && value <= 1024), GR_ERR_ARGUMENT,
             "GR_TUNE_MAXNB must be 0 (default) or in 4..1024");
        c->list_maxnb = value;
        break;
      case GR_TUNE_INV_RECORDS:
        need(value == 0 || value == 1, GR_ERR_ARGUMENT, "GR_TUNE_INV_RECORDS must be 0 or 1");
        c->tune_inv = value;
        break;
      case GR_TUNE_FUSE_SORT:
        need(value == 0 || value == 1, GR_ERR_ARGUMENT, "GR_TUNE_FUSE_SORT must be 0 or 1");
        c->tune_fuse = value;
        break;
      default:
        throw GrError(GR_ERR_ARGUMENT, "unknown tuning parameter");
    }
    c->plans.clear();  // block tables, list layout and captured graphs change
  });
}

int gr_tune_get(gr_ctx *c, int param, int *value) {
  return guard(c, [&] {
    need(value != nullptr, GR_ERR_ARGUMENT, "value is NULL");
    switch (param) {
      case GR_TUNE_NL_WIN: *value = c->tune_win; break;
      case GR_TUNE_NL_SPAN: *value = c->tune_span; break;
      case GR_TUNE_SWEEP_BLOCK: *value = c->tune_block; break;
      case GR_TUNE_MAXNB: *value = c->list_maxnb > 0 ? c->list_maxnb : (c->D == 3 ? 48 : 24); break;
      case GR_TUNE_INV_RECORDS: *value = c->tune_inv; break;
      case GR_TUNE_FUSE_SORT: *value = c->tune_fuse; break;
      default: throw GrError(GR_ERR_ARGUMENT, "unknown tuning parameter");
    }
  });
}

int gr_set_graph_mode(gr_ctx *c, int on) {
  return guard(c, [&] { c->graph_mode = on != 0; });
}
int gr_set_timing(gr_ctx *c, int on) {
  return guard(c, [&] {
    c->timing = (on & 1) != 0;
    c->count_pairs = (on & 2) != 0;
  });
}
int gr_get_timing(gr_ctx *c, double *ms_out, int nphases) {
  return guard(c, [&] {
    for (int k = 0; k < nphases && k < 9; ++k) ms_out[k] = c->phase_ms[k];
  });
}
int gr_get_pair_count(gr_ctx *c, int integ, int64_t *pairs_out) {
  return guard(c, [&] {
    (void)integ;
    *pairs_out = c->last_pairs[0] / 2 + c->last_pairs[1];  // per step
  });
}
int gr_get_pair_stats(gr_ctx *c, int64_t *self_ordered, int64_t *coupled) {
  return guard(c, [&] {
    *self_ordered = c->last_pairs[0];
    *coupled = c->last_pairs[1];
  });
}

// ---- fine-grained ---------------------------------------------------------

int gr_sort(gr_ctx *c, int sys, double cutoff) {
  return guard(c, [&] {
    validate_sys(c, sys);
    commit(c);
    Plan P;
    c->stepsys[sys].sort_cutoff = cutoff;
    upload_stepsys(c);
    if (c->sys[sys].kind == GR_SYS_GHOST) {
      // copy the ghost's pool state to the generation scratch, then sort
      plan_ghosts(c, P, {sys});
      const int64_t n = live_n(c, sys), off = c->sys[sys].off, cap = c->pool_cap;
      for (int d = 0; d < c->D; ++d)
        HIPCHK(hipMemcpyAsync(c->gxs.as<double>() + d * cap + off, c->x[c->cur].as<double>() + d * cap + off,
                              n * 8, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->gidx_s.as<int>() + off, c->gidx.as<int>() + off, n * 4, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->gbnd_s.as<int>() + off, c->gbnd.as<int>() + off, n * 4, hipMemcpyDeviceToDevice, c->stream));
      by_prec(c, [&](auto *tag) {
        using R = std::remove_pointer_t<decltype(tag)>;
        phase_ghosts<R>(c, P, false, true);
      });
    } else {
      plan_sort_real(c, P, {sys});
      by_prec(c, [&](auto *tag) {
        using R = std::remove_pointer_t<decltype(tag)>;
        phase_sort_real<R>(c, P);
      });
    }
    // floor(Int, x / cutoff) of a non-finite position raises in the
    // reference (src/Sorting.jl:122-124); here the system keeps its order
    // and the call reports it
    int flags[4];
    HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (flags[1] & (1 | 8)) {
      int sf = 0;
      HIPCHK(hipMemcpyAsync(&sf, c->d_sysfault.as<int>() + sys, sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
      int z[4] = {0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(c->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemsetAsync(c->d_sysfault.p, 0, c->sys.size() * sizeof(int), c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if (sf & GR_FAULT_NONFINITE)
        throw GrError(GR_ERR_ARGUMENT, "non-finite position: its cell key floor(x / cutoff) is undefined");
      throw GrError(GR_ERR_CAPACITY, "cell-key range exceeds 32 bits");
    }
    if (flags[1] & 16) {
      int z[4] = {0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(c->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      throw GrError(GR_ERR_HIP, "device radix sort: the tile counts did not place every key");
    }
  });
}

int gr_sort_static(gr_ctx *c, int sys, double cutoff) {
  return guard(c, [&] {
    validate_sys(c, sys);
    need(c->sys[sys].kind != GR_SYS_GHOST, GR_ERR_ARGUMENT,
         "gr_sort_static sorts real systems (ghosts: gr_sort)");
    need(!c->dirty_layout, GR_ERR_STATE,
         "systems created since the last upload are not laid out yet: upload them first");
    unsigned char *scr = (unsigned char *)scratch_of(c->device);
    need(scr != nullptr, GR_ERR_STATE, "static scratch not set up on this device");
    const int ns = (int)c->sys.size();
    const int64_t cap = c->sys[sys].cap;
    need(cap <= GR_STATIC_SORT_MAXN, GR_ERR_CAPACITY,
         "gr_sort_static: system of " + std::to_string(cap) + " particles exceeds the static "
         "scratch limit of " + std::to_string(GR_STATIC_SORT_MAXN) + " (use gr_sort)");
    need(ns <= GR_STATIC_MAXSYS, GR_ERR_CAPACITY,
         "gr_sort_static: more than " + std::to_string(GR_STATIC_MAXSYS) + " systems (use gr_sort)");
    need(radix_scratch_bytes((int)cap) <= kScrHistBytes, GR_ERR_CAPACITY,
         "gr_sort_static: radix scratch exceeds the static scratch");
    // tables (host vectors; device memory is only the static scratch)
    std::vector<Blk> bsys, ball;
    add_blocks(bsys, sys, cap, GR_BLOCK);
    for (int q = 0; q < ns; ++q)
      if (c->sys[q].kind != GR_SYS_GHOST && !c->sys[q].dead) add_blocks(ball, q, c->sys[q].cap, GR_BLOCK);
    need(ball.size() <= (size_t)GR_STATIC_MAXBLK, GR_ERR_CAPACITY,
         "gr_sort_static: the pool's real systems exceed the static block table (use gr_sort)");
    need(c->d_st.p && c->d_st.bytes >= (size_t)ns * sizeof(StepSys) && c->d_flags.p &&
             c->d_sysfault.p && c->d_sysfault.bytes >= (size_t)ns * sizeof(int),
         GR_ERR_STATE, "context tables not laid out: upload the system first");
    std::vector<int> bbase(ns, -1);
    bbase[sys] = 0;
    hipStream_t s = c->stream;
    c->stepsys[sys].sort_cutoff = cutoff;
    HIPCHK(hipMemcpyAsync(c->d_st.as<StepSys>() + sys, &c->stepsys[sys], sizeof(StepSys),
                          hipMemcpyHostToDevice, s));
    int *d_list = (int *)(scr + kScrList);
    int *d_bbase = (int *)(scr + kScrBase);
    Blk *d_bsys = (Blk *)(scr + kScrBlkSys);
    Blk *d_ball = (Blk *)(scr + kScrBlkAll);
    unsigned *keys = (unsigned *)(scr + kScrKeys), *keys2 = (unsigned *)(scr + kScrKeys2);
    int *vals = (int *)(scr + kScrVals), *vals2 = (int *)(scr + kScrVals2);
    unsigned *hist = (unsigned *)(scr + kScrHist);
    HIPCHK(hipMemcpyAsync(d_list, &sys, sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_bbase, bbase.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_bsys, bsys.data(), bsys.size() * sizeof(Blk), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_ball, ball.data(), ball.size() * sizeof(Blk), hipMemcpyHostToDevice, s));
    // the kernel sequence of phase_sort_real (sort_particles!, src/Sorting.jl:157-195)
    const XP xs = xp_pool(c, c->cur);
    launch_minmax_reset(s, d_list, 1, c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(),
                        c->bb_rs);
    launch_key_minmax(s, xs, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), d_bsys, (int)bsys.size(),
                      c->D, c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(), c->bb_rs,
                      d_list, 1, false);
    launch_key_base(s, d_list, 1, c->D, 0, c->d_n.as<int>(), c->d_st.as<StepSys>(),
                    c->d_bbmin.as<long long>(), c->d_bbmax.as<long long>(), c->bb_rs,
                    c->d_kmin.as<int>(), c->d_kmax.as<int>(), c->d_kbase.as<unsigned>(),
                    c->d_krange.as<unsigned>(), c->d_flags.as<int>(), c->d_sysfault.as<int>());
    launch_key_write(s, xs, c->d_sd.as<SysDesc>(), c->d_n.as<int>(), c->d_st.as<StepSys>(), d_bsys,
                     (int)bsys.size(), d_bbase, c->D, c->d_kmin.as<int>(),
                     c->d_krange.as<unsigned>(), c->d_kbase.as<unsigned>(), keys, vals,
                     c->d_flags.as<int>(), hist);
    launch_radix_sort(s, keys, vals, keys2, vals2, hist, (int)cap, c->d_flags.as<int>());
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      Pool<R> A = pool_view<R>(c, c->cur), B = pool_view<R>(c, 1 - c->cur);
      launch_permute_real<R>(s, A, B, c->d_sd.as<SysDesc>(), d_ball, (int)ball.size(), d_bbase,
                             vals, vals2, c->d_flags.as<int>(), c->vadj_any ? 1 : 0, c->D);
    });
    c->cur = 1 - c->cur;
    // as gr_sort: a non-finite position or a key range of 32 bits or more is
    // reported (the system keeps its order)
    int flags[4];
    HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flags[1] & (1 | 8)) {
      int sf = 0;
      HIPCHK(hipMemcpyAsync(&sf, c->d_sysfault.as<int>() + sys, sizeof(int), hipMemcpyDeviceToHost, s));
      int z[4] = {0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(c->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemsetAsync(c->d_sysfault.p, 0, (size_t)ns * sizeof(int), s));
      HIPCHK(hipStreamSynchronize(s));
      if (sf & GR_FAULT_NONFINITE)
        throw GrError(GR_ERR_ARGUMENT, "non-finite position: its cell key floor(x / cutoff) is undefined");
      throw GrError(GR_ERR_CAPACITY, "cell-key range exceeds 32 bits");
    }
    if (flags[1] & 16) {
      int z[4] = {0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(c->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      throw GrError(GR_ERR_HIP, "device radix sort: the tile counts did not place every key");
    }
  });
}

int gr_create_grid(gr_ctx *c, int inter) {
  return guard(c, [&] {
    need(inter >= 0 && inter < (int)c->inter.size(), GR_ERR_ARGUMENT, "bad interaction");
    commit(c);
    InterH &in = c->inter[inter];
    std::vector<int> syss{in.a};
    if (in.b >= 0) syss.push_back(in.b);
    std::vector<int> reals;
    for (int s : syss)
      if (c->sys[s].kind != GR_SYS_GHOST) reals.push_back(s);
    ensure_cells(c, {inter}, reals);
    // bounding boxes of the live particles
    std::vector<Blk> blk;
    for (int s : syss) add_blocks(blk, s, c->sys[s].cap, GR_BLOCK);
    DBuf db, dl;
    db.upload(blk, c->stream);
    dl.upload(syss, c->stream);
    launch_minmax_reset(c->stream, dl.as<int>(), (int)syss.size(), c->d_bbmin.as<long long>(),
                        c->d_bbmax.as<long long>(), c->bb_rs);
    launch_key_minmax(c->stream, xp_pool(c, c->cur), c->d_sd.as<SysDesc>(), c->d_n.as<int>(),
                      db.as<Blk>(), (int)blk.size(), c->D, c->d_bbmin.as<long long>(),
                      c->d_bbmax.as<long long>(), c->bb_rs, dl.as<int>(), (int)syss.size(), true);
    Plan P;
    plan_grids(c, P, {inter});
    phase_grids(c, P);
    mark_grids(c, P);
    int flags[4];
    HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (flags[1] & 2) {
      int z[4] = {0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(c->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      throw GrError(GR_ERR_CAPACITY, "cell grid outgrew its capacity");
    }
    in.grid_built = true;
  });
}

}  // extern "C"

namespace {
// One functor of one interaction over its pairs (sweep!, adjust_v!).
void sweep_one(gr_ctx *c, int inter, const gr_pfn_desc f) {
    InterH &in = c->inter[inter];
    const int fam = family_of(f.kind);
    if (fam < 0) return;
    need(in.grid_built, GR_ERR_STATE, "sweep!: create_grid! has not been called");
    std::vector<int> syss{in.a};
    if (in.b >= 0) syss.push_back(in.b);
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      refresh_shadows<R>(c, syss);
      pack_systems<R>(c, syss);
    });
    const int pk =
        f.kind == GR_PFN_XSPH ? (in.b < 0 ? P_SELF : P_TWO) : classify(c, in, f.kind);
    std::vector<Task> tv;
    std::vector<SrcDesc> sv;
    std::vector<Blk> bv;
    auto add = [&](int target, bool b_side) {
      Task t;
      memset(&t, 0, sizeof(t));
      t.sys = target;
      t.ns = c->sys[target].ns;
      t.nv = c->sys[target].nv;
      t.has_stress = f.kind == GR_PFN_CAUCHY;
      t.write_vor = f.kind == GR_PFN_SRV;
      t.assign = 0;
      t.ltask = -1;
      t.alias = -1;
      if (f.kind == GR_PFN_XSPH && c->sys[target].kind == GR_SYS_GHOST) {
        const SysH &src = c->sys[c->sys[target].src];
        need(live_n(c, target) <= src.n, GR_ERR_ARGUMENT,
             "XSPHPfn with a ghost partner: more ghosts than source particles "
             "(ps_b.v_adjustment[j] would index past the source's array)");
        t.alias = c->sys[target].src;
        t.alias_cap = (int32_t)src.n;
        t.alias_off = src.off;
      }
      if (f.kind == GR_PFN_INTERP) t.imask = f.sigma | (f.alpha != 0.0 ? 1 << 30 : 0);
      if (f.kind == GR_PFN_NCOUNT) t.imask = 1 << GR_F_COUNT;
      SrcDesc q = make_src(c, inter, target, b_side, pk);
      set_pfn(q, f, pk, c, in);
      t.src_begin = (int)sv.size();
      sv.push_back(q);
      t.src_end = (int)sv.size();
      add_blocks(bv, (int)tv.size(), c->sys[target].cap, GR_BLOCK);
      tv.push_back(t);
    };
    if (f.kind == GR_PFN_NCOUNT) {
      need(in.b >= 0 && c->sys[in.b].probe && !c->sys[in.a].virt && in.wrap == GR_WRAP_NONE,
           GR_ERR_ARGUMENT,
           "NeighborCountFn needs a real system_a and a ProbeParticleSystem system_b");
      add(in.b, true);
    } else if (f.kind == GR_PFN_INTERP) {
      need(in.b >= 0 && in.wrap == GR_WRAP_NONE && (c->sys[in.b].virt != c->sys[in.a].virt) &&
               c->sys[in.b].kind != GR_SYS_GHOST,
           GR_ERR_ARGUMENT, "InterpolateFieldFn needs a real and a virtual particle system");
      if (c->sys[in.b].virt)
        add(in.b, true);
      else
        add(in.a, false);
    } else {
      add(in.a, false);
      if (pk == P_TWO && in.b >= 0) add(in.b, true);
    }
    // a ghost-target XSPH task adds into its source's entries, which the
    // source's own task writes: it gets a launch of its own, after the rest
    std::vector<std::vector<int>> groups(1);
    for (int k = 0; k < (int)tv.size(); ++k) {
      if (tv[k].alias >= 0)
        groups.push_back({k});
      else
        groups[0].push_back(k);
    }
    (void)bv;
    DBuf ds;
    ds.upload(sv, c->stream);
    for (const auto &grp : groups) {
      if (grp.empty()) continue;
      std::vector<Task> tg;
      std::vector<Blk> bg;
      for (int k : grp) {
        add_blocks(bg, (int)tg.size(), c->sys[tv[k].sys].cap, GR_BLOCK);
        tg.push_back(tv[k]);
      }
      DBuf dt, db;
      dt.upload(tg, c->stream);
      db.upload(bg, c->stream);
      SweepArgs a;
      a.sd = c->d_sd.as<SysDesc>();
      a.n = c->d_n.as<int>();
      a.tasks = dt.as<Task>();
      a.srcs = ds.as<SrcDesc>();
      a.grids = c->d_grids.as<Grid>();
      a.carr = c->d_carr.as<CellArr>();
      a.cells = c->cells.as<int>();
      a.blk = db.as<Blk>();
      a.nblk = (int)bg.size();
      a.pair_count = c->d_pairs.as<unsigned long long>();
      by_prec(c, [&](auto *tag) {
        using R = std::remove_pointer_t<decltype(tag)>;
        launch_sweep<R>(c->stream, pool_view<R>(c, c->cur), a, c->D, fam, false);
      });
      HIPCHK(hipStreamSynchronize(c->stream));
    }
}
}  // namespace

extern "C" {

int gr_sweep(gr_ctx *c, int inter, int stage) {
  return guard(c, [&] {
    need(inter >= 0 && inter < (int)c->inter.size(), GR_ERR_ARGUMENT, "bad interaction");
    InterH &in = c->inter[inter];
    if (stage < 1 || stage > (int)in.pfns.size()) return;
    sweep_one(c, inter, in.pfns[stage - 1]);
  });
}

int gr_adjust_v(gr_ctx *c, int inter) {
  return guard(c, [&] {
    need(inter >= 0 && inter < (int)c->inter.size(), GR_ERR_ARGUMENT, "bad interaction");
    InterH &in = c->inter[inter];
    if (!in.has_xsph) return;  // velocity_adjust_pairwise_fn = nothing
    gr_pfn_desc f;
    memset(&f, 0, sizeof(f));
    f.kind = GR_PFN_XSPH;
    f.alpha = in.xsph_eps;
    sweep_one(c, inter, f);
  });
}

int gr_update_state(gr_ctx *c, int sys, int stage, double dt) {
  return guard(c, [&] {
    validate_sys(c, sys);
    commit(c);
    SysH &s = c->sys[sys];
    if (stage < 1 || stage > (int)s.upd.size()) return;
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      refresh_shadows<R>(c, {sys});
    });
    c->stepsys[sys].dt = dt;
    upload_stepsys(c);
    std::vector<UpdDesc> ud;
    int kinds = 0;
    for (auto &u : s.upd[stage - 1])
      if (u.kind != GR_UPD_NOTHING) {
        ud.push_back(u);
        kinds |= 1 << u.kind;
      }
    if (ud.empty()) return;
    std::vector<int2> ur(c->sys.size(), make_int2(0, 0));
    ur[sys] = make_int2(0, (int)ud.size());
    std::vector<Blk> bv;
    add_blocks(bv, sys, s.cap, GR_BLOCK);
    DBuf du, dr, db;
    du.upload(ud, c->stream);
    dr.upload(ur, c->stream);
    db.upload(bv, c->stream);
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      launch_updaters<R>(c->stream, pool_view<R>(c, c->cur), c->d_sd.as<SysDesc>(), c->d_st.as<StepSys>(),
                         db.as<Blk>(), (int)bv.size(), dr.as<int2>(), du.as<UpdDesc>(), c->D,
                         kinds);
    });
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int gr_measure_probes(gr_ctx *c, int integ) {
  return guard(c, [&] {
    need(integ >= 0 && integ < (int)c->integ.size() && !c->integ[integ].dead, GR_ERR_ARGUMENT,
         "bad integrator");
    const IntegH &I = c->integ[integ];
    if (I.probes.empty()) return;
    commit(c);
    auto chk = [&](int rc) {
      if (rc) throw GrError(rc, c->err);
    };
    const double cutoff = 2.0 * I.h;  // sort_cutoff (src/TimeIntegration.jl:612)
    std::vector<int> sorted;
    for (int it : I.probe_inter) {  // sources moved since the step's sort
      const int a = c->inter[it].a;
      if (std::find(sorted.begin(), sorted.end(), a) != sorted.end()) continue;
      sorted.push_back(a);
      chk(gr_sort(c, a, cutoff));
    }
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      for (int q : I.probes)
        if (c->sys[q].mirror >= 0)
          launch_probe_mirror<R>(c->stream, pool_view<R>(c, c->cur), c->d_sd.as<SysDesc>(), q,
                                 c->sys[q].mirror, c->sys[c->sys[q].mirror].n, c->D);
    });
    for (int q : I.probes) chk(gr_sort(c, q, cutoff));
    for (int it : I.probe_inter) chk(gr_create_grid(c, it));
    {  // auto_zero_probe! (src/ProbeParticles.jl:141-148): w_sum and the extras
      std::vector<Blk> bp;
      for (int q : I.probes) add_blocks(bp, q, c->sys[q].cap, GR_BLOCK);
      DBuf db;
      db.upload(bp, c->stream);
      by_prec(c, [&](auto *tag) {
        using R = std::remove_pointer_t<decltype(tag)>;
        launch_virtual<R>(c->stream, pool_view<R>(c, c->cur), c->d_sd.as<SysDesc>(),
                          c->d_st.as<StepSys>(), c->d_n.as<int>(), db.as<Blk>(),
                          (int)bp.size(), 0, c->D);
      });
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    for (int it : I.probe_inter)
      if (!c->inter[it].pfns.empty()) sweep_one(c, it, c->inter[it].pfns[0]);
    for (int q : I.probes) chk(gr_update_state(c, q, 1, 0.0));
    for (int q : I.probes) {  // _sort_probe_by_id!
      const SysH &S = c->sys[q];
      const int ncol = 2 * c->D + 4 + (S.kind >= GR_SYS_STRESS ? S.ns : 0);
      const size_t bytes = (size_t)ncol * S.n * 8;
      if (c->probe_scr.bytes < bytes) c->probe_scr.alloc(bytes);
      by_prec(c, [&](auto *tag) {
        using R = std::remove_pointer_t<decltype(tag)>;
        const Pool<R> P = pool_view<R>(c, c->cur);
        launch_probe_by_id<R>(c->stream, P, c->d_sd.as<SysDesc>(), q, S.n,
                              c->probe_scr.as<double>(), 0, c->D);
        launch_probe_by_id<R>(c->stream, P, c->d_sd.as<SysDesc>(), q, S.n,
                              c->probe_scr.as<double>(), 1, c->D);
      });
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int gr_generate_ghosts(gr_ctx *c, int ghost) {
  return guard(c, [&] {
    validate_sys(c, ghost);
    need(c->sys[ghost].kind == GR_SYS_GHOST, GR_ERR_ARGUMENT, "not a ghost system");
    commit(c);
    Plan P;
    plan_ghosts(c, P, {ghost});
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      phase_ghosts<R>(c, P, true, false);
    });
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int gr_update_ghost_kinematics(gr_ctx *c, int ghost) {
  return guard(c, [&] {
    validate_sys(c, ghost);
    need(c->sys[ghost].kind == GR_SYS_GHOST, GR_ERR_ARGUMENT, "not a ghost system");
    commit(c);
    std::vector<Blk> bv;
    add_blocks(bv, ghost, c->sys[ghost].cap, GR_BLOCK);
    DBuf db;
    db.upload(bv, c->stream);
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      refresh_shadows<R>(c, {c->sys[ghost].src});
      launch_ghost_kin<R>(c->stream, pool_view<R>(c, c->cur), c->d_sd.as<SysDesc>(), c->d_bnd.as<Bnd>(),
                          c->d_n.as<int>(), db.as<Blk>(), (int)bv.size(), c->D);
    });
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int gr_update_ghost(gr_ctx *c, int ghost, int stage) {
  return guard(c, [&] {
    validate_sys(c, ghost);
    SysH &g = c->sys[ghost];
    need(g.kind == GR_SYS_GHOST, GR_ERR_ARGUMENT, "not a ghost system");
    commit(c);
    if (stage < 1 || stage > (int)g.copy.size()) return;
    const int2 cp = g.copy[stage - 1];
    if (!cp.x && cp.y < 0) return;
    std::vector<int2> gc(c->sys.size(), make_int2(0, -1));
    gc[ghost] = cp;
    std::vector<Blk> bv;
    add_blocks(bv, ghost, g.cap, GR_BLOCK);
    DBuf dg, db;
    dg.upload(gc, c->stream);
    db.upload(bv, c->stream);
    by_prec(c, [&](auto *tag) {
      using R = std::remove_pointer_t<decltype(tag)>;
      launch_ghost_copy<R>(c->stream, pool_view<R>(c, c->cur), c->d_sd.as<SysDesc>(), c->d_bnd.as<Bnd>(),
                           c->d_n.as<int>(), db.as<Blk>(), (int)bv.size(), dg.as<int2>(), c->D);
    });
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int gr_get_grid(gr_ctx *c, int inter, double *mingridx, int64_t *ngridx, int64_t *ncells) {
  return guard(c, [&] {
    need(inter >= 0 && inter < (int)c->inter.size(), GR_ERR_ARGUMENT, "bad interaction");
    need(c->inter[inter].carr_b >= 0, GR_ERR_STATE, "grid not built");
    Grid G;
    HIPCHK(hipMemcpyAsync(&G, c->d_grids.as<Grid>() + inter, sizeof(G), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int d = 0; d < c->D; ++d) {
      mingridx[d] = G.ming[d];
      ngridx[d] = G.ng[d];
    }
    *ncells = G.ncells;
  });
}

int gr_get_cell_start(gr_ctx *c, int inter, int which, int64_t *out, int64_t count) {
  return guard(c, [&] {
    need(inter >= 0 && inter < (int)c->inter.size(), GR_ERR_ARGUMENT, "bad interaction");
    const InterH &in = c->inter[inter];
    const int ca = which ? in.carr_a : in.carr_b;
    need(ca >= 0, GR_ERR_ARGUMENT, "no such cell array");
    need(ca < (int)c->carr_valid.size() && c->carr_valid[ca], GR_ERR_STATE,
         "cell array not built by the last grid build (a step builds only the cells its "
         "functors walk): call gr_create_grid first");
    Grid G;
    HIPCHK(hipMemcpyAsync(&G, c->d_grids.as<Grid>() + inter, sizeof(G), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    need(count == (int64_t)G.ncells + 1, GR_ERR_ARGUMENT, "count mismatch");
    auto v = read_ints(c, c->cells, c->carr[ca].off, count);
    for (int64_t i = 0; i < count; ++i) out[i] = v[i];
  });
}

}  // extern "C"

// ---- calibration copy kernel ------------------------------------------------

__global__ void k_bench_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

extern "C" int gr_bench_copy(gr_ctx *c, int64_t bytes, int reps) {
  return guard(c, [&] {
    need(bytes >= 16 && reps >= 1, GR_ERR_ARGUMENT, "bytes >= 16 and reps >= 1");
    DBuf a, b;
    a.alloc(bytes);
    b.alloc(bytes);
    HIPCHK(hipMemsetAsync(a.p, 0, bytes, c->stream));
    const long long n = bytes / 16;
    for (int r = 0; r < reps; ++r)
      k_bench_copy<<<1024, 256, 0, c->stream>>>(a.as<const float4>(), b.as<float4>(), n);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}
