// gr_kernels.hip — per-particle, sort, ghost and grid kernels for gfx950.
//
// Integer-deciding floating point (cell keys, CSR cells, ghost qualification,
// positions) is evaluated in FP64 in the order the reference writes it; the
// library is compiled with -ffp-contract=off so nothing is fused unless an
// explicit fma() is written (the reference's @fastmath scalar axpy).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/grasph_hip.h"
#include "gr_device.h"
#include "gr_launch.h"

namespace gr {

// ---------------------------------------------------------------------------
// Wave / block reductions (64-lane wavefronts)
// ---------------------------------------------------------------------------

// 64-bit wave-wide minimum / maximum (all 64 lanes active): DPP exchanges
// within each row of 16 lanes, then the four row results through readlane.
template <int CTRL>
__device__ __forceinline__ long long dpp_ll(long long v) {
  const int lo = __builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(v >> 32), (int)(v >> 32), CTRL, 0xF, 0xF, false);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ long long readlane_ll(long long v, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
  return (long long)(((unsigned long long)hi << 32) | lo);
}
template <bool MAX>
__device__ inline long long wave_red_ll(long long v) {
  auto op = [](long long x, long long y) { return MAX ? (x > y ? x : y) : (x < y ? x : y); };
  v = op(v, dpp_ll<0xB1>(v));   // quad_perm 1,0,3,2
  v = op(v, dpp_ll<0x4E>(v));   // quad_perm 2,3,0,1
  v = op(v, dpp_ll<0x141>(v));  // row_half_mirror
  v = op(v, dpp_ll<0x140>(v));  // row_mirror
  return op(op(readlane_ll(v, 0), readlane_ll(v, 16)), op(readlane_ll(v, 32), readlane_ll(v, 48)));
}
__device__ inline int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) {
    int u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}
__device__ inline int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) {
    int u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}

__device__ inline long long ord_d(double d) {
  long long i = __double_as_longlong(d);
  return i < 0 ? (i ^ 0x7FFFFFFFFFFFFFFFLL) : i;
}

// Coherent (device-scope, relaxed) read of a value other workgroups update
// with atomics.  Used to skip atomics that cannot change the value: a stale
// read only costs an atomic that turns out to be redundant.
template <typename T>
__device__ __forceinline__ T ld_dev(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// floor(Int, x / cutoff) (src/Sorting.jl:122-124), clamped to int32.
__device__ inline int key_comp(double x, double cutoff) {
  double f = floor(x / cutoff);
  if (!(f > -1.0e9)) f = -1.0e9;
  if (f > 1.0e9) f = 1.0e9;
  return (int)f;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits, stable: two kernels per digit pass.
//
// The keys are cut into tiles of OS_TILE.  Each pass needs, for every tile
// and digit, the number of keys with that digit in the tiles before it.  The
// counts of pass p's tiles are taken by the kernel that writes pass p's
// input: k_key_write for pass 0, and pass p-1's scatter for pass p, which
// knows where each key lands and so which tile of the next pass holds it.
// Both count runs of equal (tile, digit) along a wavefront with one atomic
// per run (the keys are nearly in order: a wavefront holds few runs).
// k_radix_scan turns each digit's column of tile counts into exclusive
// prefixes and the digit's total; k_radix_scatter ranks its tile's keys per
// digit (stable), adds the prefixes and the digit starts, and scatters.
//
// Scratch (radix_scratch_bytes): digit totals [4][256], then two count
// arrays [tiles][256] used by alternate passes.  Every reader of a tile's
// row zeroes it, so the arrays are zero again after each sort (and are
// zeroed when allocated).
//
// Passes above the top bit of the largest key keep the order and are
// skipped; flags[3] is the number of passes that run, so the sorted pairs
// end in (keys2, vals2) when it is odd (radix_sorted_vals).  The whole sort
// is skipped when flags[0] == 0: the keys are already in order.
// ---------------------------------------------------------------------------

#ifndef OS_ITEMS
#define OS_ITEMS 16
#endif
#define OS_WAVE_KEYS (64 * OS_ITEMS)
#define OS_TILE (GR_BLOCK * OS_ITEMS)
#define RS_TOT 0      // words: digit totals [4][256]
#define RS_CNT 1024   // words: tile counts [2][tiles][256]

__device__ inline int radix_passes(unsigned bound) {
  // passes whose digit is not zero for some key below `bound`
  if (bound <= 1u) return 0;
  const int bits = 32 - __clz((int)(bound - 1u));
  return (bits + 7) >> 3;
}

// Add the lanes' bins to `cnt`: one atomic for each run of equal bins along
// the valid lanes, made by the run's first lane.
__device__ inline void radix_add_runs(unsigned bin, bool valid, unsigned *cnt) {
  const int lane = threadIdx.x & 63;
  const unsigned prev = __shfl_up(bin, 1, 64);
  const unsigned long long V = __ballot(valid);
  const bool prev_valid = lane > 0 && ((V >> (lane - 1)) & 1ull);
  const bool head = valid && (!prev_valid || prev != bin);
  const unsigned long long H = __ballot(head);
  if (head) {
    const unsigned long long ends = (H | ~V) & (lane == 63 ? 0ull : (~0ull << (lane + 1)));
    const int next = ends ? __ffsll((long long)ends) - 1 : 64;
    atomicAdd(&cnt[bin], (unsigned)(next - lane));
  }
}

// ---------------------------------------------------------------------------
// Keys: per-system key range and position bounding box
// ---------------------------------------------------------------------------

// Bounding box of the positions per dimension and system, as order-preserving
// integers (ord_d).  The cell keys' ranges follow from it (the key component
// is a monotone function of the position) and are derived in k_key_base.
// Each workgroup reduces PER consecutive entries of the block table in
// registers and flushes (workgroup reduction, then an atomic for each bound
// that improves on the current value) whenever the system changes and at the
// end.  The loads of all PER entries (table entry -> system descriptor ->
// positions) are issued before any is reduced, so their latencies overlap.
//
// The workgroups of a launch run concurrently, so they all see the initial
// bounds and most of them take their atomics; those atomics are spread over
// GR_BB_REP replicas of the bounds (stride `rs`), one for the workgroups of
// each XCD, and folded into replica 0 afterwards (k_key_base, k_bb_fold).
template <int D, int PER>
__global__ void k_key_minmax(XP xs, const SysDesc *sd, const int *n_dev, const Blk *blk, int nblk,
                             long long *bbmin, long long *bbmax, int rs) {
  bbmin += (blockIdx.x % GR_BB_REP) * rs;
  bbmax += (blockIdx.x % GR_BB_REP) * rs;
  __shared__ long long red[2][GR_BLOCK / 64][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int b0 = blockIdx.x * PER;
  int item[PER];
  bool live[PER];
  double xe[PER][D];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const Blk b = blk[b0 + k < nblk ? b0 + k : nblk - 1];
    item[k] = b.item;
    const int i = b.first + threadIdx.x;
    live[k] = b0 + k < nblk && i < n_dev[b.item];
    const int64_t gi = sd[b.item].off + i;
#pragma unroll
    for (int d = 0; d < D; ++d) xe[k][d] = live[k] ? xs[d][gi] : 0.0;
  }
  long long mn[D], mx[D];
  auto reset = [&]() {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      mn[d] = LLONG_MAX;
      mx[d] = LLONG_MIN;
    }
  };
  auto flush = [&](int s) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const long long a = wave_red_ll<false>(mn[d]), bq = wave_red_ll<true>(mx[d]);
      if (lane == 0) {
        red[0][w][d] = a;
        red[1][w][d] = bq;
      }
    }
    __syncthreads();
    if (threadIdx.x < D) {
      const int d = threadIdx.x;
      long long a = LLONG_MAX, bq = LLONG_MIN;
      for (int q = 0; q < GR_BLOCK / 64; ++q) {
        a = red[0][q][d] < a ? red[0][q][d] : a;
        bq = red[1][q][d] > bq ? red[1][q][d] : bq;
      }
      // every workgroup of a system reduces into the same few words; only
      // a partial result that improves on the current value takes the atomic
      if (a != LLONG_MAX) {
        if (a < ld_dev(&bbmin[s * 3 + d])) atomicMin(&bbmin[s * 3 + d], a);
        if (bq > ld_dev(&bbmax[s * 3 + d])) atomicMax(&bbmax[s * 3 + d], bq);
      }
    }
    __syncthreads();
  };
  reset();
  int cur = item[0];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if (b0 + k >= nblk) break;  // uniform over the workgroup
    if (item[k] != cur) {       // uniform over the workgroup
      flush(cur);
      reset();
      cur = item[k];
    }
    if (live[k]) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const long long o = ord_d(xe[k][d]);
        mn[d] = o < mn[d] ? o : mn[d];
        mx[d] = o > mx[d] ? o : mx[d];
      }
    }
  }
  flush(cur);
}

// Reset the reduction slots of the listed systems.
__global__ void k_minmax_reset(const int *list, int nlist, long long *bbmin, long long *bbmax,
                               int rs) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist * 3 * GR_BB_REP) return;
  const int r = t / (nlist * 3);
  t -= r * nlist * 3;
  int s = list[t / 3], d = t % 3;
  bbmin[r * rs + s * 3 + d] = LLONG_MAX;
  bbmax[r * rs + s * 3 + d] = LLONG_MIN;
}

// Fold the replicas of the listed systems' bounds into replica 0.
__device__ __forceinline__ void bb_fold(long long *bbmin, long long *bbmax, int rs, int e) {
  long long a = bbmin[e], b = bbmax[e];
#pragma unroll
  for (int r = 1; r < GR_BB_REP; ++r) {
    const long long u = bbmin[r * rs + e], v = bbmax[r * rs + e];
    a = u < a ? u : a;
    b = v > b ? v : b;
  }
  bbmin[e] = a;
  bbmax[e] = b;
}
__global__ void k_bb_fold(const int *list, int nlist, long long *bbmin, long long *bbmax, int rs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nlist * 3) bb_fold(bbmin, bbmax, rs, list[t / 3] * 3 + t % 3);
}

// Threads of k_key_base: one per listed system in each round.
#define KB_THREADS 1024

// Key space of each listed system: lexicographic linearisation of
// (k1 - kmin1, ..., kD - kminD), bases prefix-summed in list order.  Ghost
// systems reserve one extra key for padding slots.  flags[0] (the inversion
// flag) is cleared here; flags[1] records overflow.  kmin/kmax are the key
// components of the bounding box's corners: key_comp is monotone in the
// position, so they are the extremes of the particles' key components.  A
// NaN position leaves kmax below kmin or a range beyond 32 bits: overflow.
__device__ inline double unord_d(long long o) {
  return __longlong_as_double(o < 0 ? (o ^ 0x7FFFFFFFFFFFFFFFLL) : o);
}
__global__ void __launch_bounds__(KB_THREADS)
    k_key_base(const int *list, int nlist, int D, int pad, const int *n_dev, const StepSys *st,
               long long *bbmin, long long *bbmax, int rs, int *kmin, int *kmax,
               unsigned *kbase, unsigned *krange, int *flags, int *sysfault) {
  // one workgroup: the systems' key ranges in parallel, KB_THREADS at a time,
  // and their exclusive prefix sum (each system's keys start after the ranges
  // of the systems listed before it, plus one padding key each when `pad`)
  __shared__ unsigned long long wsum[KB_THREADS / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  unsigned long long carry = 0;
  for (int q0 = 0; q0 < nlist; q0 += KB_THREADS) {
    const int q = q0 + t;
    unsigned long long range = 0;
    int s = -1;
    if (q < nlist) {
      s = list[q];
      for (int d = 0; d < 3; ++d) bb_fold(bbmin, bbmax, rs, s * 3 + d);
      if (n_dev[s] > 0) {
        range = 1;
        const double cutoff = st[s].sort_cutoff;
        for (int d = 0; d < D; ++d) {
          const int ka = key_comp(unord_d(bbmin[s * 3 + d]), cutoff);
          const int kb = key_comp(unord_d(bbmax[s * 3 + d]), cutoff);
          const int lo = ka < kb ? ka : kb, hi = ka < kb ? kb : ka;
          kmin[s * 3 + d] = lo;
          kmax[s * 3 + d] = hi;
          long long r = (long long)hi - (long long)lo + 1;
          if (!(unord_d(bbmin[s * 3 + d]) <= unord_d(bbmax[s * 3 + d]))) r = 1LL << 32;  // NaN
          krange[s * 3 + d] = (unsigned)r;
          // any product of 2^32 or more overflows the 32-bit keys: saturate
          // there so that neither this product nor the prefix sum wraps
          range = (range > 0xFFFFFFFFull || r > 0xFFFFFFFFLL) ? (1ull << 32)
                                                               : range * (unsigned long long)r;
        }
        if (range > (1ull << 32)) range = 1ull << 32;
        // a system whose bounds are not finite, or whose own keys need 32
        // bits or more, is faulted: all its keys become its base (the
        // stable sort keeps its order), so the systems after it keep
        // their keys and the step goes on for them
        bool nonfinite = false;
        for (int d = 0; d < D; ++d)
          nonfinite = nonfinite || !(unord_d(bbmin[s * 3 + d]) <= unord_d(bbmax[s * 3 + d])) ||
                      !isfinite(unord_d(bbmin[s * 3 + d])) || !isfinite(unord_d(bbmax[s * 3 + d]));
        if (nonfinite || range >= 0xFFFFFFFFull) {
          atomicOr(&sysfault[s], nonfinite ? GR_FAULT_NONFINITE : GR_FAULT_DOMAIN);
          atomicOr(&flags[1], 8);
          for (int d = 0; d < D; ++d) {
            kmin[s * 3 + d] = 0;
            kmax[s * 3 + d] = 0;
            krange[s * 3 + d] = d == 0 ? 0u : 1u;  // krange[0] == 0 marks it
          }
          range = 1;
        }
      } else {
        for (int d = 0; d < D; ++d) krange[s * 3 + d] = 1;
      }
      range += pad ? 1 : 0;
    }
    unsigned long long inc = range;
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    unsigned long long off = 0, tot = 0;
    for (int v = 0; v < KB_THREADS / 64; ++v) {
      off += v < w ? wsum[v] : 0ull;
      tot += wsum[v];
    }
    if (q < nlist) kbase[s] = (unsigned)(carry + off + inc - range);
    carry += tot;
    __syncthreads();  // wsum is written again for the next systems
  }
  if (t == 0) {
    flags[0] = 0;
    // the running base only grows, so overflowing at some system means the
    // total overflows
    if (carry >= 0xFFFFFFFFull) flags[1] |= 1;
    // every key is below this bound: radix passes above its top bit are identities
    flags[2] = (int)(unsigned)(carry < 0xFFFFFFFFull ? carry : 0xFFFFFFFFull);
    flags[3] = radix_passes((unsigned)flags[2]);
  }
}

template <int D>
__device__ inline unsigned lin_key(XP xs, int64_t gi, double cutoff, const int *kmin,
                                   const unsigned *krange, int s) {
  unsigned k = 0;
  if (krange[s * 3] == 0u) return 0u;  // faulted system (k_key_base)
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int c = key_comp(xs[d][gi], cutoff) - kmin[s * 3 + d];
    k = k * krange[s * 3 + d] + (unsigned)c;
  }
  return k;
}

// Write (key, pool index) at each particle's compact position and flag any
// key inversion inside a system (src/Sorting.jl:179-186).
template <int D>
__global__ void k_key_write(XP xs, const SysDesc *sd, const int *n_dev,
                            const StepSys *st, const Blk *blk, const int *bbase,
                            const int *kmin, const unsigned *krange, const unsigned *kbase,
                            unsigned *keys, int *vals, int *flags, unsigned *rscr) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int lane = threadIdx.x & 63;
  const int i = b.first + threadIdx.x;
  const int cap = sd[s].cap;
  const int n = n_dev[s];
  const int64_t gi = sd[s].off + i;
  const double cutoff = st[s].sort_cutoff;
  __shared__ unsigned kw[GR_BLOCK / 64];
  const int w = threadIdx.x >> 6;
  const bool last = threadIdx.x == GR_BLOCK - 1 && i + 1 < n;
  unsigned key, kn = 0u;
  if (i < n) {
    key = kbase[s] + lin_key<D>(xs, gi, cutoff, kmin, krange, s);
    // the key after the workgroup's last particle, the only neighbour key
    // that no thread of this workgroup computes for itself
    if (last) kn = kbase[s] + lin_key<D>(xs, gi + 1, cutoff, kmin, krange, s);
  } else {
    unsigned rng = 1;
    for (int d = 0; d < D; ++d) rng *= krange[s * 3 + d];
    key = kbase[s] + (n > 0 ? rng : 0u);
  }
  // the next particle's key: from the next lane, from the next wavefront's
  // first lane (through LDS) for lane 63, and kn for the workgroup's last
  // thread (every thread of the workgroup is active here)
  if (lane == 0) kw[w] = key;
  __syncthreads();
  unsigned k2 = __shfl_down(key, 1, 64);
  if (lane == 63) k2 = w + 1 < GR_BLOCK / 64 ? kw[w + 1] : kn;
  const bool inv = i + 1 < n && k2 < key;
  // most workgroups of a moving scene hold an inversion: one atomic per
  // workgroup at most, none once the flag reads as raised
  if (__syncthreads_or(inv) && threadIdx.x == 0 && ld_dev(&flags[0]) == 0) atomicOr(&flags[0], 1);
  if (i < cap) {
    const int p = bbase[s] + i;
    keys[p] = key;
    vals[p] = (int)gi;
  }
  // the radix sort's pass-0 tile counts (its entries are the slots below cap)
  if (flags[3] > 0) {
    const unsigned p = (unsigned)(bbase[s] + i);
    radix_add_runs((p / OS_TILE) * 256u + (key & 255u), i < cap, rscr + RS_CNT);
  }
}

// The radix sort's passes (described above the key kernels).
__device__ inline unsigned block_excl_scan256(unsigned v, unsigned *wsum) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  unsigned inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  unsigned off = 0;
  for (int q = 0; q < w; ++q) off += wsum[q];
  return off + inc - v;
}

// One workgroup per digit: the digit's tile counts become exclusive prefixes
// in place (each thread sums a run of consecutive tiles, a workgroup scan
// places the runs), and its total goes to the totals of the pass.
__global__ void __launch_bounds__(GR_BLOCK) k_radix_scan(int N, int pass, unsigned *scr,
                                                         const int *flags) {
  if (pass >= flags[3]) return;
  __shared__ unsigned wsum[GR_BLOCK / 64];
  const int ntile = (N + OS_TILE - 1) / OS_TILE;
  const int t = threadIdx.x, d = blockIdx.x;
  unsigned *cnt = scr + RS_CNT + (size_t)(pass & 1) * ntile * 256 + d;
  const int per = (ntile + GR_BLOCK - 1) / GR_BLOCK;
  const int j0 = t * per, j1 = j0 + per < ntile ? j0 + per : ntile;
  if (flags[0] == 0) {
    // keys already in order: nothing runs, so clear what k_key_write counted
    if (pass == 0)
      for (int j = j0; j < j1; ++j) cnt[(size_t)j * 256] = 0u;
    return;
  }
  unsigned sum = 0;
  for (int j = j0; j < j1; ++j) sum += cnt[(size_t)j * 256];
  unsigned run = block_excl_scan256(sum, wsum);
  for (int j = j0; j < j1; ++j) {
    const unsigned c = cnt[(size_t)j * 256];
    cnt[(size_t)j * 256] = run;
    run += c;
  }
  if (t == GR_BLOCK - 1) scr[RS_TOT + pass * 256 + d] = run;
}

// Stable scatter of one tile: wavefront w holds the tile's keys
// [w·OS_WAVE_KEYS, (w+1)·OS_WAVE_KEYS) in rounds of 64, and a key's rank
// among the wavefront's keys with its digit counts the earlier rounds, then
// the lower lanes.  A round whose digits never decrease along the lanes holds
// each digit as one run (the common case for keys nearly in order); other
// rounds match digits bit by bit with ballots.
__global__ void __launch_bounds__(GR_BLOCK)
    k_radix_scatter(const unsigned *__restrict__ kin, const int *__restrict__ vin,
                    unsigned *__restrict__ kout, int *__restrict__ vout, int N, int pass,
                    unsigned *scr, int *flags) {
  const int npass = flags[3];
  if (flags[0] == 0 || pass >= npass) return;
  const int shift = 8 * pass;
  // digit bits that vary: the top pass may use fewer than 8
  const int bits = 32 - __clz((int)((unsigned)flags[2] - 1u));
  const int nb = bits - shift < 8 ? bits - shift : 8;
  const int ntile = (N + OS_TILE - 1) / OS_TILE;
  __shared__ unsigned wcnt[GR_BLOCK / 64][256];
  __shared__ unsigned wsum[GR_BLOCK / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const unsigned tile = blockIdx.x;
  const long long e0 = (long long)tile * OS_TILE + w * OS_WAVE_KEYS + lane;
  unsigned key[OS_ITEMS];
  int val[OS_ITEMS];
#pragma unroll
  for (int r = 0; r < OS_ITEMS; ++r) {
    const long long e = e0 + r * 64;
    key[r] = e < N ? kin[e] : 0u;
    val[r] = e < N ? vin[e] : 0;
  }
  // this tile's start within digit t; the row is left zeroed
  unsigned *row = scr + RS_CNT + (size_t)(pass & 1) * ntile * 256 + (size_t)tile * 256;
  const unsigned ex = row[t];
  row[t] = 0u;
#pragma unroll
  for (int q = 0; q < GR_BLOCK / 64; ++q) wcnt[q][t] = 0;
  // global start of digit t: totals of the digits below it
  const unsigned gstart = block_excl_scan256(scr[RS_TOT + pass * 256 + t], wsum);  // (synchronises)
  const unsigned long long lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const unsigned long long le_mask = (2ull << lane) - 1ull;
  unsigned rank[OS_ITEMS];
#pragma unroll
  for (int r = 0; r < OS_ITEMS; ++r) {
    const bool valid = e0 + r * 64 < N;
    const unsigned d = (key[r] >> shift) & 255u;
    const unsigned dp = __shfl_up(d, 1, 64);
    const unsigned long long V = __ballot(valid);
    unsigned rk, run;
    if (__ballot(valid && lane > 0 && d < dp) == 0ull) {
      // equal digits are adjacent: runs start where the digit changes
      const unsigned long long H = __ballot(valid && (lane == 0 || d != dp));
      const int head = 63 - __clzll((long long)(H & le_mask));
      const unsigned long long ends = (H | ~V) & ~le_mask;
      const int next = ends ? __ffsll((long long)ends) - 1 : 64;
      rk = (unsigned)(lane - head);
      run = (unsigned)(next - lane);
    } else {
      unsigned long long match = V;
      for (int b = 0; b < nb; ++b) {
        const unsigned long long bal = __ballot((d >> b) & 1u);
        match &= ((d >> b) & 1u) ? bal : ~bal;
      }
      rk = (unsigned)__popcll(match & lt_mask);
      run = (unsigned)__popcll(match);
    }
    const unsigned before = wcnt[w][d];
    rank[r] = before + rk;
    if (valid && rk == 0) wcnt[w][d] = before + run;
  }
  __syncthreads();
  // output start of (wavefront q, digit t)
  unsigned start = gstart + ex;
#pragma unroll
  for (int q = 0; q < GR_BLOCK / 64; ++q) {
    const unsigned c = wcnt[q][t];
    wcnt[q][t] = start;
    start += c;
  }
  __syncthreads();
  unsigned *cnext = scr + RS_CNT + (size_t)((pass + 1) & 1) * ntile * 256;
  const bool count_next = pass + 1 < npass;
  bool bad = false;
#pragma unroll
  for (int r = 0; r < OS_ITEMS; ++r) {
    const bool valid = e0 + r * 64 < N;
    const unsigned pos = wcnt[w][(key[r] >> shift) & 255u] + rank[r];
    // the counts place every key below N; a position beyond it is
    // reported, never written
    const bool ok = valid && pos < (unsigned)N;
    if (ok) {
      kout[pos] = key[r];
      vout[pos] = val[r];
    }
    bad = bad || (valid && !ok);
    // the next pass's tile counts
    if (count_next)
      radix_add_runs((pos / OS_TILE) * 256u + ((key[r] >> (shift + 8)) & 255u), ok, cnext);
  }
  if (bad) atomicOr(&flags[1], 16);
}

// The buffer that holds the sorted values (see the sort's header).
__device__ __forceinline__ const int *radix_sorted_vals(const int *vals, const int *vals2,
                                                        const int *flags) {
  return (flags[3] & 1) ? vals2 : vals;
}

// ---------------------------------------------------------------------------
// Permutation of real systems (src/Sorting.jl:88-105): pool A → pool B for
// every real particle; identity where the system is not sorted this step.
// ---------------------------------------------------------------------------

__device__ inline double eos_p(int kind, double rho, double rho0, double gamma, double c);

// HALF: the leapfrog half step of the systems flagged in `halfsys` is done
// here, on the values as they are moved (the arithmetic of k_halfstep with
// mult ½): one pass over the particles instead of two.  PACK (FP32, with
// HALF): their compact partner records too, as k_pack writes them from the
// half-step state (1: ρ and p, 2: 1/ρ and p/ρ²).
template <typename R, int D, bool HALF, int PACK = 0>
__global__ void k_permute_real(Pool<R> A, Pool<R> B, const SysDesc *sd, const Blk *blk,
                               const int *bbase, const int *vals, const int *vals2,
                               const int *flags, int vadj, const StepSys *st,
                               const int *halfsys) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[s];
  if (i >= S.cap) return;
  const int64_t dst = S.off + i;
  int64_t src = dst;
  if (flags[0] != 0 && bbase[s] >= 0) src = radix_sorted_vals(vals, vals2, flags)[bbase[s] + i];
  const bool half = HALF && halfsys[s] != 0;
  // fields this step overwrites before reading them are not moved (only with
  // the half step, which has read the old rates here)
  const int skip = half ? st[s].perm_skip : 0;
  if (half) {
    const StepSys T = st[s];
    const double cf = 0.5 * T.dt;
    double xn[D];
    R vsn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double v = A.v[d][src];
      const R a = A.dvdt[d][src];
      xn[d] = A.x[d][src];
      vsn[d] = (R)(v + cf * (double)a);
      B.x[d][dst] = xn[d];
      B.v[d][dst] = v;
      B.vs[d][dst] = vsn[d];
      if (!(skip & 1)) B.dvdt[d][dst] = T.reset_force ? (R)S.sv[d] : a;
      if (vadj) B.vadj[d][dst] = A.vadj[d][src];
    }
    const double rho = A.rho[src];
    const R dr = A.drhodt[src];
    const double rh = fma(cf, (double)dr, rho);
    B.rho[dst] = rho;
    B.rho_h[dst] = rh;
    B.rhos[dst] = (R)rh;
    if (!(skip & 1)) B.drhodt[dst] = T.reset_force ? (R)S.srho : dr;
    R pn = (R)0;
    if (T.fuse_eos) {
      pn = (R)eos_p(T.fuse_eos, rh, T.eos_rho0, T.eos_gamma, S.c);
      B.p[dst] = pn;
    } else if (S.kind >= GR_SYS_FLUID) {
      pn = A.p[src];
      B.p[dst] = pn;
    } else {
      pn = B.p[dst];  // kept in place, as k_pack would read it
    }
    if constexpr (PACK > 0 && sizeof(R) == 4) {
      RecC cc;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        cc.xq[d] = d < D ? (int32_t)quant_of(xn[d], B.iqs) : 0;
        cc.v[d] = d < D ? (float)vsn[d] : 0.0f;
      }
      const float rs = (float)(R)rh;
      if constexpr (PACK == 2) {
        const float ir = __builtin_amdgcn_rcpf(rs);
        cc.rho = ir;
        cc.p = pn * (ir * ir);
      } else {
        cc.rho = rs;
        cc.p = pn;
      }
      B.rc[dst] = cc;
    }
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      B.x[d][dst] = A.x[d][src];
      B.v[d][dst] = A.v[d][src];
      B.dvdt[d][dst] = A.dvdt[d][src];
      if (vadj) B.vadj[d][dst] = A.vadj[d][src];
    }
    B.rho[dst] = A.rho[src];
    B.drhodt[dst] = A.drhodt[src];
    if (S.kind >= GR_SYS_FLUID) B.p[dst] = A.p[src];
  }
  B.id[dst] = A.id[src];
  if (S.kind >= GR_SYS_STRESS) {
#pragma unroll
    for (int k = 0; k < GR_MAXNS; ++k)
      if (k < S.ns) {
        const double sg = A.stress[k][src];
        B.stress[k][dst] = sg;
        B.stress_s[k][dst] = (R)sg;
        if (!(skip & 2)) B.sr[k][dst] = A.sr[k][src];
      }
  }
  if (S.kind == GR_SYS_EP) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k < S.nv && !(skip & 4)) B.vor[k][dst] = A.vor[k][src];
#pragma unroll
    for (int k = 0; k < GR_MAXNS; ++k)
      if (k < S.ns) {
        B.strain[k][dst] = A.strain[k][src];
        B.strainp[k][dst] = A.strainp[k][src];
      }
  }
}

// ---------------------------------------------------------------------------
// Ghosts (src/GhostParticles.jl:344-421)
// ---------------------------------------------------------------------------

// Ghost generation: candidate (boundary b, source particle i) qualifies when
// abs(da) <= cutoff && da > 0, and the ghosts follow boundary-major, then
// source order.  A thread takes one source particle and every boundary of its
// ghost system: k_ghost_count counts the qualifying candidates of each
// (boundary, block of GR_BLOCK source particles) slot, an exclusive scan of
// the slots gives each its first ghost, and k_ghost_scatter writes the ghosts
// there in source order.  Slots of ghost system g start at gslot[g],
// boundary-major.
template <int D>
__device__ inline bool ghost_qualifies(const double *xi, const Bnd &B, double gcut, double &da) {
  da = (xi[0] - B.p[0]) * B.n[0];
#pragma unroll
  for (int d = 1; d < D; ++d) da = da + (xi[d] - B.p[d]) * B.n[d];
  return fabs(da) <= gcut && da > 0.0;
}

template <int D>
__global__ void __launch_bounds__(GR_BLOCK)
    k_ghost_count(XP x, const SysDesc *sd, const Bnd *bnd, const Blk *blk, const int *gslot,
                  int *cnt) {
  const Blk b = blk[blockIdx.x];
  const int g = b.item;
  const SysDesc G = sd[g];
  const SysDesc S = sd[G.src];
  const int i = b.first + threadIdx.x;
  const bool live = i < S.cap;
  const int64_t gi = S.off + (live ? i : 0);
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[d][gi];
  const int nsb = (S.cap + GR_BLOCK - 1) / GR_BLOCK, sb = b.first / GR_BLOCK;
  for (int bi = 0; bi < G.nb; ++bi) {
    double da;
    const bool q = live && ghost_qualifies<D>(xi, bnd[G.bnd + bi], G.gcut, da);
    const int n = __syncthreads_count(q);
    if (threadIdx.x == 0) cnt[gslot[g] + bi * nsb + sb] = n;
  }
}

// Mirrored position, source index, boundary index, and the boundary whose
// normal slot k keeps (normals are not permuted by the ghost sort,
// src/Sorting.jl:33-40).
template <int D>
__global__ void __launch_bounds__(GR_BLOCK)
    k_ghost_scatter(XP x, XPW gx, int *gidx, int *gbnd, int *gnrm, const SysDesc *sd,
                    const Bnd *bnd, const Blk *blk, const int *gslot, const int *scan,
                    int *n_dev) {
  __shared__ int wc[GR_BLOCK / 64];
  const Blk b = blk[blockIdx.x];
  const int g = b.item;
  const SysDesc G = sd[g];
  const SysDesc S = sd[G.src];
  const int i = b.first + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool live = i < S.cap;
  const int64_t gi = S.off + (live ? i : 0);
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[d][gi];
  const int nsb = (S.cap + GR_BLOCK - 1) / GR_BLOCK, sb = b.first / GR_BLOCK;
  const int s0 = gslot[g];
  const int first = scan[s0];
  if (sb == 0 && threadIdx.x == 0) n_dev[g] = scan[s0 + G.nb * nsb] - first;
  for (int bi = 0; bi < G.nb; ++bi) {
    const Bnd B = bnd[G.bnd + bi];
    double da;
    const bool q = live && ghost_qualifies<D>(xi, B, G.gcut, da);
    const unsigned long long bal = __ballot(q);
    if (lane == 0) wc[w] = __popcll(bal);
    __syncthreads();
    int woff = 0;
    for (int v = 0; v < w; ++v) woff += wc[v];
    __syncthreads();  // wc is written again for the next boundary
    if (q) {
      const int k = scan[s0 + bi * nsb + sb] - first + woff +
                    __popcll(bal & ((1ull << lane) - 1ull));
      const int64_t go = G.off + k;
#pragma unroll
      for (int d = 0; d < D; ++d) gx[d][go] = xi[d] - (2.0 * da) * B.n[d];
      gidx[go] = i;
      gbnd[go] = bi;
      gnrm[go] = bi;
    }
  }
}

// Ghost sort permutation: generation scratch → pool (x, idx_original,
// idx_boundary).
template <int D>
__global__ void k_permute_ghost(XP gxs, const int *gidx_s, const int *gbnd_s, XPW x, int *gidx,
                                int *gbnd, const SysDesc *sd, const int *n_dev, const Blk *blk,
                                const int *bbase, const int *vals, const int *vals2,
                                const int *flags) {
  const Blk b = blk[blockIdx.x];
  const int g = b.item;
  const int k = b.first + threadIdx.x;
  const SysDesc G = sd[g];
  if (k >= n_dev[g]) return;
  const int64_t dst = G.off + k;
  int64_t src = dst;
  if (flags[0] != 0) src = radix_sorted_vals(vals, vals2, flags)[bbase[g] + k];
#pragma unroll
  for (int d = 0; d < D; ++d) x[d][dst] = gxs[d][src];
  gidx[dst] = gidx_s[src];
  gbnd[dst] = gbnd_s[src];
}

// update_ghost_kinematics! (:404-421): reflected half-step velocity of the
// source and its density, with the normal of slot k.
template <typename R, int D>
__global__ void k_ghost_kin(Pool<R> P, const SysDesc *sd, const Bnd *bnd, const int *n_dev,
                            const Blk *blk) {
  const Blk b = blk[blockIdx.x];
  const int g = b.item;
  const int k = b.first + threadIdx.x;
  const SysDesc G = sd[g];
  if (k >= n_dev[g]) return;
  const SysDesc S = sd[G.src];
  const int64_t go = G.off + k;
  const int64_t si = S.off + P.gidx[go];
  const Bnd B = bnd[G.bnd + P.gnrm[go]];
  R vr[D];
#pragma unroll
  for (int d = 0; d < D; ++d) vr[d] = P.vs[d][si];
  R dd = vr[0] * (R)B.n[0];
#pragma unroll
  for (int d = 1; d < D; ++d) dd = dd + vr[d] * (R)B.n[d];
#pragma unroll
  for (int d = 0; d < D; ++d) P.vs[d][go] = vr[d] - (R)2 * dd * (R)B.n[d];
  P.rhos[go] = P.rhos[si];
}

// _apply_mode(σ, HouseholderReflect(), n̂) (:66-110) in FP64, one function
// per Voigt length so every index is static.
__device__ inline void householder3(double *s, const double *n) {
  const double nx = n[0], ny = n[1];
  const double tx = s[0] * nx + s[2] * ny, ty = s[2] * nx + s[1] * ny;
  const double q = tx * nx + ty * ny;
  const double o0 = s[0] - 4 * tx * nx + 4 * q * nx * nx;
  const double o1 = s[1] - 4 * ty * ny + 4 * q * ny * ny;
  const double o2 = s[2] - 2 * (tx * ny + nx * ty) + 4 * q * nx * ny;
  s[0] = o0;
  s[1] = o1;
  s[2] = o2;
}
__device__ inline void householder4(double *s, const double *n) {
  const double nx = n[0], ny = n[1];
  const double tx = s[0] * nx + s[3] * ny, ty = s[3] * nx + s[1] * ny;
  const double q = tx * nx + ty * ny;
  const double o0 = s[0] - 4 * tx * nx + 4 * q * nx * nx;
  const double o1 = s[1] - 4 * ty * ny + 4 * q * ny * ny;
  const double o3 = s[3] - 2 * (tx * ny + nx * ty) + 4 * q * nx * ny;
  s[0] = o0;
  s[1] = o1;
  s[3] = o3;
}
__device__ inline void householder6(double *s, const double *n) {
  const double nx = n[0], ny = n[1], nz = n[2];
  const double tx = s[0] * nx + s[3] * ny + s[4] * nz;
  const double ty = s[3] * nx + s[1] * ny + s[5] * nz;
  const double tz = s[4] * nx + s[5] * ny + s[2] * nz;
  const double q = tx * nx + ty * ny + tz * nz;
  const double o0 = s[0] - 4 * tx * nx + 4 * q * nx * nx;
  const double o1 = s[1] - 4 * ty * ny + 4 * q * ny * ny;
  const double o2 = s[2] - 4 * tz * nz + 4 * q * nz * nz;
  const double o3 = s[3] - 2 * (tx * ny + nx * ty) + 4 * q * nx * ny;
  const double o4 = s[4] - 2 * (tx * nz + nx * tz) + 4 * q * nx * nz;
  const double o5 = s[5] - 2 * (ty * nz + ny * tz) + 4 * q * ny * nz;
  s[0] = o0;
  s[1] = o1;
  s[2] = o2;
  s[3] = o3;
  s[4] = o4;
  s[5] = o5;
}

// GhostCopier of one stage (:112-129): copy p, copy or reflect stress.
// copy[g] = {copy_p, stress_mode (-1 none, 0 plain, 1 Householder)}.
template <typename R, int NS>
__device__ __forceinline__ void ghost_copy_stress(Pool<R> P, int64_t si, int64_t go, int mode,
                                         const double *n) {
  double s[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q] = P.stress[q][si];
  if (mode == GR_COPY_HOUSEHOLDER) {
    if (NS == 3) householder3(s, n);
    if (NS == 4) householder4(s, n);
    if (NS == 6) householder6(s, n);
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) P.stress_s[q][go] = (R)s[q];
}

template <typename R>
__global__ void k_ghost_copy(Pool<R> P, const SysDesc *sd, const Bnd *bnd, const int *n_dev,
                             const Blk *blk, const int2 *copy) {
  const Blk b = blk[blockIdx.x];
  const int g = b.item;
  const int k = b.first + threadIdx.x;
  const SysDesc G = sd[g];
  if (k >= n_dev[g]) return;
  const SysDesc S = sd[G.src];
  const int64_t go = G.off + k;
  const int64_t si = S.off + P.gidx[go];
  const int2 cp = copy[g];
  if (cp.x) P.p[go] = P.p[si];
  if (cp.y >= 0) {
    const Bnd B = bnd[G.bnd + P.gnrm[go]];
    if (S.ns == 3) ghost_copy_stress<R, 3>(P, si, go, cp.y, B.n);
    if (S.ns == 4) ghost_copy_stress<R, 4>(P, si, go, cp.y, B.n);
    if (S.ns == 6) ghost_copy_stress<R, 6>(P, si, go, cp.y, B.n);
  }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of int32 (out has N + 1 entries, out[N] = total)
// ---------------------------------------------------------------------------

#define SCAN_TILE 1024

__global__ void k_scan_local(const int *in, int *out, int *bsum, int N) {
  __shared__ int s[SCAN_TILE];
  const int t = threadIdx.x;  // 256 threads × 4
  const int base = blockIdx.x * SCAN_TILE;
  int v[4], tot = 0;
  for (int q = 0; q < 4; ++q) {
    int e = base + t * 4 + q;
    v[q] = e < N ? in[e] : 0;
    tot += v[q];
  }
  s[t] = tot;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    int a = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  int run = t > 0 ? s[t - 1] : 0;
  for (int q = 0; q < 4; ++q) {
    int e = base + t * 4 + q;
    if (e < N) out[e] = run;
    run += v[q];
  }
  if (t == 255) bsum[blockIdx.x] = s[255];
}

__global__ void k_scan_blocks(int *bsum, int nb, int *out, int N) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = t * per, b1 = min(nb, b0 + per);
  int s = 0;
  for (int k = b0; k < b1; ++k) s += bsum[k];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int k = b0; k < b1; ++k) {
    int c = bsum[k];
    bsum[k] = run;
    run += c;
  }
  if (t == 1023) out[N] = part[1023];
}

__global__ void k_scan_add(int *out, const int *bsum, int N) {
  const int e = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
  const int add = bsum[blockIdx.x];
  for (int q = 0; q < 4; ++q)
    if (e + q < N) out[e + q] += add;
}

// ---------------------------------------------------------------------------
// Grid / CSR (src/Interaction.jl:123-236)
// ---------------------------------------------------------------------------

// Grid of each listed interaction from the bounding boxes of its systems.
// inter_sys[2q] = a, inter_sys[2q+1] = b (or -1); cell arrays checked
// against their capacity.
__global__ void k_grid_params(const int *inter_list, int n, const int2 *inter_sys,
                              const double *cellsz, const long long *bbmin,
                              const long long *bbmax, const int *n_dev, Grid *grids,
                              const int2 *inter_carr, const CellArr *carr, int D, int *flags,
                              int *sysfault) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int it = inter_list[q];
  const int2 ab = inter_sys[it];
  const double cutoff = cellsz[it];
  Grid G;
  long long ncells = 1;
  for (int d = 0; d < D; ++d) {
    double mn = double_of(bbmin[ab.x * 3 + d]), mx = double_of(bbmax[ab.x * 3 + d]);
    if (ab.y >= 0 && n_dev[ab.y] > 0) {
      double mnb = double_of(bbmin[ab.y * 3 + d]), mxb = double_of(bbmax[ab.y * 3 + d]);
      mn = mnb < mn ? mnb : mn;
      mx = mxb > mx ? mxb : mx;
    }
    G.ming[d] = floor((mn - 2.0 * cutoff) / cutoff) * cutoff;  // _snap_to_grid :191
    double maxg = mx + 2.0 * cutoff;
    double ngd = floor((maxg - G.ming[d]) / cutoff) + 1.0;  // :179
    if (!(ngd >= 1.0) || ngd > 1.0e9) ngd = 1.0e9;
    G.ng[d] = (int)ngd;
    // saturate at 2^40 cells (far beyond any capacity): three factors of up
    // to 1e9 would overflow a 64-bit product
    ncells = ncells > (1LL << 40) / G.ng[d] ? (1LL << 40) : ncells * (long long)G.ng[d];
  }
  for (int d = D; d < 3; ++d) {
    G.ming[d] = 0.0;
    G.ng[d] = 1;
  }
  G.ok = 1;
  const int2 ca = inter_carr[it];
  long long cap = carr[ca.x].cap;
  if (ca.y >= 0 && carr[ca.y].cap < cap) cap = carr[ca.y].cap;
  if (ncells + 1 > cap || ncells < 1) {
    // an empty grid: the sweeps find no partner for these systems; the
    // systems are faulted, the other systems' grids are unaffected
    G.ok = 0;
    atomicOr(&flags[1], 2);
    atomicOr(&sysfault[ab.x], GR_FAULT_DOMAIN);
    if (ab.y >= 0) atomicOr(&sysfault[ab.y], GR_FAULT_DOMAIN);
    ncells = 0;
  }
  G.ncells = (int)ncells;
  G.pad = 0;
  grids[it] = G;
}

template <int D>
__device__ inline int cell_flat0(XP x, int64_t gi, const Grid &G,
                                 double cell) {  // _cell_1idx − 1 (:197-206)
  long long flat = 0, stride = 1;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    long long cd = (long long)floor((x[d][gi] - G.ming[d]) / cell);
    flat += cd * stride;
    stride *= G.ng[d];
  }
  return (int)flat;
}

// Zero the cells and write the sentinel n + 1 (:228).
__global__ void k_cells_clear(int *cells, const CellArr *carr, const Grid *grids,
                              const int *n_dev, const Blk *blk) {
  const Blk b = blk[blockIdx.x];
  const CellArr C = carr[b.item];
  const int ncells = grids[C.grid].ncells;
  for (int q = 0; q < 4; ++q) {
    int c = b.first + q * GR_BLOCK + threadIdx.x;
    if (c < C.cap && c <= ncells) cells[C.off + c] = (c == ncells) ? n_dev[C.sys] + 1 : 0;
  }
}

// First particle of each run of equal cells (forward pass :219-226); the
// last run of a cell wins, as the serial loop's last write does.
template <int D>
__global__ void k_cells_mark(XP x, int *cells, const CellArr *carr, const Grid *grids,
                             const double *cellsz, const SysDesc *sd, const int *n_dev,
                             const Blk *blk) {
  const Blk b = blk[blockIdx.x];
  const CellArr C = carr[b.item];
  const int i = b.first + threadIdx.x;
  const int n = n_dev[C.sys];
  const Grid G = grids[C.grid];
  if (G.ncells <= 0) return;  // uniform over the block
  __shared__ int cw[GR_BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool live = i < n;
  const double cell = cellsz[C.grid];
  const int64_t gi = sd[C.sys].off + i;
  const int c = live ? cell_flat0<D>(x, gi, G, cell) : -1;
  // the cell of particle i − 1: the previous lane's, the previous wavefront's
  // last lane's (through LDS), and for the workgroup's first thread its own
  // computation (the only one no other thread makes)
  const int c0 = (threadIdx.x == 0 && live && i > 0) ? cell_flat0<D>(x, gi - 1, G, cell) : -1;
  if (lane == 63) cw[w] = c;
  __syncthreads();
  int prev = __shfl_up(c, 1, 64);
  if (lane == 0) prev = w > 0 ? cw[w - 1] : c0;
  if (live && c != prev && c >= 0 && c < G.ncells) atomicMax(&cells[C.off + c], i + 1);
}

// Backward pass (:231-235): every empty cell takes the first non-empty start
// to its right.  Two kernels: block-local suffix (with the block's first
// non-empty start as its aggregate), then each block finds the first
// non-empty aggregate among the blocks after it and applies it.
#define FILL_TILE (GR_BLOCK * 4)

__device__ inline int first_nz(int a, int b) { return a != 0 ? a : b; }

__global__ void k_fill_local(int *cells, const CellArr *carr, const Grid *grids, const Blk *blk,
                             int *bagg) {
  __shared__ int s[GR_BLOCK / 64];
  const Blk b = blk[blockIdx.x];
  const CellArr C = carr[b.item];
  const int ncells = grids[C.grid].ncells;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int c0 = b.first + t * 4;
  int v[4];
  for (int q = 0; q < 4; ++q) {
    int c = c0 + q;
    v[q] = (c < C.cap && c <= ncells) ? cells[C.off + c] : 0;
  }
  // thread-local suffix, then the wavefront's (lanes lane..63) by shuffles
  for (int q = 2; q >= 0; --q) v[q] = first_nz(v[q], v[q + 1]);
  int a = v[0];
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_down(a, o, 64);
    if (lane + o < 64) a = first_nz(a, u);
  }
  if (lane == 0) s[w] = a;
  __syncthreads();
  int rw = 0;  // first non-zero in the wavefronts after this one
  for (int q = GR_BLOCK / 64 - 1; q > w; --q) rw = first_nz(s[q], rw);
  const int nx = __shfl_down(a, 1, 64);
  const int right = first_nz(lane < 63 ? nx : 0, rw);
  for (int q = 0; q < 4; ++q) {
    int c = c0 + q;
    if (c < C.cap && c <= ncells) cells[C.off + c] = first_nz(v[q], right);
  }
  if (t == 0) bagg[blockIdx.x] = first_nz(a, rw);
}

// The carry of a block is the first non-zero aggregate among the following
// blocks of its cell array (consecutive in the table, `first` restarting at 0
// for the next array), found 256 blocks at a time by each wavefront (4 per
// lane, loads independent of each other).  The array's sentinel cell (n + 1)
// is non-zero, so the search ends at the latest in the sentinel's block.
__global__ void k_fill_apply(int *cells, const CellArr *carr, const Grid *grids, const Blk *blk,
                             int nblk, const int *bagg) {
  const Blk b = blk[blockIdx.x];
  const CellArr C = carr[b.item];
  const int ncells = grids[C.grid].ncells;
  if (b.first > ncells) return;  // uniform: no cell of this block is in the grid
  const int lane = threadIdx.x & 63;
  int cr = 0;
  for (int k0 = blockIdx.x + 1; k0 < nblk; k0 += 256) {
    Blk bk[4];
    int g[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + lane * 4 + q;
      bk[q] = k < nblk ? blk[k] : Blk{-1, 0};
      g[q] = k < nblk ? bagg[k] : 0;
    }
    // this lane's first stop (end of the array or a non-zero aggregate)
    bool stop = false;
    int gs = 0;
#pragma unroll
    for (int q = 3; q >= 0; --q) {
      const bool same = bk[q].item == b.item && bk[q].first != 0;
      if (!same || g[q] != 0) {
        stop = true;
        gs = same ? g[q] : 0;
      }
    }
    const unsigned long long m = __ballot(stop);
    if (m != 0ull) {
      cr = __shfl(gs, __ffsll((long long)m) - 1, 64);  // 0 when the array ended first
      break;
    }
  }
  if (cr == 0) return;
  for (int q = 0; q < 4; ++q) {
    int c = b.first + threadIdx.x * 4 + q;
    if (c < C.cap && c <= ncells && cells[C.off + c] == 0) cells[C.off + c] = cr;
  }
}

// ---------------------------------------------------------------------------
// Leapfrog per-particle kernels (src/TimeIntegration.jl:645-685)
// ---------------------------------------------------------------------------

// Julia ^(Float64, Integer) (compensated power by squaring, fused muladd).
__device__ inline double jl_pow_int(double x, long long n) {
  if (n == 0) return 1.0;
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
  if (n == 3) return x * x * x;
  if (n < 0) {
    double rx = 1.0 / x;
    if (n == -2) return rx * rx;
    if (isfinite(x)) xnlo = -fma(x, rx, -1.0) * rx;
    x = rx;
    n = -n;
  }
  while (n > 1) {
    if (n & 1) {
      double err = fma(y, xnlo, x * ynlo);
      double hi = x * y;
      double lo = fma(x, y, -hi);
      y = hi;
      ynlo = lo + err;
    }
    double err = x * 2 * xnlo;
    double hi = x * x;
    double lo = fma(x, x, -hi);
    x = hi;
    xnlo = lo + err;
    n >>= 1;
  }
  double err = fma(y, xnlo, x * ynlo);
  return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

// Julia ^(Float64, Float64) for the EOS exponent.
__device__ inline double jl_pow(double x, double y) {
  if (x == 1.0) return 1.0;
  double t = trunc(y);
  if (t == y && fabs(y) < 1.0e18) {
    long long yi = (long long)y;
    if (yi == 0) return 1.0;
    if (yi >= -4096 && yi <= 3 * 8192) return jl_pow_int(x, yi);
  }
  return pow(x, y);
}

__device__ inline double eos_p(int kind, double rho, double rho0, double gamma, double c) {
  if (kind == GR_UPD_TAIT) return rho0 * c * c / gamma * (jl_pow(rho / rho0, gamma) - 1.0);
  return c * c * (rho - rho0);
}

// _halfstep_ps! (:158-162) writing the half-step state to the sweep shadows
// (v and ρ keep their start-of-step values for the full step), the
// FP64 half-step density, the stage-1 EOS when fused, and the source reset.
template <typename R, int D>
__global__ void k_halfstep(Pool<R> P, const SysDesc *sd, const StepSys *st, const Blk *blk,
                           double mult, int keep) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[s];
  if (i >= S.cap) return;
  const StepSys T = st[s];
  const int64_t g = S.off + i;
  // stage state q = q0 + (mult·dt)·dqdt: the leapfrog half step (mult ½) and
  // the RK4 stage advances (_rk4_advance_all!, src/TimeIntegration.jl:340-348;
  // keep: the first RK4 stage, q = q0)
  const double cf = mult * T.dt;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    double vh = keep ? P.v[d][g] : P.v[d][g] + cf * (double)P.dvdt[d][g];
    P.vs[d][g] = (R)vh;
  }
  const double rh = keep ? P.rho[g] : fma(cf, (double)P.drhodt[g], P.rho[g]);
  P.rho_h[g] = rh;
  P.rhos[g] = (R)rh;
  if (T.fuse_eos) P.p[g] = (R)eos_p(T.fuse_eos, rh, T.eos_rho0, T.eos_gamma, S.c);
  if (T.reset_force) {
#pragma unroll
    for (int d = 0; d < D; ++d) P.dvdt[d][g] = (R)S.sv[d];
    P.drhodt[g] = (R)S.srho;
  }
}

// RK4 (src/TimeIntegration.jl:350-362,801-811): damping of the stage rates,
// then acc += w·dqdt (vectors unfused, the density by the fused axpy).
template <typename R, int D>
__global__ void k_rk4_acc(Pool<R> P, const SysDesc *sd, const StepSys *st, const Blk *blk,
                          double w, int first) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[s];
  if (i >= S.cap) return;
  const StepSys T = st[s];
  const int64_t g = S.off + i;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    double a = (double)P.dvdt[d][g];
    if (T.damping) {
      a = a - T.gdt * (double)P.vs[d][g];
      P.dvdt[d][g] = (R)a;
      a = (double)P.dvdt[d][g];
    }
    P.acc[d][g] = (first ? 0.0 : P.acc[d][g]) + w * a;
  }
  P.acc[3][g] = fma(w, (double)P.drhodt[g], first ? 0.0 : P.acc[3][g]);
}

// q = q0 + dt·acc, then x += dt·v (_apply_acc_all!, _update_positions_all!).
template <typename R, int D>
__global__ void k_rk4_apply(Pool<R> P, const SysDesc *sd, const StepSys *st, const Blk *blk,
                            int pos) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[s];
  if (i >= S.cap) return;
  const StepSys T = st[s];
  const int64_t g = S.off + i;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double vn = P.v[d][g] + T.dt * P.acc[d][g];
    P.v[d][g] = vn;
    if (pos) P.x[d][g] = P.x[d][g] + T.dt * vn;
  }
  P.rho[g] = fma(T.dt, P.acc[3][g], P.rho[g]);
}

// _apply_damping_all! (:251-271), _fullstep_q_all! (:227-232),
// _update_positions_all! (:235-240).  v and x use separate products and sums
// (SVector arithmetic is outside @fastmath); ρ uses the fused axpy.
template <typename R, int D>
__global__ void k_fullstep(Pool<R> P, const SysDesc *sd, const StepSys *st, const Blk *blk,
                           int pos) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[s];
  if (i >= S.cap) return;
  const StepSys T = st[s];
  const int64_t g = S.off + i;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    double a = (double)P.dvdt[d][g];
    if (T.damping) {
      a = a - T.gdt * (double)P.vs[d][g];
      P.dvdt[d][g] = (R)a;
      a = (double)P.dvdt[d][g];
    }
    double vn = P.v[d][g] + T.dt * a;
    P.v[d][g] = vn;
    if (pos) P.x[d][g] = P.x[d][g] + T.dt * vn;
  }
  P.rho[g] = fma(T.dt, (double)P.drhodt[g], P.rho[g]);
}

// _xsph_correction! (src/TimeIntegration.jl:423-436) around the XSPH sweep.
// mode 0 (integrated systems): v ← v + (−1)·v_adjustment, v_adjustment ← 0,
//   and the sweep's view of v and ρ (the shadows) from the masters;
// mode 1 (virtual systems): shadows from the masters only;
// mode 2 (integrated systems): v ← v + 1·v_adjustment, then the position
//   update x ← x + dt·v (_update_positions_all!, :235-240).
template <typename R, int D>
__global__ void k_xsph(Pool<R> P, const SysDesc *sd, const StepSys *st, const int *n_dev,
                       const Blk *blk, int mode) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  if (i >= n_dev[s]) return;
  const int64_t g = sd[s].off + i;
  if (mode == 2) {
    const double dt = st[s].dt;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double vn = P.v[d][g] + 1.0 * P.vadj[d][g];
      P.v[d][g] = vn;
      P.x[d][g] = P.x[d][g] + dt * vn;
    }
    return;
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    double vn = P.v[d][g];
    if (mode == 0) {
      vn = vn + (-1.0) * P.vadj[d][g];
      P.v[d][g] = vn;
      P.vadj[d][g] = 0.0;
    }
    P.vs[d][g] = (R)vn;
  }
  P.rhos[g] = (R)P.rho[g];
}

// ---------------------------------------------------------------------------
// State updaters of one stage (src/StateUpdaters.jl)
// ---------------------------------------------------------------------------

// HookeLawStressUpdater (src/StateUpdaters.jl:366-405): elastic increment
// with the Jaumann correction; p = −(σxx + σyy + σzz)/3.
template <typename R>
__device__ __forceinline__ void hooke_update(Pool<R> P, int64_t g, double E, double nu, double dt) {
  const double sxx = P.stress[0][g], syy = P.stress[1][g], szz = P.stress[2][g],
               sxy = P.stress[3][g];
  const double dexx = (double)P.sr[0][g] * dt, deyy = (double)P.sr[1][g] * dt,
               dezz = (double)P.sr[2][g] * dt, dexy = (double)P.sr[3][g] * dt;
  const double D0 = E / ((1.0 + nu) * (1.0 - 2 * nu));
  double dsxx = D0 * ((1.0 - nu) * dexx + nu * deyy + nu * dezz);
  double dsyy = D0 * (nu * dexx + (1.0 - nu) * deyy + nu * dezz);
  const double dszz = D0 * (nu * dexx + nu * deyy + (1.0 - nu) * dezz);
  double dsxy = D0 * (1.0 - 2 * nu) * dexy;
  const double wdt = (double)P.vor[0][g] * dt;
  dsxx -= 2 * sxy * wdt;
  dsyy += 2 * sxy * wdt;
  dsxy += (sxx - syy) * wdt;
  const double n[4] = {sxx + dsxx, syy + dsyy, szz + dszz, sxy + dsxy};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    P.stress[k][g] = n[k];
    P.stress_s[k][g] = (R)n[k];
  }
  P.p[g] = (R)(-(n[0] + n[1] + n[2]) / 3.0);
}

// ElastoPlasticStressUpdater (:216-335), FP64.  prm = {E, nu, a_phi, k_c,
// a_psi} with the trigonometric constants evaluated on the host.
template <typename R>
__device__ __forceinline__ void ep_update(Pool<R> P, int64_t g, double E, double nu, double a_phi,
                                          double k_c, double a_psi, double dt) {
  double sxx = P.stress[0][g], syy = P.stress[1][g], szz = P.stress[2][g], sxy = P.stress[3][g];
  double dexx = (double)P.sr[0][g] * dt, deyy = (double)P.sr[1][g] * dt,
         dezz = (double)P.sr[2][g] * dt, dexy = (double)P.sr[3][g] * dt;
  double D0 = E / ((1.0 + nu) * (1.0 - 2 * nu));
  double dsxx = D0 * ((1.0 - nu) * dexx + nu * deyy + nu * dezz);
  double dsyy = D0 * (nu * dexx + (1.0 - nu) * deyy + nu * dezz);
  double dszz = D0 * (nu * dexx + nu * deyy + (1.0 - nu) * dezz);
  double dsxy = D0 * (1.0 - 2 * nu) * dexy;
  double wdt = (double)P.vor[0][g] * dt;
  dsxx -= 2 * sxy * wdt;
  dsyy += 2 * sxy * wdt;
  dsxy += (sxx - syy) * wdt;
  double txx = sxx + dsxx, tyy = syy + dsyy, tzz = szz + dszz, txy = sxy + dsxy;
  double I1 = txx + tyy + tzz;
  double p3 = I1 / 3;
  double s_xx = txx - p3, s_yy = tyy - p3, s_zz = tzz - p3, s_xy = txy;
  double J2 = 0.5 * (s_xx * s_xx + s_yy * s_yy + s_zz * s_zz + 2 * (s_xy * s_xy));
  double sqJ2 = sqrt(J2);
  double fy = a_phi * I1 + sqJ2 - k_c;
  double dpxx = 0, dpyy = 0, dpzz = 0, dpxy = 0;
  if (fy > 0.0) {
    double inv2 = 1.0 / (2 * sqJ2);
    double dfxx = a_phi + s_xx * inv2, dfyy = a_phi + s_yy * inv2, dfzz = a_phi + s_zz * inv2,
           dfxy = s_xy * inv2;
    double dgxx = a_psi + s_xx * inv2, dgyy = a_psi + s_yy * inv2, dgzz = a_psi + s_zz * inv2,
           dgxy = s_xy * inv2;
    double Dgxx = D0 * ((1.0 - nu) * dgxx + nu * dgyy + nu * dgzz);
    double Dgyy = D0 * (nu * dgxx + (1.0 - nu) * dgyy + nu * dgzz);
    double Dgzz = D0 * (nu * dgxx + nu * dgyy + (1.0 - nu) * dgzz);
    double Dgxy = D0 * (1.0 - 2 * nu) * dgxy;
    double denom = dfxx * Dgxx + dfyy * Dgyy + dfzz * Dgzz + 2 * dfxy * Dgxy;
    double dl = fy / denom;
    dpxx = dl * dgxx;
    dpyy = dl * dgyy;
    dpzz = dl * dgzz;
    dpxy = dl * dgxy;
    dsxx -= D0 * ((1.0 - nu) * dpxx + nu * dpyy + nu * dpzz);
    dsyy -= D0 * (nu * dpxx + (1.0 - nu) * dpyy + nu * dpzz);
    dszz -= D0 * (nu * dpxx + nu * dpyy + (1.0 - nu) * dpzz);
    dsxy -= D0 * (1.0 - 2 * nu) * dpxy;
  }
  double o0 = sxx + dsxx, o1 = syy + dsyy, o2 = szz + dszz, o3 = sxy + dsxy;
  if (sxx + dsxx + syy + dsyy + szz + dszz > k_c / a_phi) {
    double ap = k_c / a_phi / 3.0;
    o0 = ap;
    o1 = ap;
    o2 = ap;
    o3 = 0.0;
  }
  P.stress[0][g] = o0;
  P.stress[1][g] = o1;
  P.stress[2][g] = o2;
  P.stress[3][g] = o3;
  P.stress_s[0][g] = (R)o0;
  P.stress_s[1][g] = (R)o1;
  P.stress_s[2][g] = (R)o2;
  P.stress_s[3][g] = (R)o3;
  P.strain[0][g] += dexx;
  P.strain[1][g] += deyy;
  P.strain[2][g] += dezz;
  P.strain[3][g] += dexy;
  P.strainp[0][g] += dpxx;
  P.strainp[1][g] += dpyy;
  P.strainp[2][g] += dpzz;
  P.strainp[3][g] += dpxy;
}

// ViscoPlasticMCStressUpdater (:155-189).  prm = {eos_kind, rho0, gamma,
// tan(phi), cohesion}.
template <typename R>
__device__ __forceinline__ void vpmc_update(Pool<R> P, int64_t g, int NS, double p0, double p1,
                                            double p2, double p3, double p4, double c) {
  const double prm[5] = {p0, p1, p2, p3, p4};
  const double pval = eos_p((int)prm[0], P.rho_h[g], prm[1], prm[2], c);
  P.p[g] = (R)pval;
  double er[GR_MAXNS];
#pragma unroll
  for (int k = 0; k < GR_MAXNS; ++k) er[k] = k < NS ? (double)P.sr[k][g] : 0.0;
  double mag;
  if (NS == 3)
    mag = er[0] * er[0] + er[1] * er[1] + 2 * (er[2] * er[2]);
  else if (NS == 4)
    mag = er[0] * er[0] + er[1] * er[1] + 2 * (er[3] * er[3]);
  else
    mag = er[0] * er[0] + er[1] * er[1] + er[2] * er[2] +
          2 * (er[3] * er[3] + er[4] * er[4] + er[5] * er[5]);
  mag = sqrt(mag);
  if (!(mag >= 2.220446049250313e-16)) mag = isnan(mag) ? mag : 2.220446049250313e-16;
  const double sc = 2.0 * (prm[4] + pval * prm[3]);
  double sg[GR_MAXNS];
#pragma unroll
  for (int k = 0; k < GR_MAXNS; ++k) sg[k] = sc * er[k] / mag;
  bool zero;
  if (NS == 3) {
    sg[0] = sg[0] - pval;
    sg[1] = sg[1] - pval;
    sg[2] = sg[2] - 0.0;
    zero = sg[0] + sg[1] > 0.0;
  } else {
    sg[0] = sg[0] - pval;
    sg[1] = sg[1] - pval;
    sg[2] = sg[2] - pval;
#pragma unroll
    for (int k = 3; k < GR_MAXNS; ++k) sg[k] = sg[k] - 0.0;
    zero = sg[0] + sg[1] + sg[2] > 0.0;
  }
#pragma unroll
  for (int k = 0; k < GR_MAXNS; ++k)
    if (k < NS) {
      const double v = zero ? 0.0 : sg[k];
      P.stress[k][g] = v;
      P.stress_s[k][g] = (R)v;
    }
}

// The state updaters of one stage, in each system's order.  KM is the set of
// updater kinds (bit GR_UPD_x) the kernel handles: a stage is launched with
// the smallest instantiation that covers its kinds, so a stage of
// Drucker-Prager updaters does not carry the registers of the others.
template <typename R, int KM>
__global__ void k_updaters(Pool<R> P, const SysDesc *sd, const StepSys *st, const Blk *blk,
                           const int2 *urange, const UpdDesc *ud, int D) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[s];
  if (i >= S.cap) return;
  const int64_t g0 = S.off + i;
  const int2 r = urange[s];
  const double dt = st[s].dt;
  for (int q = r.x; q < r.y; ++q) {
    const UpdDesc u = ud[q];
    // the element address of each of the ~40 arrays would otherwise be
    // hoisted out of the loop, two registers apiece (spilled): g is made
    // opaque per iteration (an empty asm statement, no instruction)
    int64_t g = g0;
    asm volatile("" : "+v"(g));
    switch (u.kind) {
      case GR_UPD_ZERO: if constexpr ((KM >> GR_UPD_ZERO) & 1) {
        const long long m = u.mask;
#pragma unroll
        for (int k = 0; k < GR_MAXNS; ++k) {
          if (k >= S.ns) continue;
          if (m & (1 << GR_F_STRAIN_RATE)) P.sr[k][g] = (R)0;
          if (m & (1 << GR_F_STRESS)) {
            P.stress[k][g] = 0.0;
            P.stress_s[k][g] = (R)0;
          }
          if (m & (1 << GR_F_STRAIN)) P.strain[k][g] = 0.0;
          if (m & (1 << GR_F_STRAIN_P)) P.strainp[k][g] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if ((m & (1 << GR_F_VORTICITY)) && k < S.nv) P.vor[k][g] = (R)0;
          if ((m & (1 << GR_F_DVDT)) && k < D) P.dvdt[k][g] = (R)0;
        }
        if (m & (1 << GR_F_DRHODT)) P.drhodt[g] = (R)0;
        if (m & (1 << GR_F_P)) P.p[g] = (R)0;
        break;
      }
      case GR_UPD_TAIT:
      case GR_UPD_LINEAR:
        if constexpr ((KM >> GR_UPD_TAIT) & 1)
          P.p[g] = (R)eos_p(u.kind, P.rho_h[g], u.prm[0], u.prm[1], S.c);
        break;
      case GR_UPD_EP:
        if constexpr ((KM >> GR_UPD_EP) & 1)
          ep_update(P, g, u.prm[0], u.prm[1], u.prm[2], u.prm[3], u.prm[4], dt);
        break;
      case GR_UPD_VPMC:
        if constexpr ((KM >> GR_UPD_VPMC) & 1)
          vpmc_update(P, g, S.ns, u.prm[0], u.prm[1], u.prm[2], u.prm[3], u.prm[4], S.c);
        break;
      case GR_UPD_HOOKE:
        if constexpr ((KM >> GR_UPD_HOOKE) & 1) hooke_update(P, g, u.prm[0], u.prm[1], dt);
        break;
      case GR_UPD_VNORM: if constexpr ((KM >> GR_UPD_VNORM) & 1) {  // VirtualNormUpdater (src/StateUpdaters.jl:74-84)
        const long long m = u.mask;
        const double w = P.wsum[g];
        const bool z = w == 0.0;
        if (m & (1 << GR_F_V)) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            if (d >= D) continue;
            const double val = z ? 0.0 : P.v[d][g] / w;
            const double nv = val * u.prm[d] + S.pv[d];
            P.v[d][g] = nv;
            P.vs[d][g] = (R)nv;
          }
        }
        if (m & (1 << GR_F_RHO)) {
          const double val = z ? 0.0 : P.rho[g] / w;
          P.rho[g] = val;
          P.rho_h[g] = val;
          P.rhos[g] = (R)val;
        }
        if (m & (1 << GR_F_P)) P.p[g] = (R)(z ? 0.0 : (double)P.p[g] / w);
        if (m & (1 << GR_F_STRESS)) {
#pragma unroll
          for (int k = 0; k < GR_MAXNS; ++k) {
            if (k >= S.ns) continue;
            const double val = z ? 0.0 : P.stress[k][g] / w;
            P.stress[k][g] = val;
            P.stress_s[k][g] = (R)val;
          }
        }
        break;
      }
      case GR_UPD_PRESCRIBED_V: if constexpr ((KM >> GR_UPD_PRESCRIBED_V) & 1) {
        // PrescribedVelocityUpdater (:88-92)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (d >= D) continue;
          const double nv = P.v[d][g] + S.pv[d];
          P.v[d][g] = nv;
          P.vs[d][g] = (R)nv;
        }
        break;
      }
      default:
        break;
    }
  }
}

// Virtual systems: mode 0 refreshes the shadows of velocity and density
// after the sort, then clears w_sum and the zero fields before the stage
// loop (auto_zero_virtual!, src/Particles.jl:521-528); mode 1 advances the
// positions by prescribed_v·dt (_update_virtual_positions!,
// src/TimeIntegration.jl:389-397).
template <typename R, int D>
__global__ void k_virtual(Pool<R> P, const SysDesc *sd, const StepSys *st, const int *n_dev,
                          const Blk *blk, int mode) {
  const Blk b = blk[blockIdx.x];
  const int s = b.item;
  const int i = b.first + threadIdx.x;
  if (i >= n_dev[s]) return;
  const SysDesc S = sd[s];
  const int64_t g = S.off + i;
  if (mode == 1) {
    const double dt = st[s].dt;
#pragma unroll
    for (int d = 0; d < D; ++d) P.x[d][g] = P.x[d][g] + S.pv[d] * dt;
    return;
  }
  // the sort moved the masters; the single-copy shadows follow them
#pragma unroll
  for (int d = 0; d < D; ++d) P.vs[d][g] = (R)P.v[d][g];
  P.rhos[g] = (R)P.rho[g];
  P.rho_h[g] = P.rho[g];
  P.wsum[g] = 0.0;
  const long long m = S.zero_mask;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (m & (1 << GR_F_V)) {
      P.v[d][g] = 0.0;
      P.vs[d][g] = (R)0;
    }
    if (m & (1 << GR_F_DVDT)) P.dvdt[d][g] = (R)0;
  }
  if (m & (1 << GR_F_RHO)) {
    P.rho[g] = 0.0;
    P.rho_h[g] = 0.0;
    P.rhos[g] = (R)0;
  }
  if (m & (1 << GR_F_DRHODT)) P.drhodt[g] = (R)0;
  if (m & (1 << GR_F_P)) P.p[g] = (R)0;
  if (m & (1 << GR_F_COUNT)) P.pcnt[g] = 0.0;
#pragma unroll
  for (int k = 0; k < GR_MAXNS; ++k) {
    if (k >= S.ns) continue;
    if (m & (1 << GR_F_STRESS)) {
      P.stress[k][g] = 0.0;
      P.stress_s[k][g] = (R)0;
    }
    if (m & (1 << GR_F_STRAIN_RATE)) P.sr[k][g] = (R)0;
    if (m & (1 << GR_F_STRAIN)) P.strain[k][g] = 0.0;
    if (m & (1 << GR_F_STRAIN_P)) P.strainp[k][g] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if ((m & (1 << GR_F_VORTICITY)) && k < S.nv) P.vor[k][g] = (R)0;
}

// ---------------------------------------------------------------------------
// Small utilities
// ---------------------------------------------------------------------------

__global__ void k_fill_i32(int *p, int v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Partner records of the live particles of the block's system (positions,
// shadows of density and velocity, pressure), read by the pair sweeps.
// S_ONLY: the stress records alone (the compact records of the previous
// stage still hold: nothing in between wrote x, v, ρ or p).
template <typename R, int D, bool C_ONLY, bool WITH_S, bool INV = false, bool S_ONLY = false>
__global__ void k_pack(Pool<R> P, const SysDesc *sd, const int *n_dev, const Blk *blk) {
  const Blk b = blk[blockIdx.x];
  const int i = b.first + threadIdx.x;
  if (i >= n_dev[b.item]) return;
  const int64_t g = sd[b.item].off + i;
  if constexpr (!S_ONLY) {
  RecA<R> a;
  RecB<R> r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a.x[d] = d < D ? P.x[d][g] : 0.0;
    r.v[d] = d < D ? P.vs[d][g] : (R)0;
  }
  a.rho = P.rhos[g];
  if constexpr (sizeof(R) == 4) {
    a.p = P.p[g];
    r.pad = 0.0f;
  } else {
    r.p = P.p[g];
  }
  if (!C_ONLY) {
    P.ra[g] = a;
    P.rb[g] = r;
  }
  if constexpr (sizeof(R) == 4) {
    RecC cc;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cc.xq[d] = d < D ? (int32_t)quant_of(a.x[d], P.iqs) : 0;
      cc.v[d] = r.v[d];
    }
    if constexpr (INV) {  // 1/ρ and p/ρ² (the FluidPfn S = 2 force sweep)
      const float ir = __builtin_amdgcn_rcpf(a.rho);
      cc.rho = ir;
      cc.p = a.p * (ir * ir);
    } else {
      cc.rho = a.rho;
      cc.p = a.p;
    }
    P.rc[g] = cc;
  }
  }
  if constexpr (sizeof(R) == 4 && WITH_S) {  // stages with Cauchy sweeps
    const SysDesc S = sd[b.item];
    float sv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (S.kind >= GR_SYS_STRESS) {
#pragma unroll
      for (int k = 0; k < GR_MAXNS; ++k)
        if (k < S.ns) sv[k] = (float)P.stress_s[k][g];
    }
    // the whole record also in 2D, where the sweeps read only its first half:
    // writing half of every 32-byte record measured slower (partial lines)
    RecS rs;
    rs.a = make_float4(sv[0], sv[1], sv[2], sv[3]);
    rs.b = make_float4(sv[4], sv[5], sv[6], sv[7]);
    P.rs[g] = rs;
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static inline unsigned nb_of(long long n, int per) { return (unsigned)((n + per - 1) / per); }

void launch_key_minmax(hipStream_t s, XP xs, const SysDesc *sd, const int *n, const Blk *blk,
                       int nblk, int D, long long *bbmin, long long *bbmax, int rs,
                       const int *list, int nlist, bool fold) {
  if (nblk > 0) {
    // 8 table entries (2048 particles) per workgroup: enough workgroups to
    // fill the chip, few enough that the per-system atomics stay cheap
    constexpr int per = 8;
    const unsigned g = nb_of(nblk, per);
    if (D == 2)
      k_key_minmax<2, per><<<g, GR_BLOCK, 0, s>>>(xs, sd, n, blk, nblk, bbmin, bbmax, rs);
    else
      k_key_minmax<3, per><<<g, GR_BLOCK, 0, s>>>(xs, sd, n, blk, nblk, bbmin, bbmax, rs);
  }
  if (fold && nlist > 0)
    k_bb_fold<<<nb_of(nlist * 3, GR_BLOCK), GR_BLOCK, 0, s>>>(list, nlist, bbmin, bbmax, rs);
}
void launch_minmax_reset(hipStream_t s, const int *list, int nlist, long long *bbmin,
                         long long *bbmax, int rs) {
  if (nlist > 0)
    k_minmax_reset<<<nb_of(nlist * 3 * GR_BB_REP, GR_BLOCK), GR_BLOCK, 0, s>>>(list, nlist, bbmin,
                                                                               bbmax, rs);
}
void launch_key_base(hipStream_t s, const int *list, int nlist, int D, int pad, const int *n,
                     const StepSys *st, long long *bbmin, long long *bbmax, int rs, int *kmin,
                     int *kmax, unsigned *kbase, unsigned *krange, int *flags, int *sysfault) {
  k_key_base<<<1, KB_THREADS, 0, s>>>(list, nlist, D, pad, n, st, bbmin, bbmax, rs, kmin, kmax,
                                      kbase, krange, flags, sysfault);
}
void launch_key_write(hipStream_t s, XP xs, const SysDesc *sd, const int *n, const StepSys *st,
                      const Blk *blk, int nblk, const int *bbase, int D, const int *kmin,
                      const unsigned *krange, const unsigned *kbase, unsigned *keys, int *vals,
                      int *flags, unsigned *rscr) {
  if (nblk <= 0) return;
  if (D == 2)
    k_key_write<2><<<nblk, GR_BLOCK, 0, s>>>(xs, sd, n, st, blk, bbase, kmin, krange, kbase, keys,
                                             vals, flags, rscr);
  else
    k_key_write<3><<<nblk, GR_BLOCK, 0, s>>>(xs, sd, n, st, blk, bbase, kmin, krange, kbase, keys,
                                             vals, flags, rscr);
}
void launch_radix_sort(hipStream_t s, unsigned *keys, int *vals, unsigned *keys2, int *vals2,
                       unsigned *rscr, int N, int *flags) {
  if (N <= 0) return;
  const unsigned ntile = nb_of(N, OS_TILE);
  for (int pass = 0; pass < 4; ++pass) {
    const bool even = (pass & 1) == 0;
    k_radix_scan<<<256, GR_BLOCK, 0, s>>>(N, pass, rscr, flags);
    k_radix_scatter<<<ntile, GR_BLOCK, 0, s>>>(even ? keys : keys2, even ? vals : vals2,
                                               even ? keys2 : keys, even ? vals2 : vals, N, pass,
                                               rscr, flags);
  }
}
size_t radix_scratch_bytes(int N) {
  const size_t ntile = nb_of(N > 0 ? N : 1, OS_TILE);
  return ((size_t)RS_CNT + 2 * ntile * 256) * 4;
}

template <typename R>
void launch_permute_real(hipStream_t s, const Pool<R> &A, const Pool<R> &B, const SysDesc *sd,
                         const Blk *blk, int nblk, const int *bbase, const int *vals,
                         const int *vals2, const int *flags, int vadj, int D, const StepSys *st,
                         const int *halfsys, int pack) {
  if (nblk <= 0) return;
  if (halfsys) {
    auto go = [&](auto dc, auto pc) {
      constexpr int DD = decltype(dc)::value, PP = decltype(pc)::value;
      k_permute_real<R, DD, true, PP><<<nblk, GR_BLOCK, 0, s>>>(A, B, sd, blk, bbase, vals,
                                                                vals2, flags, vadj, st, halfsys);
    };
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    using P2 = std::integral_constant<int, 2>;
    if (sizeof(R) != 4) pack = 0;
    if (D == 2) {
      if (pack == 2) go(I2{}, P2{});
      else if (pack == 1) go(I2{}, P1{});
      else go(I2{}, P0{});
    } else {
      if (pack == 2) go(I3{}, P2{});
      else if (pack == 1) go(I3{}, P1{});
      else go(I3{}, P0{});
    }
    return;
  }
  if (D == 2)
    k_permute_real<R, 2, false><<<nblk, GR_BLOCK, 0, s>>>(A, B, sd, blk, bbase, vals, vals2,
                                                          flags, vadj, st, halfsys);
  else
    k_permute_real<R, 3, false><<<nblk, GR_BLOCK, 0, s>>>(A, B, sd, blk, bbase, vals, vals2,
                                                          flags, vadj, st, halfsys);
}
template void launch_permute_real<float>(hipStream_t, const Pool<float> &, const Pool<float> &,
                                         const SysDesc *, const Blk *, int, const int *,
                                         const int *, const int *, const int *, int, int,
                                         const StepSys *, const int *, int);
template void launch_permute_real<double>(hipStream_t, const Pool<double> &,
                                          const Pool<double> &, const SysDesc *, const Blk *,
                                          int, const int *, const int *, const int *, const int *,
                                          int, int, const StepSys *, const int *, int);

void launch_ghost_count(hipStream_t s, XP x, const SysDesc *sd, const Bnd *bnd, const Blk *blk,
                        int nblk, const int *gslot, int D, int *cnt) {
  if (nblk <= 0) return;
  if (D == 2)
    k_ghost_count<2><<<nblk, GR_BLOCK, 0, s>>>(x, sd, bnd, blk, gslot, cnt);
  else
    k_ghost_count<3><<<nblk, GR_BLOCK, 0, s>>>(x, sd, bnd, blk, gslot, cnt);
}
void launch_scan(hipStream_t s, const int *in, int *out, int *bsum, int N) {
  const int nb = (int)nb_of(N > 0 ? N : 1, SCAN_TILE);
  k_scan_local<<<nb, 256, 0, s>>>(in, out, bsum, N);
  k_scan_blocks<<<1, 1024, 0, s>>>(bsum, nb, out, N);
  k_scan_add<<<nb, 256, 0, s>>>(out, bsum, N);
}
int scan_block_count(int N) { return (int)nb_of(N > 0 ? N : 1, SCAN_TILE); }
void launch_ghost_scatter(hipStream_t s, XP x, XPW gx, int *gidx, int *gbnd, int *gnrm,
                          const SysDesc *sd, const Bnd *bnd, const Blk *blk, int nblk,
                          const int *gslot, const int *scan, int D, int *n) {
  if (nblk <= 0) return;
  if (D == 2)
    k_ghost_scatter<2><<<nblk, GR_BLOCK, 0, s>>>(x, gx, gidx, gbnd, gnrm, sd, bnd, blk, gslot,
                                                 scan, n);
  else
    k_ghost_scatter<3><<<nblk, GR_BLOCK, 0, s>>>(x, gx, gidx, gbnd, gnrm, sd, bnd, blk, gslot,
                                                 scan, n);
}
void launch_permute_ghost(hipStream_t s, XP gxs, const int *gidx_s, const int *gbnd_s, XPW x,
                          int *gidx, int *gbnd, const SysDesc *sd, const int *n, const Blk *blk,
                          int nblk, const int *bbase, const int *vals, const int *vals2,
                          const int *flags, int D) {
  if (nblk <= 0) return;
  if (D == 2)
    k_permute_ghost<2><<<nblk, GR_BLOCK, 0, s>>>(gxs, gidx_s, gbnd_s, x, gidx, gbnd, sd, n, blk,
                                                 bbase, vals, vals2, flags);
  else
    k_permute_ghost<3><<<nblk, GR_BLOCK, 0, s>>>(gxs, gidx_s, gbnd_s, x, gidx, gbnd, sd, n, blk,
                                                 bbase, vals, vals2, flags);
}
template <typename R>
void launch_ghost_kin(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const Bnd *bnd,
                      const int *n, const Blk *blk, int nblk, int D) {
  if (nblk <= 0) return;
  if (D == 2)
    k_ghost_kin<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, bnd, n, blk);
  else
    k_ghost_kin<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, bnd, n, blk);
}
template void launch_ghost_kin<float>(hipStream_t, const Pool<float> &, const SysDesc *,
                                      const Bnd *, const int *, const Blk *, int, int);
template void launch_ghost_kin<double>(hipStream_t, const Pool<double> &, const SysDesc *,
                                       const Bnd *, const int *, const Blk *, int, int);
template <typename R>
void launch_ghost_copy(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const Bnd *bnd,
                       const int *n, const Blk *blk, int nblk, const int2 *copy, int D) {
  (void)D;
  if (nblk > 0) k_ghost_copy<R><<<nblk, GR_BLOCK, 0, s>>>(P, sd, bnd, n, blk, copy);
}
template void launch_ghost_copy<float>(hipStream_t, const Pool<float> &, const SysDesc *,
                                       const Bnd *, const int *, const Blk *, int, const int2 *,
                                       int);
template void launch_ghost_copy<double>(hipStream_t, const Pool<double> &, const SysDesc *,
                                        const Bnd *, const int *, const Blk *, int,
                                        const int2 *, int);

void launch_grid_params(hipStream_t s, const int *inter_list, int n, const int2 *inter_sys,
                        const double *cellsz, const long long *bbmin, const long long *bbmax,
                        const int *n_dev, Grid *grids, const int2 *inter_carr,
                        const CellArr *carr, int D, int *flags, int *sysfault) {
  if (n > 0)
    k_grid_params<<<nb_of(n, 64), 64, 0, s>>>(inter_list, n, inter_sys, cellsz, bbmin, bbmax,
                                              n_dev, grids, inter_carr, carr, D, flags,
                                              sysfault);
}
void launch_cells_build(hipStream_t s, XP x, int *cells, const CellArr *carr,
                        const Grid *grids, const double *cellsz, const SysDesc *sd,
                        const int *n_dev, const Blk *cblk, int ncblk, const Blk *mblk,
                        int nmblk, int *bagg, int D) {
  if (ncblk > 0) k_cells_clear<<<ncblk, GR_BLOCK, 0, s>>>(cells, carr, grids, n_dev, cblk);
  if (nmblk > 0) {
    if (D == 2)
      k_cells_mark<2><<<nmblk, GR_BLOCK, 0, s>>>(x, cells, carr, grids, cellsz, sd, n_dev, mblk);
    else
      k_cells_mark<3><<<nmblk, GR_BLOCK, 0, s>>>(x, cells, carr, grids, cellsz, sd, n_dev, mblk);
  }
  if (ncblk > 0) {
    k_fill_local<<<ncblk, GR_BLOCK, 0, s>>>(cells, carr, grids, cblk, bagg);
    k_fill_apply<<<ncblk, GR_BLOCK, 0, s>>>(cells, carr, grids, cblk, ncblk, bagg);
  }
}

template <typename R>
void launch_halfstep(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                     const Blk *blk, int nblk, int D, double mult, int keep) {
  if (nblk <= 0) return;
  if (D == 2)
    k_halfstep<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, mult, keep);
  else
    k_halfstep<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, mult, keep);
}
template <typename R>
void launch_rk4_acc(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                    const Blk *blk, int nblk, int D, double w, int first) {
  if (nblk <= 0) return;
  if (D == 2)
    k_rk4_acc<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, w, first);
  else
    k_rk4_acc<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, w, first);
}
template <typename R>
void launch_rk4_apply(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                      const Blk *blk, int nblk, int D, int pos) {
  if (nblk <= 0) return;
  if (D == 2)
    k_rk4_apply<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, pos);
  else
    k_rk4_apply<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, pos);
}
template <typename R>
void launch_fullstep(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                     const Blk *blk, int nblk, int D, int pos) {
  if (nblk <= 0) return;
  if (D == 2)
    k_fullstep<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, pos);
  else
    k_fullstep<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, pos);
}
template <typename R>
void launch_xsph(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                 const int *n, const Blk *blk, int nblk, int D, int mode) {
  if (nblk <= 0) return;
  if (D == 2)
    k_xsph<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, n, blk, mode);
  else
    k_xsph<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, n, blk, mode);
}
template <typename R>
void launch_updaters(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                     const Blk *blk, int nblk, const int2 *urange, const UpdDesc *ud, int D,
                     int kinds) {
  if (nblk <= 0) return;
  constexpr int Z = 1 << GR_UPD_ZERO | 1 << GR_UPD_NOTHING;
  constexpr int K_EP = Z | 1 << GR_UPD_EP, K_VPMC = Z | 1 << GR_UPD_VPMC,
                K_EOS = Z | 1 << GR_UPD_TAIT | 1 << GR_UPD_LINEAR, K_HOOKE = Z | 1 << GR_UPD_HOOKE;
  auto go = [&](auto km) {
    k_updaters<R, decltype(km)::value><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, blk, urange, ud, D);
  };
  if ((kinds & ~K_EP) == 0)
    go(std::integral_constant<int, K_EP>{});
  else if ((kinds & ~K_VPMC) == 0)
    go(std::integral_constant<int, K_VPMC>{});
  else if ((kinds & ~K_EOS) == 0)
    go(std::integral_constant<int, K_EOS>{});
  else if ((kinds & ~K_HOOKE) == 0)
    go(std::integral_constant<int, K_HOOKE>{});
  else
    go(std::integral_constant<int, -1>{});
}
#define INST(R)                                                                                \
  template void launch_halfstep<R>(hipStream_t, const Pool<R> &, const SysDesc *,              \
                                   const StepSys *, const Blk *, int, int, double, int);       \
  template void launch_rk4_acc<R>(hipStream_t, const Pool<R> &, const SysDesc *,               \
                                  const StepSys *, const Blk *, int, int, double, int);        \
  template void launch_rk4_apply<R>(hipStream_t, const Pool<R> &, const SysDesc *,             \
                                    const StepSys *, const Blk *, int, int, int);              \
  template void launch_fullstep<R>(hipStream_t, const Pool<R> &, const SysDesc *,              \
                                   const StepSys *, const Blk *, int, int, int);               \
  template void launch_xsph<R>(hipStream_t, const Pool<R> &, const SysDesc *, const StepSys *, \
                               const int *, const Blk *, int, int, int);                       \
  template void launch_updaters<R>(hipStream_t, const Pool<R> &, const SysDesc *,              \
                                   const StepSys *, const Blk *, int, const int2 *,            \
                                   const UpdDesc *, int, int);
INST(float)
INST(double)
#undef INST

template <typename R>
void launch_virtual(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                    const int *n, const Blk *blk, int nblk, int mode, int D) {
  if (nblk <= 0) return;
  if (D == 2)
    k_virtual<R, 2><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, n, blk, mode);
  else
    k_virtual<R, 3><<<nblk, GR_BLOCK, 0, s>>>(P, sd, st, n, blk, mode);
}
template void launch_virtual<float>(hipStream_t, const Pool<float> &, const SysDesc *,
                                    const StepSys *, const int *, const Blk *, int, int, int);
template void launch_virtual<double>(hipStream_t, const Pool<double> &, const SysDesc *,
                                     const StepSys *, const int *, const Blk *, int, int, int);

template <typename R, bool C_ONLY, bool WITH_S>
static void launch_pack_d(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const int *n,
                          const Blk *blk, int nblk, int D) {
  if (D == 2)
    k_pack<R, 2, C_ONLY, WITH_S><<<nblk, GR_BLOCK, 0, s>>>(P, sd, n, blk);
  else
    k_pack<R, 3, C_ONLY, WITH_S><<<nblk, GR_BLOCK, 0, s>>>(P, sd, n, blk);
}

template <typename R>
void launch_pack(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const int *n, const Blk *blk,
                 int nblk, int D, bool c_only, bool with_s, bool inv, bool s_only) {
  if (nblk <= 0) return;
  c_only = c_only && sizeof(R) == 4;
  if (s_only) {  // FP32 compact records only
    if (!(c_only && with_s)) return;
    if (D == 2)
      k_pack<R, 2, true, true, false, true><<<nblk, GR_BLOCK, 0, s>>>(P, sd, n, blk);
    else
      k_pack<R, 3, true, true, false, true><<<nblk, GR_BLOCK, 0, s>>>(P, sd, n, blk);
    return;
  }
  if (inv && c_only && !with_s && sizeof(R) == 4) {
    if (D == 2)
      k_pack<R, 2, true, false, true><<<nblk, GR_BLOCK, 0, s>>>(P, sd, n, blk);
    else
      k_pack<R, 3, true, false, true><<<nblk, GR_BLOCK, 0, s>>>(P, sd, n, blk);
    return;
  }
  if (c_only && with_s)
    launch_pack_d<R, true, true>(s, P, sd, n, blk, nblk, D);
  else if (c_only)
    launch_pack_d<R, true, false>(s, P, sd, n, blk, nblk, D);
  else if (with_s)
    launch_pack_d<R, false, true>(s, P, sd, n, blk, nblk, D);
  else
    launch_pack_d<R, false, false>(s, P, sd, n, blk, nblk, D);
}
template void launch_pack<float>(hipStream_t, const Pool<float> &, const SysDesc *, const int *,
                                 const Blk *, int, int, bool, bool, bool, bool);
template void launch_pack<double>(hipStream_t, const Pool<double> &, const SysDesc *, const int *,
                                  const Blk *, int, int, bool, bool, bool, bool);

// Probe slot id[i] − 1 takes the position of mirror particle i
// (_measure_probes!, src/TimeIntegration.jl:469-479).
template <typename R, int D>
__global__ void k_probe_mirror(Pool<R> P, const SysDesc *sd, int probe, int src) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const SysDesc Sp = sd[probe], Ss = sd[src];
  if (i >= Ss.cap) return;
  const int q = P.id[Ss.off + i] - 1;
  if (q < 0 || q >= Sp.cap) return;
#pragma unroll
  for (int d = 0; d < D; ++d) P.x[d][Sp.off + q] = P.x[d][Ss.off + i];
}

// _sort_probe_by_id! (src/ProbeParticles.jl:150-158): ids are 1..n, so the
// sort is the scatter slot i -> id[i] − 1, done through a scratch copy:
// mode 0 scatters the probe's fields into scr (ncol columns of n doubles),
// mode 1 copies them back and refreshes the shadows.
template <typename R, int D>
__global__ void k_probe_by_id(Pool<R> P, const SysDesc *sd, int probe, double *scr, int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const SysDesc S = sd[probe];
  const int64_t n = S.cap;
  if (i >= n) return;
  const int64_t g = S.off + i;
  const int ns = S.kind >= GR_SYS_STRESS ? S.ns : 0;
  if (mode == 0) {
    const int64_t q = (int64_t)P.id[g] - 1;
    if (q < 0 || q >= n) return;
    int col = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) scr[(col++) * n + q] = P.x[d][g];
#pragma unroll
    for (int d = 0; d < D; ++d) scr[(col++) * n + q] = P.v[d][g];
    scr[(col++) * n + q] = P.rho[g];
    scr[(col++) * n + q] = P.wsum[g];
    scr[(col++) * n + q] = P.pcnt[g];
    scr[(col++) * n + q] = S.kind >= GR_SYS_FLUID ? (double)P.p[g] : 0.0;
    for (int k = 0; k < ns; ++k) scr[(col++) * n + q] = P.stress[k][g];
    return;
  }
  int col = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) P.x[d][g] = scr[(col++) * n + i];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double v = scr[(col++) * n + i];
    P.v[d][g] = v;
    P.vs[d][g] = (R)v;
  }
  const double rho = scr[(col++) * n + i];
  P.rho[g] = rho;
  P.rho_h[g] = rho;
  P.rhos[g] = (R)rho;
  P.wsum[g] = scr[(col++) * n + i];
  P.pcnt[g] = scr[(col++) * n + i];
  const double p = scr[(col++) * n + i];
  if (S.kind >= GR_SYS_FLUID) P.p[g] = (R)p;
  for (int k = 0; k < ns; ++k) {
    const double sg = scr[(col++) * n + i];
    P.stress[k][g] = sg;
    P.stress_s[k][g] = (R)sg;
  }
  P.id[g] = (int)(i + 1);
}

template <typename R>
void launch_probe_mirror(hipStream_t s, const Pool<R> &P, const SysDesc *sd, int probe, int src,
                         int64_t n_src, int D) {
  if (n_src <= 0) return;
  const unsigned nb = (unsigned)((n_src + GR_BLOCK - 1) / GR_BLOCK);
  if (D == 2)
    k_probe_mirror<R, 2><<<nb, GR_BLOCK, 0, s>>>(P, sd, probe, src);
  else
    k_probe_mirror<R, 3><<<nb, GR_BLOCK, 0, s>>>(P, sd, probe, src);
}
template <typename R>
void launch_probe_by_id(hipStream_t s, const Pool<R> &P, const SysDesc *sd, int probe, int64_t n,
                        double *scr, int mode, int D) {
  if (n <= 0) return;
  const unsigned nb = (unsigned)((n + GR_BLOCK - 1) / GR_BLOCK);
  if (D == 2)
    k_probe_by_id<R, 2><<<nb, GR_BLOCK, 0, s>>>(P, sd, probe, scr, mode);
  else
    k_probe_by_id<R, 3><<<nb, GR_BLOCK, 0, s>>>(P, sd, probe, scr, mode);
}
template void launch_probe_mirror<float>(hipStream_t, const Pool<float> &, const SysDesc *, int,
                                         int, int64_t, int);
template void launch_probe_mirror<double>(hipStream_t, const Pool<double> &, const SysDesc *, int,
                                          int, int64_t, int);
template void launch_probe_by_id<float>(hipStream_t, const Pool<float> &, const SysDesc *, int,
                                        int64_t, double *, int, int);
template void launch_probe_by_id<double>(hipStream_t, const Pool<double> &, const SysDesc *, int,
                                         int64_t, double *, int, int);

// A ghost-target XSPH task stores into its source's v_adjustment at the
// ghost's index: flag (bit 4 of flags[1]) a ghost count above the source's.
__global__ void k_alias_check(const int *n_dev, int ghost, int src_n, int *flags) {
  if (threadIdx.x == 0 && n_dev[ghost] > src_n) atomicOr(&flags[1], 4);
}
void launch_alias_check(hipStream_t s, const int *n, int ghost, int src_n, int *flags) {
  k_alias_check<<<1, 64, 0, s>>>(n, ghost, src_n, flags);
}

void launch_fill_i32(hipStream_t s, int *p, int v, long long n) {
  if (n > 0) k_fill_i32<<<nb_of(n, GR_BLOCK), GR_BLOCK, 0, s>>>(p, v, n);
}

}  // namespace gr
