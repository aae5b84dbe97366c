// gr_launch.h — host-callable launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gr_device.h"

namespace gr {

struct SweepArgs {
  const SysDesc *sd;
  const int *n;
  const Task *tasks;
  const SrcDesc *srcs;
  const Grid *grids;
  const CellArr *carr;
  const int *cells;
  const Blk *blk;
  int nblk;
  unsigned long long *pair_count;
  const LTask *lt = nullptr;  // non-null: use the neighbour lists
  const int *lcnt = nullptr;
  const int *lidx = nullptr;
  bool fluid_only = false;  // FORCE: every source's functor is FluidPfn / FluidSolidPfn
  int block = GR_BLOCK;     // targets per workgroup (the block table's step)
  bool inv = false;         // FP32 FluidPfn S = 2 stage whose records hold 1/ρ, p/ρ²
};

struct NlistArgs {
  const SysDesc *sd;
  const int *n;
  const LTask *lt;
  const SrcDesc *srcs;
  const Grid *grids;
  const CellArr *carr;
  const int *cells;
  const Blk *blk;
  int nblk;
  int *lcnt;
  int *lidx;
  const long long *bbmin = nullptr;  // per-system position bounds of this step
  const long long *bbmax = nullptr;  // (ordered int64, see ord_of)
  int win = 128;  // candidates staged in LDS per window (64, 128 or 256)
  int span = 2;   // lanes whose ranges start within span·(widest range) group
};

// Position bounds per system and dimension are kept in GR_BB_REP replicas
// of `rs` entries each (the key pass's atomics are spread over them); replica
// 0 holds the folded bounds that every reader uses.  launch_key_minmax folds
// them itself when `fold`, launch_key_base always does.
#define GR_BB_REP 8
void launch_key_minmax(hipStream_t s, XP xs, const SysDesc *sd, const int *n, const Blk *blk,
                       int nblk, int D, long long *bbmin, long long *bbmax, int rs,
                       const int *list, int nlist, bool fold);
void launch_minmax_reset(hipStream_t s, const int *list, int nlist, long long *bbmin,
                         long long *bbmax, int rs);
void launch_key_base(hipStream_t s, const int *list, int nlist, int D, int pad, const int *n,
                     const StepSys *st, long long *bbmin, long long *bbmax, int rs, int *kmin,
                     int *kmax, unsigned *kbase, unsigned *krange, int *flags, int *sysfault);
void launch_key_write(hipStream_t s, XP xs, const SysDesc *sd, const int *n,
                      const StepSys *st, const Blk *blk, int nblk, const int *bbase, int D,
                      const int *kmin, const unsigned *krange, const unsigned *kbase,
                      unsigned *keys, int *vals, int *flags, unsigned *rscr);
// Stable radix sort of (keys, vals) by key, N entries.  `rscr` is the sort's
// scratch (radix_scratch_bytes(N), zeroed when allocated and left zeroed by
// every sort); launch_key_write counts the first pass into it and must run
// before it with the same scratch.  The sorted values end in vals when
// flags[3] is even and in vals2 when it is odd; the permutations take both.
// flags[1] bit 16: the counts did not place every key (not expected).
void launch_radix_sort(hipStream_t s, unsigned *keys, int *vals, unsigned *keys2, int *vals2,
                       unsigned *rscr, int N, int *flags);
size_t radix_scratch_bytes(int N);
template <typename R>
void launch_permute_real(hipStream_t s, const Pool<R> &A, const Pool<R> &B, const SysDesc *sd,
                         const Blk *blk, int nblk, const int *bbase, const int *vals,
                         const int *vals2, const int *flags, int vadj, int D,
                         const StepSys *st = nullptr, const int *halfsys = nullptr,
                         int pack = 0);
// ghost generation: per-(boundary, source block) counts at gslot[g] + ...,
// then the scatter at the counts' exclusive scan
void launch_ghost_count(hipStream_t s, XP x, const SysDesc *sd, const Bnd *bnd,
                        const Blk *blk, int nblk, const int *gslot, int D, int *cnt);
void launch_scan(hipStream_t s, const int *in, int *out, int *bsum, int N);
int scan_block_count(int N);
void launch_ghost_scatter(hipStream_t s, XP x, XPW gx, int *gidx,
                          int *gbnd, int *gnrm, const SysDesc *sd, const Bnd *bnd, const Blk *blk,
                          int nblk, const int *gslot, const int *scan, int D, int *n);
void launch_permute_ghost(hipStream_t s, XP gxs, const int *gidx_s,
                          const int *gbnd_s, XPW x, int *gidx, int *gbnd,
                          const SysDesc *sd, const int *n, const Blk *blk, int nblk,
                          const int *bbase, const int *vals, const int *vals2, const int *flags,
                          int D);
template <typename R>
void launch_ghost_kin(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const Bnd *bnd,
                      const int *n, const Blk *blk, int nblk, int D);
template <typename R>
void launch_ghost_copy(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const Bnd *bnd,
                       const int *n, const Blk *blk, int nblk, const int2 *copy, int D);
void launch_grid_params(hipStream_t s, const int *inter_list, int n, const int2 *inter_sys,
                        const double *cellsz, const long long *bbmin, const long long *bbmax,
                        const int *n_dev, Grid *grids, const int2 *inter_carr,
                        const CellArr *carr, int D, int *flags, int *sysfault);
void launch_cells_build(hipStream_t s, XP x, int *cells, const CellArr *carr,
                        const Grid *grids, const double *cellsz, const SysDesc *sd,
                        const int *n_dev, const Blk *cblk, int ncblk, const Blk *mblk,
                        int nmblk, int *bagg, int D);
template <typename R>
void launch_halfstep(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                     const Blk *blk, int nblk, int D, double mult, int keep);
template <typename R>
void launch_rk4_acc(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                    const Blk *blk, int nblk, int D, double w, int first);
template <typename R>
void launch_rk4_apply(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                      const Blk *blk, int nblk, int D, int pos);
template <typename R>
void launch_fullstep(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                     const Blk *blk, int nblk, int D, int pos);
template <typename R>
void launch_xsph(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                 const int *n, const Blk *blk, int nblk, int D, int mode);
template <typename R>
// kinds: the set of updater kinds (bit GR_UPD_x) in ud
void launch_updaters(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                     const Blk *blk, int nblk, const int2 *urange, const UpdDesc *ud, int D,
                     int kinds);
template <typename R>
void launch_sweep(hipStream_t s, const Pool<R> &P, const SweepArgs &a, int D, int fam, bool cnt);
template <typename R>
void launch_virtual(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const StepSys *st,
                    const int *n, const Blk *blk, int nblk, int mode, int D);
template <typename R>
// s_only: the FP32 stress records alone (the compact records still hold)
void launch_pack(hipStream_t s, const Pool<R> &P, const SysDesc *sd, const int *n, const Blk *blk,
                 int nblk, int D, bool c_only = false, bool with_s = true, bool inv = false,
                 bool s_only = false);
template <typename R>
void launch_nlist(hipStream_t s, const Pool<R> &P, const NlistArgs &a, int D);
void launch_fill_i32(hipStream_t s, int *p, int v, long long n);
void launch_alias_check(hipStream_t s, const int *n, int ghost, int src_n, int *flags);
template <typename R>
void launch_probe_mirror(hipStream_t s, const Pool<R> &P, const SysDesc *sd, int probe, int src,
                         int64_t n_src, int D);
template <typename R>
void launch_probe_by_id(hipStream_t s, const Pool<R> &P, const SysDesc *sd, int probe, int64_t n,
                        double *scr, int mode, int D);

}  // namespace gr
