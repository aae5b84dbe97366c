// gr_sweep.hip — fused gather sweep of the pair functors (gfx950).
//
// The reference visits each unordered pair once (half-shell colour sweep,
// src/Interaction.jl:393-541) and writes to both particles; the coupled sweep
// (:600-715) writes to system_a and, for two-sided functors, system_b.  Here
// every target particle instead gathers all its partners over the 3^D
// neighbour cells of each interaction it takes part in, evaluates the
// functor from its own side (which equals the reference's update of that
// particle: the pair formulas are antisymmetric/symmetric under i <-> j),
// accumulates in registers (Kahan-compensated in FP32) and writes once.  No
// atomics; the result is independent of scheduling.
//
// Pair selection is exact: dx and r² are formed in FP64 from FP64 positions,
// in the reference's order, so the accepted pair set equals the reference's.
#include <hip/hip_runtime.h>

#include "../../include/grasph_hip.h"
#include "gr_device.h"
#include "gr_launch.h"

namespace gr {

template <typename R>
struct Acc {  // Kahan-compensated sum
  R s = 0, c = 0;
  __device__ inline void add(R v) {
    const R y = v - c;
    const R t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ inline R get() const { return s - c; }
};
template <>
struct Acc<double> {
  double s = 0;
  __device__ inline void add(double v) { s += v; }
  __device__ inline double get() const { return s; }
};

// Julia's ^(Float64, Integer) for the LJ powers in FP64; plain products in FP32.
__device__ inline double pow_i(double x, int n) {
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
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
__device__ inline float pow_i(float x, int n) {
  float x2 = x * x, x3 = x2 * x, x6 = x3 * x3;
  return n == 12 ? x6 * x6 : x6;
}

// kernel_dw_dq (src/Kernels.jl:79-81 cubic, :151-153 Wendland C2).
template <typename R>
__device__ inline R dwdq(int kk, R C, R q) {
  if (kk == GR_KERNEL_CUBIC) {
    R t2 = (R)2 - q, t1 = (R)1 - q;
    t2 = t2 > (R)0 ? t2 : (R)0;
    t1 = t1 > (R)0 ? t1 : (R)0;
    return C * ((R)-3 * ((R)0.25 * (t2 * t2) - t1 * t1));
  }
  R t = (R)2 - q;
  t = t > (R)0 ? t : (R)0;
  return ((-C * (R)10) * q) * (t * t * t);
}

template <typename R>
__device__ inline R jmin0(R v) {  // Julia min(v, 0)
  return v < (R)0 ? v : (isnan(v) ? v : (R)0);
}

// Per-source constants in the arithmetic type.
template <typename R>
struct SrcR {
  R h, norm, hf, alpha, beta, mc, h2, m_q, psi_coef, eps, ljcut, ljc2;
  int sigma, has_delta, has_eps;
};

template <int D, typename R, int FAM, bool CNT>
__global__ void __launch_bounds__(GR_BLOCK)
    k_sweep(Pool<R> P, const SysDesc *__restrict__ sd, const int *__restrict__ n_dev,
            const Task *__restrict__ tasks, const SrcDesc *__restrict__ srcs,
            const Grid *__restrict__ grids, const CellArr *__restrict__ carr,
            const int *__restrict__ cells, const Blk *__restrict__ blk,
            unsigned long long *pair_count) {
  const Blk b = blk[blockIdx.x];
  const Task T = tasks[b.item];
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[T.sys];
  const bool live = i < n_dev[T.sys];
  const int64_t gi = S.off + (live ? i : 0);
  double xi[D];
  R vi[D];
  for (int d = 0; d < D; ++d) {
    xi[d] = P.x[d][gi];
    vi[d] = P.vs[d][gi];
  }
  const R ri = P.rhos[gi];
  R pi = 0;
  R si[GR_MAXNS];
  const int NS = T.ns;
  if (FAM == FAM_FORCE) {
    if (S.kind >= GR_SYS_FLUID) pi = P.p[gi];
    if (T.has_stress)
      for (int k = 0; k < NS; ++k) si[k] = P.stress_s[k][gi];
  }
  Acc<R> adv[D], adr;
  Acc<R> asr[GR_MAXNS], avo[3];
  unsigned long long npairs_self = 0, npairs_other = 0;

  for (int q = T.src_begin; q < T.src_end; ++q) {
    const SrcDesc Q = srcs[q];
    const Grid G = grids[Q.grid];
    if (G.ncells <= 0 || !live) continue;
    const int *cs = cells + carr[Q.carr].off;
    const int64_t poff = sd[Q.partner].off;
    SrcR<R> K;
    K.h = (R)Q.h;
    K.norm = (R)Q.norm;
    K.hf = (R)Q.hf;
    K.alpha = (R)Q.alpha;
    K.beta = (R)Q.beta;
    K.mc = (R)(0.5 * (Q.c_t_piv + Q.c_q_piv));
    K.h2 = (R)(Q.hf * Q.hf * 0.01);
    K.m_q = (R)Q.mass_q;
    K.psi_coef = (R)(2.0 * Q.delta * (0.5 * (Q.c_t_psi + Q.c_q_psi)) * Q.hf);
    K.eps = (R)Q.eps;
    K.ljcut = (R)Q.ljcut;
    K.ljc2 = (R)(0.01 * Q.c_t_piv * Q.c_t_piv);
    K.sigma = Q.sigma;
    K.has_delta = Q.has_delta;
    K.has_eps = Q.has_eps;
    // target cell in this grid (same formula as the CSR build)
    int c[D];
    for (int d = 0; d < D; ++d) {
      int cd = (int)floor((xi[d] - G.ming[d]) / Q.cell);
      cd = cd < 1 ? 1 : cd;
      cd = cd > G.ng[d] - 2 ? G.ng[d] - 2 : cd;
      c[d] = cd;
    }
    // dynamic boundary: da of the target (FP64)
    double da = 0.0;
    if (Q.pk == PK_DYNAMIC) {
      da = (xi[0] - Q.wp[0]) * Q.wn[0];
      for (int d = 1; d < D; ++d) da = da + (xi[d] - Q.wp[d]) * Q.wn[d];
    }
    const int nstrip = (D == 3) ? 9 : 3;
    for (int st = 0; st < nstrip; ++st) {
      int base;
      if (D == 3) {
        const int ddx = st / 3 - 1, ddy = st % 3 - 1;
        base = ((c[0] + ddx) * G.ng[1] + (c[1] + ddy)) * G.ng[2] + (c[D - 1] - 1);
      } else {
        base = (c[0] + st - 1) * G.ng[1] + (c[1] - 1);
      }
      const int j0 = cs[base] - 1, j1 = cs[base + 3] - 1;
      for (int j = j0; j < j1; ++j) {
        if (Q.pk == PK_SELF && j == i) continue;
        const int64_t gj = poff + j;
        double dx[D];
        for (int d = 0; d < D; ++d) dx[d] = xi[d] - P.x[d][gj];
        double r2 = dx[0] * dx[0];
        for (int d = 1; d < D; ++d) r2 = r2 + dx[d] * dx[d];
        if (!(r2 < Q.cutoff2)) continue;
        if (CNT) {
          if (Q.pk == PK_SELF)
            ++npairs_self;
          else
            ++npairs_other;
        }
        R dxr[D], gx[D];
        const R r2r = (R)r2;
        const R r = sqrt(r2r);
        const R qq = r / K.h;
        const R g = dwdq<R>(Q.kk, K.norm, qq) / (r * K.h);
        for (int d = 0; d < D; ++d) {
          dxr[d] = (R)dx[d];
          gx[d] = g * dxr[d];
        }
        if (FAM == FAM_COUNT) {
          adv[0].add((R)1);
          adr.add((R)1);
          continue;
        }
        // partner state
        R vj[D], rj, m = K.m_q;
        if (Q.pk == PK_DYNAMIC) {
          double db = (P.x[0][gj] - Q.wp[0]) * Q.wn[0];
          for (int d = 1; d < D; ++d) db = db + (P.x[d][gj] - Q.wp[d]) * Q.wn[d];
          double rat = fabs(db / da);
          double mn = isnan(rat) ? rat : (rat < Q.wbeta ? rat : Q.wbeta);
          const R fac = (R)(-mn);
          for (int d = 0; d < D; ++d) vj[d] = fac * vi[d];
          rj = ri;
        } else {
          for (int d = 0; d < D; ++d) vj[d] = P.vs[d][gj];
          rj = P.rhos[gj];
        }
        if (FAM == FAM_STRAIN) {
          R dv[D];
          for (int d = 0; d < D; ++d) dv[d] = vj[d] - vi[d];
          const R f = m / rj;
          R srv[GR_MAXNS];
          if (D == 2) {
            srv[0] = dv[0] * gx[0];
            srv[1] = dv[1] * gx[1];
            if (NS == 4) {
              srv[2] = (R)0;
              srv[3] = (R)0.5 * (dv[0] * gx[1] + dv[1] * gx[0]);
            } else {
              srv[2] = (R)0.5 * (dv[0] * gx[1] + dv[1] * gx[0]);
            }
          } else {
            srv[0] = dv[0] * gx[0];
            srv[1] = dv[1] * gx[1];
            srv[2] = dv[D - 1] * gx[D - 1];
            srv[3] = (R)0.5 * (dv[0] * gx[1] + dv[1] * gx[0]);
            srv[4] = (R)0.5 * (dv[0] * gx[D - 1] + dv[D - 1] * gx[0]);
            srv[5] = (R)0.5 * (dv[1] * gx[D - 1] + dv[D - 1] * gx[1]);
          }
          for (int k = 0; k < NS; ++k) asr[k].add(srv[k] * f);
          if (Q.pfn == GR_PFN_SRV) {
            avo[0].add((R)0.5 * (dv[0] * gx[1] - dv[1] * gx[0]) * f);
            if (D == 3) {
              avo[1].add((R)0.5 * (dv[0] * gx[D - 1] - dv[D - 1] * gx[0]) * f);
              avo[2].add((R)0.5 * (dv[1] * gx[D - 1] - dv[D - 1] * gx[1]) * f);
            }
          }
          continue;
        }
        // FAM_FORCE: FluidPfn / CauchyFluidPfn
        R dv[D];
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        R vr = dxr[0] * dv[0];
        for (int d = 1; d < D; ++d) vr = vr + dxr[d] * dv[d];
        vr = jmin0(vr);
        const R muv = K.hf * vr / (r2r + K.h2);
        const R piv = (K.beta * muv - K.alpha * K.mc) * muv / ((R)0.5 * ri * rj);
        if (Q.pk == PK_STATIC) {
          const R lr = K.ljcut / r;
          R f = (r < K.ljcut ? (R)1 : (R)0) * (pow_i(lr, 12) - pow_i(lr, 6)) / r2r;
          const R rf = K.ljc2 * f;
          const R t = -m * piv;
          for (int d = 0; d < D; ++d) adv[d].add(t * gx[d] + rf * dxr[d]);
          continue;
        }
        R dr = dv[0] * gx[0];
        for (int d = 1; d < D; ++d) dr = dr + dv[d] * gx[d];
        R psi = (R)0;
        if (K.has_delta) {
          R dxg = dxr[0] * gx[0];
          for (int d = 1; d < D; ++d) dxg = dxg + dxr[d] * gx[d];
          psi = K.psi_coef * (ri - rj) * dxg / r2r;
        }
        if (Q.pfn == GR_PFN_FLUID) {
          R pj;
          if (Q.pk == PK_DYNAMIC)
            pj = pi;
          else
            pj = P.p[gj];
          R dh;
          if (K.sigma == 1)
            dh = -(pi + pj) / (ri * rj);
          else
            dh = -(pi / (ri * ri) + pj / (rj * rj));
          if (Q.pk == PK_TWO_SIDED) {
            R ast = (R)0;
            if (K.has_eps) ast = K.eps * (-(fabs(pi) + fabs(pj)) / (ri * rj));
            const R t = dh + ast - piv;
            for (int d = 0; d < D; ++d) adv[d].add(m * (t * gx[d]));
          } else {
            const R t = m * (dh - piv);
            for (int d = 0; d < D; ++d) adv[d].add(t * gx[d]);
          }
          const R coef = (K.sigma == 1) ? ri / rj : (R)1;
          adr.add(m * (dr * coef + psi / rj));
        } else {  // CauchyFluidPfn
          R sj[GR_MAXNS];
          if (Q.pk == PK_DYNAMIC) {
            for (int k = 0; k < NS; ++k) sj[k] = si[k];
          } else {
            for (int k = 0; k < NS; ++k) sj[k] = P.stress_s[k][gj];
          }
          const R ci = (R)1 / (ri * ri), cj = (R)1 / (rj * rj);
          R hv[D];
          if (D == 2) {
            const int o = (NS == 3) ? 2 : 3;
            hv[0] = (si[0] * gx[0] + si[o] * gx[1]) * ci + (sj[0] * gx[0] + sj[o] * gx[1]) * cj;
            hv[1] = (si[o] * gx[0] + si[1] * gx[1]) * ci + (sj[o] * gx[0] + sj[1] * gx[1]) * cj;
          } else {
            hv[0] = (si[0] * gx[0] + si[3] * gx[1] + si[4] * gx[D - 1]) * ci +
                    (sj[0] * gx[0] + sj[3] * gx[1] + sj[4] * gx[D - 1]) * cj;
            hv[1] = (si[3] * gx[0] + si[1] * gx[1] + si[5] * gx[D - 1]) * ci +
                    (sj[3] * gx[0] + sj[1] * gx[1] + sj[5] * gx[D - 1]) * cj;
            hv[D - 1] = (si[4] * gx[0] + si[5] * gx[1] + si[2] * gx[D - 1]) * ci +
                        (sj[4] * gx[0] + sj[5] * gx[1] + sj[2] * gx[D - 1]) * cj;
          }
          for (int d = 0; d < D; ++d) adv[d].add(m * (hv[d] - piv * gx[d]));
          adr.add(m * (dr + psi / rj));
        }
      }
    }
  }
  if (CNT) {
    // [0]: ordered self pairs (each unordered pair seen from both sides),
    // [1]: coupled pairs seen from the gathering side.
    unsigned long long a = npairs_self, o2 = npairs_other;
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      o2 += __shfl_xor(o2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      if (a) atomicAdd(&pair_count[0], a);
      if (o2) atomicAdd(&pair_count[1], o2);
    }
  }
  if (!live) return;
  if (FAM == FAM_STRAIN) {
    for (int k = 0; k < NS; ++k) {
      const R v = asr[k].get();
      P.sr[k][gi] = T.assign ? v : P.sr[k][gi] + v;
    }
    if (T.write_vor)
      for (int k = 0; k < T.nv; ++k) {
        const R v = avo[k].get();
        P.vor[k][gi] = T.assign ? v : P.vor[k][gi] + v;
      }
    return;
  }
  for (int d = 0; d < D; ++d) {
    const R v = adv[d].get();
    P.dvdt[d][gi] = T.assign ? (R)S.sv[d] + v : P.dvdt[d][gi] + v;
  }
  {
    const R v = adr.get();
    P.drhodt[gi] = T.assign ? (R)S.srho + v : P.drhodt[gi] + v;
  }
}

template <int D, typename R, int FAM>
static void launch_one(hipStream_t s, const Pool<R> &P, const SweepArgs &a, bool cnt) {
  if (cnt)
    k_sweep<D, R, FAM, true><<<a.nblk, GR_BLOCK, 0, s>>>(P, a.sd, a.n, a.tasks, a.srcs, a.grids,
                                                          a.carr, a.cells, a.blk, a.pair_count);
  else
    k_sweep<D, R, FAM, false><<<a.nblk, GR_BLOCK, 0, s>>>(P, a.sd, a.n, a.tasks, a.srcs, a.grids,
                                                           a.carr, a.cells, a.blk, a.pair_count);
}

template <typename R>
void launch_sweep(hipStream_t s, const Pool<R> &P, const SweepArgs &a, int D, int fam, bool cnt) {
  if (a.nblk <= 0) return;
  if (D == 2) {
    if (fam == FAM_FORCE) launch_one<2, R, FAM_FORCE>(s, P, a, cnt);
    else if (fam == FAM_STRAIN) launch_one<2, R, FAM_STRAIN>(s, P, a, cnt);
    else launch_one<2, R, FAM_COUNT>(s, P, a, cnt);
  } else {
    if (fam == FAM_FORCE) launch_one<3, R, FAM_FORCE>(s, P, a, cnt);
    else if (fam == FAM_STRAIN) launch_one<3, R, FAM_STRAIN>(s, P, a, cnt);
    else launch_one<3, R, FAM_COUNT>(s, P, a, cnt);
  }
}
template void launch_sweep<float>(hipStream_t, const Pool<float> &, const SweepArgs &, int, int,
                                  bool);
template void launch_sweep<double>(hipStream_t, const Pool<double> &, const SweepArgs &, int, int,
                                   bool);

}  // namespace gr
