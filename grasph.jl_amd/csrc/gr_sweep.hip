// gr_sweep.hip — neighbour search and fused pair-functor gather (gfx950).
//
// The reference visits each unordered pair once (half-shell colour sweep,
// src/Interaction.jl:393-541) and writes to both particles; the coupled sweep
// (:600-715) writes to system_a and, for two-sided functors, system_b.  Here
// every target particle instead gathers all its partners over the 3^D
// neighbour cells of each interaction it takes part in, evaluates the
// functor from its own side (which equals the reference's update of that
// particle: the pair formulas are antisymmetric/symmetric under i <-> j),
// accumulates in registers (FP64 sums of the pair terms) and writes once.  No
// atomics; the result is independent of scheduling.
//
// Two passes per step:
//  * k_nlist: neighbour search.  A wavefront stages the candidates of each
//    strip in LDS with coalesced loads and each lane keeps the candidates
//    in its own CSR ranges with r² < (2h)², evaluated in FP64 in the
//    reference's order, so the accepted pair set equals the reference's.
//  * k_sweep_list: the functors over each lane's accepted partners, once per
//    stage, so every lane does useful arithmetic on every iteration.  The
//    FP32 fluid-only force sweep reads 32-byte compact partner records
//    (fixed-point positions, see RecC) with two loads per pair.
// k_sweep walks the cells directly (fine-grained sweep! calls, and targets
// whose list overflowed).
#include <hip/hip_runtime.h>

#include "../../include/grasph_hip.h"
#include "gr_device.h"
#include "gr_launch.h"

namespace gr {

// Per-target sums of the pair terms, in FP64 whatever the arithmetic type of
// the terms: one conversion and one add per term (FP64 adds issue at the FP32
// rate), rounded to R once when stored.
template <typename R>
struct Acc {
  double s = 0;
  __device__ inline void add(R v) { s += (double)v; }
  __device__ inline R get() const { return (R)s; }
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
  if (n == 4) return x2 * x2;
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

// Reciprocals.  FP32 pair math uses the hardware reciprocal and reciprocal
// square root (1 ulp) and reuses each reciprocal within a pair; FP64 keeps
// correctly rounded division and square root.
template <typename R>
struct Fast;
template <>
struct Fast<float> {
  static constexpr bool on = true;
  __device__ static inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  __device__ static inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
};
template <>
struct Fast<double> {
  static constexpr bool on = false;
  __device__ static inline double rcp(double x) { return 1.0 / x; }
  __device__ static inline double rsq(double x) { return 1.0 / sqrt(x); }
};

// kernel_w (src/Kernels.jl:70-72 cubic, :142-144 Wendland C2).
template <typename R>
__device__ inline R wq(int kk, R C, R q) {
  if (kk == GR_KERNEL_CUBIC) {
    R t2 = (R)2 - q, t1 = (R)1 - q;
    t2 = t2 > (R)0 ? t2 : (R)0;
    t1 = t1 > (R)0 ? t1 : (R)0;
    return C * ((R)0.25 * (t2 * t2 * t2) - t1 * t1 * t1);
  }
  R t = (R)2 - q;
  t = t > (R)0 ? t : (R)0;
  return (C * pow_i(t, 4)) * ((R)2 * q + (R)1);
}

template <typename R>
__device__ inline R jmin0(R v) {  // Julia min(v, 0)
  return v < (R)0 ? v : (isnan(v) ? v : (R)0);
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of logical blocks so neighbouring cells share an L2.
__device__ inline int xcd_block(int nblk) {
  const int b = blockIdx.x;
  const int full = (nblk / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nblk / 8) + b / 8;
}

// Per-source constants in the arithmetic type.
template <typename R>
struct SrcK {
  R h, ih, norm, hf, alpha, beta, mc, h2, m_q, psi_coef, eps, ljcut, ljc2;
  int sigma, has_delta, has_eps;
  __device__ inline void load(const SrcDesc &Q) {
    h = (R)Q.h;
    ih = (R)(1.0 / Q.h);
    norm = (R)Q.norm;
    hf = (R)Q.hf;
    alpha = (R)Q.alpha;
    beta = (R)Q.beta;
    mc = (R)(0.5 * (Q.c_t_piv + Q.c_q_piv));
    h2 = (R)(Q.hf * Q.hf * 0.01);
    m_q = (R)Q.mass_q;
    psi_coef = (R)(2.0 * Q.delta * (0.5 * (Q.c_t_psi + Q.c_q_psi)) * Q.hf);
    eps = (R)Q.eps;
    ljcut = (R)Q.ljcut;
    ljc2 = (R)(0.01 * Q.c_t_piv * Q.c_t_piv);
    sigma = Q.sigma;
    has_delta = Q.has_delta;
    has_eps = Q.has_eps;
  }
};

// FORCE family specialised for stages whose force functors are all FluidPfn
// or FluidSolidPfn (no stress-based functor): the Cauchy branch and the
// target's stress are compiled out.  A template argument only, not a task
// family.
constexpr int FAM_FORCE_FLUID = NFAM + 1;
// The same for FluidPfn with S = 2, no δ-SPH and no surface tension, whose
// partner records hold 1/ρ and p/ρ² in place of ρ and p (k_pack INV): the
// reciprocal and the products the pair would compute are done once per
// particle, with the same FP32 operations, so the sums are unchanged.
constexpr int FAM_FORCE_FLUID_INV = NFAM + 2;

// Target particle: its state and its sums.
template <int D, typename R, int FAM>
struct Gather {
  static constexpr bool FORCE =
      FAM == FAM_FORCE || FAM == FAM_FORCE_FLUID || FAM == FAM_FORCE_FLUID_INV;
  static constexpr bool FLUID_ONLY = FAM == FAM_FORCE_FLUID || FAM == FAM_FORCE_FLUID_INV;
  static constexpr bool INV = FAM == FAM_FORCE_FLUID_INV;
  // FP32 fluid-only force sweeps read the compact records (RecC)
  static constexpr bool COMPACT = (FORCE || FAM == FAM_STRAIN) && sizeof(R) == 4;
  uint32_t xq[D];  // COMPACT: the target's fixed-point position
  double xi[D];
  R vi[D], ri, pi, iri;
  R si[GR_MAXNS];
  int NS;
  Acc<R> adv[D], adr;
  Acc<R> asr[GR_MAXNS], avo[3];

  __device__ inline void load(const Pool<R> P, const SysDesc &S, const Task &T, int64_t gi) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      xi[d] = P.x[d][gi];
      vi[d] = P.vs[d][gi];
    }
    if constexpr (COMPACT) {
#pragma unroll
      for (int d = 0; d < D; ++d) xq[d] = quant_of(xi[d], P.iqs);
    }
    ri = P.rhos[gi];
    iri = Fast<R>::rcp(ri);
    pi = 0;
    NS = T.ns;
    if (FORCE) {
      if (S.kind >= GR_SYS_FLUID) pi = P.p[gi];
      if (!FLUID_ONLY && T.has_stress) {
#pragma unroll
        for (int k = 0; k < GR_MAXNS; ++k) si[k] = k < NS ? P.stress_s[k][gi] : (R)0;
      }
    }
  }

  __device__ inline double da_of(const SrcDesc &Q) const {
    double da = (xi[0] - Q.wp[0]) * Q.wn[0];
#pragma unroll
    for (int d = 1; d < D; ++d) da = da + (xi[d] - Q.wp[d]) * Q.wn[d];
    return da;
  }

  // One accepted pair (target i, partner j at pool index gj), dx = x_i − x_j
  // and r² as evaluated by _pair_self!/_pair_coupled! (src/Interaction.jl:546-553).
  // DYN: 1 or 0 when the caller knows whether the source's partner is a
  // dynamic boundary (Q.pk == PK_DYNAMIC), -1 to test it here.
  template <int DYN = -1>
  __device__ __forceinline__ void pair(const Pool<R> P, const SrcDesc &Q, const SrcK<R> &K,
                                       int64_t gj, const RecA<R> &Aj, const RecB<R> &Bj,
                                       const RecC &Cj, const RecS &Sj, const double *dx,
                                       double r2, double da) {
    constexpr bool FA = Fast<R>::on;
    R dxr[D], gx[D];
    int32_t dq[D];  // COMPACT: x_i − x_j in quanta (exact)
    if constexpr (COMPACT) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dq[d] = (int32_t)(xq[d] - (uint32_t)Cj.xq[d]);
        dxr[d] = (R)dq[d] * (R)P.qs;
      }
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d) dxr[d] = (R)dx[d];
    }
    // in FP32 the squared distance is summed from the rounded components
    // (the FP64 r² has decided the acceptance already)
    R r2r;
    if constexpr (FA) {
      r2r = dxr[0] * dxr[0];
#pragma unroll
      for (int d = 1; d < D; ++d) r2r = fmaf(dxr[d], dxr[d], r2r);
    } else {
      r2r = (R)r2;
    }
    R r, g, ir = 0;
    if (FA) {
      ir = Fast<R>::rsq(r2r);
      r = r2r * ir;
      g = dwdq<R>(Q.kk, K.norm, r * K.ih) * (ir * K.ih);
    } else {
      r = sqrt(r2r);
      g = dwdq<R>(Q.kk, K.norm, r / K.h) / (r * K.h);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) gx[d] = g * dxr[d];
    if (FAM == FAM_COUNT) {
      adv[0].add((R)1);
      adr.add((R)1);
      return;
    }
    if (FAM == FAM_INTERP && Q.pfn == GR_PFN_NCOUNT) {  // NeighborCountFn (:443-460)
      avo[2].add((R)1);
      return;
    }
    if (FAM == FAM_INTERP) {  // InterpolateFieldFn (src/PairwiseFunctors.jl:424-434)
      // W from the FP64 distance in every mode: an accepted pair (r < 2h in
      // FP64) must weigh W > 0 as in the reference, or a virtual particle
      // whose partners all sit at the edge of the support would get
      // w_sum = 0, zeroed fields and a zero density while still having
      // partners in the force sweeps
      const R w = (R)wq<double>(Q.kk, Q.norm, sqrt(r2) / Q.h);
      const R rj = Aj.rho;
      const R kw = FA ? w * (K.m_q * Fast<R>::rcp(rj)) : w * (K.m_q / rj);
      const int m = Q.sigma;
      if (m & (1 << GR_F_V)) {
#pragma unroll
        for (int d = 0; d < D; ++d) adv[d].add(kw * Bj.v[d]);
      }
      if (m & (1 << GR_F_RHO)) adr.add(kw * rj);
      if (m & (1 << GR_F_P)) avo[0].add(kw * rec_pres(Aj, Bj));
      if (m & (1 << GR_F_STRESS)) {
#pragma unroll
        for (int k = 0; k < GR_MAXNS; ++k)
          if (k < NS) asr[k].add(kw * P.stress_s[k][gj]);
      }
      if (K.alpha != (R)0) avo[1].add(kw);
      return;
    }
    if (FAM == FAM_XSPH) {  // XSPHPfn (src/PairwiseFunctors.jl:364-390)
      // du = −ε·(v_i − v_j)·w/(ρ_i + ρ_j) (src/PairwisePhysics.jl:268-270);
      // the target gains du·m_partner, which is the reference's update of
      // either particle of the pair (du flips sign exactly under i <-> j)
      const R w = wq<R>(Q.kk, K.norm, FA ? r * K.ih : r / K.h);
      const R rs = ri + Aj.rho;
      const R ne = -K.alpha;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const R t = (ne * (vi[d] - Bj.v[d])) * w;
        const R du = FA ? t * Fast<R>::rcp(rs) : t / rs;
        adv[d].add(du * K.m_q);
      }
      return;
    }
    R vj[D], rj, pj = (R)0;
    const R m = K.m_q;
    const bool dyn = DYN < 0 ? Q.pk == PK_DYNAMIC : DYN == 1;
    if (dyn) {  // DynamicBoundarySystem: mirrored velocity
      double db;
      if constexpr (COMPACT) {  // (x_j − p)·n = (x_i − p)·n − (x_i − x_j)·n
        db = da;
#pragma unroll
        for (int d = 0; d < D; ++d) db = db - ((double)dq[d] * P.qs) * Q.wn[d];
      } else {
        db = (Aj.x[0] - Q.wp[0]) * Q.wn[0];
#pragma unroll
        for (int d = 1; d < D; ++d) db = db + (Aj.x[d] - Q.wp[d]) * Q.wn[d];
      }
      const double rat = fabs(db / da);
      const double mn = isnan(rat) ? rat : (rat < Q.wbeta ? rat : Q.wbeta);
      const R fac = (R)(-mn);
#pragma unroll
      for (int d = 0; d < D; ++d) vj[d] = fac * vi[d];
      rj = ri;
      pj = pi;
    } else if constexpr (COMPACT) {
#pragma unroll
      for (int d = 0; d < D; ++d) vj[d] = Cj.v[d];
      rj = Cj.rho;
      pj = Cj.p;
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d) vj[d] = Bj.v[d];
      rj = Aj.rho;
      pj = rec_pres(Aj, Bj);
    }
    R irj;
    if constexpr (INV)  // the record's 1/ρ (a dynamic partner mirrors the target)
      irj = dyn ? iri : rj;
    else  // a dynamic partner has rj = ri, so this is iri bit for bit (a select
          // of the two made the compiler keep them in scratch memory)
      irj = Fast<R>::rcp(rj);
    if (FAM == FAM_STRAIN) {  // strain_rate_tensor / vorticity_tensor
      R dv[D];
#pragma unroll
      for (int d = 0; d < D; ++d) dv[d] = vj[d] - vi[d];
      const R f = FA ? m * irj : m / rj;
      if (D == 2) {
        asr[0].add(dv[0] * gx[0] * f);
        asr[1].add(dv[1] * gx[1] * f);
        const R sh = (R)0.5 * (dv[0] * gx[1] + dv[1] * gx[0]) * f;
        if (NS == 4)
          asr[3].add(sh);
        else
          asr[2].add(sh);
      } else {
        asr[0].add(dv[0] * gx[0] * f);
        asr[1].add(dv[1] * gx[1] * f);
        asr[2].add(dv[D - 1] * gx[D - 1] * f);
        asr[3].add((R)0.5 * (dv[0] * gx[1] + dv[1] * gx[0]) * f);
        asr[4].add((R)0.5 * (dv[0] * gx[D - 1] + dv[D - 1] * gx[0]) * f);
        asr[5].add((R)0.5 * (dv[1] * gx[D - 1] + dv[D - 1] * gx[1]) * f);
      }
      if (Q.pfn == GR_PFN_SRV) {
        avo[0].add((R)0.5 * (dv[0] * gx[1] - dv[1] * gx[0]) * f);
        if (D == 3) {
          avo[1].add((R)0.5 * (dv[0] * gx[D - 1] - dv[D - 1] * gx[0]) * f);
          avo[2].add((R)0.5 * (dv[1] * gx[D - 1] - dv[D - 1] * gx[1]) * f);
        }
      }
      return;
    }
    // FAM_FORCE: FluidPfn / CauchyFluidPfn
    R dv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
    R vr = dxr[0] * dv[0];
#pragma unroll
    for (int d = 1; d < D; ++d) vr = vr + dxr[d] * dv[d];
    vr = jmin0(vr);
    const R muv = FA ? K.hf * vr * Fast<R>::rcp(r2r + K.h2) : K.hf * vr / (r2r + K.h2);
    const R piv = FA ? (K.beta * muv - K.alpha * K.mc) * muv * ((R)2 * iri * irj)
                     : (K.beta * muv - K.alpha * K.mc) * muv / ((R)0.5 * ri * rj);
    if (Q.pk == PK_STATIC) {  // lennard_jones + artificial viscosity
      const R lr = FA ? K.ljcut * ir : K.ljcut / r;
      const R lj = (r < K.ljcut ? (R)1 : (R)0) * (pow_i(lr, 12) - pow_i(lr, 6));
      const R f = FA ? lj * (ir * ir) : lj / r2r;
      const R rf = K.ljc2 * f;
      const R t = -m * piv;
#pragma unroll
      for (int d = 0; d < D; ++d) adv[d].add(t * gx[d] + rf * dxr[d]);
      return;
    }
    R dr = dv[0] * gx[0];
#pragma unroll
    for (int d = 1; d < D; ++d) dr = dr + dv[d] * gx[d];
    R psi = (R)0;
    if (K.has_delta) {
      R dxg = dxr[0] * gx[0];
#pragma unroll
      for (int d = 1; d < D; ++d) dxg = dxg + dxr[d] * gx[d];
      psi = FA ? K.psi_coef * (ri - rj) * dxg * (ir * ir) : K.psi_coef * (ri - rj) * dxg / r2r;
    }
    if (FLUID_ONLY || Q.pfn == GR_PFN_FLUID || Q.pfn == GR_PFN_FLUIDSOLID) {
      // FluidSolidPfn: the fluid's pressure on both sides (the target's on
      // the fluid side, the partner's on the solid side)
      const R pa = Q.psel == 2 ? pj : pi, pb = Q.psel == 1 ? pi : pj;
      R dh;
      if constexpr (INV) {  // S = 2; the record's p/ρ² for a real partner
        dh = -(pa * (iri * iri) + (dyn ? pb * (irj * irj) : pj));
      } else if (FA) {
        if (K.sigma == 1)
          dh = -(pa + pb) * (iri * irj);
        else
          dh = -(pa * (iri * iri) + pb * (irj * irj));
      } else {
        if (K.sigma == 1)
          dh = -(pa + pb) / (ri * rj);
        else
          dh = -(pa / (ri * ri) + pb / (rj * rj));
      }
      if (Q.pk == PK_TWO_SIDED) {
        R ast = (R)0;
        if (K.has_eps)
          ast = FA ? K.eps * (-(fabs(pi) + fabs(pj)) * (iri * irj))
                   : K.eps * (-(fabs(pi) + fabs(pj)) / (ri * rj));
        const R t = dh + ast - piv;
#pragma unroll
        for (int d = 0; d < D; ++d) adv[d].add(m * (t * gx[d]));
      } else {
        const R t = m * (dh - piv);
#pragma unroll
        for (int d = 0; d < D; ++d) adv[d].add(t * gx[d]);
      }
      const R coef = (K.sigma == 1) ? (FA ? ri * irj : ri / rj) : (R)1;
      adr.add(m * (dr * coef + (FA ? psi * irj : psi / rj)));
    } else {  // cauchy_stress_force
      R sj[GR_MAXNS];
      if (dyn) {
#pragma unroll
        for (int k = 0; k < GR_MAXNS; ++k) sj[k] = si[k];
      } else {
#pragma unroll
        for (int k = 0; k < GR_MAXNS; ++k) {
          if constexpr (COMPACT) {
            const float sk = k < 4 ? (&Sj.a.x)[k] : (&Sj.b.x)[k - 4];
            sj[k] = k < NS ? (R)sk : (R)0;
          } else {
            sj[k] = k < NS ? P.stress_s[k][gj] : (R)0;
          }
        }
      }
      const R ci = FA ? iri * iri : (R)1 / (ri * ri), cj = FA ? irj * irj : (R)1 / (rj * rj);
      R hv[D];
      if (D == 2) {
        R a0, a1;
        if (NS == 3) {
          a0 = (si[0] * gx[0] + si[2] * gx[1]) * ci + (sj[0] * gx[0] + sj[2] * gx[1]) * cj;
          a1 = (si[2] * gx[0] + si[1] * gx[1]) * ci + (sj[2] * gx[0] + sj[1] * gx[1]) * cj;
        } else {
          a0 = (si[0] * gx[0] + si[3] * gx[1]) * ci + (sj[0] * gx[0] + sj[3] * gx[1]) * cj;
          a1 = (si[3] * gx[0] + si[1] * gx[1]) * ci + (sj[3] * gx[0] + sj[1] * gx[1]) * cj;
        }
        hv[0] = a0;
        hv[D - 1] = a1;
      } else {
        hv[0] = (si[0] * gx[0] + si[3] * gx[1] + si[4] * gx[D - 1]) * ci +
                (sj[0] * gx[0] + sj[3] * gx[1] + sj[4] * gx[D - 1]) * cj;
        hv[1] = (si[3] * gx[0] + si[1] * gx[1] + si[5] * gx[D - 1]) * ci +
                (sj[3] * gx[0] + sj[1] * gx[1] + sj[5] * gx[D - 1]) * cj;
        hv[D - 1] = (si[4] * gx[0] + si[5] * gx[1] + si[2] * gx[D - 1]) * ci +
                    (sj[4] * gx[0] + sj[5] * gx[1] + sj[2] * gx[D - 1]) * cj;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) adv[d].add(m * (hv[d] - piv * gx[d]));
      adr.add(m * (dr + (FA ? psi * irj : psi / rj)));
    }
  }

  __device__ inline void store(const Pool<R> P, const SysDesc &S, const Task &T, int64_t gi) {
    if (FAM == FAM_INTERP) {  // += into the virtual system's fields (and shadows)
      const int m = T.imask;
      if (m & (1 << GR_F_V)) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const double nv = P.v[d][gi] + (double)adv[d].get();
          P.v[d][gi] = nv;
          P.vs[d][gi] = (R)nv;
        }
      }
      if (m & (1 << GR_F_RHO)) {
        const double nv = P.rho[gi] + (double)adr.get();
        P.rho[gi] = nv;
        P.rho_h[gi] = nv;
        P.rhos[gi] = (R)nv;
      }
      if (m & (1 << GR_F_P)) P.p[gi] = (R)((double)P.p[gi] + (double)avo[0].get());
      if (m & (1 << GR_F_STRESS)) {
#pragma unroll
        for (int k = 0; k < GR_MAXNS; ++k)
          if (k < NS) {
            const double nv = P.stress[k][gi] + (double)asr[k].get();
            P.stress[k][gi] = nv;
            P.stress_s[k][gi] = (R)nv;
          }
      }
      if (m & (1 << 30)) P.wsum[gi] += (double)avo[1].get();
      if (m & (1 << GR_F_COUNT)) P.pcnt[gi] += (double)avo[2].get();
      return;
    }
    if (FAM == FAM_XSPH) {
      int64_t gs = gi;
      if (T.alias >= 0) {
        const int64_t j = gi - S.off;
        if (j >= T.alias_cap) return;  // reported by k_alias_check
        gs = T.alias_off + j;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) P.vadj[d][gs] = P.vadj[d][gs] + (double)adv[d].get();
      return;
    }
    if (FAM == FAM_STRAIN) {
#pragma unroll
      for (int k = 0; k < GR_MAXNS; ++k)
        if (k < NS) {
          const R v = asr[k].get();
          P.sr[k][gi] = T.assign ? v : P.sr[k][gi] + v;
        }
      if (T.write_vor) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (k < T.nv) {
            const R v = avo[k].get();
            P.vor[k][gi] = T.assign ? v : P.vor[k][gi] + v;
          }
      }
      return;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const R v = adv[d].get();
      P.dvdt[d][gi] = T.assign ? (R)S.sv[d] + v : P.dvdt[d][gi] + v;
    }
    const R v = adr.get();
    P.drhodt[gi] = T.assign ? (R)S.srho + v : P.drhodt[gi] + v;
  }
};

// Target cell in the grid of a source, clamped so every strip lies inside.
template <int D>
__device__ inline void target_cell(const double *xi, const Grid &G, double cell, int *c) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int cd = (int)floor((xi[d] - G.ming[d]) / cell);
    cd = cd < 1 ? 1 : cd;
    cd = cd > G.ng[d] - 2 ? G.ng[d] - 2 : cd;
    c[d] = cd;
  }
}

// First cell of strip st (3 consecutive cells along the fastest dimension).
template <int D>
__device__ inline int strip_base(const int *c, const Grid &G, int st) {
  if (D == 3) {
    const int ddx = st / 3 - 1, ddy = st % 3 - 1;
    return ((c[0] + ddx) * G.ng[1] + (c[1] + ddy)) * G.ng[2] + (c[D - 1] - 1);
  }
  return (c[0] + st - 1) * G.ng[1] + (c[D - 1] - 1);
}

// Whether a pair of this family reads the partner's record B (velocity, and
// pressure in FP64): a dynamic boundary partner mirrors the target instead.
template <int FAM>
__device__ inline bool needs_b(const SrcDesc &Q) {
  if (FAM == FAM_COUNT) return false;
  if (FAM == FAM_INTERP) return Q.pfn != GR_PFN_NCOUNT;
  if (FAM == FAM_XSPH) return true;
  return Q.pk != PK_DYNAMIC;
}

// Whether a compact force pair reads the partner's stress record: Cauchy
// functors against a partner with its own stress (a dynamic boundary mirrors
// the target's).
template <int FAM>
__device__ inline bool needs_s(const SrcDesc &Q) {
  return FAM == FAM_FORCE && Q.pfn == GR_PFN_CAUCHY && Q.pk != PK_DYNAMIC;
}

// Direct cell walk of every source of the task (no list).
template <int D, typename R, int FAM>
__device__ void gather_direct(Gather<D, R, FAM> &Gt, const Pool<R> P, const SysDesc *sd,
                              const Task &T, const SrcDesc *srcs, const Grid *grids,
                              const CellArr *carr, const int *cells, int i,
                              unsigned long long &ns, unsigned long long &no) {
  for (int q = T.src_begin; q < T.src_end; ++q) {
    const SrcDesc Q = srcs[q];
    const Grid G = grids[Q.grid];
    if (G.ncells <= 0) continue;
    const int *cs = cells + carr[Q.carr].off;
    const int64_t poff = sd[Q.partner].off;
    SrcK<R> K;
    K.load(Q);
    int c[D];
    target_cell<D>(Gt.xi, G, Q.cell, c);
    const double da = (Q.pk == PK_DYNAMIC) ? Gt.da_of(Q) : 0.0;
    for (int st = 0; st < ((D == 3) ? 9 : 3); ++st) {
      const int base = strip_base<D>(c, G, st);
      const int j0 = cs[base] - 1, j1 = cs[base + 3] - 1;
      for (int j = j0; j < j1; ++j) {
        if (Q.pk == PK_SELF && j == i) continue;
        const int64_t gj = poff + j;
        constexpr bool CMP = Gather<D, R, FAM>::COMPACT;
        // compact sweeps read no RecA: the walk's FP64 test uses the positions
        RecA<R> Aj{};
        double dx[D];
        if constexpr (CMP) {
#pragma unroll
          for (int d = 0; d < D; ++d) dx[d] = Gt.xi[d] - P.x[d][gj];
        } else {
          Aj = P.ra[gj];
#pragma unroll
          for (int d = 0; d < D; ++d) dx[d] = Gt.xi[d] - Aj.x[d];
        }
        double r2 = dx[0] * dx[0];
#pragma unroll
        for (int d = 1; d < D; ++d) r2 = r2 + dx[d] * dx[d];
        if (!(r2 < Q.cutoff2)) continue;
        if (Q.pk == PK_SELF)
          ++ns;
        else
          ++no;
        Gt.pair(P, Q, K, gj, Aj, (!CMP && needs_b<FAM>(Q)) ? P.rb[gj] : RecB<R>{},
                CMP ? P.rc[gj] : RecC{}, (CMP && needs_s<FAM>(Q)) ? P.rs[gj] : RecS{}, dx,
                r2, da);
      }
    }
  }
}

template <bool CNT>
__device__ inline void count_pairs(unsigned long long *pc, unsigned long long a,
                                   unsigned long long o2) {
  if (!CNT) return;
  // [0]: ordered self pairs (each unordered pair seen from both sides),
  // [1]: coupled pairs seen from the gathering side.
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    o2 += __shfl_xor(o2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(&pc[0], a);
    if (o2) atomicAdd(&pc[1], o2);
  }
}

template <int D, typename R, int FAM, bool CNT>
__global__ void __launch_bounds__(GR_BLOCK)
    k_sweep(Pool<R> P, const SysDesc *__restrict__ sd, const int *__restrict__ n_dev,
            const Task *__restrict__ tasks, const SrcDesc *__restrict__ srcs,
            const Grid *__restrict__ grids, const CellArr *__restrict__ carr,
            const int *__restrict__ cells, const Blk *__restrict__ blk, int nblk,
            unsigned long long *pair_count) {
  const Blk b = blk[xcd_block(nblk)];
  const Task T = tasks[b.item];
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[T.sys];
  const bool live = i < n_dev[T.sys];
  const int64_t gi = S.off + (live ? i : 0);
  Gather<D, R, FAM> Gt;
  Gt.load(P, S, T, gi);
  unsigned long long ns = 0, no = 0;
  if (live) gather_direct<D, R, FAM>(Gt, P, sd, T, srcs, grids, carr, cells, i, ns, no);
  count_pairs<CNT>(pair_count, ns, no);
  if (live) Gt.store(P, S, T, gi);
}

// ---------------------------------------------------------------------------
// Neighbour lists
// ---------------------------------------------------------------------------

// Accepted partners of every target of a list task, per source, in the
// order the sources are listed, as acceptance masks: an entry, a 64-bit mask
// {lo, hi} and a base, says that candidates base + b of the source
// (source-local indices, b = 0..63) with bit b set are partners.  Entries
// follow the cell walk (strips in order, ascending candidates), so popping the
// bits in order visits the partners in the walk's order.  The masks and the
// bases are two arrays: entry e of target i has its mask at
// lidx[L.list_off + 2·(e·L.cap + i)] (an int2) and its base at
// lidx[L.list_off + 2·L.maxnb·L.cap + e·L.cap + i]; a lane writes one
// 12-byte entry per (strip, 64 candidates) that holds a partner, and a
// wavefront reads a row of entries as 512 and 256 contiguous bytes.
// lcnt[L.cnt_off + q·L.cap + i] holds the number of entries of source q.
// More than L.maxnb entries mean "overflow" (the physics pass then walks the
// cells directly).
//
// One wavefront per workgroup.  For each strip, lanes whose candidate ranges
// start within span·wmax of the lowest pending start form a group (targets
// in one cell column); the group's union of ranges is staged in LDS, WIN
// candidates at a time, with coalesced loads, and every lane walks its own
// range from LDS in ascending order, so it tests each candidate once with the
// walk's FP64 expression and sets the candidate's bit: no per-partner store.
// WIN and span change only which lanes load together, never a lane's tests
// or their order: every choice gives the same lists.
// Wave-wide minimum / maximum (all 64 lanes active): DPP within each row of
// 16 lanes, then the four row results through scalar registers.
template <bool MAX>
__device__ inline int wave_red_dpp(int v) {
  auto op = [](int x, int y) { return MAX ? (x > y ? x : y) : (x < y ? x : y); };
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
  return op(op(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
            op(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

template <int D, typename R, int WIN>
__global__ void __launch_bounds__(64)
    k_nlist(Pool<R> P, const SysDesc *__restrict__ sd, const int *__restrict__ n_dev,
            const LTask *__restrict__ lt, const SrcDesc *__restrict__ srcs,
            const Grid *__restrict__ grids, const CellArr *__restrict__ carr,
            const int *__restrict__ cells, const Blk *__restrict__ blk, int nblk, int *lcnt,
            int *lidx, const long long *__restrict__ bbmin,
            const long long *__restrict__ bbmax, int span) {
  __shared__ double sx[D][WIN];
  const Blk b = blk[xcd_block(nblk)];
  const LTask L = lt[b.item];
  const int lane = threadIdx.x;
  const int i = b.first + lane;
  const SysDesc S = sd[L.sys];
  const bool live = i < n_dev[L.sys];
  const int64_t gi = S.off + (live ? i : 0);
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = P.x[d][gi];
  int2 *emask = reinterpret_cast<int2 *>(lidx + L.list_off) + i;
  int *ebase = lidx + L.list_off + 2 * (int64_t)L.maxnb * L.cap + i;
  int ne = 0;  // entries of this target
  int cq = 0;  // entries of the current source
  for (int q = L.src_begin; q < L.src_end; ++q) {
    const SrcDesc Q = srcs[q];
    const Grid G = grids[Q.grid];
    cq = 0;
    if (G.ncells > 0) {
      const int *cs = cells + carr[Q.carr].off;
      // the target itself (a self source) passes the distance test with
      // r² = 0; its bit is cleared when a mask is stored
      const int self_j = Q.pk == PK_SELF ? i : -1;
      const double c2 = Q.cutoff2;
      const int64_t poff = sd[Q.partner].off;
      // a dynamic boundary (a sorted real system, its bounds taken this
      // step) more than a cutoff from the target along some axis holds no
      // partner: the margin keeps every FP64 r² of such a pair above c2
      bool act = live;
      if (Q.pk == PK_DYNAMIC && bbmin) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const double mn = double_of(bbmin[Q.partner * 3 + d]);
          const double mx = double_of(bbmax[Q.partner * 3 + d]);
          const double gap = mn - xi[d] > xi[d] - mx ? mn - xi[d] : xi[d] - mx;
          act = act && !(gap > Q.cell * (1.0 + 1e-9));
        }
      }
      int c[D];
      target_cell<D>(xi, G, Q.cell, c);
      constexpr int NST = (D == 3) ? 9 : 3;
      // the next strip's range is loaded while the current strip is tested;
      // a lane past the system's end keeps the empty range [0, 0) (a
      // negative base would run its candidate loop)
      int n0 = 1, n1 = 1;
      if (act) {
        const int sb = strip_base<D>(c, G, 0);
        n0 = cs[sb];
        n1 = cs[sb + 3];
      }
      for (int st = 0; st < (__builtin_amdgcn_ballot_w64(act) != 0ull ? NST : 0); ++st) {
        const int j0 = n0 - 1, j1 = n1 - 1;
        if (act && st + 1 < NST) {
          const int sb = strip_base<D>(c, G, st + 1);
          n0 = cs[sb];
          n1 = cs[sb + 3];
        }
        bool pend = j0 < j1;
        if (__builtin_amdgcn_ballot_w64(pend) == 0ull) continue;
        // mask of candidates base .. base + 63 in two words; d = j − base
        int base = j0, d = 0;
        unsigned lo = 0u, hi = 0u;
        auto flush = [&]() {
          const unsigned sdl = (unsigned)(self_j - base);
          const unsigned sbit = 1u << (sdl & 31u);
          lo &= sdl < 32u ? ~sbit : ~0u;
          hi &= sdl - 32u < 32u ? ~sbit : ~0u;
          if (live && (lo | hi) != 0u) {
            if (ne < L.maxnb) {
              emask[(int64_t)ne * L.cap] = make_int2((int)lo, (int)hi);
              ebase[(int64_t)ne * L.cap] = base;
            }
            ++ne;
            ++cq;
          }
          lo = hi = 0u;
        };
        const int wmax = wave_red_dpp<true>(j1 - j0);
        while (true) {
          const int glo = wave_red_dpp<false>(pend ? j0 : INT_MAX);
          if (glo == INT_MAX) break;
          const bool mem = pend && j0 - glo <= span * wmax;
          const int ghi = wave_red_dpp<true>(mem ? j1 : INT_MIN);
          for (int w0 = glo; w0 < ghi; w0 += WIN) {
            const int w1 = ghi - w0 < WIN ? ghi : w0 + WIN;
            __syncthreads();  // the previous window has been read
            for (int t = lane; t < w1 - w0; t += 64) {
#pragma unroll
              for (int k = 0; k < D; ++k) sx[k][t] = P.x[k][poff + w0 + t];
            }
            __syncthreads();
            // this lane's candidates in the window end before je; base + d is
            // the next one (where the previous window stopped)
            const int je = mem ? (j1 < w1 ? j1 : w1) : 0;
            while (base + d < je) {
              if (d == 64) {  // the mask is full: store it, move it on
                flush();
                base += 64;
                d = 0;
              }
              // a run of candidates within one 32-bit word of the mask
              const int wsel = d >> 5;
              const int lim = je - base, wend = (d & ~31) + 32;
              const int dend = lim < wend ? lim : wend;
              const int toff = base - w0;
              unsigned cur = 0u;
              // four, then two candidates per iteration (independent chains)
              for (; d + 3 < dend; d += 4) {
                const int t = d + toff;
                double r[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const double a0 = xi[0] - sx[0][t + u];
                  r[u] = a0 * a0;
                }
#pragma unroll
                for (int k = 1; k < D; ++k)
#pragma unroll
                  for (int u = 0; u < 4; ++u) {
                    const double ak = xi[k] - sx[k][t + u];
                    r[u] = r[u] + ak * ak;
                  }
                cur |= ((unsigned)(r[0] < c2) | ((unsigned)(r[1] < c2) << 1) |
                        ((unsigned)(r[2] < c2) << 2) | ((unsigned)(r[3] < c2) << 3))
                       << (d & 31);
              }
              for (; d + 1 < dend; d += 2) {
                const int t = d + toff;
                const double a0 = xi[0] - sx[0][t], b0 = xi[0] - sx[0][t + 1];
                double ra = a0 * a0, rb = b0 * b0;
#pragma unroll
                for (int k = 1; k < D; ++k) {
                  const double ak = xi[k] - sx[k][t], bk = xi[k] - sx[k][t + 1];
                  ra = ra + ak * ak;
                  rb = rb + bk * bk;
                }
                cur |= ((unsigned)(ra < c2) | ((unsigned)(rb < c2) << 1)) << (d & 31);
              }
              if (d < dend) {
                const int t = d + toff;
                const double a0 = xi[0] - sx[0][t];
                double r2 = a0 * a0;
#pragma unroll
                for (int k = 1; k < D; ++k) {
                  const double ak = xi[k] - sx[k][t];
                  r2 = r2 + ak * ak;
                }
                cur |= (unsigned)(r2 < c2) << (d & 31);
                ++d;
              }
              // selects, not branches (here and when the target's own bit
              // is cleared): branches make the compiler index a pair
              // {lo, hi} kept in scratch memory
              lo |= wsel == 0 ? cur : 0u;
              hi |= wsel != 0 ? cur : 0u;
            }
          }
          if (mem) pend = false;
        }
        flush();
      }
    }
    if (live) lcnt[L.cnt_off + (int64_t)(q - L.src_begin) * L.cap + i] = cq;
  }
  if (live)
    lcnt[L.cnt_off + (int64_t)(L.src_end - L.src_begin) * L.cap + i] = ne > L.maxnb ? 1 : 0;
}

// Functors over the accepted partners of each target (one stage).
template <int D, typename R, int FAM, bool CNT>
__global__ void __launch_bounds__(GR_BLOCK)
    k_sweep_list(Pool<R> P, const SysDesc *__restrict__ sd, const int *__restrict__ n_dev,
                 const Task *__restrict__ tasks, const SrcDesc *__restrict__ srcs,
                 const Grid *__restrict__ grids, const CellArr *__restrict__ carr,
                 const int *__restrict__ cells, const Blk *__restrict__ blk, int nblk,
                 unsigned long long *pair_count, const LTask *__restrict__ lt,
                 const int *__restrict__ lcnt, const int *__restrict__ lidx) {
  const Blk b = blk[xcd_block(nblk)];
  const Task T = tasks[b.item];
  // a task without a list (ltask < 0) walks the cells
  const LTask L = lt[T.ltask >= 0 ? T.ltask : 0];
  const int i = b.first + threadIdx.x;
  const SysDesc S = sd[T.sys];
  const bool live = i < n_dev[T.sys];
  const int64_t gi = S.off + (live ? i : 0);
  Gather<D, R, FAM> Gt;
  Gt.load(P, S, T, gi);
  unsigned long long ns = 0, no = 0;
  const int nls = L.src_end - L.src_begin;
  const bool overflow =
      live && (T.ltask < 0 || lcnt[L.cnt_off + (int64_t)nls * L.cap + i] != 0);
  if (overflow) {
    gather_direct<D, R, FAM>(Gt, P, sd, T, srcs, grids, carr, cells, i, ns, no);
  } else if (live) {
    const int2 *emask = reinterpret_cast<const int2 *>(lidx + L.list_off) + i;
    const int *ebase = lidx + L.list_off + 2 * (int64_t)L.maxnb * L.cap + i;
    for (int q = T.src_begin; q < T.src_end; ++q) {
      const SrcDesc Q = srcs[q];
      const int64_t poff = sd[Q.partner].off;
      SrcK<R> K;
      K.load(Q);
      int e = 0;
      for (int l = 0; l < Q.lsrc; ++l) e += lcnt[L.cnt_off + (int64_t)l * L.cap + i];
      const int eend = e + lcnt[L.cnt_off + (int64_t)Q.lsrc * L.cap + i];
      const double da = (Q.pk == PK_DYNAMIC) ? Gt.da_of(Q) : 0.0;
      // the next entry is loaded while its predecessor is popped
      const bool nb = needs_b<FAM>(Q);
      unsigned long long m = 0ull;
      int base = 0;
      int2 mpre = make_int2(0, 0);
      int bpre = 0;
      if (e < eend) {
        mpre = emask[(int64_t)e * L.cap];
        bpre = ebase[(int64_t)e * L.cap];
      }
      auto next = [&](int &j) -> bool {
        if (m == 0ull) {  // next entry (entries hold at least one partner)
          if (e >= eend) return false;
          base = bpre;
          m = ((unsigned long long)(unsigned)mpre.y << 32) | (unsigned)mpre.x;
          ++e;
          if (e < eend) {
            mpre = emask[(int64_t)e * L.cap];
            bpre = ebase[(int64_t)e * L.cap];
          }
        }
        j = base + __builtin_ctzll(m);
        m &= m - 1ull;
        return true;
      };
      constexpr bool CMP = Gather<D, R, FAM>::COMPACT;
      // prefetching the next partner pays in 3D and for the fluid-only force
      // sweeps; the 2D strain-rate and Cauchy sweeps run faster without it
      // (measured: 3D dam break, 3D visco-plastic columns, 2D EP ensemble)
      constexpr bool PREFETCH =
          D == 3 || FAM == FAM_FORCE_FLUID || FAM == FAM_FORCE_FLUID_INV;
      // compact Cauchy sources also read the partner's stress (the second
      // half only for NS > 4)
      const bool ns_a = CMP && needs_s<FAM>(Q), ns_b = ns_a && Gt.NS > 4;
      int jn = 0;
      RecA<R> An{};
      RecB<R> Bn{};
      RecC Cn{};
      RecS Sn{};
      auto fetch = [&]() {
        if constexpr (CMP) {
          Cn = P.rc[poff + jn];
          if (ns_a) Sn.a = P.rs[poff + jn].a;
          if (ns_b) Sn.b = P.rs[poff + jn].b;
        } else {
          An = P.ra[poff + jn];
          if (nb) Bn = P.rb[poff + jn];
        }
      };
      auto one = [&](auto dk, int64_t gj, const RecA<R> &Aj, const RecB<R> &Bj, const RecC &Cj,
                     const RecS &Sj) {
        double dx[D] = {};
        double r2 = 0.0;
        if constexpr (!CMP) {
#pragma unroll
          for (int d = 0; d < D; ++d) dx[d] = Gt.xi[d] - Aj.x[d];
          r2 = dx[0] * dx[0];
#pragma unroll
          for (int d = 1; d < D; ++d) r2 = r2 + dx[d] * dx[d];
        }
        if (CNT) {
          if (Q.pk == PK_SELF)
            ++ns;
          else
            ++no;
        }
        Gt.template pair<decltype(dk)::value>(P, Q, K, gj, Aj, Bj, Cj, Sj, dx, r2, da);
      };
      auto run = [&](auto dk) {
        if constexpr (PREFETCH) {
          // the next partner's records are loaded while the current pair is
          // evaluated (the copy out of the prefetch registers costs a move per
          // word; two buffers used in turn measured no faster, the compiler
          // still copied one of them)
          bool have = next(jn);
          if (have) fetch();
          while (have) {
            const int64_t gj = poff + jn;
            const RecA<R> Aj = An;
            const RecB<R> Bj = Bn;
            const RecC Cj = Cn;
            const RecS Sj = Sn;
            // the next record is loaded whether or not there is a next
            // partner (jn then still names this one): a conditional load
            // would make the compiler keep a copy of the buffer for the path
            // without it
            have = next(jn);
            fetch();
            one(dk, gj, Aj, Bj, Cj, Sj);
          }
        } else {
          // loaded at the top of each iteration: the other wavefronts of the
          // SIMD hide the latency, and no registers are copied
          while (next(jn)) {
            fetch();
            one(dk, poff + jn, An, Bn, Cn, Sn);
          }
        }
      };
      // the prefetching pair loop is compiled once for sources whose partner
      // is a dynamic boundary and once for the others (dk), so that the
      // pair's selects on the partner kind leave the loop (measured: 1 % off
      // the 3D sweeps; the 2D sweeps without prefetch gained nothing)
      if constexpr (PREFETCH && (Gather<D, R, FAM>::FORCE || FAM == FAM_STRAIN)) {
        if (Q.pk == PK_DYNAMIC)
          run(std::integral_constant<int, 1>{});
        else
          run(std::integral_constant<int, 0>{});
      } else {
        run(std::integral_constant<int, -1>{});
      }
    }
  }
  count_pairs<CNT>(pair_count, ns, no);
  if (live) Gt.store(P, S, T, gi);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <int D, typename R, int FAM>
static void launch_one(hipStream_t s, const Pool<R> &P, const SweepArgs &a, bool cnt) {
  if (a.lt) {
    // the block table steps by a.block targets (64, 128 or 256)
    if (cnt)
      k_sweep_list<D, R, FAM, true><<<a.nblk, a.block, 0, s>>>(
          P, a.sd, a.n, a.tasks, a.srcs, a.grids, a.carr, a.cells, a.blk, a.nblk, a.pair_count,
          a.lt, a.lcnt, a.lidx);
    else
      k_sweep_list<D, R, FAM, false><<<a.nblk, a.block, 0, s>>>(
          P, a.sd, a.n, a.tasks, a.srcs, a.grids, a.carr, a.cells, a.blk, a.nblk, a.pair_count,
          a.lt, a.lcnt, a.lidx);
    return;
  }
  if (cnt)
    k_sweep<D, R, FAM, true><<<a.nblk, a.block, 0, s>>>(P, a.sd, a.n, a.tasks, a.srcs, a.grids,
                                                         a.carr, a.cells, a.blk, a.nblk,
                                                         a.pair_count);
  else
    k_sweep<D, R, FAM, false><<<a.nblk, a.block, 0, s>>>(P, a.sd, a.n, a.tasks, a.srcs, a.grids,
                                                          a.carr, a.cells, a.blk, a.nblk,
                                                          a.pair_count);
}

template <typename R>
void launch_sweep(hipStream_t s, const Pool<R> &P, const SweepArgs &a, int D, int fam, bool cnt) {
  if (a.nblk <= 0) return;
  if (D == 2) {
    if (fam == FAM_FORCE && a.fluid_only && a.inv) launch_one<2, R, FAM_FORCE_FLUID_INV>(s, P, a, cnt);
    else if (fam == FAM_FORCE && a.fluid_only) launch_one<2, R, FAM_FORCE_FLUID>(s, P, a, cnt);
    else if (fam == FAM_FORCE) launch_one<2, R, FAM_FORCE>(s, P, a, cnt);
    else if (fam == FAM_STRAIN) launch_one<2, R, FAM_STRAIN>(s, P, a, cnt);
    else if (fam == FAM_INTERP) launch_one<2, R, FAM_INTERP>(s, P, a, cnt);
    else if (fam == FAM_XSPH) launch_one<2, R, FAM_XSPH>(s, P, a, cnt);
    else launch_one<2, R, FAM_COUNT>(s, P, a, cnt);
  } else {
    if (fam == FAM_FORCE && a.fluid_only && a.inv) launch_one<3, R, FAM_FORCE_FLUID_INV>(s, P, a, cnt);
    else if (fam == FAM_FORCE && a.fluid_only) launch_one<3, R, FAM_FORCE_FLUID>(s, P, a, cnt);
    else if (fam == FAM_FORCE) launch_one<3, R, FAM_FORCE>(s, P, a, cnt);
    else if (fam == FAM_STRAIN) launch_one<3, R, FAM_STRAIN>(s, P, a, cnt);
    else if (fam == FAM_INTERP) launch_one<3, R, FAM_INTERP>(s, P, a, cnt);
    else if (fam == FAM_XSPH) launch_one<3, R, FAM_XSPH>(s, P, a, cnt);
    else launch_one<3, R, FAM_COUNT>(s, P, a, cnt);
  }
}
template void launch_sweep<float>(hipStream_t, const Pool<float> &, const SweepArgs &, int, int,
                                  bool);
template void launch_sweep<double>(hipStream_t, const Pool<double> &, const SweepArgs &, int, int,
                                   bool);

template <int D, typename R, int WIN>
static void nlist_one(hipStream_t s, const Pool<R> &P, const NlistArgs &a) {
  k_nlist<D, R, WIN><<<a.nblk, 64, 0, s>>>(P, a.sd, a.n, a.lt, a.srcs, a.grids, a.carr, a.cells,
                                           a.blk, a.nblk, a.lcnt, a.lidx, a.bbmin, a.bbmax,
                                           a.span);
}

template <typename R>
void launch_nlist(hipStream_t s, const Pool<R> &P, const NlistArgs &a, int D) {
  if (a.nblk <= 0) return;
  if (D == 2) {
    if (a.win == 64) nlist_one<2, R, 64>(s, P, a);
    else if (a.win == 256) nlist_one<2, R, 256>(s, P, a);
    else if (a.win == 512) nlist_one<2, R, 512>(s, P, a);
    else nlist_one<2, R, 128>(s, P, a);
  } else {
    if (a.win == 64) nlist_one<3, R, 64>(s, P, a);
    else if (a.win == 256) nlist_one<3, R, 256>(s, P, a);
    else if (a.win == 512) nlist_one<3, R, 512>(s, P, a);
    else nlist_one<3, R, 128>(s, P, a);
  }
}
template void launch_nlist<float>(hipStream_t, const Pool<float> &, const NlistArgs &, int);
template void launch_nlist<double>(hipStream_t, const Pool<double> &, const NlistArgs &, int);

}  // namespace gr
