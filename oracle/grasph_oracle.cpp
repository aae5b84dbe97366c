// grasph_oracle.cpp — FP64 CPU restatement of Grasph.jl's particle step.
//
// TEST INFRASTRUCTURE ONLY: the parity checker of the HIP library and the
// timed CPU baseline.  Nothing shipped links this file.
//
// Follows the reference's algorithms and evaluation order in Float64:
//   src/Sorting.jl          stable lexicographic cell-key sort
//   src/Interaction.jl      CSR grid, half-shell colour sweeps, coupled sweeps
//   src/Kernels.jl          cubic spline / Wendland C2
//   src/PairwisePhysics.jl  pair formulas
//   src/PairwiseFunctors.jl FluidPfn / CauchyFluidPfn / StrainRate(Vorticity)Pfn
//   src/StateUpdaters.jl    Tait, Linear, Zero, Drucker-Prager, visco-plastic MC
//   src/GhostParticles.jl   ghost generation, kinematics, copier
//   src/TimeIntegration.jl  leapfrog step (:645-693)
// Compiled with -ffp-contract=off so that every product and sum rounds where
// Julia's does; fused operations appear only where the reference's would
// (Julia's muladd/fma inside ^(Float64, Integer), and @fastmath scalar axpy).
#include "grasph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Upd {
  int kind = GR_UPD_NOTHING;
  double prm[5] = {0, 0, 0, 0, 0};
};

struct Sys {
  int kind = GR_SYS_BASIC;
  int D = 2, NS = 0, NV = 0;
  int64_t n = 0;
  double mass = 0, c = 0, sv[3] = {0, 0, 0}, srho = 0;
  std::vector<double> x, v, vadj, rho, dvdt, drhodt, p, stress, sr, vor, strain,
      strainp;
  std::vector<int64_t> id;
  std::vector<std::vector<Upd>> upd;  // one list per updater stage
  // ghost systems
  int src = -1;
  double gcut = 0;
  int nb = 0;
  std::vector<double> bn, bp;
  std::vector<int64_t> idx_orig, idx_bnd;
  std::vector<double> normals;
  bool ex_p = false, ex_stress = false;
  std::vector<std::vector<std::pair<int, int>>> copiers;  // per stage
  // VirtualParticleSystem over this system's arrays (src/Particles.jl:476-511)
  bool virt = false;
  std::vector<double> wsum;
  int64_t zero_mask = 0;
  double pv[3] = {0, 0, 0};
  // ProbeParticleSystem (src/ProbeParticles.jl:22-89): a virtual-like
  // target measured at save cadence; cnt holds NeighborCountFn counts
  bool probe = false;
  int mirror = -1;
  std::vector<double> cnt;
};

struct Inter {
  int kk = 0, D = 2;
  double h = 0, norm = 0, cell = 0;
  int a = -1, b = -1, wrap = GR_WRAP_NONE;
  double wn[3] = {0, 0, 0}, wp[3] = {0, 0, 0}, beta = 0, ljcut = 0;
  std::vector<gr_pfn_desc> pfns;
  double ming[3] = {0, 0, 0};
  int64_t ng[3] = {0, 0, 0};
  std::vector<int64_t> cs, csa;  // 1-based values, Julia cs[c] at cs[c-1]
  double mina[3] = {0, 0, 0}, maxa[3] = {0, 0, 0};
  bool built = false;
  double xsph_eps = 0.0;  // velocity_adjust_pairwise_fn = XSPHPfn(xsph_eps)
  bool has_xsph = false;
};

struct Integ {
  std::vector<int> sys, inter, ghost, virt, probes, probe_inter;
  double gamma = 0, h = 0, c = 0;
  int nst = 0;
  int scheme = GR_SCHEME_LEAPFROG;
};

}  // namespace

struct orc_scene {
  int D = 2;
  int nthreads = 0;
  std::vector<Sys> sys;
  std::vector<Inter> inter;
  std::vector<Integ> integ;
  std::string err;
};

namespace {

constexpr double kPi = 3.141592653589793;  // Float64(π)

int fail(orc_scene *s, int code, const std::string &m) {
  s->err = m;
  return code;
}

// ---------------------------------------------------------------------------
// Julia arithmetic helpers
// ---------------------------------------------------------------------------

// Julia's ^(x::Float64, n::Integer) (Base.Math.pow_body, Julia >= 1.9):
// compensated power by squaring; muladd is fused on FMA hardware.
double jl_pow_int(double x, int64_t n) {
  if (n == 0) return 1.0;
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
  if (n == 3) return x * x * x;
  if (n < 0) {
    double rx = 1.0 / x;
    if (n == -2) return rx * rx;
    if (std::isfinite(x)) xnlo = -std::fma(x, rx, -1.0) * rx;
    x = rx;
    n = -n;
  }
  while (n > 1) {
    if (n & 1) {
      double err = std::fma(y, xnlo, x * ynlo);
      double hi = x * y;
      double lo = std::fma(x, y, -hi);
      y = hi;
      ynlo = lo + err;
    }
    double err = x * 2 * xnlo;
    double hi = x * x;
    double lo = std::fma(x, x, -hi);
    x = hi;
    xnlo = lo + err;
    n >>= 1;
  }
  double err = std::fma(y, xnlo, x * ynlo);
  return (std::isfinite(x) && std::isfinite(err)) ? std::fma(x, y, err) : x * y;
}

// Julia's ^(x::Float64, y::Float64): integer-valued y in
// [-2^12, 3*2^13] takes pow_body; otherwise the general power.
double jl_pow(double x, double y) {
  if (x == 1.0) return 1.0;
  double t = std::trunc(y);
  if (t == y && std::fabs(y) < 0x1.8p62) {
    int64_t yi = (int64_t)y;
    if (yi == 0) return 1.0;
    if (yi >= -4096 && yi <= 3 * 8192) return jl_pow_int(x, yi);
  }
  return std::pow(x, y);
}

inline double jmin(double a, double b) {  // Julia min for Float64
  if (std::isnan(a) || std::isnan(b)) return NAN;
  if (a == b) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}
inline double jmax(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NAN;
  if (a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}

inline double dotv(const double *a, const double *b, int D) {
  double s = a[0] * b[0];
  for (int d = 1; d < D; ++d) s = s + a[d] * b[d];
  return s;
}

// ---------------------------------------------------------------------------
// Kernels (src/Kernels.jl)
// ---------------------------------------------------------------------------

// norm_coeff, src/Kernels.jl:42-60 (cubic) and :116-130 (Wendland C2).
double kernel_norm(int kk, int D, double H) {
  if (kk == GR_KERNEL_CUBIC) {
    if (D == 1) return 2.0 / (3.0 * H);
    if (D == 2) return 10.0 / ((7.0 * kPi) * (H * H));
    return 1.0 / (kPi * (H * H * H));
  }
  if (D == 2) return 7.0 / ((64.0 * kPi) * (H * H));
  return 21.0 / ((256.0 * kPi) * (H * H * H));
}

// kernel_w, src/Kernels.jl:70-72, :142-144.
double kernel_w(int kk, double C, double q) {
  if (kk == GR_KERNEL_CUBIC) {
    double t2 = jmax(0.0, 2.0 - q), t1 = jmax(0.0, 1.0 - q);
    return C * (0.25 * (t2 * t2 * t2) - t1 * t1 * t1);
  }
  double t = jmax(0.0, 2.0 - q);
  return (C * jl_pow_int(t, 4)) * (2.0 * q + 1.0);
}

// kernel_dw_dq, src/Kernels.jl:79-81, :151-153.
double kernel_dw(int kk, double C, double q) {
  if (kk == GR_KERNEL_CUBIC) {
    double t2 = jmax(0.0, 2.0 - q), t1 = jmax(0.0, 1.0 - q);
    return C * (-3.0 * (0.25 * (t2 * t2) - t1 * t1));
  }
  double t = jmax(0.0, 2.0 - q);
  return ((-C * 10.0) * q) * (t * t * t);
}

// ---------------------------------------------------------------------------
// Pair physics (src/PairwisePhysics.jl)
// ---------------------------------------------------------------------------

// artificial_viscosity :29-34 (denominator 0.5·ρi·ρj as written).
double art_visc(const double *dx, const double *dv, int D, double h, double ri,
                double rj, double al, double be, double ci, double cj) {
  double vr = jmin(dotv(dx, dv, D), 0.0);
  double muv = h * vr / (dotv(dx, dx, D) + h * h * 0.01);
  double mc = 0.5 * (ci + cj);
  return (be * muv - al * mc) * muv / (0.5 * ri * rj);
}

// pressure_force_coeff :58-64.
double pfc(double pi, double pj, double ri, double rj, int S) {
  if (S == 1) return -(pi + pj) / (ri * rj);
  return -(pi / (ri * ri) + pj / (rj * rj));
}

// continuity_density_coeff :99-102.
double cdc(double ri, double rj, int S) { return S == 1 ? ri / rj : 1.0; }

// diffusion_density :238-243 (two-c/two-h form); `nothing` → 0.
double diffusion(const double *dx, double ri, double rj, double ci, double cj,
                 double hi, double hj, const double *gx, int D, double delta) {
  if (std::isnan(delta)) return 0.0;
  double rr = dotv(dx, dx, D);
  double c = 0.5 * (ci + cj);
  double h = 0.5 * (hi + hj);
  return 2.0 * delta * c * h * (ri - rj) * dotv(dx, gx, D) / rr;
}

// artificial_surface_tension_coeff :286-290.
double surf_tension(double eps, double pi, double pj, double ri, double rj) {
  if (std::isnan(eps)) return 0.0;
  return eps * pfc(std::fabs(pi), std::fabs(pj), ri, rj, 1);
}

// lennard_jones :258-263 (p1 = 12, p2 = 6 at the call sites).
double lennard_jones(const double *dx, int D, double cutoff, double c, int p1,
                     int p2) {
  double rr = dotv(dx, dx, D);
  double r = std::sqrt(rr);
  double f = (r < cutoff ? 1.0 : 0.0) *
             (jl_pow_int(cutoff / r, p1) - jl_pow_int(cutoff / r, p2)) / rr;
  return 0.01 * c * c * f;
}

// cauchy_stress_force :125-151.
void cauchy_force(const double *si, const double *sj, double ri, double rj,
                  const double *g, int NS, double *out) {
  double ci = 1.0 / (ri * ri);
  double cj = 1.0 / (rj * rj);
  if (NS == 3) {
    out[0] = (si[0] * g[0] + si[2] * g[1]) * ci + (sj[0] * g[0] + sj[2] * g[1]) * cj;
    out[1] = (si[2] * g[0] + si[1] * g[1]) * ci + (sj[2] * g[0] + sj[1] * g[1]) * cj;
  } else if (NS == 4) {
    out[0] = (si[0] * g[0] + si[3] * g[1]) * ci + (sj[0] * g[0] + sj[3] * g[1]) * cj;
    out[1] = (si[3] * g[0] + si[1] * g[1]) * ci + (sj[3] * g[0] + sj[1] * g[1]) * cj;
  } else {
    out[0] = (si[0] * g[0] + si[3] * g[1] + si[4] * g[2]) * ci +
             (sj[0] * g[0] + sj[3] * g[1] + sj[4] * g[2]) * cj;
    out[1] = (si[3] * g[0] + si[1] * g[1] + si[5] * g[2]) * ci +
             (sj[3] * g[0] + sj[1] * g[1] + sj[5] * g[2]) * cj;
    out[2] = (si[4] * g[0] + si[5] * g[1] + si[2] * g[2]) * ci +
             (sj[4] * g[0] + sj[5] * g[1] + sj[2] * g[2]) * cj;
  }
}

// strain_rate_tensor :169-195 (NS=4 plane strain pins εzz = 0).
void strain_rate(const double *dv, const double *g, int NS, double *out) {
  if (NS == 3) {
    out[0] = dv[0] * g[0];
    out[1] = dv[1] * g[1];
    out[2] = 0.5 * (dv[0] * g[1] + dv[1] * g[0]);
  } else if (NS == 4) {
    out[0] = dv[0] * g[0];
    out[1] = dv[1] * g[1];
    out[2] = 0.0;
    out[3] = 0.5 * (dv[0] * g[1] + dv[1] * g[0]);
  } else {
    out[0] = dv[0] * g[0];
    out[1] = dv[1] * g[1];
    out[2] = dv[2] * g[2];
    out[3] = 0.5 * (dv[0] * g[1] + dv[1] * g[0]);
    out[4] = 0.5 * (dv[0] * g[2] + dv[2] * g[0]);
    out[5] = 0.5 * (dv[1] * g[2] + dv[2] * g[1]);
  }
}

// vorticity_tensor :210-220.
void vorticity(const double *dv, const double *g, int D, double *out) {
  if (D == 2) {
    out[0] = 0.5 * (dv[0] * g[1] - dv[1] * g[0]);
  } else {
    out[0] = 0.5 * (dv[0] * g[1] - dv[1] * g[0]);
    out[1] = 0.5 * (dv[0] * g[2] - dv[2] * g[0]);
    out[2] = 0.5 * (dv[1] * g[2] - dv[2] * g[1]);
  }
}

// ---------------------------------------------------------------------------
// Partner views (forwarded getproperty of boundary wrappers and ghosts,
// src/BoundaryParticles.jl:13-20, src/GhostParticles.jl:226-236)
// ---------------------------------------------------------------------------

struct View {
  const Sys *s;      // arrays holder (inner system or ghost)
  double mass, c;    // forwarded scalars
  int NS;
  double *dvdt = nullptr, *drhodt = nullptr, *sr = nullptr, *vor = nullptr;
};

View make_view(orc_scene *sc, int id) {
  Sys &s = sc->sys[id];
  View v;
  v.s = &s;
  if (s.kind == GR_SYS_GHOST) {
    const Sys &src = sc->sys[s.src];
    v.mass = src.mass;
    v.c = src.c;
    v.NS = src.NS;
  } else {
    v.mass = s.mass;
    v.c = s.c;
    v.NS = s.NS;
    v.dvdt = s.dvdt.data();
    v.drhodt = s.drhodt.data();
    v.sr = s.sr.empty() ? nullptr : s.sr.data();
    v.vor = s.vor.empty() ? nullptr : s.vor.data();
  }
  return v;
}

enum PartnerKind { P_SELF, P_TWO_SIDED, P_ONE_SIDED, P_STATIC, P_DYNAMIC };

// Julia method dispatch of the coupled functors on (system_a, system_b)
// types (src/PairwiseFunctors.jl:37,53,97,111,171,193,210,222,286,303,322,334).
PartnerKind classify(const orc_scene *sc, const Inter &in, int pfn_kind) {
  if (in.b < 0) return P_SELF;
  if (in.wrap == GR_WRAP_STATIC) {
    if (pfn_kind == GR_PFN_SR || pfn_kind == GR_PFN_SRV) return P_ONE_SIDED;
    if (pfn_kind == GR_PFN_COUNT) return P_ONE_SIDED;
    return P_STATIC;
  }
  if (in.wrap == GR_WRAP_DYNAMIC) {
    if (pfn_kind == GR_PFN_COUNT) return P_ONE_SIDED;
    return P_DYNAMIC;
  }
  const Sys &b = sc->sys[in.b];
  const Sys &a = sc->sys[in.a];
  if (b.kind == GR_SYS_GHOST || b.virt || a.virt) return P_ONE_SIDED;
  switch (pfn_kind) {
    case GR_PFN_FLUID:
      return (a.kind == GR_SYS_FLUID && b.kind == GR_SYS_FLUID) ? P_TWO_SIDED
                                                                 : P_ONE_SIDED;
    case GR_PFN_CAUCHY:
      return P_TWO_SIDED;
    case GR_PFN_COUNT:
      return P_TWO_SIDED;
    default:
      return P_ONE_SIDED;
  }
}

struct PairCtx {
  const gr_pfn_desc *f;
  int D;
  PartnerKind pk;
  const Sys *A;  // system_a arrays
  View B;        // partner view (== A for self)
  double mass_a, c_a;
  int NS;
  Sys *Aw;       // writable system_a
  Sys *Bw;       // writable partner (two-sided / self), else null
  const double *wn, *wp;
  double beta, ljcut;
  int interp_side = 0;  // InterpolateFieldFn: 1 = into system_b, 2 = into system_a
  Sys *Vw = nullptr;    // the virtual system written by InterpolateFieldFn
  // writes that the colouring does not keep apart (XSPH into a ghost's source
  // at ghost indices): sweep on one thread, in the reference's serial order
  bool serial = false;
};

// f_v[t] += kw · f_r[s] for the fields of `mask` (InterpolateFieldFn
// _interp_fields_ab!/_ba!, src/PairwiseFunctors.jl:411-422).
void interp_fields(int mask, Sys &V, int64_t t, const Sys &Rl, int64_t s, double kw) {
  const int D = V.D;
  if (mask & (1 << GR_F_V))
    for (int d = 0; d < D; ++d) V.v[t * D + d] = V.v[t * D + d] + kw * Rl.v[s * D + d];
  if (mask & (1 << GR_F_RHO)) V.rho[t] = V.rho[t] + kw * Rl.rho[s];
  if ((mask & (1 << GR_F_P)) && !V.p.empty() && !Rl.p.empty()) V.p[t] = V.p[t] + kw * Rl.p[s];
  if ((mask & (1 << GR_F_STRESS)) && !V.stress.empty() && !Rl.stress.empty()) {
    const int NS = V.NS;
    for (int k = 0; k < NS; ++k)
      V.stress[t * NS + k] = V.stress[t * NS + k] + kw * Rl.stress[s * NS + k];
  }
}

// One accepted pair: the functor call of _pair_self!/_pair_coupled!
// (src/Interaction.jl:554,729) with the dispatch chosen by classify().
void apply_pair(const PairCtx &P, int64_t i, int64_t j, const double *dx,
                const double *gx, double w) {
  (void)w;
  const int D = P.D;
  const gr_pfn_desc &f = *P.f;
  const Sys &A = *P.A;
  const Sys &B = *P.B.s;
  const double *vi = &A.v[i * D];
  const double *vj = &B.v[j * D];
  double dv[3], tmp[6];
  switch (f.kind) {
    case GR_PFN_XSPH: {  // XSPHPfn :364-390, xsph_veladjust (PairwisePhysics.jl:268-270)
      const double eps = f.alpha;
      const double ri = A.rho[i], rj = B.rho[j];
      double du[3];
      for (int d = 0; d < D; ++d) du[d] = (((-eps) * (vi[d] - vj[d])) * w) / (ri + rj);
      const double mi = P.mass_a, mj = P.pk == P_SELF ? P.mass_a : P.B.mass;
      for (int d = 0; d < D; ++d) {
        P.Aw->vadj[i * D + d] += du[d] * mj;
        P.Bw->vadj[j * D + d] -= du[d] * mi;
      }
      return;
    }
    case GR_PFN_INTERP: {  // :424-434
      const int mask = f.sigma;
      if (P.interp_side == 1) {
        const double kw = w * (P.mass_a / A.rho[i]);
        interp_fields(mask, *P.Vw, j, A, i, kw);
        if (f.alpha != 0.0) P.Vw->wsum[j] += kw;
      } else if (P.interp_side == 2) {
        const double kw = w * (P.B.mass / B.rho[j]);
        interp_fields(mask, *P.Vw, i, B, j, kw);
        if (f.alpha != 0.0) P.Vw->wsum[i] += kw;
      }
      return;
    }
    case GR_PFN_NCOUNT:  // NeighborCountFn (src/PairwiseFunctors.jl:443-460)
      if (P.interp_side == 1) P.Vw->cnt[j] += 1.0;
      return;
    case GR_PFN_COUNT: {
      P.Aw->dvdt[i * D] += 1.0;
      P.Aw->drhodt[i] += 1.0;
      if (P.pk == P_SELF || P.pk == P_TWO_SIDED) {
        P.Bw->dvdt[j * D] += 1.0;
        P.Bw->drhodt[j] += 1.0;
      }
      return;
    }
    case GR_PFN_FLUID: {
      double ri = A.rho[i];
      double pi = A.p[i];
      if (P.pk == P_SELF) {  // :151-169
        double rj = A.rho[j], pj = A.p[j], m = P.mass_a;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        double dh = pfc(pi, pj, ri, rj, f.sigma);
        double t = m * (dh - piv);
        for (int d = 0; d < D; ++d) {
          double dvt = t * gx[d];
          P.Aw->dvdt[i * D + d] += dvt;
          P.Aw->dvdt[j * D + d] -= dvt;
        }
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.c_a, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += m * (dr * cdc(ri, rj, f.sigma) + psi / rj);
        P.Aw->drhodt[j] += m * (dr * cdc(rj, ri, f.sigma) - psi / ri);
      } else if (P.pk == P_TWO_SIDED) {  // :171-190
        double rj = B.rho[j], pj = B.p[j];
        double mi = P.mass_a, mj = P.B.mass;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.B.c);
        double dh = pfc(pi, pj, ri, rj, f.sigma);
        double ast = surf_tension(f.epsilon, pi, pj, ri, rj);
        double t = dh + ast - piv;
        for (int d = 0; d < D; ++d) {
          double dvt = t * gx[d];
          P.Aw->dvdt[i * D + d] += mj * dvt;
          P.Bw->dvdt[j * D + d] -= mi * dvt;
        }
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.B.c, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += mj * (dr * cdc(ri, rj, f.sigma) + psi / rj);
        P.Bw->drhodt[j] += mi * (dr * cdc(rj, ri, f.sigma) - psi / ri);
      } else if (P.pk == P_ONE_SIDED) {  // :193-207
        double rj = B.rho[j], pj = B.p[j], mj = P.B.mass;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        double dh = pfc(pi, pj, ri, rj, f.sigma);
        double t = mj * (dh - piv);
        for (int d = 0; d < D; ++d) P.Aw->dvdt[i * D + d] += t * gx[d];
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.c_a, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += mj * (dr * cdc(ri, rj, f.sigma) + psi / rj);
      } else if (P.pk == P_STATIC) {  // :210-219
        double rj = B.rho[j], mj = P.B.mass;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        double rf = lennard_jones(dx, D, P.ljcut, P.c_a, 12, 6);
        double t = -mj * piv;
        for (int d = 0; d < D; ++d) P.Aw->dvdt[i * D + d] += t * gx[d] + rf * dx[d];
      } else {  // P_DYNAMIC :222-242
        double xa[3], xb[3];
        for (int d = 0; d < D; ++d) {
          xa[d] = A.x[i * D + d] - P.wp[d];
          xb[d] = B.x[j * D + d] - P.wp[d];
        }
        double da = dotv(xa, P.wn, D), db = dotv(xb, P.wn, D);
        double fac = -jmin(P.beta, std::fabs(db / da));
        double vjd[3];
        for (int d = 0; d < D; ++d) vjd[d] = fac * vi[d];
        double rj = ri, pj = pi, m = P.mass_a;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vjd[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        double dh = pfc(pi, pj, ri, rj, f.sigma);
        double t = m * (dh - piv);
        for (int d = 0; d < D; ++d) P.Aw->dvdt[i * D + d] += t * gx[d];
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.c_a, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += m * (dr * cdc(ri, rj, f.sigma) + psi / rj);
      }
      return;
    }
    case GR_PFN_FLUIDSOLID: {  // FluidSolidPfn (src/PairwiseFunctors.jl:493-516)
      const double ri = A.rho[i], pi = A.p[i];  // the fluid's pressure, both sides
      const double rj = B.rho[j];
      const double mi = P.mass_a, mj = P.B.mass;
      for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
      const double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.B.c);
      const double dh = pfc(pi, pi, ri, rj, f.sigma);
      for (int d = 0; d < D; ++d) {
        const double dvt = (dh - piv) * gx[d];
        P.Aw->dvdt[i * D + d] += mj * dvt;
        P.Bw->dvdt[j * D + d] -= mi * dvt;
      }
      const double dr = dotv(dv, gx, D);
      const double psi = diffusion(dx, ri, rj, P.c_a, P.B.c, f.h, f.h, gx, D, f.delta);
      P.Aw->drhodt[i] += mj * (dr * cdc(ri, rj, f.sigma) + psi / rj);
      P.Bw->drhodt[j] += mi * (dr * cdc(rj, ri, f.sigma) - psi / ri);
      return;
    }
    case GR_PFN_CAUCHY: {
      const int NS = P.NS;
      double ri = A.rho[i];
      const double *si = &A.stress[i * NS];
      double hv[3];
      if (P.pk == P_SELF) {  // :265-283
        double rj = A.rho[j], m = P.mass_a;
        const double *sj = &A.stress[j * NS];
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        cauchy_force(si, sj, ri, rj, gx, NS, hv);
        for (int d = 0; d < D; ++d) {
          double dvt = m * (hv[d] - piv * gx[d]);
          P.Aw->dvdt[i * D + d] += dvt;
          P.Aw->dvdt[j * D + d] -= dvt;
        }
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.c_a, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += m * (dr + psi / rj);
        P.Aw->drhodt[j] += m * (dr - psi / ri);
      } else if (P.pk == P_ONE_SIDED) {  // :286-300 (ghost / virtual partner)
        double rj = B.rho[j], mj = P.B.mass;
        const double *sj = &B.stress[j * NS];
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.B.c);
        cauchy_force(si, sj, ri, rj, gx, NS, hv);
        for (int d = 0; d < D; ++d) P.Aw->dvdt[i * D + d] += mj * (hv[d] - piv * gx[d]);
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.c_a, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += mj * (dr + psi / rj);
      } else if (P.pk == P_TWO_SIDED) {  // :303-319
        double rj = B.rho[j], mi = P.mass_a, mj = P.B.mass;
        const double *sj = &B.stress[j * NS];
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.B.c);
        cauchy_force(si, sj, ri, rj, gx, NS, hv);
        for (int d = 0; d < D; ++d) {
          P.Aw->dvdt[i * D + d] += mj * (hv[d] - piv * gx[d]);
          P.Bw->dvdt[j * D + d] -= mi * (hv[d] - piv * gx[d]);
        }
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.B.c, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += mj * (dr + psi / rj);
        P.Bw->drhodt[j] += mi * (dr - psi / ri);
      } else if (P.pk == P_STATIC) {  // :322-331
        double rj = B.rho[j], mj = P.B.mass;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vj[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        double rf = lennard_jones(dx, D, P.ljcut, P.c_a, 12, 6);
        double t = -mj * piv;
        for (int d = 0; d < D; ++d) P.Aw->dvdt[i * D + d] += t * gx[d] + rf * dx[d];
      } else {  // P_DYNAMIC :334-355
        double xa[3], xb[3];
        for (int d = 0; d < D; ++d) {
          xa[d] = A.x[i * D + d] - P.wp[d];
          xb[d] = B.x[j * D + d] - P.wp[d];
        }
        double da = dotv(xa, P.wn, D), db = dotv(xb, P.wn, D);
        double fac = -jmin(P.beta, std::fabs(db / da));
        double vjd[3];
        for (int d = 0; d < D; ++d) vjd[d] = fac * vi[d];
        double rj = ri, m = P.mass_a;
        for (int d = 0; d < D; ++d) dv[d] = vi[d] - vjd[d];
        double piv = art_visc(dx, dv, D, f.h, ri, rj, f.alpha, f.beta, P.c_a, P.c_a);
        cauchy_force(si, si, ri, rj, gx, NS, hv);
        for (int d = 0; d < D; ++d) P.Aw->dvdt[i * D + d] += m * (hv[d] - piv * gx[d]);
        double dr = dotv(dv, gx, D);
        double psi = diffusion(dx, ri, rj, P.c_a, P.c_a, f.h, f.h, gx, D, f.delta);
        P.Aw->drhodt[i] += m * (dr + psi / rj);
      }
      return;
    }
    case GR_PFN_SR:
    case GR_PFN_SRV: {
      const int NS = P.NS;
      const int NV = (D == 2) ? 1 : 3;
      const bool vort = f.kind == GR_PFN_SRV;
      double vr[3];
      if (P.pk == P_SELF) {  // :20-34, :80-94
        double ri = A.rho[i], rj = A.rho[j], m = P.mass_a;
        for (int d = 0; d < D; ++d) dv[d] = vj[d] - vi[d];
        strain_rate(dv, gx, NS, tmp);
        if (vort) vorticity(dv, gx, D, vr);
        double fj = m / rj, fi = m / ri;
        for (int k = 0; k < NS; ++k) {
          P.Aw->sr[i * NS + k] += tmp[k] * fj;
          P.Aw->sr[j * NS + k] += tmp[k] * fi;
        }
        if (vort)
          for (int k = 0; k < NV; ++k) {
            P.Aw->vor[i * NV + k] += vr[k] * fj;
            P.Aw->vor[j * NV + k] += vr[k] * fi;
          }
      } else if (P.pk == P_DYNAMIC) {  // :53-70, :111-126
        double xa[3], xb[3];
        for (int d = 0; d < D; ++d) {
          xa[d] = A.x[i * D + d] - P.wp[d];
          xb[d] = B.x[j * D + d] - P.wp[d];
        }
        double da = dotv(xa, P.wn, D), db = dotv(xb, P.wn, D);
        double fac = -jmin(P.beta, std::fabs(db / da));
        double vjd[3];
        for (int d = 0; d < D; ++d) vjd[d] = fac * vi[d];
        double rj = A.rho[i], m = P.mass_a;
        for (int d = 0; d < D; ++d) dv[d] = vjd[d] - vi[d];
        strain_rate(dv, gx, NS, tmp);
        double fj = m / rj;
        for (int k = 0; k < NS; ++k) P.Aw->sr[i * NS + k] += tmp[k] * fj;
        if (vort) {
          vorticity(dv, gx, D, vr);
          for (int k = 0; k < NV; ++k) P.Aw->vor[i * NV + k] += vr[k] * fj;
        }
      } else {  // generic one-sided :37-50, :97-108
        double rj = B.rho[j], mj = P.B.mass;
        for (int d = 0; d < D; ++d) dv[d] = vj[d] - vi[d];
        strain_rate(dv, gx, NS, tmp);
        double fj = mj / rj;
        for (int k = 0; k < NS; ++k) P.Aw->sr[i * NS + k] += tmp[k] * fj;
        if (vort) {
          vorticity(dv, gx, D, vr);
          for (int k = 0; k < NV; ++k) P.Aw->vor[i * NV + k] += vr[k] * fj;
        }
      }
      return;
    }
    default:
      return;
  }
}

// _pair_self! / _pair_coupled! (src/Interaction.jl:543-557, 717-732).
inline void pair(const PairCtx &P, const Inter &in, double cutoff_sq,
                 const double *xa, const double *xb, int64_t i, int64_t j) {
  const int D = P.D;
  double dx[3];
  for (int d = 0; d < D; ++d) dx[d] = xa[i * D + d] - xb[j * D + d];
  double r_sq = dotv(dx, dx, D);
  if (r_sq < cutoff_sq) {
    double r = std::sqrt(r_sq);
    double q = r / in.h;
    double g = kernel_dw(in.kk, in.norm, q) / (r * in.h);
    double w = kernel_w(in.kk, in.norm, q);
    double gx[3];
    for (int d = 0; d < D; ++d) gx[d] = g * dx[d];
    apply_pair(P, i, j, dx, gx, w);
  }
}

inline int64_t jl_range_len(int64_t a, int64_t s, int64_t b) {
  if (b < a) return 0;
  return (b - a) / s + 1;
}

inline int64_t jl_floor_int(double v) { return (int64_t)std::floor(v); }

// floor(Int, x / cutoff) raises InexactError for a non-finite position
// (src/Sorting.jl:122-124); the sort is refused before any key is formed.
bool nonfinite_x(const std::vector<double> &x, int64_t n, int D) {
  for (int64_t i = 0; i < n * D; ++i)
    if (!std::isfinite(x[i])) return true;
  return false;
}

// ---------------------------------------------------------------------------
// Sorting (src/Sorting.jl)
// ---------------------------------------------------------------------------

template <typename T>
void permute(std::vector<T> &a, const std::vector<int64_t> &perm, int comps) {
  if (a.empty()) return;
  std::vector<T> tmp(a.size());
  const int64_t n = (int64_t)perm.size();
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < comps; ++k) tmp[i * comps + k] = a[perm[i] * comps + k];
  std::copy(tmp.begin(), tmp.begin() + n * comps, a.begin());
}

// sort_particles! :157-195 with _pos_to_key :122 and _lt_key :74-80.
// InsertionSort and Julia's default sortperm are both stable, so the
// permutation is the unique stable order, produced here by stable_sort.
void sort_system(orc_scene *sc, Sys &s, double cutoff) {
  const int64_t n = s.n;
  if (n <= 1) return;
  const int D = s.D;
  std::vector<int64_t> key(n * D);
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < D; ++d) key[i * D + d] = jl_floor_int(s.x[i * D + d] / cutoff);
  auto lt = [&](int64_t a, int64_t b) {
    for (int d = 0; d < D; ++d) {
      int64_t ka = key[a * D + d], kb = key[b * D + d];
      if (ka != kb) return ka < kb;
    }
    return false;
  };
  bool sorted = true;
  for (int64_t i = 0; i + 1 < n; ++i)
    if (lt(i + 1, i)) {
      sorted = false;
      break;
    }
  if (sorted) return;
  std::vector<int64_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), lt);
  if (s.kind == GR_SYS_GHOST) {
    // _particle_arrays(::GhostParticleSystem) :33-40 — normals are NOT in the
    // tuple, so they keep generation order.
    permute(s.x, perm, D);
    permute(s.v, perm, D);
    permute(s.rho, perm, 1);
    permute(s.idx_orig, perm, 1);
    permute(s.idx_bnd, perm, 1);
    if (s.ex_p) permute(s.p, perm, 1);
    if (s.ex_stress) permute(s.stress, perm, sc->sys[s.src].NS);
    return;
  }
  if (s.probe) {  // _particle_arrays(::ProbeParticleSystem) (src/ProbeParticles.jl:133-139)
    permute(s.wsum, perm, 1);
    permute(s.cnt, perm, 1);
  }
  // _particle_arrays :15-26
  permute(s.x, perm, D);
  permute(s.id, perm, 1);
  permute(s.v, perm, D);
  permute(s.vadj, perm, D);
  permute(s.rho, perm, 1);
  permute(s.dvdt, perm, D);
  permute(s.drhodt, perm, 1);
  if (s.kind >= GR_SYS_FLUID) permute(s.p, perm, 1);
  if (s.kind >= GR_SYS_STRESS) {
    permute(s.stress, perm, s.NS);
    permute(s.sr, perm, s.NS);
  }
  if (s.kind == GR_SYS_EP) {
    permute(s.vor, perm, s.NV);
    permute(s.strain, perm, s.NS);
    permute(s.strainp, perm, s.NS);
  }
}

// ---------------------------------------------------------------------------
// Grid (src/Interaction.jl:123-251)
// ---------------------------------------------------------------------------

int64_t cell_1idx(const double *x, const double *ming, double dcell,
                  const int64_t *ng, int D) {  // :197-206
  int64_t flat = 0, stride = 1;
  for (int d = D - 1; d >= 0; --d) {
    int64_t cd = jl_floor_int((x[d] - ming[d]) / dcell);
    flat += cd * stride;
    stride *= ng[d];
  }
  return flat + 1;
}

void populate(std::vector<int64_t> &cs, const Sys &s, const Inter &in) {  // :214-236
  const int64_t ncells = (int64_t)cs.size() - 1;
  const int D = in.D;
  int64_t prev = 0;
  for (int64_t i = 0; i < s.n; ++i) {
    int64_t cell = cell_1idx(&s.x[i * D], in.ming, in.cell, in.ng, D);
    if (cell != prev) {
      cs[cell - 1] = i + 1;
      prev = cell;
    }
  }
  cs[ncells] = s.n + 1;
  for (int64_t c = ncells; c >= 1; --c)
    if (cs[c - 1] == 0) cs[c - 1] = cs[c];
}

void bbox(const Sys &s, double *mn, double *mx, bool init) {
  const int D = s.D;
  int64_t i0 = 0;
  if (init) {
    for (int d = 0; d < D; ++d) mn[d] = mx[d] = s.x[d];
    i0 = 1;
  }
  for (int64_t i = i0; i < s.n; ++i)
    for (int d = 0; d < D; ++d) {
      mn[d] = jmin(mn[d], s.x[i * D + d]);
      mx[d] = jmax(mx[d], s.x[i * D + d]);
    }
}

int create_grid(orc_scene *sc, Inter &in) {
  const int D = in.D;
  const double cutoff = in.cell;
  const Sys &a = sc->sys[in.a];
  if (a.n < 1) return fail(sc, GR_ERR_ARGUMENT, "create_grid!: system_a has no particles");
  double mn[3], mx[3];
  bbox(a, mn, mx, true);
  double mna[3], mxa[3];
  for (int d = 0; d < D; ++d) {
    mna[d] = mn[d];
    mxa[d] = mx[d];
  }
  if (in.b >= 0) bbox(sc->sys[in.b], mn, mx, false);
  double maxg[3];
  for (int d = 0; d < D; ++d) {
    in.ming[d] = std::floor((mn[d] - 2.0 * cutoff) / cutoff) * cutoff;  // :191-193
    maxg[d] = mx[d] + 2.0 * cutoff;
  }
  int64_t ncells = 1;
  for (int d = 0; d < D; ++d) {
    in.ng[d] = jl_floor_int((maxg[d] - in.ming[d]) / cutoff) + 1;  // :179
    ncells *= in.ng[d];
  }
  if (ncells <= 0 || ncells > (int64_t)1 << 34)
    return fail(sc, GR_ERR_CAPACITY, "create_grid!: cell count out of range");
  in.cs.assign(ncells + 1, 0);
  if (in.b >= 0) {
    in.csa.assign(ncells + 1, 0);
    for (int d = 0; d < D; ++d) {
      in.mina[d] = mna[d];
      in.maxa[d] = mxa[d];
    }
    populate(in.csa, a, in);
    populate(in.cs, sc->sys[in.b], in);
  } else {
    populate(in.cs, a, in);
  }
  in.built = true;
  return GR_OK;
}

// ---------------------------------------------------------------------------
// Sweeps (src/Interaction.jl:393-439, 477-541, 600-645, 660-715)
// ---------------------------------------------------------------------------

int sweep_fn(orc_scene *sc, Inter &in, const gr_pfn_desc &f);

int sweep(orc_scene *sc, Inter &in, int stage) {
  if (stage < 1 || stage > (int)in.pfns.size()) return GR_OK;
  return sweep_fn(sc, in, in.pfns[stage - 1]);
}

int sweep_fn(orc_scene *sc, Inter &in, const gr_pfn_desc &f) {
  if (f.kind == GR_PFN_NONE) return GR_OK;
  if (!in.built) return fail(sc, GR_ERR_STATE, "sweep!: create_grid! has not been called");
  PairCtx P;
  P.f = &f;
  P.D = in.D;
  P.pk = classify(sc, in, f.kind);
  Sys &A = sc->sys[in.a];
  P.A = &A;
  P.Aw = &A;
  P.mass_a = A.mass;
  P.c_a = A.c;
  P.NS = A.NS;
  P.wn = in.wn;
  P.wp = in.wp;
  P.beta = in.beta;
  P.ljcut = in.ljcut;
  if (in.b < 0) {
    P.B = make_view(sc, in.a);
    P.Bw = &A;
  } else {
    P.B = make_view(sc, in.b);
    P.Bw = (P.pk == P_TWO_SIDED) ? &sc->sys[in.b] : nullptr;
  }
  if (f.kind == GR_PFN_XSPH) {  // generic two-sided method (also through wrappers)
    P.pk = in.b < 0 ? P_SELF : P_TWO_SIDED;
    P.Bw = in.b < 0 ? &A : &sc->sys[in.b];
    if (in.b >= 0) P.B = make_view(sc, in.b);
    if (in.b >= 0 && sc->sys[in.b].kind == GR_SYS_GHOST) {
      // ghost.v_adjustment is forwarded to the source system's array
      // (src/GhostParticles.jl:226-236), indexed with the ghost's own index j
      Sys &src = sc->sys[sc->sys[in.b].src];
      if (sc->sys[in.b].n > src.n)
        return fail(sc, GR_ERR_ARGUMENT,
                    "XSPHPfn with a ghost partner: more ghosts than source particles "
                    "(ps_b.v_adjustment[j] would index past the source's array)");
      P.Bw = &src;
      P.serial = true;
    }
  }
  if (f.kind == GR_PFN_FLUIDSOLID) {  // only the two-system method exists
    if (in.b < 0 || in.wrap != GR_WRAP_NONE || sc->sys[in.b].kind == GR_SYS_GHOST ||
        sc->sys[in.b].virt || A.kind < GR_SYS_FLUID)
      return fail(sc, GR_ERR_ARGUMENT,
                  "FluidSolidPfn needs a fluid system_a and a plain solid system_b");
    P.pk = P_TWO_SIDED;
    P.B = make_view(sc, in.b);
    P.Bw = &sc->sys[in.b];
  }
  if (f.kind == GR_PFN_NCOUNT) {
    if (!(in.b >= 0 && sc->sys[in.b].probe && !sc->sys[in.a].virt && in.wrap == GR_WRAP_NONE))
      return fail(sc, GR_ERR_ARGUMENT,
                  "NeighborCountFn needs a real system_a and a ProbeParticleSystem system_b");
    P.interp_side = 1;
    P.Vw = &sc->sys[in.b];
  }
  if (f.kind == GR_PFN_INTERP) {
    if (in.b >= 0 && sc->sys[in.b].virt && in.wrap == GR_WRAP_NONE) {
      P.interp_side = 1;
      P.Vw = &sc->sys[in.b];
    } else if (in.b >= 0 && A.virt && !sc->sys[in.b].virt && in.wrap == GR_WRAP_NONE &&
               sc->sys[in.b].kind != GR_SYS_GHOST) {
      P.interp_side = 2;
      P.Vw = &A;
    } else {
      return fail(sc, GR_ERR_ARGUMENT,
                  "InterpolateFieldFn needs a real and a virtual particle system");
    }
  }
  const double cutoff = in.cell;
  const double cutoff_sq = cutoff * cutoff;
  const int D = in.D;
  const double *xa = A.x.data();
  const double *xb = P.B.s->x.data();
  const int64_t *cs = in.cs.data();
  // Julia cs[c] (1-based c) lives at cs[c-1]; particle index k (1-based) at k-1.
#define CS(c) cs[(c)-1]
  if (in.b < 0) {
    if (D == 2) {
      const int64_t ny = in.ng[1], nx = in.ng[0];
      for (int colour = 0; colour <= 5; ++colour) {
        int64_t cxb = colour / 3 + 2, cyb = colour % 3 + 2;
        int64_t nax = jl_range_len(cxb, 2, nx - 1), nay = jl_range_len(cyb, 3, ny - 1);
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t flat = 1; flat <= nax * nay; ++flat) {
          int64_t sx = (flat - 1) / nay, sy = (flat - 1) % nay;
          int64_t cx = cxb + sx * 2, cy = cyb + sy * 3;
          int64_t cidx = (cx - 1) * ny + cy;
          int64_t ps = CS(cidx), pend_same = CS(cidx + 1), pend_next = CS(cidx + 2);
          int64_t ncidx = cidx + ny - 1;
          int64_t nps = CS(ncidx), npe = CS(ncidx + 3);
          for (int64_t pi = ps; pi <= pend_same - 1; ++pi) {
            for (int64_t pj = pi + 1; pj <= pend_next - 1; ++pj)
              pair(P, in, cutoff_sq, xa, xa, pi - 1, pj - 1);
            for (int64_t pj = nps; pj <= npe - 1; ++pj)
              pair(P, in, cutoff_sq, xa, xa, pi - 1, pj - 1);
          }
        }
      }
    } else {
      const int64_t nz = in.ng[2], ny = in.ng[1], nx = in.ng[0], nyz = ny * nz;
      const int64_t off[4] = {nz - 1, nyz - nz - 1, nyz - 1, nyz + nz - 1};
      for (int colour = 0; colour <= 17; ++colour) {
        int64_t cxb = colour / 9 + 2, cyb = (colour % 9) / 3 + 2, czb = colour % 3 + 2;
        int64_t nax = jl_range_len(cxb, 2, nx - 1), nay = jl_range_len(cyb, 3, ny - 1),
                naz = jl_range_len(czb, 3, nz - 1);
        int64_t nayz = nay * naz;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t flat = 1; flat <= nax * nayz; ++flat) {
          int64_t sx = (flat - 1) / nayz, rem = (flat - 1) % nayz;
          int64_t sy = rem / naz, sz = rem % naz;
          int64_t cx = cxb + sx * 2, cy = cyb + sy * 3, cz = czb + sz * 3;
          int64_t cidx = (cx - 1) * nyz + (cy - 1) * nz + cz;
          int64_t ps = CS(cidx), pend = CS(cidx + 1), pend_next = CS(cidx + 2);
          for (int64_t pi = ps; pi <= pend - 1; ++pi)
            for (int64_t pj = pi + 1; pj <= pend_next - 1; ++pj)
              pair(P, in, cutoff_sq, xa, xa, pi - 1, pj - 1);
          for (int dir = 0; dir < 4; ++dir) {
            int64_t ncidx = cidx + off[dir];
            int64_t nps = CS(ncidx), npe = CS(ncidx + 3);
            for (int64_t pi = ps; pi <= pend - 1; ++pi)
              for (int64_t pj = nps; pj <= npe - 1; ++pj)
                pair(P, in, cutoff_sq, xa, xa, pi - 1, pj - 1);
          }
        }
      }
    }
  } else {
    const int64_t *csa = in.csa.data();
#define CSA(c) csa[(c)-1]
    int64_t cmin[3], cmax[3];
    for (int d = 0; d < D; ++d) {
      cmin[d] = jl_floor_int((in.mina[d] - in.ming[d]) / cutoff) + 1;
      cmax[d] = jl_floor_int((in.maxa[d] - in.ming[d]) / cutoff) + 1;
    }
    if (D == 2) {
      const int64_t ny = in.ng[1];
      for (int colour = 0; colour <= 8; ++colour) {
        int64_t cxb = colour / 3 + cmin[0], cyb = colour % 3 + cmin[1];
        int64_t nax = jl_range_len(cxb, 3, cmax[0]), nay = jl_range_len(cyb, 3, cmax[1]);
#pragma omp parallel for schedule(dynamic, 16) if (!P.serial)
        for (int64_t flat = 1; flat <= nax * nay; ++flat) {
          int64_t sx = (flat - 1) / nay, sy = (flat - 1) % nay;
          int64_t cx = cxb + sx * 3, cy = cyb + sy * 3;
          int64_t cidx = (cx - 1) * ny + cy;
          int64_t aps = CSA(cidx), ape = CSA(cidx + 1);
          for (int64_t pi = aps; pi <= ape - 1; ++pi)
            for (int dxc = -1; dxc <= 1; ++dxc) {
              int64_t ncidx = cidx + dxc * ny - 1;
              int64_t bps = CS(ncidx), bpe = CS(ncidx + 3);
              for (int64_t pj = bps; pj <= bpe - 1; ++pj)
                pair(P, in, cutoff_sq, xa, xb, pi - 1, pj - 1);
            }
        }
      }
    } else {
      const int64_t nz = in.ng[2], ny = in.ng[1], nyz = ny * nz;
      for (int colour = 0; colour <= 26; ++colour) {
        int64_t cxb = colour / 9 + cmin[0], cyb = (colour % 9) / 3 + cmin[1],
                czb = colour % 3 + cmin[2];
        int64_t nax = jl_range_len(cxb, 3, cmax[0]), nay = jl_range_len(cyb, 3, cmax[1]),
                naz = jl_range_len(czb, 3, cmax[2]);
        int64_t nayz = nay * naz;
#pragma omp parallel for schedule(dynamic, 16) if (!P.serial)
        for (int64_t flat = 1; flat <= nax * nayz; ++flat) {
          int64_t sx = (flat - 1) / nayz, rem = (flat - 1) % nayz;
          int64_t sy = rem / naz, sz = rem % naz;
          int64_t cx = cxb + sx * 3, cy = cyb + sy * 3, cz = czb + sz * 3;
          int64_t cidx = (cx - 1) * nyz + (cy - 1) * nz + cz;
          int64_t aps = CSA(cidx), ape = CSA(cidx + 1);
          for (int64_t pi = aps; pi <= ape - 1; ++pi)
            for (int dxc = -1; dxc <= 1; ++dxc)
              for (int dyc = -1; dyc <= 1; ++dyc) {
                int64_t ncidx = cidx + dxc * nyz + dyc * nz - 1;
                int64_t bps = CS(ncidx), bpe = CS(ncidx + 3);
                for (int64_t pj = bps; pj <= bpe - 1; ++pj)
                  pair(P, in, cutoff_sq, xa, xb, pi - 1, pj - 1);
              }
        }
      }
    }
#undef CSA
  }
#undef CS
  return GR_OK;
}

// ---------------------------------------------------------------------------
// State updaters (src/StateUpdaters.jl, src/Particles.jl:542-570)
// ---------------------------------------------------------------------------

void eos(Sys &s, int kind, double rho0, double gamma, int64_t i) {
  if (kind == GR_UPD_TAIT) {  // :114-118
    const double c = s.c;
    s.p[i] = rho0 * c * c / gamma * (jl_pow(s.rho[i] / rho0, gamma) - 1.0);
  } else {  // :136-140
    const double c = s.c;
    s.p[i] = c * c * (s.rho[i] - rho0);
  }
}

// ElastoPlasticStressUpdater :216-335 (NS = 4, plane strain).
// HookeLawStressUpdater (src/StateUpdaters.jl:366-405): the elastic
// increment D:ε̇dt, the Jaumann correction with the vorticity, σ += Δσ and
// p = −(σxx + σyy + σzz)/3.
void hooke_update(Sys &s, const Upd &u, double dt, int64_t i) {
  const double E = u.prm[0], nu = u.prm[1];
  double *sg = &s.stress[i * 4];
  const double *er = &s.sr[i * 4];
  const double om = s.vor[i];
  const double sxx = sg[0], syy = sg[1], szz = sg[2], sxy = sg[3];
  const double dexx = er[0] * dt, deyy = er[1] * dt, dezz = er[2] * dt, dexy = er[3] * dt;
  const double D0 = E / ((1.0 + nu) * (1.0 - 2 * nu));
  double dsxx = D0 * ((1.0 - nu) * dexx + nu * deyy + nu * dezz);
  double dsyy = D0 * (nu * dexx + (1.0 - nu) * deyy + nu * dezz);
  const double dszz = D0 * (nu * dexx + nu * deyy + (1.0 - nu) * dezz);
  double dsxy = D0 * (1.0 - 2 * nu) * dexy;
  const double wdt = om * dt;
  dsxx -= 2 * sxy * wdt;
  dsyy += 2 * sxy * wdt;
  dsxy += (sxx - syy) * wdt;
  const double nxx = sxx + dsxx, nyy = syy + dsyy, nzz = szz + dszz, nxy = sxy + dsxy;
  sg[0] = nxx;
  sg[1] = nyy;
  sg[2] = nzz;
  sg[3] = nxy;
  s.p[i] = -(nxx + nyy + nzz) / 3.0;
}

void ep_update(Sys &s, const Upd &u, double dt, int64_t i) {
  const double E = u.prm[0], nu = u.prm[1], phi = u.prm[2], psi = u.prm[3],
               coh = u.prm[4];
  double *sg = &s.stress[i * 4];
  const double *er = &s.sr[i * 4];
  const double om = s.vor[i];
  double sxx = sg[0], syy = sg[1], szz = sg[2], sxy = sg[3];
  double dexx = er[0] * dt, deyy = er[1] * dt, dezz = er[2] * dt, dexy = er[3] * dt;
  double D0 = E / ((1.0 + nu) * (1.0 - 2 * nu));
  double dsxx = D0 * ((1.0 - nu) * dexx + nu * deyy + nu * dezz);
  double dsyy = D0 * (nu * dexx + (1.0 - nu) * deyy + nu * dezz);
  double dszz = D0 * (nu * dexx + nu * deyy + (1.0 - nu) * dezz);
  double dsxy = D0 * (1.0 - 2 * nu) * dexy;
  double wdt = om * dt;
  dsxx -= 2 * sxy * wdt;
  dsyy += 2 * sxy * wdt;
  dsxy += (sxx - syy) * wdt;
  double txx = sxx + dsxx, tyy = syy + dsyy, tzz = szz + dszz, txy = sxy + dsxy;
  double sqrt3 = std::sqrt(3.0);
  double a_phi = 2 * std::sin(phi) / (sqrt3 * (3 - std::sin(phi)));
  double k_c = 6 * coh * std::cos(phi) / (sqrt3 * (3 - std::sin(phi)));
  double I1 = txx + tyy + tzz;
  double p3 = I1 / 3;
  double s_xx = txx - p3, s_yy = tyy - p3, s_zz = tzz - p3, s_xy = txy;
  double J2 = 0.5 * (s_xx * s_xx + s_yy * s_yy + s_zz * s_zz + 2 * (s_xy * s_xy));
  double sqJ2 = std::sqrt(J2);
  double fy = a_phi * I1 + sqJ2 - k_c;
  double dpxx = 0, dpyy = 0, dpzz = 0, dpxy = 0;
  if (fy > 0.0) {
    double a_psi = 2 * std::sin(psi) / (sqrt3 * (3 - std::sin(psi)));
    double inv2 = 1.0 / (2 * sqJ2);
    double dfxx = a_phi + s_xx * inv2, dfyy = a_phi + s_yy * inv2,
           dfzz = a_phi + s_zz * inv2, dfxy = s_xy * inv2;
    double dgxx = a_psi + s_xx * inv2, dgyy = a_psi + s_yy * inv2,
           dgzz = a_psi + s_zz * inv2, dgxy = s_xy * inv2;
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
  sg[0] = sxx + dsxx;
  sg[1] = syy + dsyy;
  sg[2] = szz + dszz;
  sg[3] = sxy + dsxy;
  if (sxx + dsxx + syy + dsyy + szz + dszz > k_c / a_phi) {
    double ap = k_c / a_phi / 3.0;
    sg[0] = ap;
    sg[1] = ap;
    sg[2] = ap;
    sg[3] = 0.0;
  }
  double *st = &s.strain[i * 4], *sp = &s.strainp[i * 4];
  st[0] += dexx;
  st[1] += deyy;
  st[2] += dezz;
  st[3] += dexy;
  sp[0] += dpxx;
  sp[1] += dpyy;
  sp[2] += dpzz;
  sp[3] += dpxy;
}

// ViscoPlasticMCStressUpdater :155-189.
void vpmc_update(Sys &s, const Upd &u, int64_t i) {
  eos(s, (int)u.prm[0], u.prm[1], u.prm[2], i);
  const int NS = s.NS;
  const double phi = u.prm[3], coh = u.prm[4];
  double *sg = &s.stress[i * NS];
  const double *er = &s.sr[i * NS];
  double pval = s.p[i];
  double mag;
  if (NS == 3)
    mag = er[0] * er[0] + er[1] * er[1] + 2 * (er[2] * er[2]);
  else if (NS == 4)
    mag = er[0] * er[0] + er[1] * er[1] + 2 * (er[3] * er[3]);
  else
    mag = er[0] * er[0] + er[1] * er[1] + er[2] * er[2] +
          2 * (er[3] * er[3] + er[4] * er[4] + er[5] * er[5]);
  mag = jmax(std::sqrt(mag), 2.220446049250313e-16);
  double sc = 2.0 * (coh + pval * std::tan(phi));
  for (int k = 0; k < NS; ++k) sg[k] = sc * er[k] / mag;
  if (NS == 3) {
    sg[0] = sg[0] - pval;
    sg[1] = sg[1] - pval;
    sg[2] = sg[2] - 0.0;
    if (sg[0] + sg[1] > 0.0)
      for (int k = 0; k < 3; ++k) sg[k] = 0.0;
  } else {
    sg[0] = sg[0] - pval;
    sg[1] = sg[1] - pval;
    sg[2] = sg[2] - pval;
    for (int k = 3; k < NS; ++k) sg[k] = sg[k] - 0.0;
    if (sg[0] + sg[1] + sg[2] > 0.0)
      for (int k = 0; k < NS; ++k) sg[k] = 0.0;
  }
}

void zero_fields(Sys &s, int64_t mask) {
  if (mask & (1 << GR_F_V)) std::fill(s.v.begin(), s.v.end(), 0.0);
  if (mask & (1 << GR_F_RHO)) std::fill(s.rho.begin(), s.rho.end(), 0.0);
  if (mask & (1 << GR_F_STRAIN_RATE)) std::fill(s.sr.begin(), s.sr.end(), 0.0);
  if (mask & (1 << GR_F_VORTICITY)) std::fill(s.vor.begin(), s.vor.end(), 0.0);
  if (mask & (1 << GR_F_DVDT)) std::fill(s.dvdt.begin(), s.dvdt.end(), 0.0);
  if (mask & (1 << GR_F_DRHODT)) std::fill(s.drhodt.begin(), s.drhodt.end(), 0.0);
  if (mask & (1 << GR_F_P)) std::fill(s.p.begin(), s.p.end(), 0.0);
  if (mask & (1 << GR_F_STRESS)) std::fill(s.stress.begin(), s.stress.end(), 0.0);
  if (mask & (1 << GR_F_STRAIN)) std::fill(s.strain.begin(), s.strain.end(), 0.0);
  if (mask & (1 << GR_F_STRAIN_P)) std::fill(s.strainp.begin(), s.strainp.end(), 0.0);
  if (mask & (1 << GR_F_COUNT)) std::fill(s.cnt.begin(), s.cnt.end(), 0.0);
}

int update_state(orc_scene *sc, Sys &s, int stage, double dt) {
  if (stage < 1 || stage > (int)s.upd.size()) return GR_OK;
  for (const Upd &u : s.upd[stage - 1]) {
    switch (u.kind) {
      case GR_UPD_NOTHING:
        break;
      case GR_UPD_ZERO:  // :30-42
        zero_fields(s, (int64_t)u.prm[0]);
        break;
      case GR_UPD_TAIT:
      case GR_UPD_LINEAR:
#pragma omp parallel for
        for (int64_t i = 0; i < s.n; ++i) eos(s, u.kind, u.prm[0], u.prm[1], i);
        break;
      case GR_UPD_EP:
#pragma omp parallel for
        for (int64_t i = 0; i < s.n; ++i) ep_update(s, u, dt, i);
        break;
      case GR_UPD_VPMC:
#pragma omp parallel for
        for (int64_t i = 0; i < s.n; ++i) vpmc_update(s, u, i);
        break;
      case GR_UPD_HOOKE:
#pragma omp parallel for
        for (int64_t i = 0; i < s.n; ++i) hooke_update(s, u, dt, i);
        break;
      case GR_UPD_VNORM: {  // VirtualNormUpdater :74-84
        if (!s.virt) return fail(sc, GR_ERR_ARGUMENT, "VirtualNormUpdater needs a virtual system");
        const int64_t mask = (int64_t)u.prm[3];
        const int D = s.D, NS = s.NS;
        for (int64_t i = 0; i < s.n; ++i) {
          const double w = s.wsum[i];
          const bool z = w == 0.0;
          if (mask & (1 << GR_F_V))
            for (int d = 0; d < D; ++d) {
              const double val = z ? 0.0 : s.v[i * D + d] / w;
              s.v[i * D + d] = val * u.prm[d] + s.pv[d];
            }
          if (mask & (1 << GR_F_RHO)) s.rho[i] = z ? 0.0 : s.rho[i] / w;
          if ((mask & (1 << GR_F_P)) && !s.p.empty()) s.p[i] = z ? 0.0 : s.p[i] / w;
          if ((mask & (1 << GR_F_STRESS)) && !s.stress.empty())
            for (int k = 0; k < NS; ++k)
              s.stress[i * NS + k] = z ? 0.0 : s.stress[i * NS + k] / w;
        }
        break;
      }
      case GR_UPD_PRESCRIBED_V: {  // :88-92
        if (!s.virt) return fail(sc, GR_ERR_ARGUMENT, "PrescribedVelocityUpdater needs a virtual system");
        const int D = s.D;
        for (int64_t i = 0; i < s.n; ++i)
          for (int d = 0; d < D; ++d) s.v[i * D + d] = s.v[i * D + d] + s.pv[d];
        break;
      }
      default:
        return fail(sc, GR_ERR_ARGUMENT, "unknown updater kind");
    }
  }
  return GR_OK;
}

// ---------------------------------------------------------------------------
// Ghosts (src/GhostParticles.jl)
// ---------------------------------------------------------------------------

void generate_ghosts(orc_scene *sc, Sys &g) {  // :344-392
  const Sys &ps = sc->sys[g.src];
  const int D = g.D;
  int64_t total = 0;
  for (int b = 0; b < g.nb; ++b)
    for (int64_t i = 0; i < ps.n; ++i) {
      double t[3];
      for (int d = 0; d < D; ++d) t[d] = ps.x[i * D + d] - g.bp[b * D + d];
      double da = dotv(t, &g.bn[b * D], D);
      if (std::fabs(da) <= g.gcut && da > 0.0) ++total;
    }
  g.n = total;
  g.x.resize(total * D);
  g.v.resize(total * D);
  g.rho.resize(total);
  g.idx_orig.resize(total);
  g.idx_bnd.resize(total);
  g.normals.resize(total * D);
  if (g.ex_p) g.p.resize(total);
  if (g.ex_stress) g.stress.resize(total * ps.NS);
  int64_t cur = 0;
  for (int b = 0; b < g.nb; ++b)
    for (int64_t i = 0; i < ps.n; ++i) {
      double t[3];
      for (int d = 0; d < D; ++d) t[d] = ps.x[i * D + d] - g.bp[b * D + d];
      double da = dotv(t, &g.bn[b * D], D);
      if (std::fabs(da) <= g.gcut && da > 0.0) {
        for (int d = 0; d < D; ++d) {
          g.x[cur * D + d] = ps.x[i * D + d] - (2 * da) * g.bn[b * D + d];
          g.normals[cur * D + d] = g.bn[b * D + d];
        }
        g.idx_orig[cur] = i + 1;
        g.idx_bnd[cur] = b + 1;
        ++cur;
      }
    }
}

void ghost_kinematics(orc_scene *sc, Sys &g) {  // :404-421
  const Sys &ps = sc->sys[g.src];
  const int D = g.D;
  for (int64_t k = 0; k < g.n; ++k) {
    const double *nn = &g.normals[k * D];
    int64_t src = g.idx_orig[k] - 1;
    const double *vr = &ps.v[src * D];
    double dd = dotv(vr, nn, D);
    for (int d = 0; d < D; ++d) g.v[k * D + d] = vr[d] - 2 * dd * nn[d];
    g.rho[k] = ps.rho[src];
  }
}

// _apply_mode :66-110.
void householder(const double *s, const double *n, int NS, double *o) {
  if (NS == 3) {
    double nx = n[0], ny = n[1];
    double tx = s[0] * nx + s[2] * ny, ty = s[2] * nx + s[1] * ny;
    double q = tx * nx + ty * ny;
    o[0] = s[0] - 4 * tx * nx + 4 * q * nx * nx;
    o[1] = s[1] - 4 * ty * ny + 4 * q * ny * ny;
    o[2] = s[2] - 2 * (tx * ny + nx * ty) + 4 * q * nx * ny;
  } else if (NS == 4) {
    double nx = n[0], ny = n[1];
    double tx = s[0] * nx + s[3] * ny, ty = s[3] * nx + s[1] * ny;
    double q = tx * nx + ty * ny;
    o[0] = s[0] - 4 * tx * nx + 4 * q * nx * nx;
    o[1] = s[1] - 4 * ty * ny + 4 * q * ny * ny;
    o[2] = s[2];
    o[3] = s[3] - 2 * (tx * ny + nx * ty) + 4 * q * nx * ny;
  } else {
    double nx = n[0], ny = n[1], nz = n[2];
    double tx = s[0] * nx + s[3] * ny + s[4] * nz;
    double ty = s[3] * nx + s[1] * ny + s[5] * nz;
    double tz = s[4] * nx + s[5] * ny + s[2] * nz;
    double q = tx * nx + ty * ny + tz * nz;
    o[0] = s[0] - 4 * tx * nx + 4 * q * nx * nx;
    o[1] = s[1] - 4 * ty * ny + 4 * q * ny * ny;
    o[2] = s[2] - 4 * tz * nz + 4 * q * nz * nz;
    o[3] = s[3] - 2 * (tx * ny + nx * ty) + 4 * q * nx * ny;
    o[4] = s[4] - 2 * (tx * nz + nx * tz) + 4 * q * nx * nz;
    o[5] = s[5] - 2 * (ty * nz + ny * tz) + 4 * q * ny * nz;
  }
}

void update_ghost(orc_scene *sc, Sys &g, int stage) {  // :112-129, :261-278
  if (stage < 1 || stage > (int)g.copiers.size()) return;
  const Sys &ps = sc->sys[g.src];
  const int D = g.D, NS = ps.NS;
  for (auto &fm : g.copiers[stage - 1]) {
    for (int64_t k = 0; k < g.n; ++k) {
      int64_t src = g.idx_orig[k] - 1;
      if (fm.first == GR_F_P) {
        g.p[k] = ps.p[src];
      } else {
        if (fm.second == GR_COPY_HOUSEHOLDER)
          householder(&ps.stress[src * NS], &g.normals[k * D], NS, &g.stress[k * NS]);
        else
          for (int q = 0; q < NS; ++q) g.stress[k * NS + q] = ps.stress[src * NS + q];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Leapfrog step (src/TimeIntegration.jl:645-693)
// ---------------------------------------------------------------------------

// _sweep_all_stages! (src/TimeIntegration.jl:401-419).
int sweep_all_stages(orc_scene *sc, Integ &ig, double dt) {
  for (int v : ig.virt) {  // auto_zero_virtual! (src/Particles.jl:521-528)
    Sys &s = sc->sys[v];
    std::fill(s.wsum.begin(), s.wsum.end(), 0.0);
    zero_fields(s, s.zero_mask);
  }
  for (int stage = 1; stage <= ig.nst; ++stage) {
    for (int id : ig.sys) {
      Sys &s = sc->sys[id];
      if ((int)s.upd.size() != ig.nst) continue;
      int rc = update_state(sc, s, stage, dt);
      if (rc) return rc;
    }
    for (int g : ig.ghost) update_ghost(sc, sc->sys[g], stage);
    for (int v : ig.virt) {
      Sys &s = sc->sys[v];
      if ((int)s.upd.size() != ig.nst) continue;
      int rc = update_state(sc, s, stage, dt);
      if (rc) return rc;
    }
    for (int it : ig.inter) {
      int rc = sweep(sc, sc->inter[it], stage);
      if (rc) return rc;
    }
  }
  return GR_OK;
}

void apply_damping(orc_scene *sc, Integ &ig, double dt) {  // _apply_damping_all! :251-271
  if (ig.gamma == 0.0) return;
  const double gdt = ig.gamma / dt;
  for (int id : ig.sys) {
    Sys &s = sc->sys[id];
    for (int64_t i = 0; i < s.n * s.D; ++i) s.dvdt[i] = s.dvdt[i] - gdt * s.v[i];
  }
}

void reset_dqdt(Sys &s) {  // _reset_dqdt_pairs! :150-154
  const int D = s.D;
  for (int64_t i = 0; i < s.n; ++i) {
    for (int d = 0; d < D; ++d) s.dvdt[i * D + d] = s.sv[d];
    s.drhodt[i] = s.srho;
  }
}

void advance_virtual_positions(orc_scene *sc, Integ &ig, double dt) {  // :389-397
  for (int v : ig.virt) {
    Sys &s = sc->sys[v];
    const int D = s.D;
    for (int64_t i = 0; i < s.n; ++i)
      for (int d = 0; d < D; ++d) s.x[i * D + d] = s.x[i * D + d] + s.pv[d] * dt;
  }
}

// _advance_probe_positions! (src/TimeIntegration.jl:439-450): x += pv·dt
// when the prescribed velocity is nonzero.
void advance_probe_positions(orc_scene *sc, Integ &ig, double dt) {
  for (int q : ig.probes) {
    Sys &s = sc->sys[q];
    const int D = s.D;
    bool nz = false;
    for (int d = 0; d < D; ++d) nz = nz || s.pv[d] != 0.0;
    if (!nz) continue;
    for (int64_t i = 0; i < s.n; ++i)
      for (int d = 0; d < D; ++d) s.x[i * D + d] = s.x[i * D + d] + s.pv[d] * dt;
  }
}

// _measure_probes! (src/TimeIntegration.jl:454-502): re-sort the probe
// interactions' sources, mirror probe positions from their targets, sort the
// probes by cell, build the grids, zero w_sum and the extras, sweep stage 1,
// run the probes' updaters, sort the probes back by id.
int measure_probes(orc_scene *sc, Integ &ig) {
  const double cutoff = 2.0 * ig.h;
  std::vector<int> sorted;
  for (int it : ig.probe_inter) {
    const int a = sc->inter[it].a;
    if (std::find(sorted.begin(), sorted.end(), a) != sorted.end()) continue;
    sorted.push_back(a);
    sort_system(sc, sc->sys[a], cutoff);
  }
  for (int q : ig.probes) {
    Sys &pr = sc->sys[q];
    if (pr.mirror < 0) continue;
    const Sys &m = sc->sys[pr.mirror];
    const int D = pr.D;
    for (int64_t i = 0; i < pr.n; ++i)
      for (int d = 0; d < D; ++d) pr.x[(m.id[i] - 1) * D + d] = m.x[i * D + d];
  }
  for (int q : ig.probes) sort_system(sc, sc->sys[q], cutoff);
  for (int it : ig.probe_inter) {
    int rc = create_grid(sc, sc->inter[it]);
    if (rc) return rc;
  }
  for (int q : ig.probes) {  // auto_zero_probe! (src/ProbeParticles.jl:141-148)
    Sys &pr = sc->sys[q];
    std::fill(pr.wsum.begin(), pr.wsum.end(), 0.0);
    zero_fields(pr, pr.zero_mask);
  }
  for (int it : ig.probe_inter) {
    int rc = sweep(sc, sc->inter[it], 1);
    if (rc) return rc;
  }
  for (int q : ig.probes) {
    int rc = update_state(sc, sc->sys[q], 1, 0.0);
    if (rc) return rc;
  }
  for (int q : ig.probes) {  // _sort_probe_by_id! (:150-158)
    Sys &pr = sc->sys[q];
    std::vector<int64_t> perm(pr.n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int64_t a, int64_t b) { return pr.id[a] < pr.id[b]; });
    const int D = pr.D;
    permute(pr.x, perm, D);
    permute(pr.id, perm, 1);
    permute(pr.wsum, perm, 1);
    permute(pr.cnt, perm, 1);
    permute(pr.v, perm, D);
    permute(pr.rho, perm, 1);
    if (!pr.p.empty()) permute(pr.p, perm, 1);
    if (!pr.stress.empty()) permute(pr.stress, perm, pr.NS);
  }
  return GR_OK;
}

// _xsph_correction! (src/TimeIntegration.jl:423-436): v −= v_adjustment,
// v_adjustment ← 0, adjust_v! of every interaction, v += v_adjustment.
int xsph_correction(orc_scene *sc, Integ &ig) {
  for (int id : ig.sys) {
    Sys &s = sc->sys[id];
    for (size_t k = 0; k < s.v.size(); ++k) s.v[k] = s.v[k] + (-1.0) * s.vadj[k];
  }
  for (int id : ig.sys) {
    Sys &s = sc->sys[id];
    std::fill(s.vadj.begin(), s.vadj.end(), 0.0);
  }
  for (int it : ig.inter) {
    Inter &in = sc->inter[it];
    if (!in.has_xsph) continue;
    gr_pfn_desc f;
    std::memset(&f, 0, sizeof(f));
    f.kind = GR_PFN_XSPH;
    f.alpha = in.xsph_eps;
    int rc = sweep_fn(sc, in, f);
    if (rc) return rc;
  }
  for (int id : ig.sys) {
    Sys &s = sc->sys[id];
    for (size_t k = 0; k < s.v.size(); ++k) s.v[k] = s.v[k] + 1.0 * s.vadj[k];
  }
  return GR_OK;
}

// time_integrate!(::RK4TimeIntegrator) per-step body (src/TimeIntegration.jl:
// 773-819): one grid per step, four stage passes from q0 with the advance
// coefficients (0, ½, ½, 1)·dt, weights (1/6, 1/3, 1/3, 1/6), q = q0 + dt·acc.
// Scalars use the fused axpy (@fastmath, src/Utils.jl:19-29), vectors do not.
int step_rk4(orc_scene *sc, Integ &ig, double dt) {
  const double cutoff = 2.0 * ig.h;
  for (int id : ig.sys)
    if (nonfinite_x(sc->sys[id].x, sc->sys[id].n, sc->sys[id].D))
      return fail(sc, GR_ERR_ARGUMENT, "sort_particles!: non-finite position (InexactError)");
  for (int id : ig.sys) sort_system(sc, sc->sys[id], cutoff);
  const size_t ns = ig.sys.size();
  std::vector<std::vector<double>> v0(ns), r0(ns), av(ns), ar(ns);
  for (size_t k = 0; k < ns; ++k) {
    const Sys &s = sc->sys[ig.sys[k]];
    v0[k] = s.v;
    r0[k] = s.rho;
    av[k].assign(s.v.size(), 0.0);
    ar[k].assign(s.rho.size(), 0.0);
  }
  for (int g : ig.ghost) generate_ghosts(sc, sc->sys[g]);
  for (int g : ig.ghost) sort_system(sc, sc->sys[g], cutoff);
  for (int v : ig.virt) sort_system(sc, sc->sys[v], cutoff);
  for (int it : ig.inter) {
    int rc = create_grid(sc, sc->inter[it]);
    if (rc) return rc;
  }
  const double adv[4] = {0.0, 0.5, 0.5, 1.0};
  const double wts[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  for (int r = 0; r < 4; ++r) {
    for (size_t k = 0; k < ns; ++k) {
      Sys &s = sc->sys[ig.sys[k]];
      const int D = s.D;
      if (r > 0) {  // _rk4_advance_all!: q = q0 + (adv·dt)·dqdt
        const double cf = adv[r] * dt;
        for (int64_t i = 0; i < s.n; ++i) {
          for (int d = 0; d < D; ++d)
            s.v[i * D + d] = v0[k][i * D + d] + cf * s.dvdt[i * D + d];
          s.rho[i] = std::fma(cf, s.drhodt[i], r0[k][i]);
        }
      }
      reset_dqdt(s);
    }
    for (int g : ig.ghost) ghost_kinematics(sc, sc->sys[g]);
    int rc = sweep_all_stages(sc, ig, dt);
    if (rc) return rc;
    apply_damping(sc, ig, dt);
    for (size_t k = 0; k < ns; ++k) {  // _acc_dqdt_all!: acc += w·dqdt
      Sys &s = sc->sys[ig.sys[k]];
      for (size_t i = 0; i < av[k].size(); ++i) av[k][i] = av[k][i] + wts[r] * s.dvdt[i];
      for (size_t i = 0; i < ar[k].size(); ++i) ar[k][i] = std::fma(wts[r], s.drhodt[i], ar[k][i]);
    }
  }
  for (size_t k = 0; k < ns; ++k) {  // _apply_acc_all!
    Sys &s = sc->sys[ig.sys[k]];
    const int D = s.D;
    for (int64_t i = 0; i < s.n; ++i) {
      for (int d = 0; d < D; ++d) s.v[i * D + d] = v0[k][i * D + d] + dt * av[k][i * D + d];
      s.rho[i] = std::fma(dt, ar[k][i], r0[k][i]);
    }
  }
  int rcx = xsph_correction(sc, ig);
  if (rcx) return rcx;
  for (int id : ig.sys) {  // _update_positions_all!
    Sys &s = sc->sys[id];
    for (size_t k = 0; k < s.x.size(); ++k) s.x[k] = s.x[k] + dt * s.v[k];
  }
  advance_virtual_positions(sc, ig, dt);
  advance_probe_positions(sc, ig, dt);
  return GR_OK;
}

int step(orc_scene *sc, Integ &ig, double dt) {
  if (ig.scheme == GR_SCHEME_RK4) return step_rk4(sc, ig, dt);
  const double cutoff = 2.0 * ig.h;  // sort_cutoff :612
  for (int id : ig.sys)
    if (nonfinite_x(sc->sys[id].x, sc->sys[id].n, sc->sys[id].D))
      return fail(sc, GR_ERR_ARGUMENT, "sort_particles!: non-finite position (InexactError)");
  for (int id : ig.sys) sort_system(sc, sc->sys[id], cutoff);
  std::vector<std::vector<double>> v0(ig.sys.size()), r0(ig.sys.size());
  for (size_t k = 0; k < ig.sys.size(); ++k) {
    v0[k] = sc->sys[ig.sys[k]].v;
    r0[k] = sc->sys[ig.sys[k]].rho;
  }
  for (int g : ig.ghost) generate_ghosts(sc, sc->sys[g]);
  for (int g : ig.ghost) sort_system(sc, sc->sys[g], cutoff);
  for (int v : ig.virt) sort_system(sc, sc->sys[v], cutoff);  // _prepare_grids! :373-375
  for (int it : ig.inter) {
    int rc = create_grid(sc, sc->inter[it]);
    if (rc) return rc;
  }
  const double half = dt / 2;
  for (int id : ig.sys) {  // _halfstep_ps! :158-162
    Sys &s = sc->sys[id];
    const int D = s.D;
#pragma omp parallel for
    for (int64_t i = 0; i < s.n; ++i) {
      for (int d = 0; d < D; ++d) s.v[i * D + d] = s.v[i * D + d] + half * s.dvdt[i * D + d];
      s.rho[i] = std::fma(half, s.drhodt[i], s.rho[i]);
      for (int d = 0; d < D; ++d) s.dvdt[i * D + d] = s.sv[d];
      s.drhodt[i] = s.srho;
    }
  }
  for (int g : ig.ghost) ghost_kinematics(sc, sc->sys[g]);
  for (int v : ig.virt) {  // auto_zero_virtual! (src/Particles.jl:521-528)
    Sys &s = sc->sys[v];
    std::fill(s.wsum.begin(), s.wsum.end(), 0.0);
    zero_fields(s, s.zero_mask);
  }
  for (int stage = 1; stage <= ig.nst; ++stage) {  // _sweep_all_stages! :401-419
    for (int id : ig.sys) {
      Sys &s = sc->sys[id];
      if ((int)s.upd.size() != ig.nst) continue;
      int rc = update_state(sc, s, stage, dt);
      if (rc) return rc;
    }
    for (int g : ig.ghost) update_ghost(sc, sc->sys[g], stage);
    for (int v : ig.virt) {
      Sys &s = sc->sys[v];
      if ((int)s.upd.size() != ig.nst) continue;
      int rc = update_state(sc, s, stage, dt);
      if (rc) return rc;
    }
    for (int it : ig.inter) {
      int rc = sweep(sc, sc->inter[it], stage);
      if (rc) return rc;
    }
  }
  if (ig.gamma != 0.0) {  // _apply_damping_all! :251-271
    const double gdt = ig.gamma / dt;
    for (int id : ig.sys) {
      Sys &s = sc->sys[id];
      for (int64_t i = 0; i < s.n * s.D; ++i) s.dvdt[i] = s.dvdt[i] - gdt * s.v[i];
    }
  }
  for (size_t k = 0; k < ig.sys.size(); ++k) {  // _fullstep_q_all!, _update_positions_all!
    Sys &s = sc->sys[ig.sys[k]];
    const int D = s.D;
    const double *vv = v0[k].data(), *rr = r0[k].data();
#pragma omp parallel for
    for (int64_t i = 0; i < s.n; ++i) {
      for (int d = 0; d < D; ++d) s.v[i * D + d] = vv[i * D + d] + dt * s.dvdt[i * D + d];
      s.rho[i] = std::fma(dt, s.drhodt[i], rr[i]);
    }
  }
  {
    int rc = xsph_correction(sc, ig);
    if (rc) return rc;
  }
  for (int id : ig.sys) {  // _update_positions_all! :235-240
    Sys &s = sc->sys[id];
    for (size_t k = 0; k < s.x.size(); ++k) s.x[k] = s.x[k] + dt * s.v[k];
  }
  for (int v : ig.virt) {  // _update_virtual_positions! :389-397
    Sys &s = sc->sys[v];
    const int D = s.D;
    for (int64_t i = 0; i < s.n; ++i)
      for (int d = 0; d < D; ++d) s.x[i * D + d] = s.x[i * D + d] + s.pv[d] * dt;
  }
  advance_probe_positions(sc, ig, dt);
  return GR_OK;
}

int ncomp(const orc_scene *sc, const Sys &s, int field) {
  const int D = s.D;
  int NS = s.NS;
  if (s.kind == GR_SYS_GHOST) NS = sc->sys[s.src].NS;
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
      return s.NV;
  }
  return 0;
}

std::vector<double> *field_ptr(Sys &s, int field) {
  switch (field) {
    case GR_F_X: return &s.x;
    case GR_F_V: return &s.v;
    case GR_F_RHO: return &s.rho;
    case GR_F_DVDT: return s.kind == GR_SYS_GHOST ? nullptr : &s.dvdt;
    case GR_F_DRHODT: return s.kind == GR_SYS_GHOST ? nullptr : &s.drhodt;
    case GR_F_P: return (s.kind >= GR_SYS_FLUID && (s.kind != GR_SYS_GHOST || s.ex_p)) ? &s.p : nullptr;
    case GR_F_STRESS:
      return ((s.kind >= GR_SYS_STRESS && s.kind != GR_SYS_GHOST) || (s.kind == GR_SYS_GHOST && s.ex_stress)) ? &s.stress : nullptr;
    case GR_F_STRAIN_RATE: return (s.kind == GR_SYS_STRESS || s.kind == GR_SYS_EP) ? &s.sr : nullptr;
    case GR_F_VORTICITY: return s.kind == GR_SYS_EP ? &s.vor : nullptr;
    case GR_F_STRAIN: return s.kind == GR_SYS_EP ? &s.strain : nullptr;
    case GR_F_STRAIN_P: return s.kind == GR_SYS_EP ? &s.strainp : nullptr;
    case GR_F_V_ADJUSTMENT: return s.kind == GR_SYS_GHOST ? nullptr : &s.vadj;
    case GR_F_NORMALS: return s.kind == GR_SYS_GHOST ? &s.normals : nullptr;
    case GR_F_W_SUM: return s.virt ? &s.wsum : nullptr;
    case GR_F_COUNT: return s.probe ? &s.cnt : nullptr;
  }
  return nullptr;
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================

extern "C" {

int orc_scene_create(orc_scene **out, int ndims) {
  if (!out || (ndims != 2 && ndims != 3)) return GR_ERR_ARGUMENT;
  *out = new orc_scene();
  (*out)->D = ndims;
  return GR_OK;
}

int orc_scene_destroy(orc_scene *s) {
  delete s;
  return GR_OK;
}

const char *orc_last_error(const orc_scene *s) { return s ? s->err.c_str() : "null scene"; }

int orc_set_threads(orc_scene *s, int nthreads) {
  s->nthreads = nthreads;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  return GR_OK;
}

int orc_system_create(orc_scene *s, int kind, int64_t n, int ns, double mass,
                      double c, const double *source_v, double source_rho,
                      int *sys_out) {
  if (n <= 0) return fail(s, GR_ERR_ARGUMENT, "n must be positive, got " + std::to_string(n));
  if (kind < GR_SYS_BASIC || kind > GR_SYS_EP) return fail(s, GR_ERR_ARGUMENT, "bad system kind");
  if (kind >= GR_SYS_STRESS && ns != 3 && ns != 4 && ns != 6)
    return fail(s, GR_ERR_ARGUMENT, "ns must be 3, 4, or 6, got " + std::to_string(ns));
  Sys y;
  y.kind = kind;
  y.D = s->D;
  y.NS = kind >= GR_SYS_STRESS ? ns : 0;
  y.NV = kind == GR_SYS_EP ? (s->D == 2 ? 1 : 3) : 0;
  y.n = n;
  y.mass = mass;
  y.c = c;
  for (int d = 0; d < s->D; ++d) y.sv[d] = source_v ? source_v[d] : 0.0;
  y.srho = source_rho;
  const int D = s->D;
  y.x.assign(n * D, 0.0);
  y.v.assign(n * D, 0.0);
  y.vadj.assign(n * D, 0.0);
  y.rho.assign(n, 0.0);
  y.dvdt.assign(n * D, 0.0);
  y.drhodt.assign(n, 0.0);
  y.id.resize(n);
  for (int64_t i = 0; i < n; ++i) y.id[i] = i + 1;
  if (kind >= GR_SYS_FLUID) y.p.assign(n, 0.0);
  if (kind >= GR_SYS_STRESS) {
    y.stress.assign(n * ns, 0.0);
    y.sr.assign(n * ns, 0.0);
  }
  if (kind == GR_SYS_EP) {
    y.vor.assign(n * y.NV, 0.0);
    y.strain.assign(n * ns, 0.0);
    y.strainp.assign(n * ns, 0.0);
  }
  s->sys.push_back(std::move(y));
  *sys_out = (int)s->sys.size() - 1;
  return GR_OK;
}

int orc_virtual_create(orc_scene *s, int sys, int nzero, const int *zero_fields,
                       const double *prescribed_v) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  Sys &y = s->sys[sys];
  if (y.kind == GR_SYS_GHOST || y.virt) return fail(s, GR_ERR_ARGUMENT, "not a plain particle system");
  y.virt = true;
  y.wsum.assign(y.n, 0.0);
  y.zero_mask = 0;
  for (int k = 0; k < nzero; ++k) {
    if (zero_fields[k] < 0 || zero_fields[k] >= GR_NUM_FIELDS || !field_ptr(y, zero_fields[k]) ||
        zero_fields[k] == GR_F_X)
      return fail(s, GR_ERR_ARGUMENT, "zero_fields: field not present on the virtual system");
    y.zero_mask |= (int64_t)1 << zero_fields[k];
  }
  for (int d = 0; d < y.D; ++d) y.pv[d] = prescribed_v ? prescribed_v[d] : 0.0;
  y.upd.clear();  // the virtual system's own updaters follow
  return GR_OK;
}

int orc_probe_create(orc_scene *s, int sys, const double *prescribed_v, int mirror, int nzero,
                     const int *zero_fields) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  Sys &y = s->sys[sys];
  if (y.kind == GR_SYS_GHOST || y.virt) return fail(s, GR_ERR_ARGUMENT, "not a plain particle system");
  if (mirror >= 0) {
    if (mirror >= (int)s->sys.size() || s->sys[mirror].kind == GR_SYS_GHOST || s->sys[mirror].virt)
      return fail(s, GR_ERR_ARGUMENT, "mirror_target must be a real particle system");
    if (s->sys[mirror].n != y.n)
      return fail(s, GR_ERR_ARGUMENT, "a mirroring probe has one slot per target particle");
  }
  y.probe = true;
  y.mirror = mirror;
  y.cnt.assign(y.n, 0.0);
  int rc = orc_virtual_create(s, sys, nzero, zero_fields, prescribed_v);
  if (rc) {
    y.probe = false;
    return rc;
  }
  return GR_OK;
}

int orc_integrator_set_probes(orc_scene *s, int integ, int nprobes, const int *probes, int ninters,
                              const int *inters) {
  if (integ < 0 || integ >= (int)s->integ.size()) return fail(s, GR_ERR_ARGUMENT, "bad integrator");
  Integ &ig = s->integ[integ];
  std::vector<int> pr(probes, probes + nprobes), it(inters, inters + ninters);
  for (int q : pr)
    if (q < 0 || q >= (int)s->sys.size() || !s->sys[q].probe)
      return fail(s, GR_ERR_ARGUMENT, "probes must be ProbeParticleSystems");
  for (int t : it) {
    if (t < 0 || t >= (int)s->inter.size()) return fail(s, GR_ERR_ARGUMENT, "bad interaction");
    const Inter &in = s->inter[t];
    if (in.b < 0 || std::find(pr.begin(), pr.end(), in.b) == pr.end())
      return fail(s, GR_ERR_ARGUMENT, "a probe interaction's system_b must be one of the probes");
  }
  ig.probes = pr;
  ig.probe_inter = it;
  return GR_OK;
}

int orc_measure_probes(orc_scene *s, int integ) {
  if (integ < 0 || integ >= (int)s->integ.size()) return fail(s, GR_ERR_ARGUMENT, "bad integrator");
  return measure_probes(s, s->integ[integ]);
}

int orc_system_add_updater(orc_scene *s, int sys, int stage, int kind,
                           const double *params, int nparams) {
  if (sys < 0 || sys >= (int)s->sys.size() || stage < 1) return fail(s, GR_ERR_ARGUMENT, "bad system/stage");
  Sys &y = s->sys[sys];
  if (y.kind == GR_SYS_GHOST) return fail(s, GR_ERR_ARGUMENT, "ghost systems take copiers, not updaters");
  if (kind == GR_UPD_EP && y.NS != 4)
    return fail(s, GR_ERR_ASSERTION,
                "ElastoPlasticStressUpdater requires NS=4 (plane-strain Voigt). Got NS=" + std::to_string(y.NS) + ".");
  if (kind == GR_UPD_HOOKE && y.NS != 4)
    return fail(s, GR_ERR_ASSERTION,
                "HookeLawStressUpdater requires NS=4 (plane-strain Voigt). Got NS=" + std::to_string(y.NS) + ".");
  if (((kind == GR_UPD_EP || kind == GR_UPD_HOOKE) && y.kind != GR_SYS_EP) ||
      (kind == GR_UPD_VPMC && y.kind < GR_SYS_STRESS) ||
      ((kind == GR_UPD_TAIT || kind == GR_UPD_LINEAR) && y.kind < GR_SYS_FLUID) ||
      ((kind == GR_UPD_VNORM || kind == GR_UPD_PRESCRIBED_V) && !y.virt))
    return fail(s, GR_ERR_ARGUMENT, "updater needs fields the system does not have");
  if ((int)y.upd.size() < stage) y.upd.resize(stage);
  Upd u;
  u.kind = kind;
  for (int k = 0; k < nparams && k < 5; ++k) u.prm[k] = params[k];
  if (kind == GR_UPD_TAIT && nparams < 2) u.prm[1] = 7.0;
  y.upd[stage - 1].push_back(u);
  return GR_OK;
}

int orc_system_upload(orc_scene *s, int sys, int field, const double *data, int64_t count) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  Sys &y = s->sys[sys];
  std::vector<double> *f = field_ptr(y, field);
  if (!f) return fail(s, GR_ERR_ARGUMENT, "field not present on this system");
  int64_t need = y.n * ncomp(s, y, field);
  if (count != need) return fail(s, GR_ERR_ARGUMENT, "count mismatch");
  std::copy(data, data + count, f->begin());
  return GR_OK;
}

int orc_system_download(orc_scene *s, int sys, int field, double *data, int64_t count) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  Sys &y = s->sys[sys];
  std::vector<double> *f = field_ptr(y, field);
  if (!f) return fail(s, GR_ERR_ARGUMENT, "field not present on this system");
  int64_t need = y.n * ncomp(s, y, field);
  if (count != need) return fail(s, GR_ERR_ARGUMENT, "count mismatch");
  std::copy(f->begin(), f->begin() + count, data);
  return GR_OK;
}

int orc_system_upload_index(orc_scene *s, int sys, int field, const int64_t *data, int64_t count) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  Sys &y = s->sys[sys];
  if (count != y.n) return fail(s, GR_ERR_ARGUMENT, "count mismatch");
  std::vector<int64_t> *f = field == GR_I_ID ? &y.id : field == GR_I_IDX_ORIGINAL ? &y.idx_orig : &y.idx_bnd;
  if ((y.kind == GR_SYS_GHOST) != (field != GR_I_ID)) return fail(s, GR_ERR_ARGUMENT, "index field not present");
  std::copy(data, data + count, f->begin());
  return GR_OK;
}

int orc_system_download_index(orc_scene *s, int sys, int field, int64_t *data, int64_t count) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  Sys &y = s->sys[sys];
  if (count != y.n) return fail(s, GR_ERR_ARGUMENT, "count mismatch");
  std::vector<int64_t> *f = field == GR_I_ID ? &y.id : field == GR_I_IDX_ORIGINAL ? &y.idx_orig : &y.idx_bnd;
  if ((y.kind == GR_SYS_GHOST) != (field != GR_I_ID)) return fail(s, GR_ERR_ARGUMENT, "index field not present");
  std::copy(f->begin(), f->begin() + count, data);
  return GR_OK;
}

int orc_system_get_n(orc_scene *s, int sys, int64_t *n_out) {
  if (sys < 0 || sys >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
  *n_out = s->sys[sys].n;
  return GR_OK;
}

int orc_ghost_create(orc_scene *s, int source_sys, double cutoff, int nb,
                     const double *normals, const double *points, int *ghost_out) {
  if (source_sys < 0 || source_sys >= (int)s->sys.size() || s->sys[source_sys].kind == GR_SYS_GHOST)
    return fail(s, GR_ERR_ARGUMENT, "bad ghost source");
  if (nb < 1) return fail(s, GR_ERR_ARGUMENT, "at least one boundary (normal, point) pair is required");
  Sys g;
  g.kind = GR_SYS_GHOST;
  g.D = s->D;
  g.NS = 0;
  g.n = 0;
  g.src = source_sys;
  g.gcut = cutoff;
  g.nb = nb;
  g.bn.assign(normals, normals + nb * s->D);
  g.bp.assign(points, points + nb * s->D);
  s->sys.push_back(std::move(g));
  *ghost_out = (int)s->sys.size() - 1;
  return GR_OK;
}

int orc_ghost_set_copier(orc_scene *s, int ghost, int stage, int nfields,
                         const int *fields, const int *modes) {
  if (ghost < 0 || ghost >= (int)s->sys.size() || s->sys[ghost].kind != GR_SYS_GHOST || stage < 1)
    return fail(s, GR_ERR_ARGUMENT, "bad ghost/stage");
  Sys &g = s->sys[ghost];
  const Sys &src = s->sys[g.src];
  if ((int)g.copiers.size() < stage) g.copiers.resize(stage);
  for (int k = 0; k < nfields; ++k) {
    int f = fields[k];
    if (f == GR_F_P) {
      if (src.kind < GR_SYS_FLUID) return fail(s, GR_ERR_ARGUMENT, "source has no p");
      g.ex_p = true;
    } else if (f == GR_F_STRESS) {
      if (src.kind < GR_SYS_STRESS) return fail(s, GR_ERR_ARGUMENT, "source has no stress");
      if (modes && modes[k] == GR_COPY_HOUSEHOLDER && (src.NS == 6) != (s->D == 3))
        return fail(s, GR_ERR_ARGUMENT, "Householder reflection needs NS=6 in 3D, NS=3/4 in 2D");
      g.ex_stress = true;
    } else {
      return fail(s, GR_ERR_ARGUMENT, "GhostCopier supports :p and :stress");
    }
    g.copiers[stage - 1].push_back({f, modes ? modes[k] : GR_COPY_PLAIN});
  }
  return GR_OK;
}

int orc_interaction_create(orc_scene *s, int kernel_kind, double h, int sys_a,
                           int partner, int wrap, const double *normal,
                           const double *point, double beta, double lj_cutoff,
                           int n_stages, const gr_pfn_desc *pfns, int *inter_out) {
  if (sys_a < 0 || sys_a >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system_a");
  if (partner >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system_b");
  if (kernel_kind != GR_KERNEL_CUBIC && kernel_kind != GR_KERNEL_WENDLAND_C2)
    return fail(s, GR_ERR_ARGUMENT, "bad kernel kind");
  if (s->sys[sys_a].kind == GR_SYS_GHOST) return fail(s, GR_ERR_ARGUMENT, "a ghost system cannot be system_a");
  if (wrap != GR_WRAP_NONE && (partner < 0 || s->sys[partner].kind == GR_SYS_GHOST))
    return fail(s, GR_ERR_ARGUMENT, "boundary wrappers need a real partner system");
  Inter in;
  in.kk = kernel_kind;
  in.D = s->D;
  in.h = h;
  in.norm = kernel_norm(kernel_kind, s->D, h);
  in.cell = 2.0 * h;
  in.a = sys_a;
  in.b = partner;
  in.wrap = wrap;
  for (int d = 0; d < s->D; ++d) {
    in.wn[d] = normal ? normal[d] : 0.0;
    in.wp[d] = point ? point[d] : 0.0;
  }
  in.beta = beta;
  in.ljcut = lj_cutoff;
  in.pfns.assign(pfns, pfns + n_stages);
  const Sys &A = s->sys[sys_a];
  for (const gr_pfn_desc &f : in.pfns) {
    if (f.kind == GR_PFN_NONE) continue;
    if (f.kind == GR_PFN_FLUID && (f.sigma != 1 && f.sigma != 2))
      return fail(s, GR_ERR_ARGUMENT, "FluidPfn sigma must be 1 or 2");
    if ((f.kind == GR_PFN_FLUID) && A.kind < GR_SYS_FLUID) return fail(s, GR_ERR_ARGUMENT, "FluidPfn needs p on system_a");
    if (f.kind == GR_PFN_CAUCHY && A.kind < GR_SYS_STRESS) return fail(s, GR_ERR_ARGUMENT, "CauchyFluidPfn needs stress on system_a");
    if (f.kind == GR_PFN_SR && A.kind < GR_SYS_STRESS) return fail(s, GR_ERR_ARGUMENT, "StrainRatePfn needs strain_rate on system_a");
    if (f.kind == GR_PFN_SRV && A.kind != GR_SYS_EP) return fail(s, GR_ERR_ARGUMENT, "StrainRateVorticityPfn needs vorticity on system_a");
    if (partner >= 0) {
      const Sys &B = s->sys[partner];
      PartnerKind pk = classify(s, in, f.kind);
      if (B.kind == GR_SYS_GHOST) {
        if (f.kind == GR_PFN_FLUID && !B.ex_p)
          return fail(s, GR_ERR_ARGUMENT, "ghost partner of FluidPfn must copy :p (GhostCopier)");
        if (f.kind == GR_PFN_CAUCHY && !B.ex_stress)
          return fail(s, GR_ERR_ARGUMENT, "ghost partner of CauchyFluidPfn must copy :stress (GhostCopier)");
      } else if (pk == P_TWO_SIDED || pk == P_ONE_SIDED) {
        if (f.kind == GR_PFN_FLUID && B.kind < GR_SYS_FLUID) return fail(s, GR_ERR_ARGUMENT, "FluidPfn partner needs p");
        if (f.kind == GR_PFN_CAUCHY && B.kind < GR_SYS_STRESS) return fail(s, GR_ERR_ARGUMENT, "CauchyFluidPfn partner needs stress");
        if (f.kind == GR_PFN_CAUCHY && B.NS != A.NS) return fail(s, GR_ERR_ARGUMENT, "stress length mismatch");
      }
    }
  }
  s->inter.push_back(std::move(in));
  *inter_out = (int)s->inter.size() - 1;
  return GR_OK;
}

int orc_integrator_create(orc_scene *s, int nsys, const int *systems, int ninter,
                          const int *inters, int nghost, const int *ghosts,
                          double gamma, int *integ_out) {
  if (nsys < 1) return fail(s, GR_ERR_ARGUMENT, "systems must not be empty");
  if (ninter < 1) return fail(s, GR_ERR_ARGUMENT, "interactions must not be empty");
  Integ ig;
  for (int k = 0; k < nsys; ++k) {
    if (systems[k] < 0 || systems[k] >= (int)s->sys.size()) return fail(s, GR_ERR_ARGUMENT, "bad system");
    (s->sys[systems[k]].virt ? ig.virt : ig.sys).push_back(systems[k]);
  }
  if (ig.sys.empty()) return fail(s, GR_ERR_ARGUMENT, "systems must not be empty");
  ig.inter.assign(inters, inters + ninter);
  if (nghost > 0) ig.ghost.assign(ghosts, ghosts + nghost);
  ig.gamma = gamma;
  ig.c = -INFINITY;
  for (int id : ig.sys) ig.c = std::max(ig.c, s->sys[id].c);
  ig.h = INFINITY;
  for (int it : ig.inter) ig.h = std::min(ig.h, s->inter[it].h);
  ig.nst = (int)s->inter[ig.inter[0]].pfns.size();
  for (int it : ig.inter)
    if ((int)s->inter[it].pfns.size() != ig.nst)
      return fail(s, GR_ERR_ASSERTION, "All interactions must have the same number of stages (pfns length)");
  s->integ.push_back(std::move(ig));
  *integ_out = (int)s->integ.size() - 1;
  return GR_OK;
}

int orc_interaction_set_vadjust(orc_scene *s, int inter, double epsilon) {
  if (inter < 0 || inter >= (int)s->inter.size()) return fail(s, GR_ERR_ARGUMENT, "bad interaction");
  Inter &in = s->inter[inter];
  if (std::isnan(epsilon)) {
    in.has_xsph = false;
    return GR_OK;
  }
  in.has_xsph = true;
  in.xsph_eps = epsilon;
  return GR_OK;
}

int orc_integrator_set_scheme(orc_scene *s, int integ, int scheme) {
  if (integ < 0 || integ >= (int)s->integ.size()) return fail(s, GR_ERR_ARGUMENT, "bad integrator");
  if (scheme != GR_SCHEME_LEAPFROG && scheme != GR_SCHEME_RK4)
    return fail(s, GR_ERR_ARGUMENT, "unknown time integration scheme");
  s->integ[integ].scheme = scheme;
  return GR_OK;
}

int orc_integrator_dt(orc_scene *s, int integ, double cfl, double *dt_out) {
  const Integ &ig = s->integ[integ];
  *dt_out = cfl * ig.h / ig.c;
  return GR_OK;
}

int orc_step(orc_scene *s, int integ, int nsteps, double dt) {
  if (integ < 0 || integ >= (int)s->integ.size()) return fail(s, GR_ERR_ARGUMENT, "bad integrator");
  for (int k = 0; k < nsteps; ++k) {
    int rc = step(s, s->integ[integ], dt);
    if (rc) return rc;
  }
  return GR_OK;
}

int orc_sort(orc_scene *s, int sys, double cutoff) {
  if (nonfinite_x(s->sys[sys].x, s->sys[sys].n, s->sys[sys].D))
    return fail(s, GR_ERR_ARGUMENT, "sort_particles!: non-finite position (InexactError)");
  sort_system(s, s->sys[sys], cutoff);
  return GR_OK;
}
int orc_create_grid(orc_scene *s, int inter) { return create_grid(s, s->inter[inter]); }
int orc_sweep(orc_scene *s, int inter, int stage) { return sweep(s, s->inter[inter], stage); }
// adjust_v!(si) (src/Interaction.jl:309)
int orc_adjust_v(orc_scene *s, int inter) {
  if (inter < 0 || inter >= (int)s->inter.size()) return fail(s, GR_ERR_ARGUMENT, "bad interaction");
  Inter &in = s->inter[inter];
  if (!in.has_xsph) return GR_OK;
  gr_pfn_desc f;
  std::memset(&f, 0, sizeof(f));
  f.kind = GR_PFN_XSPH;
  f.alpha = in.xsph_eps;
  return sweep_fn(s, in, f);
}
int orc_update_state(orc_scene *s, int sys, int stage, double dt) {
  return update_state(s, s->sys[sys], stage, dt);
}
int orc_generate_ghosts(orc_scene *s, int ghost) {
  generate_ghosts(s, s->sys[ghost]);
  return GR_OK;
}
int orc_update_ghost_kinematics(orc_scene *s, int ghost) {
  ghost_kinematics(s, s->sys[ghost]);
  return GR_OK;
}
int orc_update_ghost(orc_scene *s, int ghost, int stage) {
  update_ghost(s, s->sys[ghost], stage);
  return GR_OK;
}

int orc_get_grid(orc_scene *s, int inter, double *mingridx, int64_t *ngridx, int64_t *ncells) {
  const Inter &in = s->inter[inter];
  if (!in.built) return fail(s, GR_ERR_STATE, "grid not built");
  for (int d = 0; d < in.D; ++d) {
    mingridx[d] = in.ming[d];
    ngridx[d] = in.ng[d];
  }
  *ncells = (int64_t)in.cs.size() - 1;
  return GR_OK;
}

int orc_get_cell_start(orc_scene *s, int inter, int which, int64_t *out, int64_t count) {
  const Inter &in = s->inter[inter];
  const std::vector<int64_t> &cs = which ? in.csa : in.cs;
  if (count != (int64_t)cs.size()) return fail(s, GR_ERR_ARGUMENT, "count mismatch");
  std::copy(cs.begin(), cs.end(), out);
  return GR_OK;
}

int orc_brute_force_counts(orc_scene *s, int sys_a, int sys_b, double cutoff, int64_t *counts_out) {
  const Sys &A = s->sys[sys_a];
  const Sys &B = s->sys[sys_b];
  const int D = s->D;
  const double cs = cutoff * cutoff;
#pragma omp parallel for
  for (int64_t i = 0; i < A.n; ++i) {
    int64_t cnt = 0;
    for (int64_t j = 0; j < B.n; ++j) {
      if (sys_a == sys_b && i == j) continue;
      double dx[3];
      for (int d = 0; d < D; ++d) dx[d] = A.x[i * D + d] - B.x[j * D + d];
      if (dotv(dx, dx, D) < cs) ++cnt;
    }
    counts_out[i] = cnt;
  }
  return GR_OK;
}

double orc_kernel_norm(int kernel_kind, int ndims, double h) { return kernel_norm(kernel_kind, ndims, h); }
double orc_kernel_w(int kernel_kind, int ndims, double h, double q) {
  return kernel_w(kernel_kind, kernel_norm(kernel_kind, ndims, h), q);
}
double orc_kernel_dw_dq(int kernel_kind, int ndims, double h, double q) {
  return kernel_dw(kernel_kind, kernel_norm(kernel_kind, ndims, h), q);
}
double orc_pow_int(double x, int64_t n) { return jl_pow_int(x, n); }

}  // extern "C"
