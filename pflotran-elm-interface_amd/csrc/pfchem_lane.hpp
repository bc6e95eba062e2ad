// pfchem_lane.hpp -- network-specialized batched reaction step: one cell per
// lane, persistent grid, lane-level work stealing.
//
// This file is device source compiled at run time (hiprtc) after a generated
// `struct Net` that holds one reaction network and its solver options as
// compile-time constants (species counts, species ids, stoichiometry, log K
// or its temperature coefficients, Debye-Hueckel parameters, mineral rate
// parameters, tolerances).  Every index into a cell's Jacobian is therefore a
// constant: the cell's Jacobian stays in registers, the complex and mineral
// loops are straight-line code, and no lane ever reads another lane's data
// (no atomics, no barriers; per-cell values the complex loop does not touch
// wait in the lane's own LDS slots).
//
// Algorithm: RStep (reaction.F90:3564-3738) / RReact (:3742-4055) with
// RTotalAqueous (:4665-4759), RTAccumulation(Derivative) (:5710-5848),
// RKineticMineral (reaction_mineral.F90:647-1078), RImmobileDecay
// (reaction_immobile.F90:244-294), RActivityCoefficients LAG per Newton
// iteration (reaction.F90:4555-4611) and RSolve (:5457-5516,
// back-substituting).  Reformulation of the linear solve (same Newton step,
// different rounding):
//   * in the log formulation the aqueous Jacobian of RReact is
//       J_log = jc (diag(x) + S^T diag(m) S) + sum_m g_m nu_m nu_m^T,
//     jc = phi s 1000 V rho/1000 / dt, with g_m = dIm/dQK QK V rho/1000 >= 0
//     for every TST rate law: symmetric positive definite.  It is scaled
//     symmetrically to unit diagonal, A = D^-1/2 M D^-1/2 (M = J_log / jc),
//     and factored as L D L^T without pivoting; the linear formulation solves
//     the same system and multiplies the step by x (J_lin = J_log diag(x)^-1);
//   * immobile rows are diagonal (accumulation + first-order decay);
//   * JT = float: the off-diagonal entries are accumulated pre-scaled by
//     1/sqrt(x_i x_j) in FP32 (|entry| <= sqrt(D_i D_j / x_i x_j)), scaled to
//     A in FP32 and factored in FP32; the diagonal, totals, residual, log K
//     sums, exponentials, the update and every convergence test stay FP64
//     (the Newton iteration is the iterative refinement);
//   * a pivot below the tolerance, a non-finite entry, or (FP32) a Newton
//     iteration that does not converge hands the cell, with its inputs
//     untouched, to a later pass: status kEscalate.
#pragma once

namespace pflane {

typedef long long i64;

// pflotran_constants.F90:84, 89; reactive_transport_aux.F90:19
constexpr double kLogToLn = 2.30258509299;
constexpr double kIdealGas = 8.31446;
constexpr double kRtMinSat = 1.0e-40;
constexpr int kEscalate = 16;        // per-cell status: handed to the next pass
constexpr int kClaim = 64;           // cells per refill of a wave's queue
enum : int { kOk = 0, kMaxIter = 1, kSingular = 2 };
enum : int { kFreqNewtonIter = 2 };

// kernel arguments (by value); the host defines the same layout
struct Args {
  i64 ld, ncell;
  double* pri_molal;
  double* total;
  double* immobile;
  double* sec_molal;
  double* pri_act_coef;
  double* sec_act_coef;
  double* ln_act_h2o;
  double* mnrl_volfrac;
  const double* mnrl_area;
  double* mnrl_rate;
  const double* porosity;
  const double* saturation;
  const double* den_kg;
  const double* volume;
  const double* temperature;
  int* num_iterations;
  int* num_sub_steps;
  int* status;
  unsigned long long* cursors;  // one work cursor per XCD, 128-byte stride
  double dt;
  int nxcd;
  int rescue;    // 1: only cells whose status is kEscalate
  int escalate;  // 1: hand failing cells to a later pass (status kEscalate)
  int list_min;  // with a list: the pass does nothing unless it holds more cells
  const int* list;        // NULL, or the cells to step: list[0 .. *list_count)
  const int* list_count;  // device memory, written by an earlier kernel
};
static_assert(sizeof(Args) == 208, "Args layout");

template <int V>
struct IC {
  static constexpr int value = V;
};
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(IC<B>{});
    sfor<B + 1, E>(f);
  }
}

// packed lower triangle, i > j (off-diagonal entries only)
__host__ __device__ constexpr int tri(int i, int j) {
  return i > j ? i * (i - 1) / 2 + j : j * (j - 1) / 2 + i;
}

// FP64 exp / log (inlined: out-of-line calls cost more than the
// instruction-cache misses of the larger unrolled loop, measured on MI355X)
__device__ __forceinline__ double pf_exp(double v) { return exp(v); }
__device__ __forceinline__ double pf_log(double v) { return log(v); }

// 1/g by the hardware reciprocal estimate refined with two Newton-Raphson
// steps r <- r + r (1 - g r) (within about one ulp of the quotient; g > 0,
// normal)
__device__ __forceinline__ double rcp_nr(double g) {
  double r = __builtin_amdgcn_rcp(g);
  double e = fma(-g, r, 1.0);
  r = fma(e, r, r);
  e = fma(-g, r, 1.0);
  return fma(e, r, r);
}

// ReactionInterpolateLogK (reaction_aux.F90:1285-1312)
__device__ __forceinline__ double interp_logk(const double (&c)[5], double lnT, double tk) {
  return c[0] * lnT + c[1] + c[2] * tk + c[3] / tk + c[4] / (tk * tk);
}

// Debye-Hueckel activity coefficient (reaction.F90:4586-4592)
__device__ __forceinline__ double dh_gamma(double z, double a0, double sI, double I) {
  return pf_exp((-z * z * sI * Net::DH_A / (1.0 + a0 * Net::DH_B * sI) + Net::DH_BDOT * I) *
             kLogToLn);
}

// ln K of complex k / mineral m at the cell temperature (isothermal: the
// tabulated value; the product -logK*ln10 is formed as the reference does)
template <int K>
__device__ __forceinline__ double cplx_lnk(double lnT, double tk) {
  if constexpr (Net::ISOTHERMAL || !Net::HAS_CPLX_COEF)
    return -Net::CPLX_LOGK[K] * kLogToLn;
  else
    return -interp_logk(Net::CPLX_COEF[K], lnT, tk) * kLogToLn;
}
template <int M>
__device__ __forceinline__ double mnrl_lnk(double lnT, double tk) {
  if constexpr (Net::ISOTHERMAL || !Net::HAS_MNRL_COEF)
    return -Net::MNRL_LOGK[M] * kLogToLn;
  else
    return -interp_logk(Net::MNRL_COEF[M], lnT, tk) * kLogToLn;
}

// RKineticMineral rate of mineral M (reaction_mineral.F90:729-908): Im per
// unit bulk volume and dIm/dQK per unit volume; active = the reference does
// not `cycle` past the mineral
struct MRate {
  double Im, dIm;
  bool active;
};
template <int M>
__device__ __forceinline__ MRate mineral_rate(double lnQK, double QK, double vf, double area,
                                              double temp) {
  MRate r{0.0, 0.0, false};
  constexpr bool temkin = Net::HAS_TEMKIN, mscale = Net::HAS_MINSCALE;
  double sig = 1.0;
  if constexpr (temkin && mscale) sig = Net::MNRL_MINSCALE[M] * Net::MNRL_TEMKIN[M];
  else if constexpr (temkin) sig = Net::MNRL_TEMKIN[M];
  else if constexpr (mscale) sig = Net::MNRL_MINSCALE[M];
  double aff;
  if constexpr (temkin || mscale) aff = 1.0 - pow(QK, 1.0 / sig);
  else aff = 1.0 - QK;
  const double sgn = copysign(1.0, aff);
  if (!(vf > 0.0 || sgn < 0.0)) return r;
  if constexpr (Net::HAS_IRREV) {
    if (Net::MNRL_IRREV[M] == 1 && sgn < 0.0) return r;
  }
  if constexpr (Net::HAS_THRESHOLD) {
    if (Net::MNRL_THRESHOLD[M] > 0.0 && sgn < 0.0 && QK < Net::MNRL_THRESHOLD[M]) return r;
  }
  constexpr double lim = Net::HAS_LIMITER ? Net::MNRL_LIMITER[M] : 0.0;
  if constexpr (lim > 0.0) aff = aff / (1.0 + (1.0 - aff) / lim);
  double kr = Net::MNRL_K[M];
  if constexpr (Net::HAS_EA) {
    if constexpr (Net::MNRL_EA[M] > 0.0)
      kr = kr * pf_exp(Net::MNRL_EA[M] / kIdealGas * (1.0 / (25.0 + 273.15) - 1.0 / (temp + 273.15)));
  }
  double Im_const = -area;
  if constexpr (mscale) Im_const = Im_const / Net::MNRL_MINSCALE[M];
  double Im;
  if constexpr (Net::HAS_AFFPOWER) Im = Im_const * sgn * pow(fabs(aff), Net::MNRL_AFFPOWER[M]) * kr;
  else Im = Im_const * sgn * fabs(aff) * kr;
  double d;
  if constexpr (Net::HAS_AFFPOWER) d = -Im * Net::MNRL_AFFPOWER[M] / fabs(aff);
  else d = -Im_const * kr;
  if constexpr (temkin || mscale) d = d * (1.0 / sig) / QK * (1.0 - aff);
  if constexpr (lim > 0.0) {
    const double den = 1.0 + (1.0 - aff) / lim;
    d = d * (1.0 + QK / lim / den) / den;
  }
  r.Im = Im;
  r.dIm = d;
  r.active = true;
  return r;
}

// One reaction step for cells [0, ncell) (or, with rescue, the cells whose
// status is kEscalate).  JT: precision of the off-diagonal Jacobian entries
// and of the factorization.
template <typename JT>
__device__ __forceinline__ void lane_step(const Args& A) {
  constexpr int NAQ = Net::NAQ, NIMM = Net::NIMM, N = NAQ + NIMM, NC = Net::NC, NM = Net::NM;
  constexpr int NA1 = NAQ > 0 ? NAQ : 1, NM1 = NM > 0 ? NM : 1;
  constexpr int NT = NAQ > 1 ? NAQ * (NAQ - 1) / 2 : 1;
  constexpr bool F32 = sizeof(JT) == 4;
  constexpr bool NEWTON_ACT = Net::ACT_FREQ == kFreqNewtonIter && NAQ > 0;
  // per-class inverse activity coefficients of the complexes (activity
  // coefficients fixed during the step)
  constexpr int NG = NEWTON_ACT ? 0 : Net::NGCLS;
  constexpr int NG1 = NG > 0 ? NG : 1;
  constexpr int kLdsFields = 2 * N + NAQ + 3 * NM + NG > 0 ? 2 * N + NAQ + 3 * NM + NG : 1;
  // pivot tolerance of the unit-diagonal factorization
  constexpr double kPivTol = F32 ? 1.0e-6 : 1.0e-13;
  const i64 ld = A.ld;
  const int lane = (int)__lane_id();
  const int nxcd = A.nxcd;
  const int xcd = (int)(blockIdx.x % (unsigned)nxcd);
  // cells of this pass: all, or a list an earlier kernel wrote (a short
  // list is left to a later pass: the grid returns at once)
  const bool listed = A.list != nullptr;
  const i64 nwork = listed ? (i64)*A.list_count : A.ncell;
  if (listed && nwork <= (i64)A.list_min) return;
  const i64 chunk = (nwork + nxcd - 1) / nxcd;

  // wave-uniform queue of claimed cells [qn, qe); XCD ranges found empty
  i64 qn = 0, qe = 0;
  int steal = 0;
  // lane state.  Per-cell values that the complex loop does not touch live
  // in this lane's own LDS slots (field-major, 256 lanes per field: a wave's
  // accesses are conflict free and no lane reads another's slots).
  __shared__ double lsm[kLdsFields * 256];
  const int tid = (int)threadIdx.x;
#define PF_GUESS(i) lsm[(0 + (i)) * 256 + tid]
#define PF_TINIT(i) lsm[(N + (i)) * 256 + tid]
#define PF_LNGAM(i) lsm[(2 * N + (i)) * 256 + tid]
#define PF_VF(m) lsm[(2 * N + NAQ + (m)) * 256 + tid]
#define PF_AREA(m) lsm[(2 * N + NAQ + NM + (m)) * 256 + tid]
#define PF_RLAST(m) lsm[(2 * N + NAQ + 2 * NM + (m)) * 256 + tid]
#define PF_GINV(q) lsm[(2 * N + NAQ + 3 * NM + (q)) * 256 + tid]
  bool exhausted = false;
  i64 cell = -1;
  int mode = 0;  // 0: no cell, 1: Newton iteration, 2: write-out evaluation
  double x[N];
  double por = 0.0, sat = 0.0, dkl = 0.0, vol = 0.0, temp = 25.0, ln_aw = 0.0, psv = 0.0;
  double lnT = 0.0, tk = 298.15;
  double target = 0.0, cum = 0.0, tdt = 0.0, r0 = 0.0;
  double sec_I = 0.0, sec_sum = 0.0, I_last = 0.0;
  int nconst = 0, ncuts = 0, nsub = 0, nit_total = 0, nit = 0, status = 0;
  int attempts = 0, last_why = 0;
  bool restore = false;    // outputs of a failed RStep: totals at RReact start
  bool kin = false;        // write-out: the last sub-step's kinetic update is due
  unsigned small = 0u;     // bit i: component i started at or below 1e-40
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = 1.0;

  for (;;) {
    // ---------------- work stealing: hand out cells ----------------
    {
      bool needy = cell < 0 && !exhausted;
      unsigned long long mask = __ballot(needy);
      while (mask) {
        if (qn >= qe) {
          if (steal >= nxcd) {  // every range is empty
            if (needy) exhausted = true;
            needy = false;
            break;
          }
          const int xr = xcd + steal < nxcd ? xcd + steal : xcd + steal - nxcd;
          const i64 lo = (i64)xr * chunk;
          const i64 hi = lo + chunk < nwork ? lo + chunk : nwork;
          const int first = __builtin_ctzll(__ballot(1));
          unsigned long long base = 0;
          if (lane == first) base = atomicAdd(A.cursors + 16 * xr, (unsigned long long)kClaim);
          base = __shfl(base, first, 64);
          if (lo + (i64)base >= hi) {
            ++steal;
            continue;
          }
          qn = lo + (i64)base;
          qe = qn + kClaim < hi ? qn + kClaim : hi;
        }
        const int rk = __popcll(mask & ((1ull << lane) - 1ull));
        const i64 avail = qe - qn;
        const int cnt = __popcll(mask);
        if (needy && rk < avail) {
          cell = listed ? (i64)A.list[qn + rk] : qn + rk;
          needy = false;
          // a rescue pass takes only the cells an earlier pass handed over
          if (A.rescue && A.status[cell] != kEscalate) {
            cell = -1;
            needy = true;
          }
        }
        qn += cnt < avail ? cnt : avail;
        mask = __ballot(needy);
      }
    }
    if (__all(cell < 0)) break;  // every lane of the wave is done
    if (cell < 0) continue;
    const i64 c = cell;

    // ---------------- load the cell; RStep set-up ----------------
    if (mode == 0) {
      por = A.porosity[c];
      sat = A.saturation[c];
      const double den = A.den_kg[c];
      vol = A.volume[c];
      temp = A.temperature ? A.temperature[c] : 25.0;
      tk = temp + 273.15;
      lnT = pf_log(tk);
      ln_aw = (Net::ACT_H2O && A.ln_act_h2o) ? A.ln_act_h2o[c] : 0.0;
      dkl = den * 1.0e-3;
      psv = por * sat * 1000.0 * vol;
      double tin[N];
      sfor<0, NAQ>([&](auto I) {
        constexpr int i = decltype(I)::value;
        x[i] = A.pri_molal[i * ld + c];
        tin[i] = A.total[i * ld + c];
        PF_LNGAM(i) = pf_log(A.pri_act_coef[i * ld + c]);
      });
      sfor<0, NIMM>([&](auto I) {
        constexpr int i = decltype(I)::value;
        x[NAQ + i] = A.immobile[i * ld + c];
        tin[NAQ + i] = x[NAQ + i];
      });
      if constexpr (NEWTON_ACT) {
        sec_I = sec_sum = 0.0;
        sfor<0, NC>([&](auto K) {
          constexpr int k = decltype(K)::value;
          const double mk = A.sec_molal[k * ld + c];
          sec_I += mk * Net::CPLX_Z[k] * Net::CPLX_Z[k];
          sec_sum += mk;
        });
      }
      if constexpr (!NEWTON_ACT) {
        // The Debye-Hueckel law gives complexes of equal charge magnitude
        // and ion size equal coefficients (and neutral ones exactly 1), so
        // each (Z^2, a0) class keeps one inverse coefficient for the whole
        // step.  A cell whose input coefficients break that structure is
        // handed to the team kernel, which reads them per complex.
        bool cons = true;
        double g0[NG1];
        sfor<0, NC>([&](auto K) {
          constexpr int k = decltype(K)::value;
          constexpr int q = Net::CPLX_GCLS[k];
          const double g = A.sec_act_coef[k * ld + c];
          if constexpr (q < 0) {
            cons = cons && g == 1.0;
          } else if constexpr (Net::CPLX_GFIRST[k]) {
            g0[q] = g;
            PF_GINV(q) = rcp_nr(g);
          } else {
            cons = cons && g == g0[q];
          }
        });
        if (!cons && A.escalate) {
          A.status[c] = kEscalate;
          cell = -1;
          continue;
        }
      }
      sfor<0, NM>([&](auto M) {
        constexpr int m = decltype(M)::value;
        PF_VF(m) = A.mnrl_volfrac[m * ld + c];
        PF_AREA(m) = A.mnrl_area[m * ld + c];
        PF_RLAST(m) = 0.0;
      });
      if constexpr (!Net::FULL_GEOCHEM) {
        // RStep early exit (reaction.F90:3601-3605)
        sfor<0, NAQ>([&](auto I) {
          constexpr int i = decltype(I)::value;
          A.pri_molal[i * ld + c] = tin[i] / den * 1.0e3;
        });
        if (A.num_iterations) A.num_iterations[c] = 0;
        if (A.num_sub_steps) A.num_sub_steps[c] = 0;
        if (A.status) A.status[c] = 0;
        cell = -1;
        continue;
      }
      // clamp small totals (reaction.F90:3617-3631)
      small = 0u;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        if (tin[i] <= 1.0e-40) {
          small |= 1u << i;
          tin[i] = 1.0e-40;
        }
        PF_TINIT(i) = tin[i];
        PF_GUESS(i) = (Net::TOTAL_AS_GUESS && i < NAQ) ? tin[i] : x[i];
      }
      target = A.dt;
      cum = 0.0;
      tdt = A.dt;
      nconst = ncuts = nsub = nit_total = status = attempts = last_why = 0;
      restore = false;
      kin = false;
      nit = -1;
      mode = 1;
    }

    // ---------------- RReact set-up (reaction.F90:3800-3851) ----------------
    if (mode == 1 && nit < 0) {
      if (++attempts > Net::MAX_CALLS) {
        status = last_why ? last_why : kMaxIter;
        mode = 2;
      } else {
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = PF_GUESS(i);
        nit = 0;
        r0 = 0.0;
      }
    }
    const bool iterate = mode == 1;
    if (iterate) ++nit;

    // ---------------- RTotal + RReaction at x ----------------
    const bool wet = !(sat < kRtMinSat);
    const double itdt = rcp_nr(tdt);
    const double psvd = psv * itdt;
    const double jc = psvd * dkl;
    const double wm = wet ? 1.0 : 0.0;
    // activity coefficients per Newton iteration (LAG at the current pri
    // and last secondary molalities); the write-out evaluation reuses the
    // ionic strength of the last iteration
    double I = 0.0, sI = 0.0;
    if constexpr (NEWTON_ACT) {
      if (iterate) {
        double Ip = 0.0, sp = 0.0;
        sfor<0, NAQ>([&](auto Q) {
          constexpr int i = decltype(Q)::value;
          Ip += x[i] * Net::ZP[i] * Net::ZP[i];
          if (i != Net::H2O_ID) sp += x[i];
        });
        I = 0.5 * (Ip + sec_I);
        sI = sqrt(I);
        sfor<0, NAQ>([&](auto Q) {
          constexpr int i = decltype(Q)::value;
          if constexpr (Net::ZP[i] > 1.0e-10 || Net::ZP[i] < -1.0e-10)
            PF_LNGAM(i) = pf_log(dh_gamma(Net::ZP[i], Net::A0P[i], sI, I));
          else
            PF_LNGAM(i) = 0.0;
        });
        if constexpr (Net::ACT_H2O) {
          const double a = 1.0 - 0.017 * (sp + sec_sum);
          ln_aw = a > 0.0 ? pf_log(a) : 0.0;
        }
        I_last = I;
      } else {
        I = I_last;
        sI = sqrt(I);
      }
    }
    double lnact[NA1], sig[NA1], tot[NA1], D[NA1];
    JT S[NT];
    sfor<0, NAQ>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      lnact[i] = pf_log(x[i]) + PF_LNGAM(i);
      // FP32 entries are accumulated pre-scaled by sig_i sig_j (any scale
      // works: it is divided out exactly below); FP64 ones are not scaled
      sig[i] = F32 ? __builtin_amdgcn_rsq(x[i]) : 1.0;
      tot[i] = 0.0;
      D[i] = wet ? x[i] : 1.0;
    });
#pragma unroll
    for (int q = 0; q < NT; ++q) S[q] = (JT)0;
    const bool wout = mode == 2;
    double nsI = 0.0, nss = 0.0;
    double gi[NG1];
    sfor<0, NG>([&](auto Q) {
      constexpr int q = decltype(Q)::value;
      gi[q] = PF_GINV(q);
    });
    // complexes (RTotalAqueous, reaction.F90:4665-4759)
    sfor<0, NC>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr int ns = Net::CPLX_N[k];
      double lnQK = cplx_lnk<k>(lnT, tk);
      if constexpr (Net::CPLX_H2O[k] != 0.0) lnQK += Net::CPLX_H2O[k] * ln_aw;
      sfor<0, ns>([&](auto Q) {
        constexpr int a = decltype(Q)::value;
        lnQK += Net::CPLX_NU[k][a] * lnact[Net::CPLX_ID[k][a]];
      });
      double m;
      double g = 1.0;
      if constexpr (NEWTON_ACT) {
        if constexpr (Net::CPLX_Z[k] > 1.0e-10 || Net::CPLX_Z[k] < -1.0e-10)
          g = dh_gamma(Net::CPLX_Z[k], Net::CPLX_A0[k], sI, I);
        m = pf_exp(lnQK) / g;
      } else if constexpr (Net::CPLX_GCLS[k] >= 0) {
        m = pf_exp(lnQK) * gi[Net::CPLX_GCLS[k]];
      } else {
        m = pf_exp(lnQK);
      }
      if (wout) {
        A.sec_molal[k * ld + c] = m;
        if constexpr (NEWTON_ACT) A.sec_act_coef[k * ld + c] = g;
      }
      if constexpr (NEWTON_ACT) {
        nsI += m * (Net::CPLX_Z[k] * Net::CPLX_Z[k]);
        nss += m;
      }
      const double mj = m * wm;
      sfor<0, ns>([&](auto Q) {
        constexpr int a = decltype(Q)::value;
        constexpr int ia = Net::CPLX_ID[k][a];
        constexpr double na = Net::CPLX_NU[k][a];
        tot[ia] += na * m;
        D[ia] += (na * na) * mj;
        const double pa = mj * sig[ia];
        sfor<a + 1, ns>([&](auto Q2) {
          constexpr int b = decltype(Q2)::value;
          constexpr int ib = Net::CPLX_ID[k][b];
          constexpr double nab = na * Net::CPLX_NU[k][b];
          if constexpr (ia == ib) {
            D[ia] += (2.0 * nab) * mj;
          } else {
            S[tri(ia, ib)] += (JT)nab * (JT)(pa * sig[ib]);
          }
        });
      });
    });
    // residual (RTAccumulation, reaction.F90:3899-3916) and its norms
    double res[N], totc[NA1];
    double mx = 0.0, nr2 = 0.0;
    sfor<0, NAQ>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      totc[i] = (x[i] + tot[i]) * dkl;
      res[i] = wet ? (psv * totc[i] - psv * PF_TINIT(i)) * itdt : 0.0;
    });
    sfor<0, NIMM>([&](auto Q) {
      constexpr int i = NAQ + decltype(Q)::value;
      double r = wet ? (x[i] * vol - PF_TINIT(i) * vol) * itdt : 0.0;
      sfor<0, Net::NDECAY>([&](auto Dq) {
        constexpr int d = decltype(Dq)::value;
        if constexpr (Net::DECAY_ID[d] == i - NAQ) {
          if (wet) r += Net::DECAY_K[d] * vol * x[i];
        }
      });
      res[i] = r;
    });
    // kinetic minerals (RKineticMineral, after the accumulation terms as in
    // RReaction): rates, residual and Jacobian terms
    double rt[NM1];
    const double ijc = wet ? rcp_nr(jc) : 0.0;
    sfor<0, NM>([&](auto Mi) {
      constexpr int mi = decltype(Mi)::value;
      constexpr int ns = Net::MNRL_N[mi];
      double lnQK = mnrl_lnk<mi>(lnT, tk);
      if constexpr (Net::MNRL_H2O[mi] != 0.0) lnQK += Net::MNRL_H2O[mi] * ln_aw;
      sfor<0, ns>([&](auto Q) {
        constexpr int a = decltype(Q)::value;
        lnQK += Net::MNRL_NU[mi][a] * lnact[Net::MNRL_ID[mi][a]];
      });
      rt[mi] = 0.0;
      const double ar = PF_AREA(mi);
      if (ar != 0.0) {  // zero area gives Im = 0
        const double QK = pf_exp(lnQK);
        const MRate r = mineral_rate<mi>(lnQK, QK, PF_VF(mi), ar, temp);
        if (r.active) {
          rt[mi] = r.Im;
          // dry cells: RReaction adds nothing (reaction.F90:4083)
          const double Rm = r.Im * vol * wm;
          // d res / d ln m_j = nu_i (dIm/dQK QK V rho/1000) nu_j, in units of jc
          const double gm = r.dIm * vol * QK * dkl * ijc;
          sfor<0, ns>([&](auto Q) {
            constexpr int a = decltype(Q)::value;
            constexpr int ia = Net::MNRL_ID[mi][a];
            constexpr double na = Net::MNRL_NU[mi][a];
            res[ia] += na * Rm;
            D[ia] += (na * na) * gm;
            const double pa = gm * sig[ia];
            sfor<a + 1, ns>([&](auto Q2) {
              constexpr int b = decltype(Q2)::value;
              constexpr int ib = Net::MNRL_ID[mi][b];
              constexpr double nab = na * Net::MNRL_NU[mi][b];
              if constexpr (ia == ib) {
                D[ia] += (2.0 * nab) * gm;
              } else {
                S[tri(ia, ib)] += (JT)nab * (JT)(pa * sig[ib]);
              }
            });
          });
        }
      }
    });
#pragma unroll
    for (int i = 0; i < N; ++i) {
      mx = fmax(mx, fabs(res[i]));
      nr2 += res[i] * res[i];
    }

    if (wout) {
      // ---------------- write the cell back ----------------
      const bool ok = status == 0;
      if (kin) {
        // RUpdateKineticState of the last sub-step (reaction_mineral.F90:
        // 1456-1525): rates at the final state, volume fractions of the
        // sub-step's start
        sfor<0, NM>([&](auto M) {
          constexpr int m = decltype(M)::value;
          double v = PF_VF(m) + rt[m] * Net::MNRL_VM[m] * tdt;
          if (v < 0.0) v = 0.0;
          PF_VF(m) = v;
          PF_RLAST(m) = rt[m];
        });
      }
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        A.pri_molal[i * ld + c] = x[i];
        const bool keep = ok && ((small >> i) & 1u);  // original value restored
        if (!keep) A.total[i * ld + c] = (!ok && restore) ? PF_TINIT(i) : totc[i];
        if constexpr (NEWTON_ACT) A.pri_act_coef[i * ld + c] = pf_exp(PF_LNGAM(i));
      });
      sfor<0, NIMM>([&](auto Q) {
        constexpr int i = NAQ + decltype(Q)::value;
        const bool keep = ok && ((small >> i) & 1u);
        if (!keep) A.immobile[(i - NAQ) * ld + c] = (!ok && restore) ? PF_TINIT(i) : x[i];
      });
      sfor<0, NM>([&](auto M) {
        constexpr int m = decltype(M)::value;
        A.mnrl_volfrac[m * ld + c] = PF_VF(m);
        if (A.mnrl_rate) A.mnrl_rate[m * ld + c] = PF_RLAST(m);
      });
      if constexpr (NEWTON_ACT && Net::ACT_H2O) {
        if (A.ln_act_h2o) A.ln_act_h2o[c] = ln_aw;
      }
      if (A.num_iterations) A.num_iterations[c] = nit_total;
      if (A.num_sub_steps) A.num_sub_steps[c] = nsub;
      if (A.status) A.status[c] = status;
      cell = -1;
      mode = 0;
      continue;
    }

    if constexpr (NEWTON_ACT) {
      sec_I = nsI;
      sec_sum = nss;
    }
    sfor<0, NM>([&](auto M) {
      constexpr int m = decltype(M)::value;
      PF_RLAST(m) = rt[m];
    });
    const double nr = sqrt(nr2);
    if (nit == 1) r0 = nr;
    // 0 continue, 1 converged, 2 max iterations, 3 hand over
    int outcome = 0;
    // FP32: when the convergence test first passes on the last sub-step, one
    // more FP32-preconditioned correction of the FP64 residual is applied
    // before the cell exits (the write-out evaluates at the corrected
    // state): the exit test bounds the residual, not the error, and the
    // linear convergence of the FP32-preconditioned iteration would leave
    // up to about 1e-9 of relative error at the exit
    bool polish = false;
    if (nit > Net::MAX_IT) {
      outcome = 2;
    } else if (mx < Net::TOL_ABS || nr / r0 < Net::TOL_REL) {
      if (F32 && cum + tdt >= target) polish = true;
      else outcome = 1;
    }
    if (outcome == 0) {
      // ---------------- linear solve (RSolve) ----------------
      // aqueous block: A = D^-1/2 M D^-1/2, L D L^T, A y = D^-1/2 res / jc
      double s[NA1];
      JT rr[NA1], b[NA1], dg0[NA1];
      const double ijs = wet ? ijc : 1.0;
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        // symmetric scaling to (about) unit diagonal: any positive s gives
        // the same step, so the hardware estimate of 1/sqrt(D_i) is used
        s[i] = __builtin_amdgcn_rsq(D[i]);
        rr[i] = F32 ? (JT)(s[i] * rcp_nr(sig[i])) : (JT)s[i];  // ~sqrt(x_i / D_i)
        dg0[i] = (JT)(D[i] * s[i] * s[i]);
        b[i] = (JT)(res[i] * s[i] * ijs);
      });
      JT a[NT], dinv[NA1];
#pragma unroll
      for (int q = 0; q < NT; ++q) a[q] = (JT)0;
      sfor<1, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        sfor<0, i>([&](auto Q2) {
          constexpr int j = decltype(Q2)::value;
          a[tri(i, j)] = S[tri(i, j)] * rr[i] * rr[j];
        });
      });
      bool bad = false;
      // right-looking L D L^T on the unit-diagonal matrix
      JT dg[NA1];
#pragma unroll
      for (int i = 0; i < NA1; ++i) dg[i] = dg0[i];
      sfor<0, NAQ>([&](auto P) {
        constexpr int p = decltype(P)::value;
        const JT dp = dg[p];
        if (!((double)dp > kPivTol)) bad = true;
        const JT inv = (JT)1 / dp;
        dinv[p] = inv;
        JT l[NA1];
        sfor<p + 1, NAQ>([&](auto Q) {
          constexpr int i = decltype(Q)::value;
          l[i] = a[tri(i, p)] * inv;
        });
        sfor<p + 1, NAQ>([&](auto Q) {
          constexpr int i = decltype(Q)::value;
          dg[i] -= l[i] * a[tri(i, p)];
          sfor<p + 1, i>([&](auto Q2) {
            constexpr int j = decltype(Q2)::value;
            a[tri(i, j)] -= l[i] * a[tri(j, p)];
          });
        });
        sfor<p + 1, NAQ>([&](auto Q) {
          constexpr int i = decltype(Q)::value;
          a[tri(i, p)] = l[i];
        });
      });
      // forward (L z = b), diagonal, backward (L^T y = z)
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        sfor<0, i>([&](auto Q2) {
          constexpr int j = decltype(Q2)::value;
          b[i] -= a[tri(i, j)] * b[j];
        });
      });
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        b[i] *= dinv[i];
      });
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = NAQ - 1 - decltype(Q)::value;
        sfor<i + 1, NAQ>([&](auto Q2) {
          constexpr int j = decltype(Q2)::value;
          b[i] -= a[tri(j, i)] * b[j];
        });
      });
      double delta[N];
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        delta[i] = (double)b[i] * s[i];
        if (!(delta[i] == delta[i])) bad = true;
      });
      // immobile rows: accumulation + decay on the diagonal
      sfor<0, NIMM>([&](auto Q) {
        constexpr int i = NAQ + decltype(Q)::value;
        double d = vol * itdt * x[i];
        sfor<0, Net::NDECAY>([&](auto Dq) {
          constexpr int dd = decltype(Dq)::value;
          if constexpr (Net::DECAY_ID[dd] == i - NAQ) d += Net::DECAY_K[dd] * vol * x[i];
        });
        if (!wet) d = 1.0;
        delta[i] = res[i] * rcp_nr(d);
        if (!(d > 0.0)) bad = true;
      });
      if (!Net::LOG_FORM) {
#pragma unroll
        for (int i = 0; i < N; ++i) delta[i] *= x[i];
      }
      if (bad) {
        outcome = 3;
      } else {
        // ---------------- update (reaction.F90:3977-4049) ----------------
        double latest[N];
        if constexpr (Net::LOG_FORM) {
#pragma unroll
          for (int i = 0; i < N; ++i) {
            const double u = copysign(1.0, delta[i]) * fmin(fabs(delta[i]), Net::MAX_DLNC);
            latest[i] = x[i] * pf_exp(-u);
          }
        } else {
          double ratio = 1.0e308;
#pragma unroll
          for (int i = 0; i < N; ++i)
            if (x[i] <= delta[i]) ratio = fmin(ratio, fabs(x[i] / delta[i]));
#pragma unroll
          for (int i = 0; i < N; ++i) latest[i] = x[i] - (ratio < 1.0 ? delta[i] * ratio * 0.99 : delta[i]);
        }
        double rc = -1.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double q = fabs((latest[i] - x[i]) * rcp_nr(x[i]));
          if (q == q) rc = fmax(rc, q);
        }
        // The small-step exit holds for the FP32 factorization too: the
        // scaled matrix has unit diagonal and |A_ij| <= 1 (positive
        // semi-definite terms), so |A| <= n and a step below the tolerance
        // certifies a scaled FP64 residual |b| <= n |y| of the same size.
        if (polish) {
#pragma unroll
          for (int i = 0; i < N; ++i) x[i] = latest[i];
          outcome = 1;
        } else if (rc >= 0.0 && rc < Net::TOL_CHANGE) {
          outcome = 1;
        } else {
#pragma unroll
          for (int i = 0; i < N; ++i) x[i] = latest[i];
        }
      }
    }

    // ---------------- RReact / RStep transitions ----------------
    if (outcome == 1) {
      // RUpdateKineticState (reaction_mineral.F90:1456-1525); the last
      // sub-step's is done by the write-out evaluation at the final state
      const bool last = cum + tdt >= target;
      if (!last) {
        sfor<0, NM>([&](auto M) {
          constexpr int m = decltype(M)::value;
          double v = PF_VF(m) + rt[m] * Net::MNRL_VM[m] * tdt;
          if (v < 0.0) v = 0.0;
          PF_VF(m) = v;
        });
      }
      nit_total += nit;
      cum += tdt;
      ++nsub;
      ++nconst;
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        PF_TINIT(i) = totc[i];
      });
#pragma unroll
      for (int i = NAQ; i < N; ++i) PF_TINIT(i) = x[i];
#pragma unroll
      for (int i = 0; i < N; ++i) PF_GUESS(i) = x[i];
      nit = -1;
      if (cum >= target) {
        status = 0;
        kin = true;
        mode = 2;  // tdt stays the last sub-step's (the kinetic update uses it)
      } else if (nconst >= 4) {
        --ncuts;
        tdt = fmin(2.0 * tdt, target - cum);
      }
    } else if (outcome == 3 || (outcome == 2 && F32 && A.escalate)) {
      // hand the untouched cell to the next pass
      if (A.escalate) {
        A.status[c] = kEscalate;
        cell = -1;
        mode = 0;
        continue;
      }
      // no later pass: a singular system (the reference's zero-row failure)
      nit_total += nit;
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        PF_TINIT(i) = totc[i];
      });
#pragma unroll
      for (int i = NAQ; i < N; ++i) PF_TINIT(i) = x[i];
      restore = false;
      last_why = kSingular;
      ++ncuts;
      if (ncuts > Net::MAX_CUTS) {
        status = last_why;
        mode = 2;
      } else {
        tdt *= 0.5;
        nconst = 0;
        nit = -1;
      }
    } else if (outcome == 2) {
      // maximum iterations: totals / immobile restored (reaction.F90:3869-3895)
      nit_total += nit;
      restore = true;
      last_why = kMaxIter;
      ++ncuts;
      if (ncuts > Net::MAX_CUTS) {
        status = last_why;
        mode = 2;
      } else {
        tdt *= 0.5;
        nconst = 0;
        nit = -1;
      }
    }
  }
#undef PF_GUESS
#undef PF_TINIT
#undef PF_LNGAM
#undef PF_VF
#undef PF_AREA
#undef PF_RLAST
#undef PF_GINV
}

}  // namespace pflane

// one module holds either the step kernels or (PFLANE_ADJOINT defined) the
// adjoint kernel, so a context that never asks for sensitivities does not
// compile it
#ifndef PFLANE_ADJOINT
extern "C" __global__ void __launch_bounds__(256) pfchem_lane_step_f32(pflane::Args a) {
  pflane::lane_step<float>(a);
}
extern "C" __global__ void __launch_bounds__(256) pfchem_lane_step_f64(pflane::Args a) {
  pflane::lane_step<double>(a);
}
#endif

// ---------------------------------------------------------------------------
// Adjoint of one reaction step, one cell per lane, FP64.
//
// The step is RStep's accepted sub-steps s of size h_s, each the converged
// RReact system (reaction.F90:3742-4055)
//   F(m; T0, VF0, p) = psv (T(m; p) - T0) / h + vol sum_q nu_q Im_q(m; p) = 0
// followed by VF = max(0, VF0 + Im V_m h) (reaction_mineral.F90:1456-1525);
// the next sub-step starts from T(m), VF.  Objective per cell:
// f = sum_q wv_q VF_q(end) + sum_k ws_k m_k(end); parameters log10 k_q
// (rate constants) and log10 K_k (complexes).  The forward pass repeats the
// step (the FP64 solve of pfchem_lane_step_f64) and records each accepted
// sub-step (m, VF0, h) in a per-cell log; the reverse pass applies the
// implicit-function theorem at each recorded solution, one solve per
// sub-step.  In the log form J diag(m) = jc M with M symmetric
// (M = diag(m) + S^T diag(m_k) S + sum_q g_q nu_q nu_q^T), so the transposed
// solve J^T lambda = b is M lambda = diag(m) b / jc with the same L D L^T.
// The mineral block of M carries rho/1000 in the Newton iteration (the
// reference's dQK_dmj, reaction_mineral.F90:945) and in the sensitivities
// only when asked (jacobian = 0); otherwise it is the exact derivative.
// Scope: aqueous networks without immobile species, TST rates with optional
// Arrhenius factor, activity coefficients fixed during the step.  A cell the
// pass cannot take (a pivot below tolerance, activity coefficients that do
// not follow the charge/size classes) is left with status kEscalate for the
// general kernel.
namespace pflane {

constexpr int kSensSub = 8;  // sub-steps one cell may record

struct SensArgs {
  i64 ld, ncell;
  const double* pri_molal;
  const double* total;
  const double* pri_act_coef;
  const double* sec_act_coef;
  const double* ln_act_h2o;
  const double* mnrl_volfrac;
  const double* mnrl_area;
  const double* porosity;
  const double* saturation;
  const double* den_kg;
  const double* volume;
  const double* temperature;
  const double* wv;    // [NM] weights of the end-of-step volume fractions
  const double* ws;    // [NC] weights of the end-of-step secondary molalities
  double* grad;        // [(NM + NC) * ldo] per-cell gradient
  i64 ldo;
  double* fcell;       // [ncell] objective per cell, or null
  int* status;         // [ncell]
  double* log;         // [kSensSub * (NAQ + NM + 1) * ldl] sub-step log
  i64 ldl;
  double dt;
  int jacobian;        // 1: exact dF/dm; 0: the reference's mineral block
  int pad;
};
static_assert(sizeof(SensArgs) == 192, "SensArgs layout");

// L D L^T solve of the symmetrically scaled M (diagonal dg, strict lower
// packed a; both overwritten by the factors): M y = b in place; false if a
// pivot is below tolerance
template <int NA>
__device__ __forceinline__ bool sens_ldlt(double (&dg)[NA], double (&a)[NA > 1 ? NA * (NA - 1) / 2 : 1],
                                          double (&b)[NA]) {
  double s[NA], dinv[NA];
  sfor<0, NA>([&](auto Q) {
    constexpr int i = decltype(Q)::value;
    s[i] = __builtin_amdgcn_rsq(dg[i]);
    dg[i] = dg[i] * s[i] * s[i];
    b[i] *= s[i];
  });
  sfor<1, NA>([&](auto Q) {
    constexpr int i = decltype(Q)::value;
    sfor<0, i>([&](auto Q2) {
      constexpr int j = decltype(Q2)::value;
      a[tri(i, j)] *= s[i] * s[j];
    });
  });
  bool bad = false;
  sfor<0, NA>([&](auto P) {
    constexpr int p = decltype(P)::value;
    const double dp = dg[p];
    if (!(dp > 1.0e-13)) bad = true;
    const double inv = 1.0 / dp;
    dinv[p] = inv;
    double l[NA];
    sfor<p + 1, NA>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      l[i] = a[tri(i, p)] * inv;
    });
    sfor<p + 1, NA>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      dg[i] -= l[i] * a[tri(i, p)];
      sfor<p + 1, i>([&](auto Q2) {
        constexpr int j = decltype(Q2)::value;
        a[tri(i, j)] -= l[i] * a[tri(j, p)];
      });
    });
    sfor<p + 1, NA>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      a[tri(i, p)] = l[i];
    });
  });
  sfor<0, NA>([&](auto Q) {
    constexpr int i = decltype(Q)::value;
    sfor<0, i>([&](auto Q2) {
      constexpr int j = decltype(Q2)::value;
      b[i] -= a[tri(i, j)] * b[j];
    });
  });
  sfor<0, NA>([&](auto Q) {
    constexpr int i = decltype(Q)::value;
    b[i] *= dinv[i];
  });
  sfor<0, NA>([&](auto Q) {
    constexpr int i = NA - 1 - decltype(Q)::value;
    sfor<i + 1, NA>([&](auto Q2) {
      constexpr int j = decltype(Q2)::value;
      b[i] -= a[tri(j, i)] * b[j];
    });
  });
  sfor<0, NA>([&](auto Q) {
    constexpr int i = decltype(Q)::value;
    b[i] *= s[i];
    if (!(b[i] == b[i])) bad = true;
  });
  return !bad;
}

__device__ inline void lane_adjoint(const SensArgs& A) {
  constexpr int NAQ = Net::NAQ, NC = Net::NC, NM = Net::NM;
  constexpr int NA1 = NAQ > 0 ? NAQ : 1, NM1 = NM > 0 ? NM : 1;
  constexpr int NT = NA1 > 1 ? NA1 * (NA1 - 1) / 2 : 1;
  constexpr int NG = Net::NGCLS, NG1 = NG > 0 ? NG : 1;
  constexpr int NF = NAQ + NM + 1;  // log fields per sub-step
  const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.ncell) return;
  const i64 ld = A.ld, ldo = A.ldo, ldl = A.ldl;
  if constexpr (Net::NIMM > 0 || Net::ACT_FREQ == kFreqNewtonIter || !Net::FULL_GEOCHEM) {
    A.status[c] = kEscalate;  // outside this pass's scope
    return;
  } else {
    // ---------------- cell constants ----------------
    const double por = A.porosity[c], sat = A.saturation[c];
    const double den = A.den_kg[c], vol = A.volume[c];
    const double temp = A.temperature ? A.temperature[c] : 25.0;
    const double tk = temp + 273.15, lnT = pf_log(tk);
    const double ln_aw = (Net::ACT_H2O && A.ln_act_h2o) ? A.ln_act_h2o[c] : 0.0;
    const double dkl = den * 1.0e-3, psv = por * sat * 1000.0 * vol;
    auto zero_grad = [&]() {
      sfor<0, NM + NC>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        A.grad[q * ldo + c] = 0.0;
      });
    };
    if (sat < kRtMinSat) {  // dry: no sensitivity
      zero_grad();
      A.status[c] = 4;  // PFCHEM_CELL_SENS_UNSUPPORTED
      return;
    }
    double lng[NA1], gi[NG1], area[NM1], vf[NM1];
    sfor<0, NAQ>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      lng[i] = pf_log(A.pri_act_coef[i * ld + c]);
    });
    bool cons = true;
    {
      double g0[NG1];
      sfor<0, NC>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int q = Net::CPLX_GCLS[k];
        const double g = A.sec_act_coef[k * ld + c];
        if constexpr (q < 0) {
          cons = cons && g == 1.0;
        } else if constexpr (Net::CPLX_GFIRST[k]) {
          g0[q] = g;
          gi[q] = 1.0 / g;
        } else {
          cons = cons && g == g0[q];
        }
      });
    }
    if (!cons) {
      A.status[c] = kEscalate;
      return;
    }
    sfor<0, NM>([&](auto M) {
      constexpr int m = decltype(M)::value;
      vf[m] = A.mnrl_volfrac[m * ld + c];
      area[m] = A.mnrl_area[m * ld + c];
    });
    // secondary molality of complex K at ln activities lna
    auto cplx_m = [&](auto K, const double(&lna)[NA1]) {
      constexpr int k = decltype(K)::value;
      double lnQK = cplx_lnk<k>(lnT, tk);
      if constexpr (Net::CPLX_H2O[k] != 0.0) lnQK += Net::CPLX_H2O[k] * ln_aw;
      sfor<0, Net::CPLX_N[k]>([&](auto Q) {
        constexpr int a = decltype(Q)::value;
        lnQK += Net::CPLX_NU[k][a] * lna[Net::CPLX_ID[k][a]];
      });
      double m = pf_exp(lnQK);
      if constexpr (Net::CPLX_GCLS[k] >= 0) m *= gi[Net::CPLX_GCLS[k]];
      return m;
    };
    // rate of mineral M at lna with volume fraction v: Im, dIm/dlnQK
    auto mnrl_r = [&](auto Mq, const double(&lna)[NA1], double v, double& Im, double& dI) {
      constexpr int mi = decltype(Mq)::value;
      double lnQK = mnrl_lnk<mi>(lnT, tk);
      if constexpr (Net::MNRL_H2O[mi] != 0.0) lnQK += Net::MNRL_H2O[mi] * ln_aw;
      sfor<0, Net::MNRL_N[mi]>([&](auto Q) {
        constexpr int a = decltype(Q)::value;
        lnQK += Net::MNRL_NU[mi][a] * lna[Net::MNRL_ID[mi][a]];
      });
      Im = 0.0;
      dI = 0.0;
      if (area[mi] == 0.0) return false;
      const double QK = pf_exp(lnQK);
      const MRate r = mineral_rate<mi>(lnQK, QK, v, area[mi], temp);
      if (!r.active) return false;
      Im = r.Im;
      dI = r.dIm * QK;
      return true;
    };
    // the system at molalities x: totals (mol/L), residual, and M's diagonal
    // D / strict lower S (in units of jc) with the mineral block times mf
    auto system = [&](const double(&x)[NA1], const double(&T0)[NA1], const double(&v0)[NM1],
                      double h, double mf, double(&T)[NA1], double(&res)[NA1], double(&D)[NA1],
                      double(&S)[NT]) {
      double lna[NA1], tot[NA1];
      const double jc = psv * dkl / h;
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        lna[i] = pf_log(x[i]) + lng[i];
        tot[i] = 0.0;
        D[i] = x[i];
      });
#pragma unroll
      for (int q = 0; q < NT; ++q) S[q] = 0.0;
      sfor<0, NC>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int ns = Net::CPLX_N[k];
        const double m = cplx_m(K, lna);
        sfor<0, ns>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          constexpr int ia = Net::CPLX_ID[k][a];
          constexpr double na = Net::CPLX_NU[k][a];
          tot[ia] += na * m;
          D[ia] += (na * na) * m;
          sfor<a + 1, ns>([&](auto Q2) {
            constexpr int b = decltype(Q2)::value;
            constexpr int ib = Net::CPLX_ID[k][b];
            constexpr double nab = na * Net::CPLX_NU[k][b];
            if constexpr (ia == ib) D[ia] += (2.0 * nab) * m;
            else S[tri(ia, ib)] += nab * m;
          });
        });
      });
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        T[i] = (x[i] + tot[i]) * dkl;
        res[i] = (psv * T[i] - psv * T0[i]) / h;
      });
      sfor<0, NM>([&](auto Mq) {
        constexpr int mi = decltype(Mq)::value;
        constexpr int ns = Net::MNRL_N[mi];
        double Im, dI;
        if (!mnrl_r(Mq, lna, v0[mi], Im, dI)) return;
        const double g = dI * vol * mf / jc;
        sfor<0, ns>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          constexpr int ia = Net::MNRL_ID[mi][a];
          constexpr double na = Net::MNRL_NU[mi][a];
          res[ia] += na * Im * vol;
          D[ia] += (na * na) * g;
          sfor<a + 1, ns>([&](auto Q2) {
            constexpr int b = decltype(Q2)::value;
            constexpr int ib = Net::MNRL_ID[mi][b];
            constexpr double nab = na * Net::MNRL_NU[mi][b];
            if constexpr (ia == ib) D[ia] += (2.0 * nab) * g;
            else S[tri(ia, ib)] += nab * g;
          });
        });
      });
    };

    // ---------------- forward: RStep with the sub-step log ----------------
    double x[NA1], T0[NA1], guess[NA1], T[NA1];
    sfor<0, NAQ>([&](auto Q) {
      constexpr int i = decltype(Q)::value;
      x[i] = A.pri_molal[i * ld + c];
      T0[i] = A.total[i * ld + c];
      if (T0[i] <= 1.0e-40) T0[i] = 1.0e-40;  // reaction.F90:3617-3631
      guess[i] = Net::TOTAL_AS_GUESS ? T0[i] : x[i];
    });
    int nlog = 0, st = 0, nconst = 0, ncuts = 0, attempts = 0;
    double cum = 0.0, tdt = A.dt;
    bool handover = false;
    while (cum < A.dt) {
      if (++attempts > Net::MAX_CALLS) {
        st = kMaxIter;
        break;
      }
      double r0 = 0.0;
      int why = -1;
#pragma unroll
      for (int i = 0; i < NAQ; ++i) x[i] = guess[i];
      for (int it = 1; why < 0; ++it) {
        double res[NA1], D[NA1], S[NT];
        system(x, T0, vf, tdt, dkl, T, res, D, S);
        double mx = 0.0, nr2 = 0.0;
#pragma unroll
        for (int i = 0; i < NAQ; ++i) {
          mx = fmax(mx, fabs(res[i]));
          nr2 += res[i] * res[i];
        }
        const double nr = sqrt(nr2);
        if (it == 1) r0 = nr;
        if (it > Net::MAX_IT) {
          why = kMaxIter;
          break;
        }
        if (mx < Net::TOL_ABS || nr / r0 < Net::TOL_REL) {
          why = kOk;
          break;
        }
        const double jc = psv * dkl / tdt;
        double delta[NA1];
#pragma unroll
        for (int i = 0; i < NAQ; ++i) delta[i] = res[i] / jc;
        if (!sens_ldlt<NA1>(D, S, delta)) {
          handover = true;
          break;
        }
        if (!Net::LOG_FORM) {
#pragma unroll
          for (int i = 0; i < NAQ; ++i) delta[i] *= x[i];
        }
        double latest[NA1];
        if constexpr (Net::LOG_FORM) {
#pragma unroll
          for (int i = 0; i < NAQ; ++i) {
            const double u = copysign(1.0, delta[i]) * fmin(fabs(delta[i]), Net::MAX_DLNC);
            latest[i] = x[i] * pf_exp(-u);
          }
        } else {
          double ratio = 1.0e308;
#pragma unroll
          for (int i = 0; i < NAQ; ++i)
            if (x[i] <= delta[i]) ratio = fmin(ratio, fabs(x[i] / delta[i]));
#pragma unroll
          for (int i = 0; i < NAQ; ++i)
            latest[i] = x[i] - (ratio < 1.0 ? delta[i] * ratio * 0.99 : delta[i]);
        }
        double rc = -1.0;
#pragma unroll
        for (int i = 0; i < NAQ; ++i) {
          const double q = fabs((latest[i] - x[i]) / x[i]);
          if (q == q) rc = fmax(rc, q);
        }
        if (rc >= 0.0 && rc < Net::TOL_CHANGE) {
          why = kOk;
          break;
        }
#pragma unroll
        for (int i = 0; i < NAQ; ++i) x[i] = latest[i];
      }
      if (handover) break;
      if (why != kOk) {
        // maximum iterations: the totals stay those of the sub-step's start
        if (++ncuts > Net::MAX_CUTS) {
          st = kMaxIter;
          break;
        }
        tdt *= 0.5;
        nconst = 0;
        continue;
      }
      if (nlog == kSensSub) {
        st = 5;  // PFCHEM_CELL_TOO_MANY_SUBSTEPS
        break;
      }
      double* lg = A.log + (i64)nlog * NF * ldl + c;
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        lg[i * ldl] = x[i];
      });
      // kinetic update with the rates at the sub-step's solution
      {
        double lna[NA1];
        sfor<0, NAQ>([&](auto Q) {
          constexpr int i = decltype(Q)::value;
          lna[i] = pf_log(x[i]) + lng[i];
        });
        sfor<0, NM>([&](auto Mq) {
          constexpr int mi = decltype(Mq)::value;
          lg[(NAQ + mi) * ldl] = vf[mi];
          double Im, dI;
          mnrl_r(Mq, lna, vf[mi], Im, dI);
          double v = vf[mi] + Im * Net::MNRL_VM[mi] * tdt;
          vf[mi] = v < 0.0 ? 0.0 : v;
        });
      }
      lg[(NAQ + NM) * ldl] = tdt;
      ++nlog;
#pragma unroll
      for (int i = 0; i < NAQ; ++i) {
        T0[i] = T[i];
        guess[i] = x[i];
      }
      cum += tdt;
      if (++nconst >= 4) {
        --ncuts;
        tdt = fmin(2.0 * tdt, A.dt - cum);
      }
    }
    if (handover) {
      A.status[c] = kEscalate;
      return;
    }
    if (st) {
      zero_grad();
      A.status[c] = st;
      return;
    }

    // ---------------- objective and its explicit terms ----------------
    double gm[NA1], gT[NA1], gV[NM1];
    {
      double lna[NA1];
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        lna[i] = pf_log(x[i]) + lng[i];
        gm[i] = 0.0;
        gT[i] = 0.0;
      });
      double f = 0.0;
      sfor<0, NM>([&](auto M) {
        constexpr int m = decltype(M)::value;
        gV[m] = A.wv[m];
        f += gV[m] * vf[m];
      });
      sfor<0, NC>([&](auto K) {
        constexpr int k = decltype(K)::value;
        const double wk = A.ws[k];
        if (wk == 0.0) return;
        const double mk = cplx_m(K, lna);
        f += wk * mk;
        sfor<0, Net::CPLX_N[k]>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          constexpr int ia = Net::CPLX_ID[k][a];
          gm[ia] += wk * mk * Net::CPLX_NU[k][a] / x[ia];
        });
      });
      if (A.fcell) A.fcell[c] = f;
    }

    // ---------------- reverse sweep over the recorded sub-steps ----------------
    const double mf = A.jacobian ? 1.0 : dkl;
    for (int k = nlog - 1; k >= 0; --k) {
      const bool first = k == nlog - 1;
      const double* lg = A.log + (i64)k * NF * ldl + c;
      double m[NA1], v0[NM1], lna[NA1];
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        m[i] = lg[i * ldl];
        lna[i] = pf_log(m[i]) + lng[i];
      });
      sfor<0, NM>([&](auto M) {
        constexpr int mi = decltype(M)::value;
        v0[mi] = lg[(NAQ + mi) * ldl];
      });
      const double h = lg[(NAQ + NM) * ldl];
      const double jc = psv * dkl / h;
      // kinetic-state update of this sub-step
      double gkin[NM1], Imq[NM1];
      bool actq[NM1];
      sfor<0, NM>([&](auto Mq) {
        constexpr int mi = decltype(Mq)::value;
        double Im, dI;
        actq[mi] = mnrl_r(Mq, lna, v0[mi], Im, dI);
        Imq[mi] = Im;
        gkin[mi] = 0.0;
        if (v0[mi] + Im * Net::MNRL_VM[mi] * h < 0.0) gV[mi] = 0.0;
        if (gV[mi] == 0.0 || !actq[mi]) return;
        gkin[mi] = gV[mi] * kLogToLn * Im * Net::MNRL_VM[mi] * h;
        sfor<0, Net::MNRL_N[mi]>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          constexpr int ia = Net::MNRL_ID[mi][a];
          gm[ia] += gV[mi] * Net::MNRL_VM[mi] * h * dI * Net::MNRL_NU[mi][a] / m[ia];
        });
      });
      // the sub-step's end totals feed the next sub-step's accumulation:
      // gm_j += dkl (gT_j + sum_k nu_jk m_k (nu_k . gT) / m_j)
      double D[NA1], S[NT], Tt[NA1], rs[NA1];
      {
        double T0d[NA1];
#pragma unroll
        for (int i = 0; i < NAQ; ++i) T0d[i] = 0.0;
        system(m, T0d, v0, h, mf, Tt, rs, D, S);
      }
      double gt[NA1];
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        gt[i] = gT[i] * m[i];
      });
      sfor<0, NC>([&](auto K) {
        constexpr int kk = decltype(K)::value;
        constexpr int ns = Net::CPLX_N[kk];
        double dotT = 0.0;
        sfor<0, ns>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          dotT += Net::CPLX_NU[kk][a] * gT[Net::CPLX_ID[kk][a]];
        });
        if (dotT == 0.0) return;
        const double mk = cplx_m(K, lna);
        sfor<0, ns>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          gt[Net::CPLX_ID[kk][a]] += Net::CPLX_NU[kk][a] * mk * dotT;
        });
      });
      // J^T lambda = gm  <=>  M lambda = diag(m) gm / jc
      double lam[NA1];
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        gm[i] += dkl * gt[i] / m[i];
        lam[i] = m[i] * gm[i] / jc;
      });
      if (!sens_ldlt<NA1>(D, S, lam)) {
        A.status[c] = kEscalate;
        return;
      }
      // log K of the complexes: through the sub-step's totals (gT) and
      // through the residual (lambda); at the end of the step also the
      // secondary molality itself
      sfor<0, NC>([&](auto K) {
        constexpr int kk = decltype(K)::value;
        constexpr int ns = Net::CPLX_N[kk];
        double accT = 0.0, accL = 0.0;
        sfor<0, ns>([&](auto Q) {
          constexpr int a = decltype(Q)::value;
          constexpr int ia = Net::CPLX_ID[kk][a];
          accT += gT[ia] * Net::CPLX_NU[kk][a];
          accL += lam[ia] * Net::CPLX_NU[kk][a];
        });
        const double wk = first ? A.ws[kk] : 0.0;
        double* o = A.grad + (i64)(NM + kk) * ldo + c;
        if (accT == 0.0 && accL == 0.0 && wk == 0.0) {
          if (first) *o = 0.0;
          return;
        }
        const double mk = cplx_m(K, lna);
        const double v = (accT - psv / h * accL) * (-kLogToLn * mk * dkl) + wk * (-kLogToLn * mk);
        *o = first ? v : *o + v;
      });
      // rate constants: the rate scales with k
      sfor<0, NM>([&](auto Mq) {
        constexpr int mi = decltype(Mq)::value;
        double v = gkin[mi];
        if (actq[mi]) {
          double acc = 0.0;
          sfor<0, Net::MNRL_N[mi]>([&](auto Q) {
            constexpr int a = decltype(Q)::value;
            acc += lam[Net::MNRL_ID[mi][a]] * Net::MNRL_NU[mi][a];
          });
          v -= acc * vol * kLogToLn * Imq[mi];
        }
        double* o = A.grad + (i64)mi * ldo + c;
        *o = first ? v : *o + v;
      });
      sfor<0, NAQ>([&](auto Q) {
        constexpr int i = decltype(Q)::value;
        gT[i] = lam[i] * psv / h;
        gm[i] = 0.0;
      });
    }
    A.status[c] = 0;
  }
}

}  // namespace pflane

#ifdef PFLANE_ADJOINT
extern "C" __global__ void __launch_bounds__(256) pfchem_lane_adjoint(pflane::SensArgs a) {
  pflane::lane_adjoint(a);
}
#endif
