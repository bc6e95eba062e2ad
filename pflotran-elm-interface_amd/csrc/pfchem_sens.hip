// pfchem_sens.hip -- tangent and adjoint of one reaction step with respect to
// the log10 rate constants of the kinetic minerals and the log10 K of the
// aqueous complexes.  One cell per thread, FP64 throughout.
//
// The step is RStep (reaction.F90:3564-3738): accepted sub-steps s of size
// dt_s, each the converged RReact system (reaction.F90:3742-4055)
//
//   F(m; T0, VF0, p) = psv (T(m; p) - T0) / dt_s + vol sum_n nu_n Im_n(m; p) = 0
//
// followed by VF = max(0, VF0 + Im V_m dt_s) (MineralUpdateKineticState,
// reaction_mineral.F90:1456-1525); the next sub-step starts from T(m), VF.
// Per cell the objective is f = sum_n w_vf[n] VF_n(end) + sum_k w_sec[k]
// m_k(end).  The kernel runs the forward step, recording each accepted
// sub-step (T0, VF0, m, dt_s), then either
//   * tangent: dm_s = -J^-1 (dF/dp dp + dF/dT0 dT0) sub-step by sub-step,
//     df = w . d(VF, m_k); or
//   * adjoint: the reverse sweep, one transposed solve J^T lambda = mbar per
//     sub-step (implicit-function theorem at each converged sub-step),
//     giving the whole gradient df/dp per cell.
// J = dF/dm is rebuilt at each recorded solution.  The reference assembles
// its mineral block with an extra rho/1000 (dQK_dmj, reaction_mineral.F90:
// 945); the Newton iteration here keeps that form (same iterates as RReact)
// and the sensitivities use it only when asked (jacobian = 0), otherwise the
// exact derivative.
//
// Scope (checked by the API): aqueous networks (no immobile species),
// kinetic minerals under plain TST with optional Arrhenius temperature
// dependence, activity coefficients held fixed during the step.  Dry cells
// get PFCHEM_CELL_SENS_UNSUPPORTED.
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"
#include "pfchem_lu.hpp"

namespace pfchem {
namespace {

constexpr int kSensMaxMnrl = 16;
constexpr int kSensMaxSub = 8;

template <int MAXN>
struct SensCell {
  int64_t c;
  double lng[MAXN];  // ln gamma of the primaries
  double vf[kSensMaxMnrl], area[kSensMaxMnrl];
  double psv, den, vol, temp, ln_aw;
};

template <int MAXN>
struct SubStepLog {
  int n;
  double dt[kSensMaxSub];
  double t0[kSensMaxSub][MAXN];
  double m[kSensMaxSub][MAXN];
  double vf0[kSensMaxSub][kSensMaxMnrl];
};

__device__ __forceinline__ double sens_cplx_logk(const DevNet& n, int k, double temp) {
  return n.isothermal || !n.has_cplx_logKcoef ? gld(n.cplx_logK + k)
                                              : interp_logk(n.cplx_logKcoef + 5 * k, temp);
}

// secondary molality of complex k at ln activities lna
template <int MAXN>
__device__ double sens_mk(const DevNet& n, const DevState& s, const SensCell<MAXN>& w, int k,
                          const double* lna) {
  const int ms = n.cplx_ms;
  double lnQK = -sens_cplx_logk(n, k, w.temp) * kLogToLn;
  if (n.has_cplx_h2o) lnQK += gld(n.cplx_h2o + k) * w.ln_aw;
  const int* ids = n.cplx_ids + k * (ms + 1);
  const double* st = n.cplx_st + k * ms;
  const int ncs = gld(ids);
  for (int a = 0; a < ncs; ++a) lnQK += gld(st + a) * lna[gld(ids + 1 + a)];
  return exp(lnQK) / gld(s.sec_act_coef + k * s.ld + w.c);
}

// totals (mol/L) and, if dT != nullptr, dtotal (RTotalAqueous,
// reaction.F90:4665-4759) at molalities m
template <int MAXN>
__device__ void sens_totals(const DevNet& n, const DevState& s, const SensCell<MAXN>& w,
                            const double* m, double* T, double* dT) {
  const int naq = n.naq, ms = n.cplx_ms;
  double lna[MAXN];
  for (int i = 0; i < naq; ++i) {
    lna[i] = log(m[i]) + w.lng[i];
    T[i] = m[i];
  }
  if (dT)
    for (int i = 0; i < naq; ++i)
      for (int j = 0; j < naq; ++j) dT[i * naq + j] = i == j ? 1.0 : 0.0;
  for (int k = 0; k < n.nc; ++k) {
    const double mk = sens_mk(n, s, w, k, lna);
    const int* ids = n.cplx_ids + k * (ms + 1);
    const double* st = n.cplx_st + k * ms;
    const int ncs = gld(ids);
    for (int a = 0; a < ncs; ++a) {
      const int i = gld(ids + 1 + a);
      const double nu = gld(st + a);
      T[i] += nu * mk;
      if (dT)
        for (int b = 0; b < ncs; ++b) {
          const int j = gld(ids + 1 + b);
          dT[i * naq + j] += nu * gld(st + b) * mk / m[j];
        }
    }
  }
  const double dkl = w.den * 1.0e-3;
  for (int i = 0; i < naq; ++i) T[i] *= dkl;
  if (dT)
    for (int i = 0; i < naq * naq; ++i) dT[i] *= dkl;
}

// plain TST rate of mineral q (RKineticMineral, reaction_mineral.F90:
// 760-840): Im (mol/m^3 bulk/s) and dIm/dlnQK; 0 if the mineral is inactive
template <int MAXN>
__device__ int sens_tst(const DevNet& n, const SensCell<MAXN>& w, int q, const double* m,
                        const double* vf, double* Im, double* dI) {
  const int ms = n.mnrl_ms;
  double logK = n.isothermal || !n.has_mnrl_logKcoef ? gld(n.mnrl_logK + q)
                                                     : interp_logk(n.mnrl_logKcoef + 5 * q, w.temp);
  double lnQK = -logK * kLogToLn;
  if (n.has_mnrl_h2o) lnQK += gld(n.mnrl_h2o + q) * w.ln_aw;
  const int* ids = n.mnrl_ids + q * (ms + 1);
  const double* st = n.mnrl_st + q * ms;
  const int ncs = gld(ids);
  for (int a = 0; a < ncs; ++a) {
    const int j = gld(ids + 1 + a);
    lnQK += gld(st + a) * (log(m[j]) + w.lng[j]);
  }
  const double QK = exp(lnQK);
  const double aff = 1.0 - QK;
  *Im = 0.0;
  *dI = 0.0;
  if (w.area[q] == 0.0) return 0;
  if (!(vf[q] > 0.0 || aff < 0.0)) return 0;
  double arr = 1.0;
  if (n.has_Ea) {
    const double Ea = gld(n.mnrl_Ea + q);
    if (Ea > 0.0) arr = exp(Ea / kIdealGas * (1.0 / (25.0 + 273.15) - 1.0 / (w.temp + 273.15)));
  }
  const double kr = gld(n.mnrl_k + q) * arr;
  *Im = -w.area[q] * copysign(1.0, aff) * fabs(aff) * kr;
  *dI = w.area[q] * kr * QK;
  return 1;
}

// dF/dm at m (row-major naq x naq); factor on the mineral block
template <int MAXN>
__device__ void sens_jacobian(const DevNet& n, const SensCell<MAXN>& w, const double* m,
                              const double* dT, const double* dI, const int* act, double dt,
                              double mnrl_factor, double* J) {
  const int naq = n.naq, ms = n.mnrl_ms;
  for (int i = 0; i < naq * naq; ++i) J[i] = w.psv / dt * dT[i];
  for (int q = 0; q < n.nm; ++q) {
    if (!act[q]) continue;
    const int* ids = n.mnrl_ids + q * (ms + 1);
    const double* st = n.mnrl_st + q * ms;
    const int ncs = gld(ids);
    for (int a = 0; a < ncs; ++a)
      for (int b = 0; b < ncs; ++b) {
        const int i = gld(ids + 1 + a), j = gld(ids + 1 + b);
        J[i * naq + j] += w.vol * gld(st + a) * dI[q] * gld(st + b) / m[j] * mnrl_factor;
      }
  }
}

// one RReact solve from start totals T0 (mol/L) and volume fractions vf0;
// m holds the guess on entry and the iterate on exit.  Returns 0, kMaxIter
// or kSingular; T receives the totals of the last iterate.
template <int MAXN>
__device__ int sens_rreact(const DevNet& n, const DevState& s, const SensCell<MAXN>& w,
                           double* m, const double* T0, const double* vf0, double dt, double* T,
                           double* dT, double* J) {
  const int naq = n.naq;
  double res[MAXN], Im[kSensMaxMnrl], dI[kSensMaxMnrl];
  int act[kSensMaxMnrl];
  double r0 = 0.0;
  for (int it = 1;; ++it) {
    sens_totals(n, s, w, m, T, dT);
    for (int i = 0; i < naq; ++i) res[i] = (w.psv * T[i] - w.psv * T0[i]) / dt;
    for (int q = 0; q < n.nm; ++q) {
      act[q] = sens_tst(n, w, q, m, vf0, &Im[q], &dI[q]);
      if (!act[q]) continue;
      const int* ids = n.mnrl_ids + q * (n.mnrl_ms + 1);
      const double* st = n.mnrl_st + q * n.mnrl_ms;
      for (int a = 0; a < gld(ids); ++a) res[gld(ids + 1 + a)] += gld(st + a) * Im[q] * w.vol;
    }
    double mx = 0.0, nr = 0.0;
    for (int i = 0; i < naq; ++i) {
      mx = fmax(mx, fabs(res[i]));
      nr += res[i] * res[i];
    }
    nr = sqrt(nr);
    if (it == 1) r0 = nr;
    if (it > n.max_it) return kMaxIter;
    if (mx < n.tol_abs || nr / r0 < n.tol_rel) return 0;
    // RSolve (reaction.F90:5457-5516): row scaling, log formulation
    sens_jacobian(n, w, m, dT, dI, act, dt, w.den * 1.0e-3, J);
    for (int i = 0; i < naq; ++i) {
      double nrm = 0.0;
      for (int j = 0; j < naq; ++j) nrm = fmax(nrm, fabs(J[i * naq + j]));
      nrm = 1.0 / fmax(nrm, 1.0);
      res[i] *= nrm;
      for (int j = 0; j < naq; ++j) J[i * naq + j] *= nrm;
    }
    if (n.log_form)
      for (int i = 0; i < naq; ++i)
        for (int j = 0; j < naq; ++j) J[i * naq + j] *= m[j];
    if (crout_lu_solve<MAXN>(J, naq, res)) return kSingular;
    double latest[MAXN], rc = -1.0;
    if (n.log_form) {
      for (int i = 0; i < naq; ++i)
        latest[i] = m[i] * exp(-copysign(1.0, res[i]) * fmin(fabs(res[i]), n.max_dlnC));
    } else {
      double ratio = 1.0e308;
      for (int i = 0; i < naq; ++i)
        if (m[i] <= res[i]) ratio = fmin(ratio, fabs(m[i] / res[i]));
      for (int i = 0; i < naq; ++i) latest[i] = m[i] - (ratio < 1.0 ? res[i] * ratio * 0.99 : res[i]);
    }
    for (int i = 0; i < naq; ++i) {
      const double q = fabs((latest[i] - m[i]) / m[i]);
      if (q == q) rc = fmax(rc, q);
    }
    if (rc >= 0.0 && rc < n.tol_change) return 0;
    for (int i = 0; i < naq; ++i) m[i] = latest[i];
  }
}

template <int MAXN>
__global__ void __launch_bounds__(64)
sens_kernel(const DevNet* np, const DevState* sp, int64_t ncell, double dt,
            const double* __restrict__ wv, const double* __restrict__ ws,
            const double* __restrict__ dp, int mode, int jacobian, double* __restrict__ fcell,
            double* __restrict__ out, int64_t ldo, int32_t* __restrict__ status, int rescue) {
  const DevNet& n = *np;
  const DevState& s = *sp;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  // rescue: only the cells an earlier pass handed over
  if (rescue && gld(status + c) != kTeamEscalate) return;
  const int naq = n.naq, nm = n.nm, nc = n.nc, npar = nm + nc;
  const int64_t ld = s.ld;
  SensCell<MAXN> w;
  w.c = c;
  const double por = gld(s.porosity + c), sat = gld(s.saturation + c);
  w.den = gld(s.den_kg + c);
  w.vol = gld(s.volume + c);
  w.temp = s.temperature ? gld(s.temperature + c) : 25.0;
  w.ln_aw = s.ln_act_h2o ? gld(s.ln_act_h2o + c) : 0.0;
  w.psv = por * sat * 1000.0 * w.vol;
  double m[MAXN], T0[MAXN], T[MAXN], guess[MAXN];
  for (int i = 0; i < naq; ++i) {
    m[i] = gld(s.pri_molal + i * ld + c);
    T0[i] = gld(s.total + i * ld + c);
    w.lng[i] = log(gld(s.pri_act_coef + i * ld + c));
    if (T0[i] <= 1.0e-40) T0[i] = 1.0e-40;  // reaction.F90:3617-3631
    guess[i] = n.total_as_guess ? T0[i] : m[i];
  }
  for (int q = 0; q < nm; ++q) {
    w.vf[q] = gld(s.mnrl_volfrac + q * ld + c);
    w.area[q] = gld(s.mnrl_area + q * ld + c);
  }
  const bool fill = mode == 1 && out;
  if (fill)
    for (int q = 0; q < npar; ++q) gst(out + q * ldo + c, 0.0);
  if (sat < kRtMinSaturation) {
    if (status) gst(status + c, (int32_t)PFCHEM_CELL_SENS_UNSUPPORTED);
    return;
  }
  // ---------------- forward RStep with the sub-step log ----------------
  double dTm[MAXN * MAXN], J[MAXN * MAXN];
  SubStepLog<MAXN> lg;
  lg.n = 0;
  double target = dt, cum = 0.0, tdt = dt;
  int nconst = 0, ncuts = 0, attempts = 0, st = 0;
  while (cum < target) {
    if (++attempts > n.max_calls) {
      st = PFCHEM_CELL_MAX_CUTS_ITERATIONS;
      break;
    }
    for (int i = 0; i < naq; ++i) m[i] = guess[i];
    const int why = sens_rreact<MAXN>(n, s, w, m, T0, w.vf, tdt, T, dTm, J);
    if (why) {
      // max iterations: totals restored; singular: the retry starts from
      // the totals of the failed iterate (reaction.F90:3963-3967, 4041)
      if (why == kSingular)
        for (int i = 0; i < naq; ++i) T0[i] = T[i];
      if (++ncuts > n.max_cuts) {
        st = why == kSingular ? PFCHEM_CELL_MAX_CUTS_SINGULAR : PFCHEM_CELL_MAX_CUTS_ITERATIONS;
        break;
      }
      tdt *= 0.5;
      nconst = 0;
      continue;
    }
    if (lg.n == kSensMaxSub) {
      st = PFCHEM_CELL_TOO_MANY_SUBSTEPS;
      break;
    }
    const int k = lg.n++;
    lg.dt[k] = tdt;
    for (int i = 0; i < naq; ++i) {
      lg.t0[k][i] = T0[i];
      lg.m[k][i] = m[i];
      T0[i] = T[i];
      guess[i] = m[i];
    }
    for (int q = 0; q < nm; ++q) {
      lg.vf0[k][q] = w.vf[q];
      double Im, dI;
      sens_tst(n, w, q, m, lg.vf0[k], &Im, &dI);
      w.vf[q] = fmax(0.0, w.vf[q] + Im * gld(n.mnrl_Vm + q) * tdt);
    }
    cum += tdt;
    if (++nconst >= 4) {
      --ncuts;
      tdt = fmin(2.0 * tdt, target - cum);
    }
  }
  if (st) {
    if (status) gst(status + c, (int32_t)st);
    return;
  }
  // objective at the end of the step
  double lna[MAXN];
  for (int i = 0; i < naq; ++i) lna[i] = log(m[i]) + w.lng[i];
  double f = 0.0;
  for (int q = 0; q < nm; ++q) f += gld(wv + q) * w.vf[q];
  for (int k = 0; k < nc; ++k) {
    const double wk = gld(ws + k);
    if (wk != 0.0) f += wk * sens_mk(n, s, w, k, lna);
  }
  if (fcell) gst(fcell + c, f);
  const double mf = jacobian ? 1.0 : w.den * 1.0e-3;
  const double dkl = w.den * 1.0e-3;
  const int cms = n.cplx_ms, mms = n.mnrl_ms;
  double Im[kSensMaxMnrl], dI[kSensMaxMnrl];
  int act[kSensMaxMnrl];
  if (mode == 0) {
    // ---------------- tangent ----------------
    double dT[MAXN], dV[kSensMaxMnrl], dm[MAXN];
    for (int i = 0; i < naq; ++i) dT[i] = dm[i] = 0.0;
    for (int q = 0; q < nm; ++q) dV[q] = 0.0;
    for (int k = 0; k < lg.n; ++k) {
      const double h = lg.dt[k];
      double* mk_ = lg.m[k];
      sens_totals(n, s, w, mk_, T, dTm);
      for (int q = 0; q < nm; ++q) act[q] = sens_tst(n, w, q, mk_, lg.vf0[k], &Im[q], &dI[q]);
      for (int i = 0; i < naq; ++i) lna[i] = log(mk_[i]) + w.lng[i];
      double dTp[MAXN], rhs[MAXN];
      for (int i = 0; i < naq; ++i) dTp[i] = 0.0;
      for (int kk = 0; kk < nc; ++kk) {
        const double v = gld(dp + nm + kk);
        if (v == 0.0) continue;
        const double mk = sens_mk(n, s, w, kk, lna);
        const int* ids = n.cplx_ids + kk * (cms + 1);
        const double* stc = n.cplx_st + kk * cms;
        for (int a = 0; a < gld(ids); ++a)
          dTp[gld(ids + 1 + a)] += gld(stc + a) * (-kLogToLn * mk) * dkl * v;
      }
      for (int i = 0; i < naq; ++i) rhs[i] = -(w.psv / h * (dTp[i] - dT[i]));
      for (int q = 0; q < nm; ++q) {
        const double v = gld(dp + q);
        if (!act[q] || v == 0.0) continue;
        const int* ids = n.mnrl_ids + q * (mms + 1);
        const double* stm = n.mnrl_st + q * mms;
        for (int a = 0; a < gld(ids); ++a)
          rhs[gld(ids + 1 + a)] -= gld(stm + a) * w.vol * kLogToLn * Im[q] * v;
      }
      sens_jacobian(n, w, mk_, dTm, dI, act, h, mf, J);
      for (int i = 0; i < naq; ++i) dm[i] = rhs[i];
      if (crout_lu_solve<MAXN>(J, naq, dm)) {
        if (status) gst(status + c, (int32_t)PFCHEM_CELL_MAX_CUTS_SINGULAR);
        return;
      }
      for (int i = 0; i < naq; ++i) {
        double acc = dTp[i];
        for (int j = 0; j < naq; ++j) acc += dTm[i * naq + j] * dm[j];
        dT[i] = acc;
      }
      for (int q = 0; q < nm; ++q) {
        const double Vm = gld(n.mnrl_Vm + q);
        if (lg.vf0[k][q] + Im[q] * Vm * h < 0.0) {
          dV[q] = 0.0;
          continue;
        }
        if (!act[q]) continue;
        const int* ids = n.mnrl_ids + q * (mms + 1);
        const double* stm = n.mnrl_st + q * mms;
        double dl = 0.0;
        for (int a = 0; a < gld(ids); ++a) {
          const int j = gld(ids + 1 + a);
          dl += gld(stm + a) * dm[j] / mk_[j];
        }
        dV[q] += Vm * h * (dI[q] * dl + kLogToLn * Im[q] * gld(dp + q));
      }
    }
    double df = 0.0;
    for (int q = 0; q < nm; ++q) df += gld(wv + q) * dV[q];
    for (int i = 0; i < naq; ++i) lna[i] = log(m[i]) + w.lng[i];
    for (int kk = 0; kk < nc; ++kk) {
      const double wk = gld(ws + kk);
      if (wk == 0.0) continue;
      const int* ids = n.cplx_ids + kk * (cms + 1);
      const double* stc = n.cplx_st + kk * cms;
      double dl = -kLogToLn * gld(dp + nm + kk);
      for (int a = 0; a < gld(ids); ++a) {
        const int j = gld(ids + 1 + a);
        dl += gld(stc + a) * dm[j] / m[j];
      }
      df += wk * sens_mk(n, s, w, kk, lna) * dl;
    }
    if (out) gst(out + c, df);
    if (status) gst(status + c, (int32_t)0);
    return;
  }
  // ---------------- adjoint ----------------
  double gm[MAXN], gT[MAXN], gV[kSensMaxMnrl];
  for (int i = 0; i < naq; ++i) gm[i] = gT[i] = 0.0;
  for (int q = 0; q < nm; ++q) gV[q] = gld(wv + q);
  // explicit dependence of the end-of-step secondary molalities
  for (int kk = 0; kk < nc; ++kk) {
    const double wk = gld(ws + kk);
    if (wk == 0.0) continue;
    const double mk = sens_mk(n, s, w, kk, lna);
    if (fill) gst(out + (nm + kk) * ldo + c, gld(out + (nm + kk) * ldo + c) + wk * (-kLogToLn * mk));
    const int* ids = n.cplx_ids + kk * (cms + 1);
    const double* stc = n.cplx_st + kk * cms;
    for (int a = 0; a < gld(ids); ++a) {
      const int j = gld(ids + 1 + a);
      gm[j] += wk * mk * gld(stc + a) / m[j];
    }
  }
  for (int k = lg.n - 1; k >= 0; --k) {
    const double h = lg.dt[k];
    double* mk_ = lg.m[k];
    sens_totals(n, s, w, mk_, T, dTm);
    for (int q = 0; q < nm; ++q) act[q] = sens_tst(n, w, q, mk_, lg.vf0[k], &Im[q], &dI[q]);
    for (int i = 0; i < naq; ++i) lna[i] = log(mk_[i]) + w.lng[i];
    // kinetic-state update
    for (int q = 0; q < nm; ++q) {
      const double Vm = gld(n.mnrl_Vm + q);
      if (lg.vf0[k][q] + Im[q] * Vm * h < 0.0) gV[q] = 0.0;
      if (gV[q] == 0.0 || !act[q]) continue;
      if (fill) gst(out + q * ldo + c, gld(out + q * ldo + c) + gV[q] * kLogToLn * Im[q] * Vm * h);
      const int* ids = n.mnrl_ids + q * (mms + 1);
      const double* stm = n.mnrl_st + q * mms;
      for (int a = 0; a < gld(ids); ++a) {
        const int j = gld(ids + 1 + a);
        gm[j] += gV[q] * Vm * h * dI[q] * gld(stm + a) / mk_[j];
      }
    }
    // totals at the end of the sub-step
    for (int j = 0; j < naq; ++j) {
      double acc = 0.0;
      for (int i = 0; i < naq; ++i) acc += gT[i] * dTm[i * naq + j];
      gm[j] += acc;
    }
    // J^T lambda = gm
    sens_jacobian(n, w, mk_, dTm, dI, act, h, mf, J);
    for (int i = 0; i < naq; ++i)
      for (int j = i + 1; j < naq; ++j) {
        const double t = J[i * naq + j];
        J[i * naq + j] = J[j * naq + i];
        J[j * naq + i] = t;
      }
    double lam[MAXN];
    for (int i = 0; i < naq; ++i) lam[i] = gm[i];
    if (crout_lu_solve<MAXN>(J, naq, lam)) {
      if (status) gst(status + c, (int32_t)PFCHEM_CELL_MAX_CUTS_SINGULAR);
      return;
    }
    // complexes: T depends on log K (end-of-sub-step totals and the solve)
    for (int kk = 0; kk < nc; ++kk) {
      const int* ids = n.cplx_ids + kk * (cms + 1);
      const double* stc = n.cplx_st + kk * cms;
      double accT = 0.0, accL = 0.0;
      for (int a = 0; a < gld(ids); ++a) {
        const int i = gld(ids + 1 + a);
        accT += gT[i] * gld(stc + a);
        accL += lam[i] * gld(stc + a);
      }
      if ((accT != 0.0 || accL != 0.0) && fill) {
        const double dTdp = -kLogToLn * sens_mk(n, s, w, kk, lna) * dkl;
        gst(out + (nm + kk) * ldo + c,
            gld(out + (nm + kk) * ldo + c) + (accT - w.psv / h * accL) * dTdp);
      }
    }
    // minerals: the rate constant scales the rate
    for (int q = 0; q < nm; ++q) {
      if (!act[q] || !fill) continue;
      const int* ids = n.mnrl_ids + q * (mms + 1);
      const double* stm = n.mnrl_st + q * mms;
      double acc = 0.0;
      for (int a = 0; a < gld(ids); ++a) acc += lam[gld(ids + 1 + a)] * gld(stm + a);
      gst(out + q * ldo + c, gld(out + q * ldo + c) - acc * w.vol * kLogToLn * Im[q]);
    }
    for (int i = 0; i < naq; ++i) {
      gT[i] = lam[i] * w.psv / h;
      gm[i] = 0.0;
    }
  }
  if (status) gst(status + c, (int32_t)0);
}

// deterministic sum over cells of each parameter row of the per-cell
// gradients (cells whose status is not 0 are skipped)
__global__ void __launch_bounds__(256)
sens_reduce_kernel(const double* __restrict__ g, int64_t ldo, int64_t ncell,
                   const int32_t* __restrict__ status, double* __restrict__ sum) {
  __shared__ double part[256];
  const int q = blockIdx.x;
  double acc = 0.0;
  for (int64_t c = threadIdx.x; c < ncell; c += 256)
    if (gld(status + c) == 0) acc += gld(g + q * ldo + c);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) gst(sum + q, part[0]);
}

}  // namespace

int sens_supported(const DevNet& n, const pfchem_options& opt) {
  if (n.variable || n.nimm > 0 || n.nm > kSensMaxMnrl || n.ncomp > 32) return 0;
  if (opt.act_coef_update_frequency == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER) return 0;
  if (n.has_temkin || n.has_minscale || n.has_affpower || n.has_limiter || n.has_threshold ||
      n.has_irrev || n.has_pref)
    return 0;
  return 1;
}

hipError_t launch_sens(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                       double dt, const double* wv, const double* ws, const double* dp, int mode,
                       int jacobian, double* fcell, double* out, int64_t ldo, int32_t* status,
                       double* sum, hipStream_t stream, int rescue) {
  const int block = 64;
  const dim3 grid((unsigned)((ncell + block - 1) / block));
#define PFCHEM_SENS(MAXN)                                                                   \
  sens_kernel<MAXN><<<grid, dim3(block), 0, stream>>>(dn, ds, ncell, dt, wv, ws, dp, mode, \
                                                      jacobian, fcell, out, ldo, status, rescue)
  if (n.naq <= 4)
    PFCHEM_SENS(4);
  else if (n.naq <= 8)
    PFCHEM_SENS(8);
  else if (n.naq <= 16)
    PFCHEM_SENS(16);
  else
    PFCHEM_SENS(32);
#undef PFCHEM_SENS
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || mode != 1 || !sum) return e;
  sens_reduce_kernel<<<dim3((unsigned)(n.nm + n.nc)), dim3(256), 0, stream>>>(out, ldo, ncell,
                                                                               status, sum);
  return hipGetLastError();
}

}  // namespace pfchem
