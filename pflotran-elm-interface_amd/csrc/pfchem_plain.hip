// pfchem_plain.hip -- plain FP64 reaction step: one cell per thread.
//
// The straightforward GPU form of RStep (reaction.F90:3564-3738): every
// thread runs the whole sub-step / Newton loop of one cell with runtime-sized
// loops over the network and per-thread arrays (compiler-placed in scratch).
// No work stealing: a wavefront runs as long as its slowest cell.  This is
// the FP64 reference point the batched kernels are measured against.
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"
#include "pfchem_lu.hpp"
#include "pfchem_somdec.hpp"

namespace pfchem {
namespace {

constexpr int kMaxMnrl = 16;

template <int MAXN>
struct PlainCell {
  double x[MAXN];        // pri_molal (aq) then immobile
  double total[MAXN];    // rt_auxvar%total(:,1) (aq part used)
  double imm[MAXN];      // rt_auxvar%immobile
  double gp[MAXN];       // pri_act_coef
  double J[MAXN * MAXN]; // row-major J(i,j); dtotal shares the aq block
  double res[MAXN];
  double vf[kMaxMnrl], rate[kMaxMnrl], logKm[kMaxMnrl];
  double ln_act_h2o;
  double por, sat, den, vol, temp;
  int naq, N;  // the cell's aqueous primaries and unknowns (per-cell active
               // counts with options.variable_ncomp)
};

// complex k / mineral m takes part in the cell's system: with per-cell
// active primaries, only those all of whose species are active
template <int MAXN>
__device__ __forceinline__ bool plain_cplx_on(const DevNet& n, const PlainCell<MAXN>& w, int k) {
  return !n.variable || n.cplx_maxid[k] < w.naq;
}
template <int MAXN>
__device__ __forceinline__ bool plain_mnrl_on(const DevNet& n, const PlainCell<MAXN>& w, int m) {
  return !n.variable || n.mnrl_maxid[m] < w.naq;
}

// RTotalAqueous (reaction.F90:4665-4759): totals, secondary molalities
// (written to the SoA state) and dtotal into the aq block of J (not yet
// multiplied by the accumulation factor).
template <int MAXN>
__device__ void plain_total(const DevNet& n, const DevState& s, int64_t c,
                            PlainCell<MAXN>& w, int want_dtotal) {
  const int naq = w.naq, N = w.N, ms = n.cplx_ms;
  double lnc[MAXN], lna[MAXN];
  for (int i = 0; i < naq; ++i) {
    lnc[i] = log(w.x[i]);
    lna[i] = lnc[i] + log(w.gp[i]);
    w.total[i] = w.x[i];
  }
  if (want_dtotal) {
    for (int i = 0; i < naq; ++i)
      for (int j = 0; j < naq; ++j) w.J[i * N + j] = (i == j) ? 1.0 : 0.0;
  }
  for (int k = 0; k < n.nc; ++k) {
    if (!plain_cplx_on(n, w, k)) continue;
    double logK = n.cplx_logK[k];
    if (!n.isothermal && n.has_cplx_logKcoef)
      logK = interp_logk(n.cplx_logKcoef + 5 * k, w.temp);
    double lnQK = -logK * kLogToLn;
    if (n.has_cplx_h2o && n.cplx_h2o[k] != 0.0) lnQK += n.cplx_h2o[k] * w.ln_act_h2o;
    const int* ids = n.cplx_ids + k * (ms + 1);
    const double* st = n.cplx_st + k * ms;
    const int ncs = ids[0];
    for (int i = 0; i < ncs; ++i) lnQK += st[i] * lna[ids[1 + i]];
    const double gk = s.sec_act_coef[k * s.ld + c];
    const double mk = exp(lnQK) / gk;
    s.sec_molal[k * s.ld + c] = mk;
    for (int i = 0; i < ncs; ++i) w.total[ids[1 + i]] += st[i] * mk;
    if (want_dtotal) {
      for (int j = 0; j < ncs; ++j) {
        const int jc = ids[1 + j];
        const double t = st[j] * exp(lnQK - lnc[jc]) / gk;
        for (int i = 0; i < ncs; ++i) w.J[ids[1 + i] * N + jc] += st[i] * t;
      }
    }
  }
  const double dkl = w.den * 1.0e-3;
  for (int i = 0; i < naq; ++i) w.total[i] *= dkl;
  if (want_dtotal)
    for (int i = 0; i < naq; ++i)
      for (int j = 0; j < naq; ++j) w.J[i * N + j] *= dkl;
}

// RKineticMineral (reaction_mineral.F90:647-1078), default path
template <int MAXN>
__device__ void plain_minerals(const DevNet& n, const DevState& s, int64_t c,
                               PlainCell<MAXN>& w, double* res, double* J,
                               int deriv) {
  const int naq = w.naq, N = w.N, ms = n.mnrl_ms;
  double lnc[MAXN], lna[MAXN];
  for (int i = 0; i < naq; ++i) {
    lnc[i] = log(w.x[i]);
    lna[i] = lnc[i] + log(w.gp[i]);
  }
  for (int m = 0; m < n.nm; ++m) w.rate[m] = 0.0;
  for (int m = 0; m < n.nm; ++m) {
    if (!plain_mnrl_on(n, w, m)) continue;
    const int* ids = n.mnrl_ids + m * (ms + 1);
    const double* st = n.mnrl_st + m * ms;
    const int ncs = ids[0];
    double lnQK = -w.logKm[m] * kLogToLn;
    if (n.has_mnrl_h2o && n.mnrl_h2o[m] != 0.0) lnQK += n.mnrl_h2o[m] * w.ln_act_h2o;
    for (int i = 0; i < ncs; ++i) lnQK += st[i] * lna[ids[1 + i]];
    const double QK = exp(lnQK);
    double aff;
    if (n.has_temkin) {
      aff = n.has_minscale ? 1.0 - pow(QK, 1.0 / (n.mnrl_minscale[m] * n.mnrl_temkin[m]))
                           : 1.0 - pow(QK, 1.0 / n.mnrl_temkin[m]);
    } else if (n.has_minscale) {
      aff = 1.0 - pow(QK, 1.0 / n.mnrl_minscale[m]);
    } else {
      aff = 1.0 - QK;
    }
    const double sgn = copysign(1.0, aff);
    if (!(w.vf[m] > 0.0 || sgn < 0.0)) continue;
    if (n.has_irrev && n.mnrl_irrev[m] == 1 && sgn < 0.0) continue;
    if (n.has_threshold && n.mnrl_threshold[m] > 0.0 && sgn < 0.0 &&
        QK < n.mnrl_threshold[m])
      continue;
    const double lim = n.has_limiter ? n.mnrl_limiter[m] : 0.0;
    if (lim > 0.0) aff = aff / (1.0 + (1.0 - aff) / lim);
    const bool pref = n.has_pref && n.pref_n[m] > 0;
    double kr;
    if (pref) {
      kr = prefactor_sum(n, m, lna, w.temp);  // reaction_mineral.F90:820-860
    } else {
      double arr = 1.0;
      if (n.has_Ea && n.mnrl_Ea[m] > 0.0)
        arr = exp(n.mnrl_Ea[m] / kIdealGas * (1.0 / (25.0 + 273.15) - 1.0 / (w.temp + 273.15)));
      kr = n.mnrl_k[m] * arr;
    }
    double Im_const = -s.mnrl_area[m * s.ld + c];
    if (n.has_minscale) Im_const = Im_const / n.mnrl_minscale[m];
    double Im = n.has_affpower
                    ? Im_const * sgn * pow(fabs(aff), n.mnrl_affpower[m]) * kr
                    : Im_const * sgn * fabs(aff) * kr;
    w.rate[m] = Im;
    Im_const *= w.vol;
    Im *= w.vol;
    for (int i = 0; i < ncs; ++i) res[ids[1 + i]] += st[i] * Im;
    if (!deriv) continue;
    double dIm_dQK = n.has_affpower ? -Im * n.mnrl_affpower[m] / fabs(aff) : -Im_const * kr;
    if (n.has_temkin) {
      const double e = n.has_minscale ? n.mnrl_minscale[m] * n.mnrl_temkin[m] : n.mnrl_temkin[m];
      dIm_dQK = dIm_dQK * (1.0 / e) / QK * (1.0 - aff);
    } else if (n.has_minscale) {
      dIm_dQK = dIm_dQK * (1.0 / n.mnrl_minscale[m]) / QK * (1.0 - aff);
    }
    double fac = 1.0;
    if (lim > 0.0) {
      const double den = 1.0 + (1.0 - aff) / lim;
      fac = (1.0 + QK / lim / den) / den;
    }
    for (int j = 0; j < ncs; ++j) {
      const int jc = ids[1 + j];
      const double dQK_dmj = st[j] * QK * exp(-lnc[jc]) * w.den * 1.0e-3;
      for (int i = 0; i < ncs; ++i) {
        if (lim > 0.0)
          J[ids[1 + i] * N + jc] += st[i] * dIm_dQK * fac * dQK_dmj;
        else
          J[ids[1 + i] * N + jc] += st[i] * dIm_dQK * dQK_dmj;
      }
    }
    // rate prefactors (reaction_mineral.F90:964-1033), d/dm = d/dln m / m
    if (pref)
      prefactor_jacobian(n, m, lna, w.temp, Im / kr, [&](int j, double v) {
        for (int i = 0; i < ncs; ++i) J[ids[1 + i] * N + j] += st[i] * v / w.x[j];
      });
  }
}

// RActivityCoefficients LAG (reaction.F90:4555-4611) with the cell's
// current pri/sec molalities; sec_act_coef written to the SoA state.
template <int MAXN>
__device__ void plain_activity_lag(const DevNet& n, const DevState& s, int64_t c,
                                   PlainCell<MAXN>& w) {
  double I = 0.0, sum_pri = 0.0, sum_sec = 0.0;
  for (int i = 0; i < w.naq; ++i) {
    I += w.x[i] * n.Zp[i] * n.Zp[i];
    if (i != n.h2o_aq_id) sum_pri += w.x[i];
  }
  for (int k = 0; k < n.nc; ++k) {
    if (!plain_cplx_on(n, w, k)) continue;
    const double m = s.sec_molal[k * s.ld + c];
    I += m * n.cplx_Z[k] * n.cplx_Z[k];
  }
  I *= 0.5;
  const double sI = sqrt(I);
  for (int i = 0; i < w.naq; ++i)
    w.gp[i] = fabs(n.Zp[i]) > 1.0e-10 ? dh_gamma(n.Zp[i], n.a0p[i], sI, I, n) : 1.0;
  for (int k = 0; k < n.nc; ++k) {
    if (!plain_cplx_on(n, w, k)) continue;
    s.sec_act_coef[k * s.ld + c] =
        fabs(n.cplx_Z[k]) > 1.0e-10 ? dh_gamma(n.cplx_Z[k], n.cplx_a0[k], sI, I, n) : 1.0;
    sum_sec += s.sec_molal[k * s.ld + c];
  }
  if (n.act_h2o) {
    const double a = 1.0 - 0.017 * (sum_pri + sum_sec);
    w.ln_act_h2o = a > 0.0 ? log(a) : 0.0;
  }
}

// RActivityCoefficients NEWTON (reaction.F90:4403-4551): fixed point on the
// ionic strength I with the d I / d I correction of the secondary
// molalities; the activities of the primaries stay those at entry, and every
// pass recomputes the secondary molalities (written to the SoA state, as the
// reference writes rt_auxvar%sec_molal).  More than 50 passes, or a negative
// corrected I, is the reference's failure (:4420-4431, 4477-4486): free
// molalities and coefficients become NaN and 1 is returned (the reference
// keeps looping on NaN; this stops).
template <int MAXN>
__device__ int plain_activity_newton(const DevNet& n, const DevState& s, int64_t c,
                                     PlainCell<MAXN>& w) {
  const int naq = w.naq, ms = n.cplx_ms;
  double lna[MAXN];
  double sum_pri = 0.0, fpri = 0.0;
  for (int i = 0; i < naq; ++i) {
    lna[i] = log(w.x[i]) + log(w.gp[i]);
    fpri += w.x[i] * n.Zp[i] * n.Zp[i];
    if (i != n.h2o_aq_id) sum_pri += w.x[i];
  }
  double II = 0.0;
  int err = 0;
  for (int it = 1;; ++it) {
    if (it > 50) err = 1;
    double I = fpri;
    for (int k = 0; k < n.nc; ++k) {
      if (!plain_cplx_on(n, w, k)) continue;
      I += s.sec_molal[k * s.ld + c] * n.cplx_Z[k] * n.cplx_Z[k];
    }
    I = 0.5 * I;
    const double f = I;
    if (fabs(I - II) < 1.0e-6 * I) break;
    if (err) break;
    if (n.nc > 0) {
      double didi = 0.0;
      const double sI = sqrt(I);
      for (int k = 0; k < n.nc; ++k) {
        if (!plain_cplx_on(n, w, k)) continue;
        const double zk = n.cplx_Z[k];
        if (fabs(zk) > 0.0) {
          const double t = 1.0 + n.B * n.cplx_a0[k] * sI;
          double sum = 0.5 * n.A * zk * zk / (sI * t * t) - n.Bdot;
          const int* ids = n.cplx_ids + k * (ms + 1);
          const double* st = n.cplx_st + k * ms;
          for (int q = 0; q < ids[0]; ++q) {
            const int j = ids[1 + q];
            const double zj = n.Zp[j];
            if (fabs(zj) > 0.0) {
              const double tj = 1.0 + n.B * n.a0p[j] * sI;
              sum += st[q] * (-0.5 * n.A * zj * zj / (sI * tj * tj) + n.Bdot);
            }
          }
          didi += 0.5 * zk * zk * s.sec_molal[k * s.ld + c] * kLogToLn * sum;
        }
      }
      const double den = 1.0 - didi;
      II = fabs(den) > 0.0 ? (f - I * didi) / den : f;
    } else {
      II = f;
    }
    if (II < 0.0) err = 1;
    I = II;
    const double sI = sqrt(I);
    for (int i = 0; i < naq; ++i)
      w.gp[i] = fabs(n.Zp[i]) > 0.0 ? dh_gamma(n.Zp[i], n.a0p[i], sI, I, n) : 1.0;
    double sum_sec = 0.0;
    for (int k = 0; k < n.nc; ++k) {
      if (!plain_cplx_on(n, w, k)) continue;
      const double g = fabs(n.cplx_Z[k]) > 0.0 ? dh_gamma(n.cplx_Z[k], n.cplx_a0[k], sI, I, n) : 1.0;
      s.sec_act_coef[k * s.ld + c] = g;
      double logK = n.cplx_logK[k];
      if (!n.isothermal && n.has_cplx_logKcoef) logK = interp_logk(n.cplx_logKcoef + 5 * k, w.temp);
      double lnQK = -logK * kLogToLn;
      if (n.has_cplx_h2o && n.cplx_h2o[k] != 0.0) lnQK += n.cplx_h2o[k] * w.ln_act_h2o;
      const int* ids = n.cplx_ids + k * (ms + 1);
      const double* st = n.cplx_st + k * ms;
      for (int q = 0; q < ids[0]; ++q) lnQK += st[q] * lna[ids[1 + q]];
      const double mk = exp(lnQK) / g;
      s.sec_molal[k * s.ld + c] = mk;
      sum_sec += mk;
    }
    if (n.act_h2o) {
      const double a = 1.0 - 0.017 * (sum_pri + sum_sec);
      w.ln_act_h2o = a > 0.0 ? log(a) : 0.0;
    }
  }
  if (err) {
    const double nan = __builtin_nan("");
    for (int i = 0; i < naq; ++i) {
      w.x[i] = nan;
      w.gp[i] = nan;
    }
    for (int k = 0; k < n.nc; ++k)
      if (plain_cplx_on(n, w, k)) s.sec_act_coef[k * s.ld + c] = nan;
  }
  return err;
}

// RActivityCoefficients with the network's algorithm; 1 on failure
template <int MAXN>
__device__ int plain_activity(const DevNet& n, const DevState& s, int64_t c, PlainCell<MAXN>& w) {
  if (n.act_alg == PFCHEM_ACT_COEF_ALGORITHM_NEWTON) return plain_activity_newton(n, s, c, w);
  plain_activity_lag(n, s, c, w);
  return 0;
}

// Per-cell SomDec stoichiometry of one step: upstream N:C, mineralization C
// and N coefficients, downstream N:C of every reaction (reset from the
// descriptor's Setup values at the start of the step).
struct PlainSdState {
  double unc[kSdMaxRxn], mcs[kSdMaxRxn], mns[kSdMaxRxn], dnc[kSdMaxRxn * kSdDown];
  double f_t, f_w;
};

// SomDec accumulator over the thread's own residual and linear Jacobian
// (d res / d m for aqueous, d res / d c for immobile columns).  The SomDec
// networks carry no complexes, so d total_i / d m_j = rho_w / 1000 on the
// diagonal only.
template <int MAXN>
struct PlainSdEnv {
  static constexpr bool kExt = true;
  PlainCell<MAXN>& w;
  PlainSdState& sd;
  int naq, N;
  double dkl, theta, vol, sat, temp, f_t, f_w;
  bool has_f_t, has_f_w;
  __device__ double total(int i) const { return w.total[i]; }
  __device__ double imm(int i) const { return w.imm[i]; }
  __device__ double lnact(int i) const { return log(w.x[i]) + log(w.gp[i]); }
  __device__ double dtot(int i, int j) const { return i == j ? dkl : 0.0; }
  __device__ void R(int row, double v) { w.res[row] += v; }
  __device__ void J(int row, int col, double v) { w.J[row * N + col] += v; }
  __device__ double& unc(int r) { return sd.unc[r]; }
  __device__ double& mcs(int r) { return sd.mcs[r]; }
  __device__ double& mns(int r) { return sd.mns[r]; }
  __device__ double& dnc(int r, int j) { return sd.dnc[r * kSdDown + j]; }
};

// SomDecReact (reaction_sandbox_somdec.F90:1504-1910) then SomDecNemission
// (:3477-3640) on the thread's cell
template <int MAXN>
__device__ void plain_somdec(const DevNet& n, const DevState& s, PlainCell<MAXN>& w,
                             PlainSdState& sd, double dt) {
  PlainSdEnv<MAXN> e{w,     sd,     w.naq, w.N,     w.den * 1.0e-3, w.por * w.sat, w.vol,
                     w.sat, w.temp, sd.f_t, sd.f_w,  s.somdec_f_t != nullptr,
                     s.somdec_f_w != nullptr};
  double nmin = 0.0;
  for (int r = 0; r < n.sd.nrxn; ++r) nmin += sd_reaction(n.sd, e, r, dt, true);
  sd_nemission(n.sd, e, nmin, dt, true);
}

// RReact (reaction.F90:3742-4055) for the plain kernel.  Returns the fail
// reason (0 = converged); *nit receives the Newton iterations.
template <int MAXN, bool SD>
__device__ int plain_rreact(const DevNet& n, const DevState& s, int64_t c,
                            PlainCell<MAXN>& w, PlainSdState* sd, const double* guess,
                            double dt, int* nit) {
  const int naq = w.naq, nimm = n.nimm, N = w.N;
  double fixed[MAXN], initial[MAXN], upd[MAXN], prev[MAXN];
  double r0 = 0.0;
  *nit = 0;
  if (!n.isothermal && n.has_mnrl_logKcoef)
    for (int m = 0; m < n.nm; ++m) w.logKm[m] = interp_logk(n.mnrl_logKcoef + 5 * m, w.temp);
  // RTAccumulation of total^k / immobile^k
  const bool wet = !(w.sat < kRtMinSaturation);
  const double psv = w.por * w.sat * 1000.0 * w.vol;
  for (int i = 0; i < N; ++i) fixed[i] = 0.0;
  if (wet) {
    for (int i = 0; i < naq; ++i) fixed[i] = psv * w.total[i];
    for (int i = 0; i < nimm; ++i) fixed[naq + i] += w.imm[i] * w.vol;
  }
  for (int i = 0; i < naq; ++i) initial[i] = w.total[i];
  for (int i = 0; i < nimm; ++i) initial[naq + i] = w.imm[i];
  for (int i = 0; i < naq; ++i) w.x[i] = guess[i];
  for (int i = 0; i < nimm; ++i) {
    w.imm[i] = guess[naq + i];
    w.x[naq + i] = guess[naq + i];
  }
  for (;;) {
    ++*nit;
    int act_err = 0;
    if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER) act_err = plain_activity(n, s, c, w);
    plain_total(n, s, c, w, 1);
    if (*nit > n.max_it) {
      for (int i = 0; i < naq; ++i) w.total[i] = initial[i];
      for (int i = 0; i < nimm; ++i) {
        w.imm[i] = initial[naq + i];
        w.x[naq + i] = initial[naq + i];
      }
      return kMaxIter;
    }
    // RTAccumulation / RTAccumulationDerivative
    for (int i = 0; i < N; ++i) w.res[i] = 0.0;
    if (wet) {
      const double psvd = psv / dt;
      for (int i = 0; i < naq; ++i) {
        w.res[i] = psv * w.total[i];
        for (int j = 0; j < naq; ++j) w.J[i * N + j] *= psvd;
        for (int j = naq; j < N; ++j) w.J[i * N + j] = 0.0;
      }
      for (int i = 0; i < nimm; ++i) {
        const int d = naq + i;
        w.res[d] += w.imm[i] * w.vol;
        for (int j = 0; j < N; ++j) w.J[d * N + j] = 0.0;
        w.J[d * N + d] = w.vol / dt;
      }
    } else {
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) w.J[i * N + j] = (i == j) ? 1.0 : 0.0;
    }
    for (int i = 0; i < N; ++i) w.res[i] = (w.res[i] - fixed[i]) / dt;
    // RReaction (reaction.F90:4059-4130)
    if (wet) {
      if (n.nm > 0) plain_minerals(n, s, c, w, w.res, w.J, 1);
      for (int r = 0; r < n.ndecay; ++r) {
        const int ic = n.decay_id[r];
        const double kc = n.decay_k[r] * w.vol;
        const int d = naq + ic;
        w.res[d] += kc * w.imm[ic];
        w.J[d * N + d] += kc;
      }
      if constexpr (SD) plain_somdec(n, s, w, *sd, dt);
    }
    // an activity-coefficient failure ends the RReact call (reaction.F90:
    // 3922-3936)
    if (act_err) return kActivity;
    double nr = 0.0, mx = 0.0;
    for (int i = 0; i < N; ++i) {
      nr += w.res[i] * w.res[i];
      mx = fmax(mx, fabs(w.res[i]));
    }
    nr = sqrt(nr);
    if (*nit == 1) r0 = nr;
    if (mx < n.tol_abs) break;
    if (nr / r0 < n.tol_rel) break;
    // RSolve (reaction.F90:5457-5516), back-substituting form
    for (int i = 0; i < N; ++i) prev[i] = w.x[i];
    for (int i = 0; i < N; ++i) {
      double nm = 0.0;
      for (int j = 0; j < N; ++j) nm = fmax(nm, fabs(w.J[i * N + j]));
      nm = 1.0 / fmax(nm, 1.0);
      upd[i] = w.res[i] * nm;
      for (int j = 0; j < N; ++j) w.J[i * N + j] *= nm;
    }
    if (n.log_form)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) w.J[i * N + j] *= prev[j];
    if (crout_lu_solve<MAXN>(w.J, N, upd)) return kSingular;
    double mrc = -1.0;
    if (n.log_form) {
      for (int i = 0; i < N; ++i) {
        const double u = copysign(1.0, upd[i]) * fmin(fabs(upd[i]), n.max_dlnC);
        upd[i] = prev[i] * exp(-u);
      }
    } else {
      double mr = 1.0e308;
      for (int i = 0; i < N; ++i)
        if (prev[i] <= upd[i]) mr = fmin(mr, fabs(prev[i] / upd[i]));
      for (int i = 0; i < N; ++i)
        upd[i] = prev[i] - (mr < 1.0 ? upd[i] * mr * 0.99 : upd[i]);
    }
    for (int i = 0; i < N; ++i) {
      const double rc = fabs((upd[i] - prev[i]) / prev[i]);
      if (rc == rc) mrc = fmax(mrc, rc);
    }
    if (mrc >= 0.0 && mrc < n.tol_change) break;
    for (int i = 0; i < N; ++i) w.x[i] = upd[i];
    for (int i = 0; i < nimm; ++i) w.imm[i] = upd[naq + i];
  }
  plain_total(n, s, c, w, 0);
  return kOk;
}

template <int MAXN, bool SD>
__global__ void __launch_bounds__(128)
plain_step_kernel(const DevNet* np, DevState s, int64_t ncell, double dt) {
  const DevNet& n = *np;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int naq = n.variable ? min(max(s.ncomp_active[c], 1), n.naq) : n.naq;
  const int nimm = n.nimm, N = naq + nimm;
  PlainCell<MAXN> w;
  w.naq = naq;
  w.N = N;
  PlainSdState sd;
  if constexpr (SD) {
    for (int r = 0; r < n.sd.nrxn; ++r) {
      sd.unc[r] = n.sd.upstream_nc[r];
      sd.mcs[r] = n.sd.mineral_c_stoich[r];
      sd.mns[r] = n.sd.mineral_n_stoich[r];
      for (int j = 0; j < kSdDown; ++j) sd.dnc[r * kSdDown + j] = n.sd.downstream_nc[r * kSdDown + j];
    }
    sd.f_t = s.somdec_f_t ? s.somdec_f_t[c] : 1.0;
    sd.f_w = s.somdec_f_w ? s.somdec_f_w[c] : 1.0;
  }
  for (int i = 0; i < naq; ++i) {
    w.x[i] = s.pri_molal[i * s.ld + c];
    w.total[i] = s.total[i * s.ld + c];
    w.gp[i] = s.pri_act_coef[i * s.ld + c];
  }
  for (int i = 0; i < nimm; ++i) {
    w.imm[i] = s.immobile[i * s.ld + c];
    w.x[naq + i] = w.imm[i];
  }
  for (int m = 0; m < n.nm; ++m) {
    w.vf[m] = s.mnrl_volfrac[m * s.ld + c];
    w.logKm[m] = n.mnrl_logK[m];
    w.rate[m] = 0.0;
  }
  w.ln_act_h2o = s.ln_act_h2o ? s.ln_act_h2o[c] : 0.0;
  w.por = s.porosity[c];
  w.sat = s.saturation[c];
  w.den = s.den_kg[c];
  w.vol = s.volume[c];
  w.temp = s.temperature ? s.temperature[c] : 25.0;

  int nsub = 0, nit_total = 0, status = 0;
  if (!n.full_geochem) {
    for (int i = 0; i < naq; ++i) w.x[i] = w.total[i] / w.den * 1.0e3;
  } else {
    // RStep (reaction.F90:3564-3738)
    double guess[MAXN], small_val[MAXN];
    uint64_t small = 0;
    for (int i = 0; i < naq; ++i)
      if (w.total[i] <= 1.0e-40) {
        small |= 1ull << i;
        small_val[i] = w.total[i];
        w.total[i] = 1.0e-40;
      }
    for (int i = 0; i < nimm; ++i)
      if (w.imm[i] <= 1.0e-40) {
        small |= 1ull << (naq + i);
        small_val[naq + i] = w.imm[i];
        w.imm[i] = 1.0e-40;
      }
    for (int i = 0; i < N; ++i) guess[i] = w.x[i];
    if (n.total_as_guess)
      for (int i = 0; i < naq; ++i) guess[i] = w.total[i];
    double target = dt, cum = 0.0, tdt = dt;
    int nconst = 0, ncuts = 0, attempts = 0, last_why = 0;
    for (;;) {
      if (cum >= target) break;
      if (++attempts > n.max_calls) {
        status = last_why ? last_why : kMaxIter;
        break;
      }
      int inner = 0;
      const int why = plain_rreact<MAXN, SD>(n, s, c, w, &sd, guess, tdt, &inner);
      nit_total += inner;
      if (why) {
        last_why = why;
        ++ncuts;
        if (ncuts > n.max_cuts) {
          status = why;
          break;
        }
        tdt *= 0.5;
        nconst = 0;
      } else {
        // RUpdateKineticState -> MineralUpdateKineticState
        if (n.nm > 0) {
          double rdum[MAXN];
          for (int i = 0; i < N; ++i) rdum[i] = 0.0;
          plain_minerals(n, s, c, w, rdum, (double*)nullptr, 0);
          for (int m = 0; m < n.nm; ++m) {
            w.vf[m] += w.rate[m] * n.mnrl_Vm[m] * tdt;
            if (w.vf[m] < 0.0) w.vf[m] = 0.0;
          }
        }
        cum += tdt;
        ++nsub;
        ++nconst;
        for (int i = 0; i < N; ++i) guess[i] = w.x[i];
        if (nconst >= 4) {
          --ncuts;
          tdt = fmin(2.0 * tdt, target - cum);
        }
      }
    }
    if (!status) {
      for (int i = 0; i < naq; ++i)
        if (small & (1ull << i)) w.total[i] = small_val[i];
      for (int i = 0; i < nimm; ++i)
        if (small & (1ull << (naq + i))) w.imm[i] = small_val[naq + i];
    }
  }
  for (int i = 0; i < naq; ++i) {
    s.pri_molal[i * s.ld + c] = w.x[i];
    s.total[i * s.ld + c] = w.total[i];
  }
  for (int i = 0; i < nimm; ++i) s.immobile[i * s.ld + c] = w.imm[i];
  for (int m = 0; m < n.nm; ++m) {
    if (!plain_mnrl_on(n, w, m)) continue;
    s.mnrl_volfrac[m * s.ld + c] = w.vf[m];
    if (s.mnrl_rate) s.mnrl_rate[m * s.ld + c] = w.rate[m];
  }
  if (s.num_iterations) s.num_iterations[c] = nit_total;
  if (s.num_sub_steps) s.num_sub_steps[c] = nsub;
  if (s.status) s.status[c] = status;
}

template <int MAXN>
__global__ void __launch_bounds__(256)
activity_kernel(const DevNet* np, DevState s, int64_t ncell) {
  const DevNet& n = *np;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  PlainCell<MAXN> w;
  w.naq = n.variable ? min(max(s.ncomp_active[c], 1), n.naq) : n.naq;
  w.N = w.naq + n.nimm;
  for (int i = 0; i < w.naq; ++i) w.x[i] = s.pri_molal[i * s.ld + c];
  w.ln_act_h2o = s.ln_act_h2o ? s.ln_act_h2o[c] : 0.0;
  for (int i = 0; i < w.naq; ++i) w.gp[i] = s.pri_act_coef[i * s.ld + c];
  const int err = plain_activity(n, s, c, w);
  for (int i = 0; i < w.naq; ++i) s.pri_act_coef[i * s.ld + c] = w.gp[i];
  if (err)
    for (int i = 0; i < w.naq; ++i) s.pri_molal[i * s.ld + c] = w.x[i];
  if (s.ln_act_h2o && n.act_h2o) s.ln_act_h2o[c] = w.ln_act_h2o;
}


// ln Q/K of a constraint reaction row (mineral reaction.F90:1756-1771, gas
// :1780-1793) at the cell's molalities and activity coefficients
template <int MAXN>
__device__ double con_lnqk(const DevNet& n, const DevConstraint& k, const PlainCell<MAXN>& w,
                           int r) {
  double logK = k.logK[r];
  if (!n.isothermal && k.has_coef) logK = interp_logk(k.coef + 5 * r, w.temp);
  double lnQK = -logK * kLogToLn;
  if (k.has_h2o && k.h2o[r] != 0.0) lnQK += k.h2o[r] * w.ln_act_h2o;
  const int* ids = k.ids + r * (k.ms + 1);
  for (int j = 0; j < ids[0]; ++j) {
    const int id = ids[1 + j];
    lnQK += k.st[r * k.ms + j] * log(w.x[id] * w.gp[id]);
  }
  return lnQK;
}

// ReactionEquilibrateConstraint (reaction.F90:1328-2117) for one cell per
// thread: constraint set-up, then the Newton loop with the two-stage
// convergence test (activity coefficients switched on once the unit-
// activity problem has converged).  J holds the naq x naq constraint
// Jacobian (row stride naq).
template <int MAXN>
__global__ void __launch_bounds__(128)
equilibrate_kernel(const DevNet* np, const DevConstraint* kp, DevState s, const double* conc,
                   int64_t ncell) {
  const DevNet& n = *np;
  const DevConstraint& k = *kp;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int naq = n.naq;
  PlainCell<MAXN> w;
  w.naq = naq;
  w.N = naq;
  w.den = s.den_kg[c];
  w.temp = s.temperature ? s.temperature[c] : 25.0;
  double cv[MAXN], total_conc[MAXN], free_conc[MAXN], prev[MAXN];
  for (int i = 0; i < naq; ++i) cv[i] = conc[i * s.ld + c];
  // :1465-1472 (xmass = 1)
  const double to_molar = k.with_molality ? w.den / 1000.0 : 1.0;
  const double to_molal = k.with_molality ? 1.0 : 1000.0 / w.den;
  int status = PFCHEM_CELL_OK, it = 0;
  if (!n.full_geochem) {  // :1474-1478
    for (int i = 0; i < naq; ++i) {
      s.pri_molal[i * s.ld + c] = cv[i] * to_molal;
      s.total[i * s.ld + c] = cv[i] * to_molar;
    }
    if (s.num_iterations) s.num_iterations[c] = 0;
    if (s.status) s.status[c] = status;
    return;
  }
  // a fresh auxvar has unit activity coefficients; with the previous
  // solution as guess the state's coefficients are kept
  if (k.use_prev) {
    for (int i = 0; i < naq; ++i) w.gp[i] = s.pri_act_coef[i * s.ld + c];
    w.ln_act_h2o = s.ln_act_h2o ? s.ln_act_h2o[c] : 0.0;
  } else {
    for (int i = 0; i < naq; ++i) w.gp[i] = 1.0;
    for (int q = 0; q < n.nc; ++q) s.sec_act_coef[q * s.ld + c] = 1.0;
    w.ln_act_h2o = 0.0;
  }
  // :1486-1599
  for (int i = 0; i < naq; ++i) {
    free_conc[i] = k.use_prev ? s.pri_molal[i * s.ld + c] : (k.has_guess ? k.guess[i] : 1.0e-9);
    total_conc[i] = 0.0;
    switch (k.type[i]) {
      case PFCHEM_CONSTRAINT_NULL:
      case PFCHEM_CONSTRAINT_TOTAL: total_conc[i] = cv[i] * to_molar; break;
      case PFCHEM_CONSTRAINT_FREE: free_conc[i] = cv[i] * to_molal; break;
      case PFCHEM_CONSTRAINT_LOG: free_conc[i] = pow(10.0, cv[i]) * to_molal; break;
      case PFCHEM_CONSTRAINT_CHARGE_BAL:
      case PFCHEM_CONSTRAINT_MINERAL:
        if (!k.use_prev) free_conc[i] = cv[i] * to_molal;
        break;
      case PFCHEM_CONSTRAINT_PH: free_conc[i] = pow(10.0, -cv[i]); break;
      case PFCHEM_CONSTRAINT_GAS:
        if (cv[i] <= 0.0) cv[i] = pow(10.0, cv[i]);
        break;
      default: break;
    }
  }
  for (int i = 0; i < naq; ++i) w.x[i] = free_conc[i];
  int it_act_on = 0;
  bool charge_warned = false, compute_act = k.use_prev != 0;
  for (;;) {
    for (int i = 0; i < naq; ++i)  // :1625-1631
      if (k.type[i] == PFCHEM_CONSTRAINT_FREE || k.type[i] == PFCHEM_CONSTRAINT_LOG)
        w.x[i] = free_conc[i];
    if (n.act_freq != PFCHEM_ACT_COEF_FREQUENCY_OFF && compute_act)  // :1632-1638
      (void)plain_activity(n, s, c, w);
    plain_total(n, s, c, w, 1);  // :1639; dtotal (kg water / L) in J
    // rows of J: TOTAL keeps its dtotal row; the others are rebuilt.  The
    // charge-balance row needs every dtotal row, so it is built first.
    double cb_row[MAXN];
    bool has_cb = false;
    for (int i = 0; i < naq; ++i) has_cb |= k.type[i] == PFCHEM_CONSTRAINT_CHARGE_BAL;
    double cb_res = 0.0;
    if (has_cb) {
      for (int j = 0; j < naq; ++j) {
        cb_res += n.Zp[j] * w.total[j];
        double v = 0.0;
        for (int q = 0; q < naq; ++q) v += n.Zp[q] * w.J[q * naq + j];
        cb_row[j] = v;
      }
    }
    for (int i = 0; i < naq; ++i) {
      double* row = w.J + i * naq;
      const int t = k.type[i];
      if (t == PFCHEM_CONSTRAINT_NULL || t == PFCHEM_CONSTRAINT_TOTAL) {  // :1648-1655
        w.res[i] = w.total[i] - total_conc[i];
        continue;
      }
      for (int j = 0; j < naq; ++j) row[j] = 0.0;
      w.res[i] = 0.0;
      if (t == PFCHEM_CONSTRAINT_FREE || t == PFCHEM_CONSTRAINT_LOG) {  // :1680-1685
        row[i] = 1.0;
      } else if (t == PFCHEM_CONSTRAINT_CHARGE_BAL) {  // :1687-1714
        w.res[i] = cb_res;
        for (int j = 0; j < naq; ++j) row[j] = cb_row[j];
        if (w.x[i] < 1.0e-20 && !charge_warned) {
          const double z = n.Zp[i];
          if ((cb_res > 0.0 && z > 0.0) || (cb_res < 0.0 && z < 0.0)) {
            charge_warned = true;
            w.x[i] = 1.0e-3;
          }
        }
      } else if (t == PFCHEM_CONSTRAINT_PH) {  // :1716-1753
        if (k.h_ion_id >= 0) {
          w.x[i] = pow(10.0, -cv[i]) / w.gp[i];
          row[i] = 1.0;
        } else {
          const int q = -k.h_ion_id - 1, ms = n.cplx_ms;
          double logK = n.cplx_logK[q];
          if (!n.isothermal && n.has_cplx_logKcoef)
            logK = interp_logk(n.cplx_logKcoef + 5 * q, w.temp);
          double lnQK = -logK * kLogToLn;
          if (n.has_cplx_h2o && n.cplx_h2o[q] != 0.0) lnQK += n.cplx_h2o[q] * w.ln_act_h2o;
          const int* ids = n.cplx_ids + q * (ms + 1);
          for (int j = 0; j < ids[0]; ++j)
            lnQK += n.cplx_st[q * ms + j] * log(w.x[ids[1 + j]] * w.gp[ids[1 + j]]);
          lnQK += cv[i] * kLogToLn;
          const double QK = exp(lnQK);
          w.res[i] = 1.0 - QK;
          for (int j = 0; j < ids[0]; ++j)
            row[ids[1 + j]] = -QK / w.x[ids[1 + j]] * n.cplx_st[q * ms + j];
        }
      } else {  // MINERAL / GAS, :1754-1804
        const int r = k.rxn[i];
        const double lnQK = con_lnqk(n, k, w, r);
        w.res[i] = t == PFCHEM_CONSTRAINT_MINERAL ? lnQK : lnQK - log(cv[i]);
        const int* ids = k.ids + r * (k.ms + 1);
        for (int j = 0; j < ids[0]; ++j) row[ids[1 + j]] = k.st[r * k.ms + j] / w.x[ids[1 + j]];
      }
    }
    double max_res = 0.0;
    for (int i = 0; i < naq; ++i) max_res = fmax(max_res, fabs(w.res[i]));
    // :1870-1881
    const bool use_log = n.log_form ? ((it > 3 && it < 9) ? (it % 2 == 0) : true) : false;
    // RSolve (:5457-5516): row scaling, log-variable columns, LU
    double upd[MAXN];
    for (int i = 0; i < naq; ++i) {
      double nrm = 0.0;
      for (int j = 0; j < naq; ++j) nrm = fmax(nrm, fabs(w.J[i * naq + j]));
      nrm = 1.0 / fmax(1.0, nrm);
      upd[i] = w.res[i] * nrm;
      for (int j = 0; j < naq; ++j) w.J[i * naq + j] *= nrm;
    }
    if (use_log)
      for (int i = 0; i < naq; ++i)
        for (int j = 0; j < naq; ++j) w.J[i * naq + j] *= w.x[j];
    if (crout_lu_solve<MAXN>(w.J, naq, upd)) {
      status = PFCHEM_CELL_CONSTRAINT_SINGULAR;
      break;
    }
    for (int i = 0; i < naq; ++i) prev[i] = w.x[i];
    if (use_log) {  // :1930-1932
      for (int i = 0; i < naq; ++i) {
        const double u = copysign(1.0, upd[i]) * fmin(fabs(upd[i]), k.max_dlnC);
        w.x[i] *= exp(-u);
      }
    } else {  // :1933-1950
      double min_ratio = 1.7976931348623157e308;
      for (int i = 0; i < naq; ++i)
        if (prev[i] <= upd[i]) min_ratio = fmin(min_ratio, fabs(prev[i] / upd[i]));
      if (min_ratio <= 1.0)
        for (int i = 0; i < naq; ++i) upd[i] *= min_ratio * 0.99;
      for (int i = 0; i < naq; ++i) w.x[i] = prev[i] - upd[i];
    }
    double min_m = w.x[0], max_rel = 0.0;
    for (int i = 1; i < naq; ++i) min_m = fmin(min_m, w.x[i]);
    ++it;
    if (min_m <= 0.0) {  // :1954-1975
      status = PFCHEM_CELL_CONSTRAINT_NONPOSITIVE;
      break;
    }
    for (int i = 0; i < naq; ++i) max_rel = fmax(max_rel, fabs((w.x[i] - prev[i]) / prev[i]));
    if (it >= k.max_it) {  // :2002-2016
      status = PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS;
      break;
    }
    if (max_res < n.tol_abs && max_rel < n.tol_change) {  // :2017-2030
      if (compute_act && it - it_act_on > 1) break;
      if (!compute_act) it_act_on = it;
      compute_act = true;
    }
  }
  for (int i = 0; i < naq; ++i) {
    s.pri_molal[i * s.ld + c] = w.x[i];
    s.total[i * s.ld + c] = w.total[i];
    s.pri_act_coef[i * s.ld + c] = w.gp[i];
  }
  if (s.ln_act_h2o) s.ln_act_h2o[c] = w.ln_act_h2o;
  if (s.num_iterations) s.num_iterations[c] = it;
  if (s.status) s.status[c] = status;
  if (status != PFCHEM_CELL_OK) atomicAdd(k.nfail, 1);
}

}  // namespace

int plain_max_ncomp() { return 64; }

hipError_t launch_plain_step(const DevNet& n, const DevNet* dn, const DevState& s, int64_t ncell,
                             double dt, hipStream_t stream) {
  const int block = 128;
  const dim3 grid((unsigned)((ncell + block - 1) / block));
  if (n.has_somdec) {
    return hipErrorInvalidValue;  // refused by pfchem_step (SomDec runs in the team kernels)
  } else if (n.ncomp <= 4) {
    plain_step_kernel<4, false><<<grid, block, 0, stream>>>(dn, s, ncell, dt);
  } else if (n.ncomp <= 8) {
    plain_step_kernel<8, false><<<grid, block, 0, stream>>>(dn, s, ncell, dt);
  } else if (n.ncomp <= 16) {
    plain_step_kernel<16, false><<<grid, block, 0, stream>>>(dn, s, ncell, dt);
  } else if (n.ncomp <= 32) {
    plain_step_kernel<32, false><<<grid, block, 0, stream>>>(dn, s, ncell, dt);
  } else {
    plain_step_kernel<64, false><<<grid, block, 0, stream>>>(dn, s, ncell, dt);
  }
  return hipGetLastError();
}

hipError_t launch_activity_lag(const DevNet& n, const DevNet* dn, const DevState& s, int64_t ncell,
                               hipStream_t stream) {
  const int block = 256;
  const dim3 grid((unsigned)((ncell + block - 1) / block));
  if (n.naq <= 16)
    activity_kernel<16><<<grid, block, 0, stream>>>(dn, s, ncell);
  else
    activity_kernel<64><<<grid, block, 0, stream>>>(dn, s, ncell);
  return hipGetLastError();
}

hipError_t launch_equilibrate(const DevNet& n, const DevNet* dn, const DevConstraint* dk,
                              const DevState& s, const double* conc, int64_t ncell,
                              hipStream_t stream) {
  if (equil_team_lanes(n) > 0) return launch_equil_team(n, dn, dk, s, conc, ncell, stream);
  const int block = 128;
  const dim3 grid((unsigned)((ncell + block - 1) / block));
  if (n.naq <= 4)
    equilibrate_kernel<4><<<grid, block, 0, stream>>>(dn, dk, s, conc, ncell);
  else if (n.naq <= 8)
    equilibrate_kernel<8><<<grid, block, 0, stream>>>(dn, dk, s, conc, ncell);
  else if (n.naq <= 16)
    equilibrate_kernel<16><<<grid, block, 0, stream>>>(dn, dk, s, conc, ncell);
  else if (n.naq <= 32)
    equilibrate_kernel<32><<<grid, block, 0, stream>>>(dn, dk, s, conc, ncell);
  else
    equilibrate_kernel<64><<<grid, block, 0, stream>>>(dn, dk, s, conc, ncell);
  return hipGetLastError();
}

}  // namespace pfchem
