/*
 * pfchem_oracle_adjoint.c -- FP64 CPU reference of the tangent and adjoint
 * of one reaction step (test infrastructure, like the rest of oracle/).
 *
 * One step is RStep (reaction.F90:3564-3738): accepted sub-steps s = 1..S of
 * size dt_s, each solving the RReact system (reaction.F90:3742-4055)
 *
 *   F(m; T0, VF0, p) = psv (T(m; p) - T0) / dt_s + vol sum_n nu_n Im_n(m; p) = 0
 *
 * for the free molalities m, from the start totals T0 (mol/L) and mineral
 * volume fractions VF0 of the sub-step, followed by the kinetic-state update
 * VF = max(0, VF0 + Im(m) V_m dt_s) (reaction_mineral.F90:1456-1525).  The
 * next sub-step starts from T(m) and VF.  Parameters are the log10 rate
 * constants of the kinetic minerals and the log10 K of the complexes:
 *
 *   dT_i / dlog10 K_k = -ln10 nu_ki m_k rho/1000 ,
 *   dIm_n / dlog10 k_n = ln10 Im_n .
 *
 * Tangent: dm_s = -J^-1 (dF/dp dp + dF/dT0 dT0_s), propagated through the
 * sub-steps.  Adjoint: one transposed solve J^T lambda = dbar_m per
 * sub-step, swept in reverse (implicit-function theorem at each converged
 * sub-step).  J = dF/dm at the converged m: psv/dt dtotal + the TST rate
 * derivatives.  The reference assembles its mineral block with an extra
 * rho/1000 (dQK_dmj, reaction_mineral.F90:945); exact_jacobian = 0 keeps
 * that factor (the tangent "from PFLOTRAN's analytic Jacobian"), 1 drops it
 * (the exact derivative, which finite differences of the step reproduce).
 *
 * Scope: aqueous networks with kinetic minerals under plain TST (optional
 * Arrhenius), activity coefficients held fixed during the step (LAG or
 * OFF / TIMESTEP frequency), wet cells.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pfchem_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define LN10 2.30258509299 /* LOG_TO_LN, pflotran_constants.F90:84 */

static int sens_supported(const pfchem_network* net, const pfchem_options* opt) {
  if (net->nimmobile > 0 || net->somdec) return 0;
  if (opt->act_coef_update_frequency == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER) return 0;
  if (net->kinmnrl_Temkin_const || net->kinmnrl_min_scale_factor ||
      net->kinmnrl_affinity_power || net->kinmnrl_rate_limiter ||
      net->kinmnrl_affinity_threshold || net->kinmnrl_irreversible)
    return 0;
  return 1;
}

/* plain TST rate of mineral n at the cell's molalities: Im (mol/m^3 bulk/s,
   as mnrl_rate) and dIm/dlnQK; returns 0 for an inactive mineral
   (reaction_mineral.F90:760-840) */
static int tst_rate(const pfchem_network* net, const orc_cell* c, int n, double* Im,
                    double* dIm_dlnQK) {
  const int ms = net->kinmnrl_maxsp;
  const int32_t* ids = net->kinmnrlspecid + (size_t)n * (ms + 1);
  const double* st = net->kinmnrlstoich + (size_t)n * ms;
  double lnQK = -c->kinmnrl_logK[n] * LN10;
  if (net->kinmnrlh2ostoich && net->kinmnrlh2ostoich[n] != 0.0)
    lnQK += net->kinmnrlh2ostoich[n] * c->ln_act_h2o;
  for (int a = 0; a < ids[0]; ++a) {
    const int j = ids[1 + a];
    lnQK += st[a] * (log(c->pri_molal[j]) + log(c->pri_act_coef[j]));
  }
  const double QK = exp(lnQK);
  const double aff = 1.0 - QK;
  *Im = 0.0;
  *dIm_dlnQK = 0.0;
  if (!(c->mnrl_volfrac[n] > 0.0 || aff < 0.0)) return 0;
  double arr = 1.0;
  const double Ea = net->kinmnrl_activation_energy ? net->kinmnrl_activation_energy[n] : 0.0;
  if (Ea > 0.0)
    arr = exp(Ea / 8.31446 * (1.0 / (25.0 + 273.15) - 1.0 / (c->temp + 273.15)));
  const double kr = net->kinmnrl_rate_constant[n] * arr;
  *Im = -c->mnrl_area[n] * copysign(1.0, aff) * fabs(aff) * kr;
  *dIm_dlnQK = c->mnrl_area[n] * kr * QK;
  return 1;
}

typedef struct {
  int naq, nc, nm, ms, mms;
  double psv, rho;
} sens_dims;

/* the cell at the converged molalities m of a sub-step: totals, secondary
   molalities, dtotal; mineral rates with volume fractions vf0 */
static void at_state(const pfchem_network* net, orc_cell* c, const double* m, const double* vf0,
                     double* Im, double* dI, int* act) {
  for (int i = 0; i < net->naqcomp; ++i) c->pri_molal[i] = m[i];
  for (int n = 0; n < net->nkinmnrl; ++n) c->mnrl_volfrac[n] = vf0[n];
  orc_total(net, c);
  for (int n = 0; n < net->nkinmnrl; ++n) act[n] = tst_rate(net, c, n, &Im[n], &dI[n]);
}

/* J = dF/dm (row-major, naq x naq) */
static void jacobian(const pfchem_network* net, const orc_cell* c, const sens_dims* d,
                     double dt, const double* dI, const int* act, int exact, double* J) {
  const int naq = d->naq;
  for (int i = 0; i < naq; ++i)
    for (int j = 0; j < naq; ++j) J[i * naq + j] = d->psv / dt * c->dtotal[i * naq + j];
  const double f = exact ? 1.0 : d->rho * 1.0e-3;
  for (int n = 0; n < d->nm; ++n) {
    if (!act[n]) continue;
    const int32_t* ids = net->kinmnrlspecid + (size_t)n * (d->mms + 1);
    const double* st = net->kinmnrlstoich + (size_t)n * d->mms;
    for (int a = 0; a < ids[0]; ++a)
      for (int b = 0; b < ids[0]; ++b) {
        const int i = ids[1 + a], j = ids[1 + b];
        J[i * naq + j] += c->volume * st[a] * dI[n] * st[b] / c->pri_molal[j] * f;
      }
  }
}

static int solve(double* A, int n, double* b) {
  int indx[ORC_MAXCOMP], d;
  if (orc_lu_decomposition(A, n, indx, &d)) return 1;
  orc_lu_back_substitution(A, n, indx, b);
  return 0;
}

/* forward step with the sub-step log; returns orc_rstep's error or 2/3 */
static int forward(const pfchem_network* net, const pfchem_options* opt, orc_cell* c,
                   double dt, orc_substep_log* log, sens_dims* d) {
  if (!sens_supported(net, opt)) return 2;
  if (c->sat < 1.0e-40) return 2;
  d->naq = net->naqcomp;
  d->nc = net->neqcplx;
  d->nm = net->nkinmnrl;
  d->ms = net->eqcplx_maxsp;
  d->mms = net->kinmnrl_maxsp;
  d->psv = c->porosity * c->sat * 1000.0 * c->volume;
  d->rho = c->den_kg;
  double guess[ORC_MAXCOMP];
  for (int i = 0; i < net->naqcomp; ++i) guess[i] = c->pri_molal[i];
  int ns, ni, nk;
  orc_substep_log* saved = orc_log;
  orc_log = log;
  const int err = orc_rstep(net, opt, c, guess, dt, &ns, &ni, &nk, NULL);
  orc_log = saved;
  if (err) return 1;
  if (log->overflow) return 3;
  return 0;
}

static double objective(const orc_cell* c, const sens_dims* d, const double* wv,
                        const double* ws) {
  double f = 0.0;
  for (int n = 0; n < d->nm; ++n) f += wv[n] * c->mnrl_volfrac[n];
  for (int k = 0; k < d->nc; ++k) f += ws[k] * c->sec_molal[k];
  return f;
}

int orc_cell_adjoint(const pfchem_network* net, const pfchem_options* opt, orc_cell* c,
                     double dt, const double* wv, const double* ws, int exact_jacobian,
                     double* f, double* grad) {
  orc_substep_log* log = (orc_substep_log*)malloc(sizeof(orc_substep_log));
  sens_dims d;
  const int err = forward(net, opt, c, dt, log, &d);
  const int np = net->nkinmnrl + net->neqcplx;
  for (int q = 0; q < np; ++q) grad[q] = 0.0;
  if (err) {
    free(log);
    return err;
  }
  const int naq = d.naq, nm = d.nm;
  double vf_end[ORC_MAXMNRL];
  for (int n = 0; n < nm; ++n) vf_end[n] = c->mnrl_volfrac[n];
  /* the sub-step log holds m of the last sub-step; secondaries at it */
  orc_cell* w = (orc_cell*)malloc(sizeof(orc_cell));
  *w = *c;
  double Im[ORC_MAXMNRL], dI[ORC_MAXMNRL];
  int act[ORC_MAXMNRL];
  at_state(net, w, log->m[log->n - 1], log->vf0[log->n - 1], Im, dI, act);
  for (int n = 0; n < nm; ++n) w->mnrl_volfrac[n] = vf_end[n];
  *f = objective(w, &d, wv, ws);

  double gm[ORC_MAXCOMP], gT[ORC_MAXCOMP], gV[ORC_MAXMNRL];
  for (int i = 0; i < naq; ++i) gm[i] = gT[i] = 0.0;
  for (int n = 0; n < nm; ++n) gV[n] = wv[n];
  /* explicit dependence of the secondary molalities at the end */
  for (int k = 0; k < d.nc; ++k) {
    if (ws[k] == 0.0) continue;
    const double mk = w->sec_molal[k];
    grad[nm + k] += ws[k] * (-LN10 * mk);
    const int32_t* ids = net->eqcplxspecid + (size_t)k * (d.ms + 1);
    const double* st = net->eqcplxstoich + (size_t)k * d.ms;
    for (int a = 0; a < ids[0]; ++a) {
      const int j = ids[1 + a];
      gm[j] += ws[k] * mk * st[a] / w->pri_molal[j];
    }
  }
  double* J = (double*)malloc(sizeof(double) * naq * naq);
  double* JT = (double*)malloc(sizeof(double) * naq * naq);
  int rc = 0;
  for (int s = log->n - 1; s >= 0 && !rc; --s) {
    const double h = log->dt[s];
    at_state(net, w, log->m[s], log->vf0[s], Im, dI, act);
    /* kinetic-state update VF = max(0, VF0 + Im V_m h) */
    for (int n = 0; n < nm; ++n) {
      const double vf1 = log->vf0[s][n] + Im[n] * net->kinmnrl_molar_vol[n] * h;
      if (vf1 < 0.0) gV[n] = 0.0;
      if (gV[n] == 0.0 || !act[n]) continue;
      grad[n] += gV[n] * LN10 * Im[n] * net->kinmnrl_molar_vol[n] * h;
      const int32_t* ids = net->kinmnrlspecid + (size_t)n * (d.mms + 1);
      const double* st = net->kinmnrlstoich + (size_t)n * d.mms;
      for (int a = 0; a < ids[0]; ++a) {
        const int j = ids[1 + a];
        gm[j] += gV[n] * net->kinmnrl_molar_vol[n] * h * dI[n] * st[a] / w->pri_molal[j];
      }
    }
    /* totals at the end of the sub-step, T(m; p) */
    for (int j = 0; j < naq; ++j) {
      double acc = 0.0;
      for (int i = 0; i < naq; ++i) acc += gT[i] * w->dtotal[i * naq + j];
      gm[j] += acc;
    }
    for (int k = 0; k < d.nc; ++k) {
      const int32_t* ids = net->eqcplxspecid + (size_t)k * (d.ms + 1);
      const double* st = net->eqcplxstoich + (size_t)k * d.ms;
      double acc = 0.0;
      for (int a = 0; a < ids[0]; ++a) acc += gT[ids[1 + a]] * st[a];
      grad[nm + k] += acc * (-LN10 * w->sec_molal[k]) * d.rho * 1.0e-3;
    }
    /* implicit solve: J^T lambda = gm */
    jacobian(net, w, &d, h, dI, act, exact_jacobian, J);
    for (int i = 0; i < naq; ++i)
      for (int j = 0; j < naq; ++j) JT[i * naq + j] = J[j * naq + i];
    double lam[ORC_MAXCOMP];
    for (int i = 0; i < naq; ++i) lam[i] = gm[i];
    if (solve(JT, naq, lam)) {
      rc = 1;
      break;
    }
    for (int k = 0; k < d.nc; ++k) {
      const int32_t* ids = net->eqcplxspecid + (size_t)k * (d.ms + 1);
      const double* st = net->eqcplxstoich + (size_t)k * d.ms;
      double acc = 0.0;
      for (int a = 0; a < ids[0]; ++a) acc += lam[ids[1 + a]] * st[a];
      grad[nm + k] -= d.psv / h * acc * (-LN10 * w->sec_molal[k]) * d.rho * 1.0e-3;
    }
    for (int n = 0; n < nm; ++n) {
      if (!act[n]) continue;
      const int32_t* ids = net->kinmnrlspecid + (size_t)n * (d.mms + 1);
      const double* st = net->kinmnrlstoich + (size_t)n * d.mms;
      double acc = 0.0;
      for (int a = 0; a < ids[0]; ++a) acc += lam[ids[1 + a]] * st[a];
      grad[n] -= acc * w->volume * LN10 * Im[n];
    }
    for (int i = 0; i < naq; ++i) {
      gT[i] = lam[i] * d.psv / h;
      gm[i] = 0.0;
    }
  }
  free(J);
  free(JT);
  free(w);
  free(log);
  return rc;
}

int orc_cell_tangent(const pfchem_network* net, const pfchem_options* opt, orc_cell* c,
                     double dt, const double* wv, const double* ws, const double* dp,
                     int exact_jacobian, double* f, double* df) {
  orc_substep_log* log = (orc_substep_log*)malloc(sizeof(orc_substep_log));
  sens_dims d;
  const int err = forward(net, opt, c, dt, log, &d);
  *df = 0.0;
  if (err) {
    free(log);
    return err;
  }
  const int naq = d.naq, nm = d.nm;
  double vf_end[ORC_MAXMNRL];
  for (int n = 0; n < nm; ++n) vf_end[n] = c->mnrl_volfrac[n];
  orc_cell* w = (orc_cell*)malloc(sizeof(orc_cell));
  *w = *c;
  double Im[ORC_MAXMNRL], dI[ORC_MAXMNRL];
  int act[ORC_MAXMNRL];
  double dT[ORC_MAXCOMP], dV[ORC_MAXMNRL], dm[ORC_MAXCOMP];
  for (int i = 0; i < naq; ++i) dT[i] = dm[i] = 0.0;
  for (int n = 0; n < nm; ++n) dV[n] = 0.0;
  double* J = (double*)malloc(sizeof(double) * naq * naq);
  int rc = 0;
  for (int s = 0; s < log->n; ++s) {
    const double h = log->dt[s];
    at_state(net, w, log->m[s], log->vf0[s], Im, dI, act);
    /* dF/dp dp + dF/dT0 dT0 */
    double rhs[ORC_MAXCOMP], dTp[ORC_MAXCOMP];
    for (int i = 0; i < naq; ++i) dTp[i] = 0.0;
    for (int k = 0; k < d.nc; ++k) {
      const double v = dp[nm + k];
      if (v == 0.0) continue;
      const int32_t* ids = net->eqcplxspecid + (size_t)k * (d.ms + 1);
      const double* st = net->eqcplxstoich + (size_t)k * d.ms;
      for (int a = 0; a < ids[0]; ++a)
        dTp[ids[1 + a]] += st[a] * (-LN10 * w->sec_molal[k]) * d.rho * 1.0e-3 * v;
    }
    for (int i = 0; i < naq; ++i) rhs[i] = -(d.psv / h * (dTp[i] - dT[i]));
    for (int n = 0; n < nm; ++n) {
      if (!act[n] || dp[n] == 0.0) continue;
      const int32_t* ids = net->kinmnrlspecid + (size_t)n * (d.mms + 1);
      const double* st = net->kinmnrlstoich + (size_t)n * d.mms;
      for (int a = 0; a < ids[0]; ++a)
        rhs[ids[1 + a]] -= st[a] * w->volume * LN10 * Im[n] * dp[n];
    }
    jacobian(net, w, &d, h, dI, act, exact_jacobian, J);
    for (int i = 0; i < naq; ++i) dm[i] = rhs[i];
    if (solve(J, naq, dm)) {
      rc = 1;
      break;
    }
    /* totals and volume fractions at the end of the sub-step */
    for (int i = 0; i < naq; ++i) {
      double acc = dTp[i];
      for (int j = 0; j < naq; ++j) acc += w->dtotal[i * naq + j] * dm[j];
      dT[i] = acc;
    }
    for (int n = 0; n < nm; ++n) {
      const double vf1 = log->vf0[s][n] + Im[n] * net->kinmnrl_molar_vol[n] * h;
      if (vf1 < 0.0) {
        dV[n] = 0.0;
        continue;
      }
      if (!act[n]) continue;
      const int32_t* ids = net->kinmnrlspecid + (size_t)n * (d.mms + 1);
      const double* st = net->kinmnrlstoich + (size_t)n * d.mms;
      double dlnQK = 0.0;
      for (int a = 0; a < ids[0]; ++a) dlnQK += st[a] * dm[ids[1 + a]] / w->pri_molal[ids[1 + a]];
      dV[n] += net->kinmnrl_molar_vol[n] * h * (dI[n] * dlnQK + LN10 * Im[n] * dp[n]);
    }
  }
  if (!rc) {
    for (int n = 0; n < nm; ++n) w->mnrl_volfrac[n] = vf_end[n];
    *f = objective(w, &d, wv, ws);
    double acc = 0.0;
    for (int n = 0; n < nm; ++n) acc += wv[n] * dV[n];
    for (int k = 0; k < d.nc; ++k) {
      if (ws[k] == 0.0) continue;
      const int32_t* ids = net->eqcplxspecid + (size_t)k * (d.ms + 1);
      const double* st = net->eqcplxstoich + (size_t)k * d.ms;
      double dln = -LN10 * dp[nm + k];
      for (int a = 0; a < ids[0]; ++a) dln += st[a] * dm[ids[1 + a]] / w->pri_molal[ids[1 + a]];
      acc += ws[k] * w->sec_molal[k] * dln;
    }
    *df = acc;
  }
  free(J);
  free(w);
  free(log);
  return rc;
}

int64_t orc_adjoint_batch(const pfchem_network* net, const pfchem_options* opt,
                          const pfchem_state* s, int64_t ncell, double dt, const double* wv,
                          const double* ws, int exact_jacobian, int nthreads,
                          double* grad_sum, double* f, int32_t* status) {
  const int np = net->nkinmnrl + net->neqcplx;
  int64_t nfail = 0;
  for (int q = 0; q < np; ++q) grad_sum[q] = 0.0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : nfail)
  {
    orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
    double* g = (double*)malloc(sizeof(double) * (np > 0 ? np : 1));
    double* acc = (double*)calloc(np > 0 ? np : 1, sizeof(double));
#pragma omp for schedule(dynamic, 64)
    for (int64_t cell = 0; cell < ncell; ++cell) {
      orc_gather(net, s, cell, c);
      double fc = 0.0;
      const int rc = orc_cell_adjoint(net, opt, c, dt, wv, ws, exact_jacobian, &fc, g);
      if (f) f[cell] = fc;
      if (status) status[cell] = rc;
      if (rc) {
        ++nfail;
        continue;
      }
      for (int q = 0; q < np; ++q) acc[q] += g[q];
    }
#pragma omp critical
    for (int q = 0; q < np; ++q) grad_sum[q] += acc[q];
    free(acc);
    free(g);
    free(c);
  }
  return nfail;
}

int64_t orc_tangent_batch(const pfchem_network* net, const pfchem_options* opt,
                          const pfchem_state* s, int64_t ncell, double dt, const double* wv,
                          const double* ws, const double* dp, int exact_jacobian, int nthreads,
                          double* df, double* f, int32_t* status) {
  int64_t nfail = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : nfail)
  {
    orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
#pragma omp for schedule(dynamic, 64)
    for (int64_t cell = 0; cell < ncell; ++cell) {
      orc_gather(net, s, cell, c);
      double fc = 0.0, dfc = 0.0;
      const int rc = orc_cell_tangent(net, opt, c, dt, wv, ws, dp, exact_jacobian, &fc, &dfc);
      if (df) df[cell] = dfc;
      if (f) f[cell] = fc;
      if (status) status[cell] = rc;
      if (rc) ++nfail;
    }
    free(c);
  }
  return nfail;
}
