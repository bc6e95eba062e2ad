/*
 * pfchem_oracle.c -- FP64 CPU restatement of PFLOTRAN's operator-split
 * per-cell chemistry step.  TEST INFRASTRUCTURE ONLY: the checker for the
 * HIP kernels (tests/, __graft_entry__.smoke, bench.py cpu_baseline).  The
 * product library never links or calls this file.
 *
 * Each routine names the reference routine (file:line, relative to
 * src/pflotran) whose arithmetic it restates, in the same order of
 * operations, so that it agrees with the reference to rounding.
 *
 * Divergences from the reference, all intended:
 *  - RSolve always back-substitutes (the intended branch,
 *    reaction.F90:5506-5512); the reference's solve_error branch
 *    (:5501-5505, taken by RReact) never writes `update`.
 *  - RActivityCoefficients NEWTON gives up after 50 fixed-point iterations
 *    (reaction.F90:4420-4431 sets NaNs but keeps looping).
 *  - SomDec stoichiometry that the reference stores in the shared sandbox
 *    object per cell visit (reaction_sandbox_somdec.F90:1817-1858) is kept
 *    per cell, starting from the Setup values.
 *  - RStep gives up after options.maximum_rreact_calls RReact calls
 *    (0: ORC_MAX_RREACT_ATTEMPTS); the
 *    reference's sub-step loop has no bound (its cut counter can go
 *    negative after runs of good sub-steps, reaction.F90:3721-3725).
 *  - Mineral rate prefactors (reaction_mineral.F90:806-870) and sorption,
 *    gas, surface complexation, general/microbial/radioactive-decay
 *    reactions are not part of this path.
 */
#include "pfchem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* pflotran_constants.F90:84, 89; reactive_transport_aux.F90:19 */
#define LOG_TO_LN 2.30258509299
#define IDEAL_GAS_CONSTANT 8.31446
#define RT_MIN_SATURATION 1.0e-40
#define PERTURBATION_TOLERANCE 1.0e-5 /* reaction.F90:41 */

static inline int ncomp_of(const pfchem_network* n) {
  return n->naqcomp + n->nimmobile;
}

void orc_cell_defaults(const pfchem_network* net, orc_cell* c) {
  memset(c, 0, sizeof(*c));
  for (int i = 0; i < net->naqcomp; ++i) c->pri_act_coef[i] = 1.0;
  for (int k = 0; k < net->neqcplx; ++k) {
    c->sec_act_coef[k] = 1.0;
    c->eqcplx_logK[k] = net->eqcplx_logK[k];
  }
  for (int m = 0; m < net->nkinmnrl; ++m) c->kinmnrl_logK[m] = net->kinmnrl_logK[m];
  c->porosity = 1.0;
  c->sat = 1.0;
  c->den_kg = 1000.0;
  c->volume = 1.0;
  c->temp = 25.0;
  if (net->somdec) {
    const pfchem_somdec* sd = net->somdec;
    for (int r = 0; r < sd->nrxn; ++r) {
      c->sd_upstream_nc[r] = sd->upstream_nc[r];
      c->sd_mineral_c_stoich[r] = sd->mineral_c_stoich[r];
      c->sd_mineral_n_stoich[r] = sd->mineral_n_stoich[r];
      for (int j = 0; j < PFCHEM_SOMDEC_MAX_DOWN; ++j)
        c->sd_downstream_nc[r * PFCHEM_SOMDEC_MAX_DOWN + j] =
            sd->downstream_nc[r * PFCHEM_SOMDEC_MAX_DOWN + j];
    }
  }
}

/* ReactionInterpolateLogK reaction_aux.F90:1285-1312 via
   RUpdateTempDependentCoefs reaction.F90:5976-6067 (per cell here). */
static void interpolate_logk(const double* coefs, double* logk, double temp,
                             int n) {
  double tk = temp + 273.15;
  for (int i = 0; i < n; ++i) {
    const double* b = coefs + 5 * i;
    logk[i] = b[0] * log(tk) + b[1] + b[2] * tk + b[3] / tk + b[4] / (tk * tk);
  }
}

void orc_update_temp_dependent_coefs(const pfchem_network* net,
                                     const pfchem_options* opt, orc_cell* c) {
  if (opt->isothermal) return;
  if (net->eqcplx_logKcoef)
    interpolate_logk(net->eqcplx_logKcoef, c->eqcplx_logK, c->temp, net->neqcplx);
  if (net->kinmnrl_logKcoef)
    interpolate_logk(net->kinmnrl_logKcoef, c->kinmnrl_logK, c->temp,
                     net->nkinmnrl);
}

/* RActivityCoefficients reaction.F90:4368-4614 */
void orc_activity_coefficients(const pfchem_network* net,
                               const pfchem_options* opt, orc_cell* c) {
  const int naq = net->naqcomp, nc = net->neqcplx;
  const double A = net->debyeA, B = net->debyeB, Bdot = net->debyeBdot;
  const int ms = net->eqcplx_maxsp;
  double sum_pri_molal = 0.0, sum_sec_molal;
  c->activity_error = 0;

  if (opt->use_activity_h2o) {
    for (int j = 0; j < naq; ++j)
      if (j != net->h2o_aq_id) sum_pri_molal += c->pri_molal[j];
  }

  if (opt->act_coef_update_algorithm == PFCHEM_ACT_COEF_ALGORITHM_NEWTON) {
    double ln_act[ORC_MAXCOMP];
    for (int j = 0; j < naq; ++j)
      ln_act[j] = log(c->pri_molal[j]) + log(c->pri_act_coef[j]);
    double fpri = 0.0;
    for (int j = 0; j < naq; ++j)
      fpri += c->pri_molal[j] * net->primary_spec_Z[j] * net->primary_spec_Z[j];
    int it = 0;
    double II = 0.0, I, f, sqrt_I, den = 0.0;
    for (;;) {
      ++it;
      if (it > 50) { /* reaction.F90:4420-4431 (stop here, see header) */
        for (int j = 0; j < naq; ++j) {
          c->pri_molal[j] = NAN;
          c->pri_act_coef[j] = NAN;
        }
        for (int k = 0; k < nc; ++k) c->sec_act_coef[k] = NAN;
        c->activity_error = 1;
        return;
      }
      I = fpri;
      for (int k = 0; k < nc; ++k)
        I += c->sec_molal[k] * net->eqcplx_Z[k] * net->eqcplx_Z[k];
      I = 0.5 * I;
      f = I;
      if (fabs(I - II) < 1.0e-6 * I) break;
      if (nc > 0) {
        double didi = 0.0;
        sqrt_I = sqrt(I);
        for (int k = 0; k < nc; ++k) {
          double zk = net->eqcplx_Z[k];
          if (fabs(zk) > 0.0) {
            double t = 1.0 + B * net->eqcplx_a0[k] * sqrt_I;
            double sum = 0.5 * A * zk * zk / (sqrt_I * t * t) - Bdot;
            const int32_t* ids = net->eqcplxspecid + (size_t)k * (ms + 1);
            const double* st = net->eqcplxstoich + (size_t)k * ms;
            for (int jj = 0; jj < ids[0]; ++jj) {
              int j = ids[1 + jj];
              double zj = net->primary_spec_Z[j];
              if (fabs(zj) > 0.0) {
                double tj = 1.0 + B * net->primary_spec_a0[j] * sqrt_I;
                double dgamdi = -0.5 * A * zj * zj / (sqrt_I * tj * tj) + Bdot;
                sum += st[jj] * dgamdi;
              }
            }
            double dcdi = c->sec_molal[k] * LOG_TO_LN * sum;
            didi += 0.5 * zk * zk * dcdi;
          }
        }
        den = 1.0 - didi;
        if (fabs(den) > 0.0)
          II = (f - I * didi) / den;
        else
          II = f;
      } else {
        II = f;
      }
      if (II < 0.0) { /* reaction.F90:4477-4486 */
        for (int j = 0; j < naq; ++j) {
          c->pri_molal[j] = NAN;
          c->pri_act_coef[j] = NAN;
        }
        for (int k = 0; k < nc; ++k) c->sec_act_coef[k] = NAN;
        c->activity_error = 1;
      }
      I = II;
      sqrt_I = sqrt(I);
      for (int i = 0; i < naq; ++i) {
        double z = net->primary_spec_Z[i];
        if (fabs(z) > 0.0)
          c->pri_act_coef[i] =
              exp((-z * z * sqrt_I * A / (1.0 + net->primary_spec_a0[i] * B * sqrt_I) +
                   Bdot * I) * LOG_TO_LN);
        else
          c->pri_act_coef[i] = 1.0;
      }
      sum_sec_molal = 0.0;
      for (int k = 0; k < nc; ++k) {
        double z = net->eqcplx_Z[k];
        if (fabs(z) > 0.0)
          c->sec_act_coef[k] =
              exp((-z * z * sqrt_I * A / (1.0 + net->eqcplx_a0[k] * B * sqrt_I) +
                   Bdot * I) * LOG_TO_LN);
        else
          c->sec_act_coef[k] = 1.0;
        double lnQK = -c->eqcplx_logK[k] * LOG_TO_LN;
        if (net->eqcplxh2ostoich && net->eqcplxh2ostoich[k] != 0.0)
          lnQK += net->eqcplxh2ostoich[k] * c->ln_act_h2o;
        const int32_t* ids = net->eqcplxspecid + (size_t)k * (ms + 1);
        const double* st = net->eqcplxstoich + (size_t)k * ms;
        for (int jj = 0; jj < ids[0]; ++jj) lnQK += st[jj] * ln_act[ids[1 + jj]];
        c->sec_molal[k] = exp(lnQK) / c->sec_act_coef[k];
        sum_sec_molal += c->sec_molal[k];
      }
      if (opt->use_activity_h2o) {
        double a = 1.0 - 0.017 * (sum_pri_molal + sum_sec_molal);
        c->ln_act_h2o = a > 0.0 ? log(a) : 0.0;
      }
    }
  } else {
    /* LAG, reaction.F90:4555-4611 */
    double I = 0.0;
    for (int i = 0; i < naq; ++i)
      I += c->pri_molal[i] * net->primary_spec_Z[i] * net->primary_spec_Z[i];
    for (int k = 0; k < nc; ++k)
      I += c->sec_molal[k] * net->eqcplx_Z[k] * net->eqcplx_Z[k];
    I = 0.5 * I;
    double sqrt_I = sqrt(I);
    for (int i = 0; i < naq; ++i) {
      double z = net->primary_spec_Z[i];
      if (fabs(z) > 1.0e-10)
        c->pri_act_coef[i] =
            exp((-z * z * sqrt_I * A / (1.0 + net->primary_spec_a0[i] * B * sqrt_I) +
                 Bdot * I) * LOG_TO_LN);
      else
        c->pri_act_coef[i] = 1.0;
    }
    sum_sec_molal = 0.0;
    for (int k = 0; k < nc; ++k) {
      double z = net->eqcplx_Z[k];
      if (fabs(z) > 1.0e-10)
        c->sec_act_coef[k] =
            exp((-z * z * sqrt_I * A / (1.0 + net->eqcplx_a0[k] * B * sqrt_I) +
                 Bdot * I) * LOG_TO_LN);
      else
        c->sec_act_coef[k] = 1.0;
      sum_sec_molal += c->sec_molal[k];
    }
    if (opt->use_activity_h2o) {
      double a = 1.0 - 0.017 * (sum_pri_molal + sum_sec_molal);
      c->ln_act_h2o = a > 0.0 ? log(a) : 0.0;
    }
  }
}

/* RTotal reaction.F90:4618-4661 -> RTotalAqueous :4665-4759 */
void orc_total(const pfchem_network* net, orc_cell* c) {
  const int naq = net->naqcomp, ms = net->eqcplx_maxsp;
  double ln_conc[ORC_MAXCOMP], ln_act[ORC_MAXCOMP];
  const double den_kg_per_L = c->den_kg * 1.0e-3; /* xmass = 1 */
  for (int i = 0; i < naq; ++i) {
    ln_conc[i] = log(c->pri_molal[i]);
    ln_act[i] = ln_conc[i] + log(c->pri_act_coef[i]);
    c->total[i] = c->pri_molal[i];
  }
  memset(c->dtotal, 0, sizeof(double) * naq * naq);
  for (int i = 0; i < naq; ++i) c->dtotal[i * naq + i] = 1.0;
  for (int k = 0; k < net->neqcplx; ++k) {
    double lnQK = -c->eqcplx_logK[k] * LOG_TO_LN;
    if (net->eqcplxh2ostoich && net->eqcplxh2ostoich[k] != 0.0)
      lnQK += net->eqcplxh2ostoich[k] * c->ln_act_h2o;
    const int32_t* ids = net->eqcplxspecid + (size_t)k * (ms + 1);
    const double* st = net->eqcplxstoich + (size_t)k * ms;
    const int ncs = ids[0];
    for (int i = 0; i < ncs; ++i) lnQK += st[i] * ln_act[ids[1 + i]];
    c->sec_molal[k] = exp(lnQK) / c->sec_act_coef[k];
    for (int i = 0; i < ncs; ++i) c->total[ids[1 + i]] += st[i] * c->sec_molal[k];
    for (int j = 0; j < ncs; ++j) {
      int jc = ids[1 + j];
      double tempreal = st[j] * exp(lnQK - ln_conc[jc]) / c->sec_act_coef[k];
      for (int i = 0; i < ncs; ++i)
        c->dtotal[ids[1 + i] * naq + jc] += st[i] * tempreal;
    }
  }
  for (int i = 0; i < naq; ++i) c->total[i] *= den_kg_per_L;
  for (int i = 0; i < naq * naq; ++i) c->dtotal[i] *= den_kg_per_L;
}

/* RTAccumulation reaction.F90:5710-5771 (liquid phase only) */
void orc_accumulation(const pfchem_network* net, const orc_cell* c,
                      double* res) {
  const int n = ncomp_of(net);
  for (int i = 0; i < n; ++i) res[i] = 0.0;
  if (c->sat < RT_MIN_SATURATION) return;
  double psv_t = c->porosity * c->sat * 1000.0 * c->volume;
  for (int i = 0; i < net->naqcomp; ++i) res[i] = psv_t * c->total[i];
  for (int i = 0; i < net->nimmobile; ++i)
    res[net->naqcomp + i] += c->immobile[i] * c->volume;
}

/* RTAccumulationDerivative reaction.F90:5775-5848 */
void orc_accumulation_derivative(const pfchem_network* net, const orc_cell* c,
                                 double dt, double* jac) {
  const int n = ncomp_of(net), naq = net->naqcomp;
  memset(jac, 0, sizeof(double) * n * n);
  if (c->sat < RT_MIN_SATURATION) {
    for (int i = 0; i < n; ++i) jac[i * n + i] = 1.0;
    return;
  }
  double psvd_t = c->porosity * c->sat * 1000.0 * c->volume / dt;
  for (int i = 0; i < naq; ++i)
    for (int j = 0; j < naq; ++j) jac[i * n + j] = c->dtotal[i * naq + j] * psvd_t;
  for (int i = 0; i < net->nimmobile; ++i) {
    int d = naq + i;
    jac[d * n + d] = c->volume / dt;
  }
}

/* RKineticMineral reaction_mineral.F90:647-1078 (no solid solution;
   prefactor species primary only) */
#define ORC_MAXPREF 8
static void kinetic_mineral(const pfchem_network* net, orc_cell* c, double* res,
                            double* jac, int compute_derivative) {
  const int naq = net->naqcomp, n = ncomp_of(net), ms = net->kinmnrl_maxsp;
  double ln_conc[ORC_MAXCOMP], ln_act[ORC_MAXCOMP];
  for (int i = 0; i < naq; ++i) {
    ln_conc[i] = log(c->pri_molal[i]);
    ln_act[i] = ln_conc[i] + log(c->pri_act_coef[i]);
  }
  for (int m = 0; m < net->nkinmnrl; ++m) c->mnrl_rate[m] = 0.0;
  for (int m = 0; m < net->nkinmnrl; ++m) {
    const int32_t* ids = net->kinmnrlspecid + (size_t)m * (ms + 1);
    const double* st = net->kinmnrlstoich + (size_t)m * ms;
    const int ncs = ids[0];
    double lnQK = -c->kinmnrl_logK[m] * LOG_TO_LN;
    if (net->kinmnrlh2ostoich && net->kinmnrlh2ostoich[m] != 0.0)
      lnQK += net->kinmnrlh2ostoich[m] * c->ln_act_h2o;
    for (int i = 0; i < ncs; ++i) lnQK += st[i] * ln_act[ids[1 + i]];
    double QK = exp(lnQK);
    const double* temkin = net->kinmnrl_Temkin_const;
    const double* mscale = net->kinmnrl_min_scale_factor;
    double aff;
    if (temkin) {
      if (mscale)
        aff = 1.0 - pow(QK, 1.0 / (mscale[m] * temkin[m]));
      else
        aff = 1.0 - pow(QK, 1.0 / temkin[m]);
    } else if (mscale) {
      aff = 1.0 - pow(QK, 1.0 / mscale[m]);
    } else {
      aff = 1.0 - QK;
    }
    double sgn = copysign(1.0, aff);
    double Im, Im_const, sum_prefactor_rate;
    const int npref = net->kinmnrl_num_prefactors ? net->kinmnrl_num_prefactors[m] : 0;
    const int mp = net->kinmnrl_max_pref, msp = net->kinmnrl_max_pref_spec;
    double prefactor[ORC_MAXPREF];
    if (c->mnrl_volfrac[m] > 0.0 || sgn < 0.0) {
      if (net->kinmnrl_irreversible && net->kinmnrl_irreversible[m] == 1 && sgn < 0.0)
        continue;
      if (net->kinmnrl_affinity_threshold && net->kinmnrl_affinity_threshold[m] > 0.0) {
        if (sgn < 0.0 && QK < net->kinmnrl_affinity_threshold[m]) continue;
      }
      double limiter = net->kinmnrl_rate_limiter ? net->kinmnrl_rate_limiter[m] : 0.0;
      if (limiter > 0.0) aff = aff / (1.0 + (1.0 - aff) / limiter);
      if (npref > 0) {
        /* parallel prefactor rates (:820-860) */
        sum_prefactor_rate = 0.0;
        for (int p = 0; p < npref; ++p) {
          const int32_t* pid = net->kinmnrl_prefactor_id + ((size_t)m * mp + p) * (msp + 1);
          double ln_prefactor = 0.0;
          for (int q = 0; q < pid[0]; ++q) {
            const size_t k = ((size_t)m * mp + p) * msp + q;
            const double ln_spec_act = ln_act[pid[1 + q]];
            const double ln_numerator = net->kinmnrl_pref_alpha[k] * ln_spec_act;
            const double ln_denominator =
                log(1.0 + exp(log(net->kinmnrl_pref_atten_coef[k]) +
                              net->kinmnrl_pref_beta[k] * ln_spec_act));
            ln_prefactor = ln_prefactor + ln_numerator;
            ln_prefactor = ln_prefactor - ln_denominator;
          }
          prefactor[p] = exp(ln_prefactor);
          double arrhenius = 1.0;
          const double Ea = net->kinmnrl_pref_activation_energy
                                ? net->kinmnrl_pref_activation_energy[(size_t)m * mp + p]
                                : 0.0;
          if (Ea > 0.0)
            arrhenius = exp(Ea / IDEAL_GAS_CONSTANT *
                            (1.0 / (25.0 + 273.15) - 1.0 / (c->temp + 273.15)));
          sum_prefactor_rate = sum_prefactor_rate +
                               prefactor[p] * net->kinmnrl_pref_rate[(size_t)m * mp + p] * arrhenius;
        }
      } else {
        double arrhenius = 1.0;
        double Ea = net->kinmnrl_activation_energy ? net->kinmnrl_activation_energy[m] : 0.0;
        if (Ea > 0.0)
          arrhenius = exp(Ea / IDEAL_GAS_CONSTANT *
                          (1.0 / (25.0 + 273.15) - 1.0 / (c->temp + 273.15)));
        sum_prefactor_rate = net->kinmnrl_rate_constant[m] * arrhenius;
      }
      Im_const = -c->mnrl_area[m];
      if (mscale) Im_const = Im_const / mscale[m];
      if (net->kinmnrl_affinity_power)
        Im = Im_const * sgn * pow(fabs(aff), net->kinmnrl_affinity_power[m]) *
             sum_prefactor_rate;
      else
        Im = Im_const * sgn * fabs(aff) * sum_prefactor_rate;
      c->mnrl_rate[m] = Im;
    } else {
      continue;
    }
    Im_const = Im_const * c->volume;
    Im = Im * c->volume;
    for (int i = 0; i < ncs; ++i) res[ids[1 + i]] += st[i] * Im;
    if (!compute_derivative) continue;
    double dIm_dQK;
    if (net->kinmnrl_affinity_power)
      dIm_dQK = -Im * net->kinmnrl_affinity_power[m] / fabs(aff);
    else
      dIm_dQK = -Im_const * sum_prefactor_rate;
    if (temkin) {
      if (mscale)
        dIm_dQK = dIm_dQK * (1.0 / (mscale[m] * temkin[m])) / QK * (1.0 - aff);
      else
        dIm_dQK = dIm_dQK * (1.0 / temkin[m]) / QK * (1.0 - aff);
    } else if (mscale) {
      dIm_dQK = dIm_dQK * (1.0 / mscale[m]) / QK * (1.0 - aff);
    }
    double limiter = net->kinmnrl_rate_limiter ? net->kinmnrl_rate_limiter[m] : 0.0;
    if (limiter <= 0.0) {
      for (int j = 0; j < ncs; ++j) {
        int jc = ids[1 + j];
        double dQK_dCj = st[j] * QK * exp(-ln_conc[jc]);
        double dQK_dmj = dQK_dCj * c->den_kg * 1.0e-3;
        for (int i = 0; i < ncs; ++i)
          jac[ids[1 + i] * n + jc] += st[i] * dIm_dQK * dQK_dmj;
      }
    } else {
      double den = 1.0 + (1.0 - aff) / limiter;
      for (int j = 0; j < ncs; ++j) {
        int jc = ids[1 + j];
        double dQK_dCj = st[j] * QK * exp(-ln_conc[jc]);
        double dQK_dmj = dQK_dCj * c->den_kg * 1.0e-3;
        for (int i = 0; i < ncs; ++i)
          jac[ids[1 + i] * n + jc] +=
              st[i] * dIm_dQK * (1.0 + QK / limiter / den) * dQK_dmj / den;
      }
    }
    if (npref > 0) {
      /* derivative through the prefactors (:964-1033, primary species) */
      const double dIm_dsum_prefactor_rate = Im / sum_prefactor_rate;
      for (int p = 0; p < npref; ++p) {
        double arrhenius = 1.0;
        const double Ea = net->kinmnrl_pref_activation_energy
                              ? net->kinmnrl_pref_activation_energy[(size_t)m * mp + p]
                              : 0.0;
        if (Ea > 0.0)
          arrhenius = exp(Ea / IDEAL_GAS_CONSTANT *
                          (1.0 / (25.0 + 273.15) - 1.0 / (c->temp + 273.15)));
        const double ln_prefactor = log(prefactor[p]);
        const int32_t* pid = net->kinmnrl_prefactor_id + ((size_t)m * mp + p) * (msp + 1);
        for (int q = 0; q < pid[0]; ++q) {
          const size_t k = ((size_t)m * mp + p) * msp + q;
          const int icomp = pid[1 + q];
          const double ln_spec_act = ln_act[icomp];
          const double ln_prefactor_spec =
              net->kinmnrl_pref_alpha[k] * ln_spec_act -
              log(1.0 + exp(log(net->kinmnrl_pref_atten_coef[k]) +
                            net->kinmnrl_pref_beta[k] * ln_spec_act));
          const double dprefactor_dprefactor_spec = exp(ln_prefactor - ln_prefactor_spec);
          const double dnum = net->kinmnrl_pref_alpha[k] * exp(ln_prefactor_spec - ln_spec_act);
          const double ln_gam_m_beta = net->kinmnrl_pref_beta[k] * ln_spec_act;
          const double denominator =
              1.0 + exp(log(net->kinmnrl_pref_atten_coef[k]) + ln_gam_m_beta);
          const double dden = -1.0 * exp(ln_prefactor_spec) / denominator *
                              net->kinmnrl_pref_atten_coef[k] * net->kinmnrl_pref_beta[k] *
                              exp(ln_gam_m_beta - ln_spec_act);
          const double dspec = (dnum + dden) * c->pri_act_coef[icomp];
          const double dIm_dspec = dIm_dsum_prefactor_rate * dprefactor_dprefactor_spec * dspec *
                                   net->kinmnrl_pref_rate[(size_t)m * mp + p] * arrhenius;
          for (int i = 0; i < ncs; ++i) jac[ids[1 + i] * n + icomp] += st[i] * dIm_dspec;
        }
      }
    }
  }
}

/* RImmobileDecay reaction_immobile.F90:244-294 */
static void immobile_decay(const pfchem_network* net, const orc_cell* c,
                           double* res, double* jac, int compute_derivative) {
  const int n = ncomp_of(net);
  for (int r = 0; r < net->ndecay_rxn; ++r) {
    int icomp = net->decayspecid[r];
    double rate_constant = net->decay_rate_constant[r] * c->volume;
    double rate = rate_constant * c->immobile[icomp];
    int id = net->naqcomp + icomp;
    res[id] += rate;
    if (!compute_derivative) continue;
    jac[id * n + id] += rate_constant;
  }
}

void orc_somdec_react(const pfchem_network* net, orc_cell* c, double dt,
                      double* res, double* jac, int compute_derivative);

/* RReaction reaction.F90:4059-4130 */
void orc_reaction(const pfchem_network* net, orc_cell* c, double dt,
                  double* res, double* jac, int compute_derivative) {
  if (c->sat < RT_MIN_SATURATION) return;
  if (net->nkinmnrl > 0) kinetic_mineral(net, c, res, jac, compute_derivative);
  if (net->ndecay_rxn > 0) immobile_decay(net, c, res, jac, compute_derivative);
  if (net->somdec) orc_somdec_react(net, c, dt, res, jac, compute_derivative);
}

/* RReactionDerivative reaction.F90:4134-4239, numerical branch: forward
   differences with relative perturbation 1e-5 (reaction.F90:41). */
void orc_reaction_derivative_numerical(const pfchem_network* net,
                                       const orc_cell* c0, double dt,
                                       double* res, double* jac) {
  const int n = ncomp_of(net), naq = net->naqcomp;
  double res_orig[ORC_MAXCOMP], res_pert[ORC_MAXCOMP];
  static __thread orc_cell base, pert;
  double* jdummy = (double*)calloc((size_t)n * n, sizeof(double));
  memset(res, 0, sizeof(double) * n);
  memset(jac, 0, sizeof(double) * n * n);
  if (c0->sat < RT_MIN_SATURATION) { free(jdummy); return; }
  base = *c0;
  for (int i = 0; i < n; ++i) res_orig[i] = 0.0;
  orc_reaction(net, &base, dt, res_orig, jdummy, 0);
  for (int j = 0; j < naq; ++j) {
    pert = *c0;
    for (int i = 0; i < n; ++i) res_pert[i] = 0.0;
    double p = pert.pri_molal[j] * PERTURBATION_TOLERANCE;
    pert.pri_molal[j] += p;
    orc_total(net, &pert);
    orc_reaction(net, &pert, dt, res_pert, jdummy, 0);
    for (int i = 0; i < n; ++i) jac[i * n + j] += (res_pert[i] - res_orig[i]) / p;
  }
  for (int j = 0; j < net->nimmobile; ++j) {
    pert = *c0;
    for (int i = 0; i < n; ++i) res_pert[i] = 0.0;
    double p = pert.immobile[j] * PERTURBATION_TOLERANCE;
    pert.immobile[j] += p;
    orc_reaction(net, &pert, dt, res_pert, jdummy, 0);
    int jo = naq + j;
    for (int i = 0; i < n; ++i) jac[i * n + jo] += (res_pert[i] - res_orig[i]) / p;
  }
  for (int i = 0; i < n * n; ++i)
    if (fabs(jac[i]) < 1.0e-40) jac[i] = 0.0;
  for (int i = 0; i < n; ++i) res[i] = res_orig[i];
  free(jdummy);
}

/* LUDecomposition1 utility.F90:597-688 (Crout, implicit partial pivoting).
   a is row-major a[i*n+j] == A(i,j).  Returns 1 on a zero row. */
int orc_lu_decomposition(double* a, int n, int* indx, int* d) {
  const double tiny = 1.0e-20;
  double vv[ORC_MAXCOMP];
  int imax = 0;
  *d = 1;
  for (int i = 0; i < n; ++i) {
    double aamax = 0.0;
    for (int j = 0; j < n; ++j)
      if (fabs(a[i * n + j]) > aamax) aamax = fabs(a[i * n + j]);
    if (aamax <= 0.0) return 1;
    vv[i] = 1.0 / aamax;
  }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < j; ++i) {
      double sum = a[i * n + j];
      for (int k = 0; k < i; ++k) sum -= a[i * n + k] * a[k * n + j];
      a[i * n + j] = sum;
    }
    double aamax = 0.0;
    for (int i = j; i < n; ++i) {
      double sum = a[i * n + j];
      for (int k = 0; k < j; ++k) sum -= a[i * n + k] * a[k * n + j];
      a[i * n + j] = sum;
      double dum = vv[i] * fabs(sum);
      if (dum >= aamax) {
        imax = i;
        aamax = dum;
      }
    }
    if (j != imax) {
      for (int k = 0; k < n; ++k) {
        double dum = a[imax * n + k];
        a[imax * n + k] = a[j * n + k];
        a[j * n + k] = dum;
      }
      *d = -*d;
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (a[j * n + j] == 0.0) a[j * n + j] = tiny;
    if (j != n - 1) {
      double dum = 1.0 / a[j * n + j];
      for (int i = j + 1; i < n; ++i) a[i * n + j] *= dum;
    }
  }
  return 0;
}

/* LUBackSubstitution utility.F90:692-735 (with the leading-zero skip) */
void orc_lu_back_substitution(const double* a, int n, const int* indx,
                              double* b) {
  int ii = -1;
  for (int i = 0; i < n; ++i) {
    int ll = indx[i];
    double sum = b[ll];
    b[ll] = b[i];
    if (ii >= 0) {
      for (int j = ii; j < i; ++j) sum -= a[i * n + j] * b[j];
    } else if (sum != 0.0) {
      ii = i;
    }
    b[i] = sum;
  }
  for (int i = n - 1; i >= 0; --i) {
    double sum = b[i];
    for (int j = i + 1; j < n; ++j) sum -= a[i * n + j] * b[j];
    b[i] = sum / a[i * n + i];
  }
}

/* RSolve reaction.F90:5457-5516, back-substituting form (:5506-5512).
   jac is overwritten by its scaled LU.  Returns the LU error flag. */
int orc_rsolve(const double* res, double* jac, const double* conc,
               double* update, int n, int use_log_formulation) {
  double rhs[ORC_MAXCOMP];
  int indx[ORC_MAXCOMP], d;
  for (int i = 0; i < n; ++i) {
    double norm = 0.0;
    for (int j = 0; j < n; ++j)
      if (fabs(jac[i * n + j]) > norm) norm = fabs(jac[i * n + j]);
    norm = norm > 1.0 ? norm : 1.0;
    norm = 1.0 / norm;
    rhs[i] = res[i] * norm;
    for (int j = 0; j < n; ++j) jac[i * n + j] *= norm;
  }
  if (use_log_formulation)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) jac[i * n + j] *= conc[j];
  if (orc_lu_decomposition(jac, n, indx, &d)) return 1;
  orc_lu_back_substitution(jac, n, indx, rhs);
  for (int i = 0; i < n; ++i) update[i] = rhs[i];
  return 0;
}

/* optional per-iteration trace (test diagnostics): iteration, ||r||_2,
   ||r||/||r0||, max relative change (NaN before the update) */
void (*orc_trace)(int it, double two_norm, double rel, double max_rel_change) = 0;

/* accepted sub-steps of the current RStep (per thread; see orc_substep_log) */
__thread orc_substep_log* orc_log = 0;

/* body of one RReact iterate: reaction.F90:3899-3916 */
void orc_rreact_system(const pfchem_network* net, const pfchem_options* opt,
                       orc_cell* c, const double* fixed_accum, double dt,
                       double* res, double* jac, int compute_derivative) {
  (void)opt;
  const int n = ncomp_of(net);
  orc_accumulation(net, c, res);
  orc_accumulation_derivative(net, c, dt, jac);
  for (int i = 0; i < n; ++i) res[i] = (res[i] - fixed_accum[i]) / dt;
  orc_reaction(net, c, dt, res, jac, compute_derivative);
}

/* RReact reaction.F90:3742-4055.  Returns ierror; *fail_reason in
   {1 max iterations, 2 singular, 3 activity}. */
int orc_rreact(const pfchem_network* net, const pfchem_options* opt,
               orc_cell* c, double* guess, double dt, int* num_iterations,
               orc_substep_record* rec, int* fail_reason) {
  const int n = ncomp_of(net), naq = net->naqcomp, nimm = net->nimmobile;
  double fixed_accum[ORC_MAXCOMP], initial_total[ORC_MAXCOMP];
  double residual[ORC_MAXCOMP], update[ORC_MAXCOMP], conc[ORC_MAXCOMP];
  double prev[ORC_MAXCOMP], latest[ORC_MAXCOMP];
  double* J = (double*)malloc(sizeof(double) * n * n);
  double two_norm_r0 = 0.0;
  int ierror = 0;
  *num_iterations = 0;
  *fail_reason = 0;
  c->activity_error = 0;

  orc_update_temp_dependent_coefs(net, opt, c);
  orc_accumulation(net, c, fixed_accum);
  if (rec) {
    rec->dt = dt;
    for (int i = 0; i < naq; ++i) rec->fixed_total[i] = c->total[i];
    for (int i = 0; i < nimm; ++i) rec->fixed_immobile[i] = c->immobile[i];
    for (int i = 0; i < n; ++i) rec->guess[i] = guess[i];
    for (int m = 0; m < net->nkinmnrl; ++m) rec->volfrac[m] = c->mnrl_volfrac[m];
  }
  for (int i = 0; i < naq; ++i) initial_total[i] = c->total[i];
  for (int i = 0; i < nimm; ++i) initial_total[naq + i] = c->immobile[i];
  for (int i = 0; i < naq; ++i) c->pri_molal[i] = guess[i];
  for (int i = 0; i < nimm; ++i) c->immobile[i] = guess[naq + i];

  for (;;) {
    ++*num_iterations;
    if (opt->act_coef_update_frequency == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
      orc_activity_coefficients(net, opt, c);
    orc_total(net, c);
    if (*num_iterations > opt->maximum_reaction_iterations) {
      for (int i = 0; i < naq; ++i) c->total[i] = initial_total[i];
      for (int i = 0; i < nimm; ++i) c->immobile[i] = initial_total[naq + i];
      *fail_reason = 1;
      ierror = 1;
      break;
    }
    orc_rreact_system(net, opt, c, fixed_accum, dt, residual, J, 1);
    if (c->activity_error) { /* reaction.F90:3922-3936 */
      *fail_reason = 3;
      ierror = 1;
      break;
    }
    double two_norm_r = 0.0, maxabs = 0.0;
    for (int i = 0; i < n; ++i) {
      two_norm_r += residual[i] * residual[i];
      if (fabs(residual[i]) > maxabs) maxabs = fabs(residual[i]);
    }
    two_norm_r = sqrt(two_norm_r);
    if (*num_iterations == 1) two_norm_r0 = two_norm_r;
    double rel_residual = two_norm_r / two_norm_r0;
    if (rec) {
      if (*num_iterations == 1) rec->r0_norm = two_norm_r0;
      rec->final_rel_residual = rel_residual;
    }
    if (orc_trace) orc_trace(*num_iterations, two_norm_r, rel_residual, NAN);
    if (maxabs < opt->max_residual_tolerance) break;
    if (rel_residual < opt->max_rel_residual_tolerance) break;
    for (int i = 0; i < naq; ++i) conc[i] = c->pri_molal[i];
    for (int i = 0; i < nimm; ++i) conc[naq + i] = c->immobile[i];
    if (orc_rsolve(residual, J, conc, update, n, opt->use_log_formulation)) {
      *fail_reason = 2; /* reaction.F90:3963-3967: totals not restored */
      ierror = 1;
      break;
    }
    for (int i = 0; i < n; ++i) prev[i] = conc[i];
    if (opt->use_log_formulation) {
      for (int i = 0; i < n; ++i) {
        double u = copysign(1.0, update[i]) * fmin(fabs(update[i]), opt->max_dlnC_rreact);
        latest[i] = prev[i] * exp(-u);
      }
    } else {
      double min_ratio = 1.0e308;
      for (int i = 0; i < n; ++i)
        if (prev[i] <= update[i]) {
          double ratio = fabs(prev[i] / update[i]);
          if (ratio < min_ratio) min_ratio = ratio;
        }
      if (min_ratio < 1.0)
        for (int i = 0; i < n; ++i) update[i] = update[i] * min_ratio * 0.99;
      for (int i = 0; i < n; ++i) latest[i] = prev[i] - update[i];
    }
    double max_rel_change = 0.0;
    int have = 0;
    for (int i = 0; i < n; ++i) {
      double rc = fabs((latest[i] - prev[i]) / prev[i]);
      if (rc != rc) continue; /* maxval skips NaN unless all are NaN */
      if (!have || rc > max_rel_change) max_rel_change = rc;
      have = 1;
    }
    if (!have) max_rel_change = NAN;
    if (orc_trace) orc_trace(*num_iterations, NAN, NAN, max_rel_change);
    if (max_rel_change < opt->max_relative_change_tolerance) break;
    for (int i = 0; i < naq; ++i) c->pri_molal[i] = latest[i];
    for (int i = 0; i < nimm; ++i) c->immobile[i] = latest[naq + i];
  }
  free(J);
  if (!ierror) orc_total(net, c); /* one last update, reaction.F90:4052 */
  return ierror;
}

/* MineralUpdateKineticState reaction_mineral.F90:1456-1525 via
   RUpdateKineticState reaction.F90:5935-5972 */
void orc_update_kinetic_state(const pfchem_network* net, orc_cell* c,
                              double dt) {
  if (net->nkinmnrl == 0) return;
  double res[ORC_MAXCOMP];
  double jac[1];
  memset(res, 0, sizeof(res));
  kinetic_mineral(net, c, res, jac, 0);
  for (int m = 0; m < net->nkinmnrl; ++m) {
    double dv = c->mnrl_rate[m] * net->kinmnrl_molar_vol[m] * dt;
    c->mnrl_volfrac[m] += dv;
    if (c->mnrl_volfrac[m] < 0.0) c->mnrl_volfrac[m] = 0.0;
  }
}

/* RStep reaction.F90:3564-3738.  Returns ierror (0/1). */
int orc_rstep(const pfchem_network* net, const pfchem_options* opt,
              orc_cell* c, double* guess, double dt, int* num_sub_steps,
              int* num_iterations, int* num_kinetic_state_updates,
              orc_substep_record* rec) {
  const int naq = net->naqcomp, nimm = net->nimmobile, n = naq + nimm;
  int small[ORC_MAXCOMP];
  double small_val[ORC_MAXCOMP];
  *num_sub_steps = 0;
  *num_iterations = 0;
  *num_kinetic_state_updates = 0;
  if (!opt->use_full_geochemistry) {
    for (int i = 0; i < naq; ++i) c->pri_molal[i] = c->total[i] / c->den_kg * 1.0e3;
    return 0;
  }
  for (int i = 0; i < n; ++i) small[i] = 0;
  for (int i = 0; i < naq; ++i)
    if (c->total[i] <= 1.0e-40) {
      small[i] = 1;
      small_val[i] = c->total[i];
      c->total[i] = 1.0e-40;
    }
  for (int i = 0; i < nimm; ++i)
    if (c->immobile[i] <= 1.0e-40) {
      small[naq + i] = 1;
      small_val[naq + i] = c->immobile[i];
      c->immobile[i] = 1.0e-40;
    }
  if (opt->use_total_as_guess)
    for (int i = 0; i < naq; ++i) guess[i] = c->total[i];

  double target_time = dt, cumulative_time = 0.0, tran_dt = dt;
  int num_const = 0, num_cuts = 0, last_reason = 0, attempts = 0;
  orc_substep_record tmp;
  double t_start[ORC_MAXCOMP];
  if (orc_log) orc_log->n = orc_log->overflow = 0;
  for (;;) {
    if (cumulative_time >= target_time) break;
    if (++attempts > (opt->maximum_rreact_calls > 0 ? opt->maximum_rreact_calls
                                                    : ORC_MAX_RREACT_ATTEMPTS))
      return last_reason ? last_reason : 1;
    int inner = 0, reason = 0;
    if (orc_log)
      for (int i = 0; i < naq; ++i) t_start[i] = c->total[i];
    int ierr = orc_rreact(net, opt, c, guess, tran_dt, &inner, rec ? &tmp : NULL,
                          &reason);
    *num_iterations += inner;
    if (ierr) {
      last_reason = reason;
      ++num_cuts;
      if (num_cuts > opt->maximum_reaction_cuts) return last_reason;
      tran_dt = 0.5 * tran_dt;
      num_const = 0;
    } else {
      if (rec) *rec = tmp;
      if (orc_log) {
        if (orc_log->n < ORC_LOG_MAX_SUBSTEPS) {
          const int k = orc_log->n++;
          orc_log->dt[k] = tran_dt;
          for (int i = 0; i < naq; ++i) {
            orc_log->t0[k][i] = t_start[i];
            orc_log->m[k][i] = c->pri_molal[i];
          }
          for (int m = 0; m < net->nkinmnrl; ++m) orc_log->vf0[k][m] = c->mnrl_volfrac[m];
        } else {
          orc_log->overflow = 1;
        }
      }
      orc_update_kinetic_state(net, c, tran_dt);
      cumulative_time += tran_dt;
      ++*num_sub_steps;
      ++num_const;
      if (net->nkinmnrl > 0 || net->somdec) ++*num_kinetic_state_updates;
      for (int i = 0; i < naq; ++i) guess[i] = c->pri_molal[i];
      for (int i = 0; i < nimm; ++i) guess[naq + i] = c->immobile[i];
      if (num_const >= 4) {
        --num_cuts;
        double a = 2.0 * tran_dt, b = target_time - cumulative_time;
        tran_dt = a < b ? a : b;
      }
    }
  }
  for (int i = 0; i < naq; ++i)
    if (small[i]) c->total[i] = small_val[i];
  for (int i = 0; i < nimm; ++i)
    if (small[naq + i]) c->immobile[i] = small_val[naq + i];
  return 0;
}

/* ---------------------------------------------------------------- */
/* SoA gather / scatter and batch drivers                            */
/* ---------------------------------------------------------------- */
#define AT(p, k, cell) ((p)[(int64_t)(k) * s->ld + (cell)])

void orc_gather(const pfchem_network* net, const pfchem_state* s, int64_t cell,
                orc_cell* c) {
  orc_cell_defaults(net, c);
  for (int i = 0; i < net->naqcomp; ++i) {
    c->pri_molal[i] = AT(s->pri_molal, i, cell);
    c->total[i] = AT(s->total, i, cell);
    if (s->pri_act_coef) c->pri_act_coef[i] = AT(s->pri_act_coef, i, cell);
  }
  for (int i = 0; i < net->nimmobile; ++i) c->immobile[i] = AT(s->immobile, i, cell);
  for (int k = 0; k < net->neqcplx; ++k) {
    if (s->sec_molal) c->sec_molal[k] = AT(s->sec_molal, k, cell);
    if (s->sec_act_coef) c->sec_act_coef[k] = AT(s->sec_act_coef, k, cell);
  }
  c->ln_act_h2o = s->ln_act_h2o ? s->ln_act_h2o[cell] : 0.0;
  for (int m = 0; m < net->nkinmnrl; ++m) {
    c->mnrl_volfrac[m] = AT(s->mnrl_volfrac, m, cell);
    c->mnrl_area[m] = AT(s->mnrl_area, m, cell);
    c->mnrl_rate[m] = s->mnrl_rate ? AT(s->mnrl_rate, m, cell) : 0.0;
  }
  c->porosity = s->porosity[cell];
  c->sat = s->saturation[cell];
  c->den_kg = s->den_kg[cell];
  c->volume = s->volume[cell];
  c->temp = s->temperature ? s->temperature[cell] : 25.0;
  c->has_f_t = s->somdec_f_t != NULL;
  c->has_f_w = s->somdec_f_w != NULL;
  c->f_t = c->has_f_t ? s->somdec_f_t[cell] : 1.0;
  c->f_w = c->has_f_w ? s->somdec_f_w[cell] : 1.0;
}

void orc_scatter(const pfchem_network* net, const orc_cell* c,
                 const pfchem_state* s, int64_t cell) {
  for (int i = 0; i < net->naqcomp; ++i) {
    AT(s->pri_molal, i, cell) = c->pri_molal[i];
    AT(s->total, i, cell) = c->total[i];
    if (s->pri_act_coef) AT(s->pri_act_coef, i, cell) = c->pri_act_coef[i];
  }
  for (int i = 0; i < net->nimmobile; ++i) AT(s->immobile, i, cell) = c->immobile[i];
  for (int k = 0; k < net->neqcplx; ++k) {
    if (s->sec_molal) AT(s->sec_molal, k, cell) = c->sec_molal[k];
    if (s->sec_act_coef) AT(s->sec_act_coef, k, cell) = c->sec_act_coef[k];
  }
  if (s->ln_act_h2o) s->ln_act_h2o[cell] = c->ln_act_h2o;
  for (int m = 0; m < net->nkinmnrl; ++m) {
    AT(s->mnrl_volfrac, m, cell) = c->mnrl_volfrac[m];
    if (s->mnrl_rate) AT(s->mnrl_rate, m, cell) = c->mnrl_rate[m];
  }
}

int64_t orc_step_batch(const pfchem_network* net, const pfchem_options* opt,
                       const pfchem_state* s, int64_t ncell, double dt,
                       int nthreads, orc_substep_record* rec) {
  int64_t nfail = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : nfail)
#endif
  {
    orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
    double guess[ORC_MAXCOMP];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t cell = 0; cell < ncell; ++cell) {
      orc_gather(net, s, cell, c);
      for (int i = 0; i < net->naqcomp; ++i) guess[i] = c->pri_molal[i];
      for (int i = 0; i < net->nimmobile; ++i) guess[net->naqcomp + i] = c->immobile[i];
      int nsub, nit, nkin;
      int ierr = orc_rstep(net, opt, c, guess, dt, &nsub, &nit, &nkin,
                           rec ? &rec[cell] : NULL);
      orc_scatter(net, c, s, cell);
      if (s->num_iterations) s->num_iterations[cell] = nit;
      if (s->num_sub_steps) s->num_sub_steps[cell] = nsub;
      if (s->status) s->status[cell] = ierr;
      if (ierr) ++nfail;
    }
    free(c);
  }
  return nfail;
}

void orc_update_activity_batch(const pfchem_network* net,
                               const pfchem_options* opt,
                               const pfchem_state* s, int64_t ncell,
                               int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t cell = 0; cell < ncell; ++cell) {
      orc_gather(net, s, cell, c);
      orc_activity_coefficients(net, opt, c);
      orc_scatter(net, c, s, cell);
    }
    free(c);
  }
}

double orc_record_rel_residual(const pfchem_network* net,
                               const pfchem_options* opt,
                               const pfchem_state* s, int64_t cell,
                               const orc_substep_record* rec) {
  const int n = ncomp_of(net), naq = net->naqcomp;
  orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
  double fixed[ORC_MAXCOMP], res[ORC_MAXCOMP];
  double* J = (double*)malloc(sizeof(double) * n * n);
  orc_gather(net, s, cell, c);
  orc_update_temp_dependent_coefs(net, opt, c);
  /* fixed accumulation of the record's sub-step */
  for (int i = 0; i < naq; ++i) c->total[i] = rec->fixed_total[i];
  double save_imm[ORC_MAXCOMP];
  for (int i = 0; i < net->nimmobile; ++i) {
    save_imm[i] = c->immobile[i];
    c->immobile[i] = rec->fixed_immobile[i];
  }
  orc_accumulation(net, c, fixed);
  for (int i = 0; i < net->nimmobile; ++i) c->immobile[i] = save_imm[i];
  for (int m = 0; m < net->nkinmnrl; ++m) c->mnrl_volfrac[m] = rec->volfrac[m];
  orc_total(net, c);
  orc_rreact_system(net, opt, c, fixed, rec->dt, res, J, 0);
  double nr = 0.0;
  for (int i = 0; i < n; ++i) nr += res[i] * res[i];
  free(J);
  free(c);
  return sqrt(nr) / rec->r0_norm;
}

size_t orc_sizeof_record(void) { return sizeof(orc_substep_record); }
size_t orc_sizeof_cell(void) { return sizeof(orc_cell); }

/* totals / secondary molalities / dtotal of one cell at its stored free
   molality (RTotal), for tests */
void orc_cell_total(const pfchem_network* net, const pfchem_options* opt,
                    const pfchem_state* s, int64_t cell, double* total,
                    double* sec_molal, double* dtotal) {
  orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
  orc_gather(net, s, cell, c);
  orc_update_temp_dependent_coefs(net, opt, c);
  orc_total(net, c);
  for (int i = 0; i < net->naqcomp; ++i) total[i] = c->total[i];
  if (sec_molal)
    for (int k = 0; k < net->neqcplx; ++k) sec_molal[k] = c->sec_molal[k];
  if (dtotal)
    for (int i = 0; i < net->naqcomp * net->naqcomp; ++i) dtotal[i] = c->dtotal[i];
  free(c);
}
