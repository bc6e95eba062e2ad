/* pfchem_oracle_constraint.c -- FP64 restatement of the constraint
 * speciation, ReactionEquilibrateConstraint (reaction.F90:1328-2117).
 *
 * TEST INFRASTRUCTURE: the checker for pfchem_equilibrate.  Only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Follows the reference's order of operations: constraint set-up
 * (:1487-1608), then per iteration the free-ion reset (:1625-1631), activity
 * coefficients once switched on (:1632-1638), RTotal (:1639), the residual
 * and Jacobian row of every constraint type (:1646-1866), RSolve
 * (:1868-1916), the log / linear update (:1930-1950), the positivity check
 * (:1954-1975) and the two-stage convergence test (:2017-2030).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "pfchem_oracle.h"

/* pflotran_constants.F90:84 */
#define LOG_TO_LN 2.30258509299

/* ReactionInterpolateLogK (reaction_aux.F90:1285-1312) for one row */
static double con_logk(const pfchem_constraint* con, const pfchem_options* opt, int r,
                       double temp) {
  if (!opt->isothermal && con->rxn_logKcoef) {
    const double* b = con->rxn_logKcoef + 5 * r;
    const double tk = temp + 273.15;
    return b[0] * log(tk) + b[1] + b[2] * tk + b[3] / tk + b[4] / (tk * tk);
  }
  return con->rxn_logK[r];
}

/* ln Q/K of constraint reaction r (mineral :1756-1771, gas :1780-1793) */
static double con_lnqk(const pfchem_network* net, const pfchem_options* opt,
                       const pfchem_constraint* con, const orc_cell* c, int r) {
  const int ms = con->rxn_maxsp;
  double lnQK = -con_logk(con, opt, r, c->temp) * LOG_TO_LN;
  if (con->rxnh2ostoich && con->rxnh2ostoich[r] != 0.0)
    lnQK += con->rxnh2ostoich[r] * c->ln_act_h2o;
  const int32_t* ids = con->rxnspecid + r * (ms + 1);
  for (int j = 0; j < ids[0]; ++j) {
    const int id = ids[1 + j];
    lnQK += con->rxnstoich[r * ms + j] * log(c->pri_molal[id] * c->pri_act_coef[id]);
  }
  return lnQK;
}

int orc_equilibrate(const pfchem_network* net, const pfchem_options* opt,
                    const pfchem_constraint* con, const double* conc_in, orc_cell* c,
                    int* num_iterations) {
  const int naq = net->naqcomp;
  double conc[ORC_MAXCOMP], total_conc[ORC_MAXCOMP], free_conc[ORC_MAXCOMP];
  double res[ORC_MAXCOMP], update[ORC_MAXCOMP], prev[ORC_MAXCOMP];
  double jac[ORC_MAXCOMP * ORC_MAXCOMP];
  const int32_t* type = con->constraint_type;
  *num_iterations = 0;
  for (int i = 0; i < naq; ++i) conc[i] = conc_in[i];

  /* :1465-1472 (xmass = 1) */
  double molal_to_molar, molar_to_molal;
  if (con->initialize_with_molality) {
    molal_to_molar = c->den_kg / 1000.0;
    molar_to_molal = 1.0;
  } else {
    molal_to_molar = 1.0;
    molar_to_molal = 1000.0 / c->den_kg;
  }
  /* :1474-1478 */
  if (!opt->use_full_geochemistry) {
    for (int i = 0; i < naq; ++i) {
      c->pri_molal[i] = conc[i] * molar_to_molal;
      c->total[i] = conc[i] * molal_to_molar;
    }
    return PFCHEM_CELL_OK;
  }
  /* :1482-1484 */
  orc_update_temp_dependent_coefs(net, opt, c);

  /* :1486-1599 */
  for (int i = 0; i < naq; ++i) {
    if (con->use_prev_soln_as_guess) free_conc[i] = c->pri_molal[i];
    else if (con->free_ion_guess) free_conc[i] = con->free_ion_guess[i];
    else free_conc[i] = 1.0e-9;
    total_conc[i] = 0.0;
  }
  for (int i = 0; i < naq; ++i) {
    switch (type[i]) {
      case PFCHEM_CONSTRAINT_NULL:
      case PFCHEM_CONSTRAINT_TOTAL:
        total_conc[i] = conc[i] * molal_to_molar;
        break;
      case PFCHEM_CONSTRAINT_FREE:
        free_conc[i] = conc[i] * molar_to_molal;
        break;
      case PFCHEM_CONSTRAINT_LOG:
        free_conc[i] = pow(10.0, conc[i]) * molar_to_molal;
        break;
      case PFCHEM_CONSTRAINT_CHARGE_BAL:
      case PFCHEM_CONSTRAINT_MINERAL:
        if (!con->use_prev_soln_as_guess) free_conc[i] = conc[i] * molar_to_molal;
        break;
      case PFCHEM_CONSTRAINT_PH:
        free_conc[i] = pow(10.0, -conc[i]);
        break;
      case PFCHEM_CONSTRAINT_GAS:
        if (conc[i] <= 0.0) conc[i] = pow(10.0, conc[i]);
        break;
      default:
        return -1;
    }
  }
  for (int i = 0; i < naq; ++i) c->pri_molal[i] = free_conc[i];

  int it = 0, it_act_on = 0, charge_warned = 0;
  int compute_act = con->use_prev_soln_as_guess != 0;
  for (;;) {
    /* :1625-1631 */
    for (int i = 0; i < naq; ++i)
      if (type[i] == PFCHEM_CONSTRAINT_FREE || type[i] == PFCHEM_CONSTRAINT_LOG)
        c->pri_molal[i] = free_conc[i];
    /* :1632-1638 */
    if (opt->act_coef_update_frequency != PFCHEM_ACT_COEF_FREQUENCY_OFF && compute_act)
      orc_activity_coefficients(net, opt, c);
    /* :1639 */
    orc_total(net, c);

    memset(jac, 0, sizeof(double) * naq * naq);
    for (int i = 0; i < naq; ++i) {
      double* row = jac + i * naq;
      switch (type[i]) {
        case PFCHEM_CONSTRAINT_NULL:
        case PFCHEM_CONSTRAINT_TOTAL: /* :1648-1655 */
          res[i] = c->total[i] - total_conc[i];
          for (int j = 0; j < naq; ++j) row[j] = c->dtotal[i * naq + j];
          break;
        case PFCHEM_CONSTRAINT_FREE:
        case PFCHEM_CONSTRAINT_LOG: /* :1680-1685 */
          res[i] = 0.0;
          row[i] = 1.0;
          break;
        case PFCHEM_CONSTRAINT_CHARGE_BAL: /* :1687-1714 */
          res[i] = 0.0;
          for (int j = 0; j < naq; ++j) {
            res[i] += net->primary_spec_Z[j] * c->total[j];
            for (int k = 0; k < naq; ++k)
              row[j] += net->primary_spec_Z[k] * c->dtotal[k * naq + j];
          }
          if (c->pri_molal[i] < 1.0e-20 && !charge_warned) {
            const double z = net->primary_spec_Z[i];
            if ((res[i] > 0.0 && z > 0.0) || (res[i] < 0.0 && z < 0.0)) {
              charge_warned = 1;
              c->pri_molal[i] = 1.0e-3;
            }
          }
          break;
        case PFCHEM_CONSTRAINT_PH: /* :1716-1753 */
          res[i] = 0.0;
          if (con->h_ion_id >= 0) {
            c->pri_molal[i] = pow(10.0, -conc[i]) / c->pri_act_coef[i];
            row[i] = 1.0;
          } else {
            const int k = -con->h_ion_id - 1, ms = net->eqcplx_maxsp;
            double lnQK = -c->eqcplx_logK[k] * LOG_TO_LN;
            if (net->eqcplxh2ostoich && net->eqcplxh2ostoich[k] != 0.0)
              lnQK += net->eqcplxh2ostoich[k] * c->ln_act_h2o;
            const int32_t* ids = net->eqcplxspecid + k * (ms + 1);
            for (int j = 0; j < ids[0]; ++j) {
              const int id = ids[1 + j];
              lnQK += net->eqcplxstoich[k * ms + j] *
                      log(c->pri_molal[id] * c->pri_act_coef[id]);
            }
            lnQK += conc[i] * LOG_TO_LN;
            const double QK = exp(lnQK);
            res[i] = 1.0 - QK;
            for (int j = 0; j < ids[0]; ++j) {
              const int id = ids[1 + j];
              row[id] = -QK / c->pri_molal[id] * net->eqcplxstoich[k * ms + j];
            }
          }
          break;
        case PFCHEM_CONSTRAINT_MINERAL:
        case PFCHEM_CONSTRAINT_GAS: { /* :1754-1804 */
          const int r = con->constraint_rxn[i], ms = con->rxn_maxsp;
          const double lnQK = con_lnqk(net, opt, con, c, r);
          res[i] = type[i] == PFCHEM_CONSTRAINT_MINERAL ? lnQK : lnQK - log(conc[i]);
          const int32_t* ids = con->rxnspecid + r * (ms + 1);
          for (int j = 0; j < ids[0]; ++j) {
            const int id = ids[1 + j];
            row[id] = con->rxnstoich[r * ms + j] / c->pri_molal[id];
          }
          break;
        }
      }
    }
    /* :1868 */
    double max_res = 0.0;
    for (int i = 0; i < naq; ++i) max_res = fmax(max_res, fabs(res[i]));
    /* :1870-1881 */
    int use_log = 0;
    if (opt->use_log_formulation)
      use_log = (it > 3 && it < 9) ? (it % 2 == 0) : 1;
    /* :1883-1926 */
    if (orc_rsolve(res, jac, c->pri_molal, update, naq, use_log)) {
      *num_iterations = it;
      return PFCHEM_CELL_CONSTRAINT_SINGULAR;
    }
    for (int i = 0; i < naq; ++i) prev[i] = c->pri_molal[i];
    if (use_log) { /* :1930-1932 */
      for (int i = 0; i < naq; ++i) {
        const double u = copysign(1.0, update[i]) * fmin(fabs(update[i]), con->max_dlnC);
        c->pri_molal[i] *= exp(-u);
      }
    } else { /* :1933-1950 */
      double min_ratio = DBL_MAX;
      for (int i = 0; i < naq; ++i)
        if (prev[i] <= update[i]) {
          const double ratio = fabs(prev[i] / update[i]);
          if (ratio < min_ratio) min_ratio = ratio;
        }
      if (min_ratio <= 1.0)
        for (int i = 0; i < naq; ++i) update[i] *= min_ratio * 0.99;
      for (int i = 0; i < naq; ++i) c->pri_molal[i] = prev[i] - update[i];
    }
    /* :1954-1975 */
    double min_m = c->pri_molal[0];
    for (int i = 1; i < naq; ++i) min_m = fmin(min_m, c->pri_molal[i]);
    if (min_m <= 0.0) {
      *num_iterations = it + 1;
      return PFCHEM_CELL_CONSTRAINT_NONPOSITIVE;
    }
    /* :1997-2000 */
    double max_rel = 0.0;
    for (int i = 0; i < naq; ++i)
      max_rel = fmax(max_rel, fabs((c->pri_molal[i] - prev[i]) / prev[i]));
    ++it;
    *num_iterations = it;
    /* :2002-2016 (the reference stops the run at 10000) */
    if (it >= con->max_iterations) return PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS;
    /* :2017-2030 */
    if (max_res < opt->max_residual_tolerance &&
        max_rel < opt->max_relative_change_tolerance) {
      if (compute_act && it - it_act_on > 1) break;
      if (!compute_act) it_act_on = it;
      compute_act = 1;
    }
  }
  return PFCHEM_CELL_OK;
}

int64_t orc_equilibrate_batch(const pfchem_network* net, const pfchem_options* opt,
                              const pfchem_constraint* con, const pfchem_state* s,
                              const double* conc, int64_t ncell, int nthreads) {
  int64_t nfail = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : nfail)
#endif
  {
    orc_cell* c = (orc_cell*)malloc(sizeof(orc_cell));
    double cv[ORC_MAXCOMP];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t cell = 0; cell < ncell; ++cell) {
      orc_gather(net, s, cell, c);
      if (!con->use_prev_soln_as_guess) {
        /* a fresh auxvar (RTAuxVarInit): unit activity coefficients */
        for (int i = 0; i < net->naqcomp; ++i) c->pri_act_coef[i] = 1.0;
        for (int k = 0; k < net->neqcplx; ++k) c->sec_act_coef[k] = 1.0;
        c->ln_act_h2o = 0.0;
      }
      for (int i = 0; i < net->naqcomp; ++i) cv[i] = conc[(int64_t)i * s->ld + cell];
      int nit = 0;
      const int ierr = orc_equilibrate(net, opt, con, cv, c, &nit);
      orc_scatter(net, c, s, cell);
      if (s->num_iterations) s->num_iterations[cell] = nit;
      if (s->status) s->status[cell] = ierr;
      if (ierr) ++nfail;
    }
    free(c);
  }
  return nfail;
}
