/*
 * pfchem_oracle_somdec.c -- FP64 restatement of the SomDec decomposition
 * sandbox (reaction_sandbox_somdec.F90) for the oracle.  TEST
 * INFRASTRUCTURE ONLY: used by tests/, smoke() and bench.py's CPU leg.
 *
 * Follows SomDecReact (:1504-1910), SomDecReact1 (:1914-2418, N
 * mineralization), SomDecReact2 (:2423-3472, N immobilization) and
 * SomDecNemission (:3477-3640) statement by statement for the features the
 * descriptor carries:
 *   * upstream / downstream pools are immobile species;
 *   * aqueous NH4+ / NO3- / CO2(aq) / N2O(aq) (and H+ for the N2O pH factor);
 *   * tracking species (HR, Nmin, Nimm, Nimp, NGASmin and per-pool ids);
 *   * Monod and inhibition terms, the NH4:NO3 preference, the N-rate cap;
 *   * temperature responses (elm_rspfuncs.F90:61-120), LOGTHETA moisture,
 *     WFPS aerobic factor, and the ELM scalars f_t / f_w per cell (the
 *     ELM_PFLOTRAN branch, used when the state carries them).
 * The reference's residual and Jacobian entries are kept as written,
 * including entries that are not the exact derivative (for example d
 * rate / d NH4 columns without the theta*1000 factor of the concentration
 * conversion, and dnh4_duc = dco2_duc * mineral_n_stoich in React1): the
 * Jacobian steers the Newton iteration only, not its solution.
 * Downstream entries without a species (id < 0) are skipped in the
 * Jacobian columns of NH4+ / NO3- too (the reference checks them only in
 * the residual).  Two intended differences: the variable C:N stoichiometry lives in the
 * cell (orc_cell.sd_*), starting from the Setup values at each step,
 * instead of in the shared sandbox object (:1817-1858), so results do not
 * depend on cell order; and a reaction skipped by its abiotic factors
 * (:1744-1746) does not shift the Monod / inhibition lists of the
 * reactions after it.
 */
#include <math.h>
#include <string.h>

#include "pfchem_oracle.h"

#define RPI 3.14159265358979323846 /* elm_rspfuncs.F90 rpi */

/* HfunctionSmooth utility.F90:2542-2597 */
static void hsmooth(double x, double x1, double x0, double* H, double* dH) {
  if (fabs(x1 - x0) < 1.0e-50) {
    *H = copysign(0.5, x - x1) + 0.5;
    *dH = 0.0;
    return;
  }
  const double u = (x - x0) / (x1 - x0);
  if (u < 0.0) {
    *H = 0.0;
    *dH = 0.0;
  } else if (u > 1.0) {
    *H = 1.0;
    *dH = 0.0;
  } else {
    const double xs = 1.0 - (x - x0) * (x - x0) / (x1 - x0) / (x1 - x0);
    *H = 1.0 - xs * xs;
    *dH = 4.0 * xs * (x - x0) / (x1 - x0) / (x1 - x0);
  }
}

/* FuncMonod elm_rspfuncs.F90:320-335 */
static double monod(double c, double k, int deriv) {
  return deriv ? k / (c + k) / (c + k) : c / (c + k);
}

/* FuncInhibition elm_rspfuncs.F90:340-375 */
static double inhibition(int deriv, double c, double k, double k2, int has_k2) {
  if (!has_k2) return deriv ? -k / (c + k) / (c + k) : k / (c + k);
  if (!deriv) return 0.5 + atan((c - k) * k2) / M_PI;
  const double t = (c - k) * k2;
  return (k2 / (1.0 + t * t)) / M_PI;
}

/* GetTemperatureResponse elm_rspfuncs.F90:61-120 */
static double temperature_response(double tc, int itype, double q) {
  switch (itype) {
    case 2: /* Q10 */
      if (tc > 0.0) return pow(q, (tc - 25.0) / 10.0);
      return pow(q, -25.0 / 10.0) * pow(2.0, tc / 10.0);
    case 1: { /* CLMCN */
      const double tk = tc + 273.15;
      if (tk > 227.15) return exp(308.56 * (1.408054069e-2 - 1.0 / (tk - 227.13)));
      return 0.0;
    }
    case 3: /* DLEM */
      if (tc < -5.0) return 0.0;
      if (tc >= 30.0) return 1.0;
      return pow(q, (tc - 30.0) / 10.0);
    case 4: /* ARRHENIUS */
      return exp(q / 8.31446 * (1.0 / 298.15 - 1.0 / (tc + 273.15)));
    default:
      return 1.0;
  }
}

/* GetAerobicCondition elm_rspfuncs.F90:282-316, WFPS branch */
static double wfps_factor(double sat, double half_sat) {
  (void)half_sat;
  const double a = 1.27 - sat, b = sat - 0.0012;
  if (a <= 0.0 || b <= 0.0) return 0.0;
  double f = pow(a / 0.67, 3.1777) * pow(b / 0.5988, 2.84);
  return f < 1.0 ? f : 1.0;
}

typedef struct {
  const pfchem_network* net;
  const pfchem_somdec* sd;
  orc_cell* c;
  double* res;
  double* jac;
  int n, naq;
  double theta, vol;
} sd_ctx;

static inline int imm(const sd_ctx* x, int id) { return x->naq + id; }
static inline void R(sd_ctx* x, int i, double v) { x->res[i] += v; }
static inline void J(sd_ctx* x, int i, int j, double v) { x->jac[i * x->n + j] += v; }
static inline double dtot(const sd_ctx* x, int i, int j) {
  return x->c->dtotal[i * x->naq + j];
}

/* SomDec_SpeciesConc :3809-3870 with the descriptor's id encoding:
   >= 0 aqueous id (total, mol/L), < 0 immobile id -(id)-1 */
static double species_conc(const sd_ctx* x, int id) {
  if (id >= 0) return x->c->total[id];
  return x->c->immobile[-id - 1];
}

/* a 1-based species code (aqueous id + 1 > 0, -(immobile id + 1) < 0):
   SomDec_SpeciesConc's value and the residual row (-1 for code 0) */
static double code_conc(const sd_ctx* x, int code) {
  return code > 0 ? x->c->total[code - 1] : x->c->immobile[-code - 1];
}
static int code_row(const sd_ctx* x, int code) {
  return code > 0 ? code - 1 : (code < 0 ? x->naq - code - 1 : -1);
}

/* Monod oxygen response of reaction r (:2129-2147, 2764-2784):
   GetAerobicCondition (elm_rspfuncs.F90:282-316) = FuncMonod of the
   oxidizer, and its derivative in the oxidizer, which the reference adds to
   d crate / d uc */
static void ox_response(const sd_ctx* x, int r, double* f, double* df) {
  const pfchem_somdec* sd = x->sd;
  *f = 1.0;
  *df = 0.0;
  if (sd->Ox_response_function[r] != 1 || !sd->Ox_specid || sd->Ox_specid[r] == 0) return;
  const double ox = code_conc(x, sd->Ox_specid[r]), k = sd->Ox_half_saturation[r];
  *f = monod(ox, k, 0);
  *df = monod(ox, k, 1);
}

#define MD PFCHEM_SOMDEC_MAX_DOWN
#define MM PFCHEM_SOMDEC_MAX_MONOD
#define MI PFCHEM_SOMDEC_MAX_INHIB

/* product of the general Monod / inhibition terms of reaction r
   (:2080-2139 / :2590-2660); skip_nh4_no3 for React2's own NH4/NO3 Monods */
static void monod_inhib(sd_ctx* x, int r, int skip_nh4_no3, int uc, double* fmb,
                        double* dfmb, double* fnh4, double* dfnh4, double* fno3,
                        double* dfno3, double c_nh4, double c_no3) {
  const pfchem_somdec* sd = x->sd;
  for (int q = 0; q < sd->n_monod[r]; ++q) {
    const int id = sd->monod_specid[r * MM + q];
    const double k = sd->monod_half_saturation[r * MM + q];
    const double th = sd->monod_threshold[r * MM + q];
    const int pn = sd->monod_pool_normalized[r * MM + q];
    if (skip_nh4_no3 && id >= 0 && id == sd->species_id_nh4) {
      double t = fmax(0.0, c_nh4 - th);
      if (pn) t = t / x->c->immobile[uc];
      *fnh4 = monod(t, k, 0);
      *dfnh4 = monod(t, k, 1);
      continue;
    }
    if (skip_nh4_no3 && id >= 0 && sd->species_id_no3 >= 0 && id == sd->species_id_no3) {
      double t = fmax(0.0, c_no3 - th);
      if (pn) t = t / x->c->immobile[uc];
      *fno3 = monod(t, k, 0);
      *dfno3 = monod(t, k, 1);
      continue;
    }
    double t = fmax(0.0, species_conc(x, id) - th);
    if (pn) {
      t = t / x->c->immobile[uc];
      if (id >= 0) t = t * x->theta * 1000.0;
    }
    const double fx = monod(t, k, 0);
    double dfx = monod(t, k, 1);
    if (id != -uc - 1) dfx = 0.0;
    *dfmb = *dfmb * fx + *fmb * dfx;
    *fmb = *fmb * fx;
  }
  for (int q = 0; q < sd->n_inhibition[r]; ++q) {
    const int id = sd->inhibition_specid[r * MI + q];
    const int ty = sd->inhibition_type[r * MI + q];
    const double k = sd->inhibition_constant[r * MI + q];
    const double k2 = sd->inhibition_constant2[r * MI + q];
    const double t = species_conc(x, id);
    double fx = 1.0, dfx = 0.0;
    if (ty != 1 /* INHIBITION_THRESHOLD */) {
      if (ty == 3) { /* MONOD: k/(k+s) */
        fx = inhibition(0, t, k, 0.0, 0);
        dfx = inhibition(1, t, k, 0.0, 0);
      } else if (ty == 4) { /* INVERSE_MONOD: s/(k+s) */
        fx = monod(t, k, 0);
        dfx = monod(t, k, 1);
      }
    } else {
      fx = inhibition(0, t, k, k2, 1);
      dfx = inhibition(1, t, k, k2, 1);
    }
    if (id != -uc - 1) dfx = 0.0;
    *dfmb = *dfmb * fx + *fmb * dfx;
    *fmb = *fmb * fx;
  }
}

/* SomDecReact1 :1914-2418 (N mineralization); returns nmin */
static double react1(sd_ctx* x, int r, double crate_uc, double dcrate_uc_duc, int deriv) {
  const pfchem_somdec* sd = x->sd;
  orc_cell* c = x->c;
  const double n_st = c->sd_mineral_n_stoich[r], c_st = c->sd_mineral_c_stoich[r];
  if (n_st < 0.0) return 0.0;
  const int uc = sd->upstream_c_id[r], un = sd->upstream_n_id[r];
  const int varycn = sd->upstream_is_varycn[r];
  const int ruc = imm(x, uc), run = varycn ? imm(x, un) : -1;
  const int rco2 = sd->species_id_co2, rnh4 = sd->species_id_nh4;
  double fmb = 1.0, dfmb = 0.0, d1 = 0, d2 = 0, d3 = 0, d4 = 0;
  monod_inhib(x, r, 0, uc, &fmb, &dfmb, &d1, &d2, &d3, &d4, 0.0, 0.0);
  dcrate_uc_duc = dcrate_uc_duc * fmb + crate_uc * dfmb;
  crate_uc = crate_uc * fmb;
  double f_ox, df_ox; /* Monod oxygen response (WFPS is applied to f_w) */
  ox_response(x, r, &f_ox, &df_ox);
  dcrate_uc_duc = dcrate_uc_duc * f_ox + crate_uc * df_ox;
  crate_uc = crate_uc * f_ox;
  const int rox = code_row(x, sd->species_o2);
  const double crate = crate_uc;
  R(x, ruc, crate);
  R(x, rco2, -c_st * crate);
  if (rox >= 0) R(x, rox, c_st * crate); /* O2 consumed as CO2 is released (:2176-2186) */
  if (sd->upstream_hr_id[r] >= 0) R(x, imm(x, sd->upstream_hr_id[r]), -c_st * crate);
  if (sd->species_id_hr >= 0) R(x, imm(x, sd->species_id_hr), -c_st * crate);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    const int dc = sd->downstream_c_id[r * MD + j];
    if (dc >= 0) R(x, imm(x, dc), -sd->downstream_stoich[r * MD + j] * crate);
  }
  if (varycn) R(x, run, c->sd_upstream_nc[r] * crate);
  double nmin = 0.0;
  R(x, rnh4, -n_st * crate);
  nmin = n_st * crate;
  if (sd->upstream_nmin_id[r] >= 0) R(x, imm(x, sd->upstream_nmin_id[r]), -n_st * crate);
  if (sd->species_id_nmin >= 0) R(x, imm(x, sd->species_id_nmin), -n_st * crate);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    if (!sd->downstream_is_varycn[r * MD + j]) continue;
    const int dn = sd->downstream_n_id[r * MD + j];
    if (dn >= 0)
      R(x, imm(x, dn),
        -sd->downstream_stoich[r * MD + j] * crate * c->sd_downstream_nc[r * MD + j]);
  }
  if (!deriv) return nmin;
  const double dcrate_dx = dcrate_uc_duc;
  const double dco2_duc = dcrate_dx * c_st;
  const double duc_duc = -1.0 * dcrate_dx;
  const double dnh4_duc = dco2_duc * n_st; /* as written (:2273) */
  const double dun_duc = c->sd_upstream_nc[r] * duc_duc;
  J(x, rco2, ruc, -dco2_duc);
  if (rox >= 0) J(x, rox, ruc, dco2_duc); /* :2300-2308 */
  if (sd->upstream_hr_id[r] >= 0) J(x, imm(x, sd->upstream_hr_id[r]), ruc, -dco2_duc);
  if (sd->species_id_hr >= 0) J(x, imm(x, sd->species_id_hr), ruc, -dco2_duc);
  J(x, ruc, ruc, -duc_duc);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    const int dc = sd->downstream_c_id[r * MD + j];
    const double ddc_duc = sd->downstream_stoich[r * MD + j] * (-1.0 * duc_duc);
    if (dc >= 0) J(x, imm(x, dc), ruc, -ddc_duc);
  }
  J(x, rnh4, ruc, -dnh4_duc);
  if (sd->upstream_nmin_id[r] >= 0) J(x, imm(x, sd->upstream_nmin_id[r]), ruc, -dnh4_duc);
  if (sd->species_id_nmin >= 0) J(x, imm(x, sd->species_id_nmin), ruc, -dnh4_duc);
  if (varycn) J(x, run, ruc, -dun_duc);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    if (!sd->downstream_is_varycn[r * MD + j]) continue;
    const int dn = sd->downstream_n_id[r * MD + j];
    const double ddn_duc =
        sd->downstream_stoich[r * MD + j] * (-1.0 * duc_duc) * c->sd_downstream_nc[r * MD + j];
    if (dn >= 0) J(x, imm(x, dn), ruc, -ddn_duc);
  }
  return nmin;
}

/* SomDecReact2 :2423-3472 (N immobilization); returns nimm (<= 0) */
static double react2(sd_ctx* x, int r, double crate_uc, double dcrate_uc_duc, double dt,
                     int deriv) {
  const pfchem_somdec* sd = x->sd;
  orc_cell* c = x->c;
  const double n_st = c->sd_mineral_n_stoich[r], c_st = c->sd_mineral_c_stoich[r];
  if (n_st >= 0.0) return 0.0;
  const int uc = sd->upstream_c_id[r], un = sd->upstream_n_id[r];
  const int varycn = sd->upstream_is_varycn[r];
  const int ruc = imm(x, uc), run = varycn ? imm(x, un) : -1;
  const int rco2 = sd->species_id_co2, rnh4 = sd->species_id_nh4, rno3 = sd->species_id_no3;
  const int has_nh4 = rnh4 >= 0, has_no3 = rno3 >= 0;
  double c_nh4 = 0.0, c_no3 = 0.0;
  double fin = 1.0, dfin_dnh4 = 0.0, dfin_dno3 = 0.0; /* fnh4_inhibit_no3 */
  if (has_nh4) c_nh4 = c->total[rnh4] * x->theta * 1000.0;
  if (has_no3) c_no3 = c->total[rno3] * x->theta * 1000.0;
  if (sd->inhibition_nh4_no3 > 0.0) {
    if (c_nh4 > sd->x0eps && c_no3 > sd->x0eps) {
      fin = monod(c_nh4 / c_no3, 1.0 / sd->inhibition_nh4_no3, 0);
    } else {
      if (c_nh4 > sd->x0eps && c_no3 <= sd->x0eps)
        fin = 1.0;
      else if (c_nh4 <= sd->x0eps && c_no3 > sd->x0eps)
        fin = 0.0;
      else
        return 0.0; /* both tiny (:2547) */
    }
  }
  double fnh4 = 1.0, dfnh4 = 0.0, fno3 = 1.0, dfno3 = 0.0, fmb = 1.0, dfmb = 0.0;
  monod_inhib(x, r, 1, uc, &fmb, &dfmb, &fnh4, &dfnh4, &fno3, &dfno3, c_nh4, c_no3);
  dcrate_uc_duc = dcrate_uc_duc * fmb + crate_uc * dfmb;
  crate_uc = crate_uc * fmb;
  double f_ox, df_ox; /* Monod oxygen response (:2764-2784) */
  ox_response(x, r, &f_ox, &df_ox);
  dcrate_uc_duc = dcrate_uc_duc * f_ox + crate_uc * df_ox;
  crate_uc = crate_uc * f_ox;
  const int rox = code_row(x, sd->species_o2);
  double feps, dfeps;
  if (has_nh4) {
    if (sd->x0eps > 0.0)
      hsmooth(c_nh4, sd->x0eps * 10.0, sd->x0eps, &feps, &dfeps);
    else {
      feps = 1.0;
      dfeps = 0.0;
    }
    dfnh4 = dfnh4 * feps + fnh4 * dfeps;
    fnh4 = fnh4 * feps;
  }
  if (has_no3) {
    if (sd->x0eps > 0.0)
      hsmooth(c_no3, sd->x0eps * 10.0, sd->x0eps, &feps, &dfeps);
    else {
      feps = 1.0;
      dfeps = 0.0;
    }
    dfno3 = dfno3 * feps + fno3 * dfeps;
    fno3 = fno3 * feps;
  }
  /* N-rate cap (:2829-2858) */
  const double nratecap = -crate_uc * n_st * dt / 0.45;
  if (has_nh4) {
    double fcap = 1.0, dfcap = 0.0;
    if (nratecap * fin > c_nh4 * x->vol) {
      fcap = monod(c_nh4 * x->vol, nratecap * fin - c_nh4 * x->vol, 0);
      dfcap = monod(c_nh4 * x->vol, nratecap * fin - c_nh4 * x->vol, 1);
    }
    dfnh4 = dfnh4 * fcap + fnh4 * dfcap;
    fnh4 = fnh4 * fcap;
  }
  if (has_no3) {
    double fcap = 1.0, dfcap = 0.0;
    if (nratecap * (1.0 - fin) > c_no3 * x->vol) {
      fcap = monod(c_no3 * x->vol, nratecap * (1.0 - fin) - c_no3 * x->vol, 0);
      dfcap = monod(c_no3 * x->vol, nratecap * (1.0 - fin) - c_no3 * x->vol, 1);
    }
    dfno3 = dfno3 * fcap + fno3 * dfcap;
    fno3 = fno3 * fcap;
  }
  const double crate_nh4 = crate_uc * fnh4 * fin;
  const double crate_no3 = crate_uc * fno3 * (1.0 - fin);
  const double crate = crate_nh4 + crate_no3;
  R(x, ruc, crate);
  R(x, rco2, -c_st * crate);
  if (rox >= 0) R(x, rox, c_st * crate); /* :2892-2900 */
  if (sd->upstream_hr_id[r] >= 0) R(x, imm(x, sd->upstream_hr_id[r]), -c_st * crate);
  if (sd->species_id_hr >= 0) R(x, imm(x, sd->species_id_hr), -c_st * crate);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    const int dc = sd->downstream_c_id[r * MD + j];
    if (dc >= 0) R(x, imm(x, dc), -sd->downstream_stoich[r * MD + j] * crate);
  }
  if (varycn) R(x, run, c->sd_upstream_nc[r] * crate);
  double nimm = 0.0;
  if (has_nh4) {
    R(x, rnh4, -n_st * crate_nh4);
    nimm += n_st * crate_nh4;
  }
  if (has_no3) {
    R(x, rno3, -n_st * crate_no3);
    nimm += n_st * crate_no3;
  }
  if (sd->upstream_nimm_id[r] >= 0) R(x, imm(x, sd->upstream_nimm_id[r]), n_st * crate);
  if (sd->upstream_nimp_id[r] >= 0) R(x, imm(x, sd->upstream_nimp_id[r]), n_st * crate_uc);
  if (sd->species_id_nimm >= 0) R(x, imm(x, sd->species_id_nimm), n_st * crate);
  if (sd->species_id_nimp >= 0) R(x, imm(x, sd->species_id_nimp), n_st * crate_uc);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    if (!sd->downstream_is_varycn[r * MD + j]) continue;
    const int dn = sd->downstream_n_id[r * MD + j];
    if (dn >= 0)
      R(x, imm(x, dn),
        -sd->downstream_stoich[r * MD + j] * crate * c->sd_downstream_nc[r * MD + j]);
  }
  if (!deriv) return nimm;
  const double unc = c->sd_upstream_nc[r];
  /* d/duc */
  double dcrate_dx = dcrate_uc_duc * (fnh4 * fin + fno3 - fno3 * fin);
  const double dco2_duc = dcrate_dx * c_st;
  const double duc_duc = -1.0 * dcrate_dx;
  const double dnh4_duc = dcrate_uc_duc * n_st * fnh4 * fin;
  const double dno3_duc = dcrate_uc_duc * n_st * fno3 * (1.0 - fin);
  const double dun_duc = unc * duc_duc;
  /* d/dnh4 */
  dcrate_dx = (dfnh4 * fin + (fnh4 - fno3) * dfin_dnh4) * crate_uc;
  const double dco2_dnh4 = dcrate_dx * c_st;
  const double duc_dnh4 = -1.0 * dcrate_dx;
  const double dnh4_dnh4 = (fnh4 * dfin_dnh4 + dfnh4 * fin) * crate_uc * n_st;
  const double dno3_dnh4 = (-1.0 * fno3 * dfin_dnh4) * crate_uc * n_st;
  const double dun_dnh4 = unc * duc_dnh4;
  /* d/dno3 */
  dcrate_dx = ((fnh4 - fno3) * dfin_dno3 + dfno3 * (1.0 - fnh4 * fin)) * crate_uc;
  const double dco2_dno3 = dcrate_dx * c_st;
  const double duc_dno3 = -1.0 * dcrate_dx;
  const double dnh4_dno3 = fnh4 * dfin_dno3 * crate_uc * n_st;
  const double dno3_dno3 = (-1.0 * fno3 * dfin_dno3 + dfno3 * (1.0 - fin)) * crate_uc * n_st;
  const double dun_dno3 = unc * duc_dno3;

  J(x, rco2, ruc, -dco2_duc);
  if (rox >= 0) J(x, rox, ruc, dco2_duc); /* :3104-3111 */
  if (sd->upstream_hr_id[r] >= 0) J(x, imm(x, sd->upstream_hr_id[r]), ruc, -dco2_duc);
  if (sd->species_id_hr >= 0) J(x, imm(x, sd->species_id_hr), ruc, -dco2_duc);
  J(x, ruc, ruc, -duc_duc);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    const int dc = sd->downstream_c_id[r * MD + j];
    const double ddc = sd->downstream_stoich[r * MD + j] * (-1.0 * duc_duc);
    if (dc >= 0) J(x, imm(x, dc), ruc, -ddc);
  }
  if (has_nh4) {
    J(x, rnh4, ruc, -dnh4_duc);
    if (sd->upstream_nimm_id[r] >= 0) J(x, imm(x, sd->upstream_nimm_id[r]), ruc, dnh4_duc);
    if (sd->species_id_nimm >= 0) J(x, imm(x, sd->species_id_nimm), ruc, dnh4_duc);
  }
  if (has_no3) {
    J(x, rno3, ruc, -dno3_duc);
    if (sd->upstream_nimm_id[r] >= 0) J(x, imm(x, sd->upstream_nimm_id[r]), ruc, dno3_duc);
    if (sd->species_id_nimm >= 0) J(x, imm(x, sd->species_id_nimm), ruc, dno3_duc);
  }
  if (varycn) J(x, run, ruc, -dun_duc);
  for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
    if (!sd->downstream_is_varycn[r * MD + j]) continue;
    const int dn = sd->downstream_n_id[r * MD + j];
    const double ddn =
        sd->downstream_stoich[r * MD + j] * (-1.0 * duc_duc) * c->sd_downstream_nc[r * MD + j];
    if (dn >= 0) J(x, imm(x, dn), ruc, -ddn);
  }
  if (has_nh4) {
    J(x, rco2, rnh4, -dco2_dnh4 * dtot(x, rco2, rnh4));
    if (sd->upstream_hr_id[r] >= 0) J(x, imm(x, sd->upstream_hr_id[r]), rnh4, -dco2_dnh4);
    if (sd->species_id_hr >= 0) J(x, imm(x, sd->species_id_hr), rnh4, -dco2_dnh4);
    J(x, ruc, rnh4, -duc_dnh4);
    for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
      const int dc = sd->downstream_c_id[r * MD + j];
      const double ddc = sd->downstream_stoich[r * MD + j] * (-1.0 * duc_dnh4);
      if (dc >= 0) J(x, imm(x, dc), rnh4, -ddc);
    }
    J(x, rnh4, rnh4, -dnh4_dnh4 * dtot(x, rnh4, rnh4));
    if (sd->upstream_nimm_id[r] >= 0) J(x, imm(x, sd->upstream_nimm_id[r]), rnh4, dnh4_dnh4);
    if (sd->species_id_nimm >= 0) J(x, imm(x, sd->species_id_nimm), rnh4, dnh4_dnh4);
    if (has_no3) {
      J(x, rno3, rnh4, -dno3_dnh4 * dtot(x, rno3, rnh4));
      if (sd->upstream_nimm_id[r] >= 0) J(x, imm(x, sd->upstream_nimm_id[r]), rnh4, dno3_dnh4);
      if (sd->species_id_nimm >= 0) J(x, imm(x, sd->species_id_nimm), rnh4, dno3_dnh4);
    }
    if (varycn) J(x, run, rnh4, -dun_dnh4);
    for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
      if (!sd->downstream_is_varycn[r * MD + j]) continue;
      const int dn = sd->downstream_n_id[r * MD + j];
      const double ddn =
          sd->downstream_stoich[r * MD + j] * (-1.0 * duc_dnh4) * c->sd_downstream_nc[r * MD + j];
      if (dn >= 0) J(x, imm(x, dn), rnh4, -ddn);
    }
  }
  if (has_no3) {
    J(x, rco2, rno3, -dco2_dno3 * dtot(x, rco2, rno3));
    if (sd->upstream_hr_id[r] >= 0) J(x, imm(x, sd->upstream_hr_id[r]), rno3, -dco2_dno3);
    if (sd->species_id_hr >= 0) J(x, imm(x, sd->species_id_hr), rno3, -dco2_dno3);
    J(x, ruc, rno3, -duc_dno3);
    for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
      const int dc = sd->downstream_c_id[r * MD + j];
      const double ddc = sd->downstream_stoich[r * MD + j] * (-1.0 * duc_dno3);
      if (dc >= 0) J(x, imm(x, dc), rno3, -ddc);
    }
    if (has_nh4) {
      J(x, rnh4, rno3, -dnh4_dno3 * dtot(x, rnh4, rno3));
      if (sd->upstream_nimm_id[r] >= 0) J(x, imm(x, sd->upstream_nimm_id[r]), rno3, dnh4_dno3);
      if (sd->species_id_nimm >= 0) J(x, imm(x, sd->species_id_nimm), rno3, dnh4_dno3);
    }
    J(x, rno3, rno3, -dno3_dno3 * dtot(x, rno3, rno3));
    if (sd->upstream_nimm_id[r] >= 0) J(x, imm(x, sd->upstream_nimm_id[r]), rno3, dno3_dno3);
    if (sd->species_id_nimm >= 0) J(x, imm(x, sd->species_id_nimm), rno3, dno3_dno3);
    if (varycn) J(x, run, rno3, -dun_dno3);
    for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
      if (!sd->downstream_is_varycn[r * MD + j]) continue;
      const int dn = sd->downstream_n_id[r * MD + j];
      const double ddn =
          sd->downstream_stoich[r * MD + j] * (-1.0 * duc_dno3) * c->sd_downstream_nc[r * MD + j];
      if (dn >= 0) J(x, imm(x, dn), rno3, -ddn);
    }
  }
  return nimm;
}

/* SomDecNemission :3477-3640 */
static void nemission(sd_ctx* x, double net_nmin_rate, double dt, int deriv) {
  const pfchem_somdec* sd = x->sd;
  orc_cell* c = x->c;
  const int rnh4 = sd->species_id_nh4, rn2o = sd->species_id_n2o;
  if (rn2o < 0 || !(net_nmin_rate > sd->x0eps)) return;
  const double c_nh4 = c->total[rnh4] * x->theta * 1000.0;
  double f_t = -0.06 + 0.13 * exp(0.07 * c->temp);
  double f_w = pow((1.27 - c->sat) / 0.67, 3.1777) * pow((c->sat - 0.0012) / 0.5988, 2.84);
  double ph = 6.5;
  if (sd->species_id_proton >= 0) {
    const int h = sd->species_id_proton;
    ph = -log10(c->pri_molal[h] * c->pri_act_coef[h]);
  }
  double f_ph = 0.56 + atan(RPI * 0.45 * (-5.0 + ph)) / RPI;
  if (!(f_t > sd->x0eps && f_w > sd->x0eps && f_ph > sd->x0eps)) return;
  f_t = fmin(f_t, 1.0);
  f_w = fmin(f_w, 1.0);
  f_ph = fmin(f_ph, 1.0);
  const double tr = f_t * f_w * f_ph;
  double feps, dfeps;
  if (sd->x0eps > 0.0)
    hsmooth(c_nh4, sd->x0eps * 10.0, sd->x0eps, &feps, &dfeps);
  else {
    feps = 1.0;
    dfeps = 0.0;
  }
  const double nratecap = tr * sd->n2o_frac_mineralization * net_nmin_rate * dt;
  double fcap = 1.0, dfcap = 0.0;
  if (nratecap > c_nh4 * x->vol) {
    fcap = monod(c_nh4 * x->vol, nratecap - c_nh4 * x->vol, 0);
    dfcap = monod(c_nh4 * x->vol, nratecap - c_nh4 * x->vol, 1);
  }
  dfeps = dfeps * fcap + feps * dfcap;
  feps = feps * fcap;
  const double rate = tr * sd->n2o_frac_mineralization * net_nmin_rate * feps;
  R(x, rnh4, rate);
  R(x, rn2o, -0.5 * rate);
  if (sd->species_id_ngasmin >= 0) R(x, imm(x, sd->species_id_ngasmin), -rate);
  if (!deriv) return;
  const double d = tr * sd->n2o_frac_mineralization * net_nmin_rate * dfeps;
  J(x, rnh4, rnh4, d * dtot(x, rnh4, rnh4));
  J(x, rn2o, rnh4, -0.5 * d * dtot(x, rn2o, rnh4));
  if (sd->species_id_ngasmin >= 0) J(x, imm(x, sd->species_id_ngasmin), rnh4, -d);
}

/* SomDecReact :1504-1910 */
void orc_somdec_react(const pfchem_network* net, orc_cell* c, double dt, double* res,
                      double* jac, int compute_derivative) {
  const pfchem_somdec* sd = net->somdec;
  sd_ctx x;
  x.net = net;
  x.sd = sd;
  x.c = c;
  x.res = res;
  x.jac = jac;
  x.naq = net->naqcomp;
  x.n = net->naqcomp + net->nimmobile;
  x.theta = c->sat * c->porosity;
  x.vol = c->volume;
  double net_nmin_rate = 0.0;
  for (int r = 0; r < sd->nrxn; ++r) {
    /* moisture: the ELM scalar when given, else LOGTHETA or 1 (:1636-1665) */
    double f_w;
    if (c->has_f_w) {
      f_w = c->f_w;
    } else if (sd->moisture_response_function[r] == 3 /* LOGTHETA */) {
      f_w = x.theta <= 0.08 ? 0.01 : log(x.theta / 0.08) / log(1.0 / 0.08);
    } else {
      f_w = 1.0;
    }
    if (sd->Ox_response_function[r] == 2 /* WFPS */)
      f_w = f_w * wfps_factor(c->sat, sd->Ox_half_saturation[r]);
    /* temperature (:1685-1715) */
    const int tf = sd->temperature_response_function[r];
    double f_t;
    if (tf == 4)
      f_t = temperature_response(c->temp, 4, sd->Ea[r]);
    else if (tf == 1)
      f_t = temperature_response(c->temp, 1, 0.0);
    else if (tf == 2 || tf == 3)
      f_t = temperature_response(c->temp, tf, sd->Q10[r]);
    else
      f_t = c->has_f_t ? c->f_t : 1.0;
    const double f_depth = 1.0;
    if (f_t < 1.0e-20 || f_w < 1.0e-20 || f_depth < 1.0e-20) continue;
    /* rate constant (:1750-1765) */
    double k = 0.0;
    if (sd->rate_constant[r] >= 0.0)
      k = sd->rate_constant[r];
    else if (sd->rate_decomposition[r] >= 0.0)
      k = (1.0 - exp(-sd->rate_decomposition[r] * dt)) / dt;
    k = sd->rate_ad_factor[r] * k;
    k = fmin(k, 1.0 / dt);
    const double scaled = k * x.vol * f_t * f_w * f_depth;
    const int uc = sd->upstream_c_id[r];
    const double c_uc = c->immobile[uc];
    double feps, dfeps;
    if (sd->x0eps > 0.0) {
      hsmooth(c_uc, sd->x0eps * 10.0, sd->x0eps, &feps, &dfeps);
    } else {
      feps = 1.0;
      dfeps = 0.0;
      if (c_uc <= sd->x0eps) continue;
    }
    const double crate_uc = scaled * c_uc * feps;
    const double dcrate = scaled * (feps + c_uc * dfeps);
    /* variable C:N stoichiometry (:1805-1860) */
    for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
      if (!sd->downstream_is_varycn[r * MD + j]) continue;
      const int dc = sd->downstream_c_id[r * MD + j], dn = sd->downstream_n_id[r * MD + j];
      if (dc >= 0 && dn >= 0) {
        const double c_dc = c->immobile[dc], c_dn = c->immobile[dn];
        if (c_dn >= sd->x0eps && c_dc >= sd->x0eps) c->sd_downstream_nc[r * MD + j] = c_dn / c_dc;
      }
    }
    if (sd->upstream_is_varycn[r]) {
      const int un = sd->upstream_n_id[r];
      if (un >= 0) {
        const double c_un = c->immobile[un];
        if (c_un >= sd->x0eps && c_uc >= sd->x0eps) c->sd_upstream_nc[r] = c_un / c_uc;
        double st_c = 1.0;
        for (int j = 0; j < sd->n_downstream_pools[r]; ++j) st_c -= sd->downstream_stoich[r * MD + j];
        c->sd_mineral_c_stoich[r] = st_c;
        double st_n = c->sd_upstream_nc[r];
        for (int j = 0; j < sd->n_downstream_pools[r]; ++j)
          st_n -= sd->downstream_stoich[r * MD + j] * c->sd_downstream_nc[r * MD + j];
        c->sd_mineral_n_stoich[r] = st_n;
      }
    }
    if (c->sd_mineral_n_stoich[r] >= 0.0)
      net_nmin_rate += react1(&x, r, crate_uc, dcrate, compute_derivative);
    else
      net_nmin_rate += react2(&x, r, crate_uc, dcrate, dt, compute_derivative);
  }
  if (net_nmin_rate > sd->x0eps) nemission(&x, net_nmin_rate, dt, compute_derivative);
}
