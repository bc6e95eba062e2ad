// pfchem_somdec.hpp -- device side of the SomDec decomposition sandbox
// (reaction_sandbox_somdec.F90): residual and Jacobian contributions of one
// cell, shared by the one-cell-per-thread and the team kernels through an
// accumulator `Env`:
//   double total(i), imm(i), lnact(i), dtot(i, j)   cell values
//   double theta, vol, sat, temp, f_t, f_w; bool has_f_t, has_f_w
//   void R(row, v), J(row, col, v)                  residual / d res / d c
//   double &unc(r), &dnc(r, j), &mcs(r), &mns(r)    per-cell stoichiometry
//   int naq                                         first immobile dof
//   static constexpr bool kExt                      oxygen / complexes
// Kernels instantiate the sandbox twice: Env::kExt = false is the sandbox
// without the Monod oxygen response, O2 tracking and aqueous complexes (the
// code a plain ELM-CN network executes, kept as small as before those were
// added: the kernels are large enough for their instruction-cache footprint
// to show), Env::kExt = true carries them; DevSomdec::ext selects.
// Statements follow SomDecReact (:1504-1910), SomDecReact1 (:1914-2418),
// SomDecReact2 (:2423-3472) and SomDecNemission (:3477-3640), including the
// reference's Jacobian entries that are not exact derivatives (they steer
// the Newton iteration, not its solution).  The variable C:N stoichiometry
// is per cell, starting from the Setup values at each step (the reference
// keeps it in the shared sandbox object, :1817-1858).
#pragma once
#include "pfchem_device.hpp"

namespace pfchem {

constexpr int kSdDown = PFCHEM_SOMDEC_MAX_DOWN;
constexpr int kSdMonod = PFCHEM_SOMDEC_MAX_MONOD;
constexpr int kSdInhib = PFCHEM_SOMDEC_MAX_INHIB;
constexpr int kSdMaxRxn = 32;
constexpr int kSdPairs = 8;  // the dtotal pair sums (see sd_pair_sums)
// LDS doubles of a team's per-cell SomDec state: the stoichiometry of
// every reaction (upstream N:C, mineral C and N, downstream N:C) and then
// the dtotal pair sums
__host__ __device__ constexpr int sd_lds_doubles(int nrxn) {
  return nrxn * (3 + kSdDown) + kSdPairs;
}
// the same for a kernel level: 2 (oxygen / complexes) keeps the pair sums,
// 1 only the stoichiometry, 0 nothing
__host__ __device__ constexpr int sd_lds_doubles(int nrxn, int level) {
  return level >= 2 ? sd_lds_doubles(nrxn) : level == 1 ? nrxn * (3 + kSdDown) : 0;
}

// HfunctionSmooth (utility.F90:2542-2597)
__device__ inline void sd_hsmooth(double x, double x1, double x0, double* H, double* dH) {
  if (fabs(x1 - x0) < 1.0e-50) {
    *H = copysign(0.5, x - x1) + 0.5;
    *dH = 0.0;
    return;
  }
  const double u = (x - x0) / (x1 - x0);
  if (u < 0.0) {
    *H = *dH = 0.0;
  } else if (u > 1.0) {
    *H = 1.0;
    *dH = 0.0;
  } else {
    const double xs = 1.0 - (x - x0) * (x - x0) / (x1 - x0) / (x1 - x0);
    *H = 1.0 - xs * xs;
    *dH = 4.0 * xs * (x - x0) / (x1 - x0) / (x1 - x0);
  }
}

// FuncMonod / FuncInhibition (elm_rspfuncs.F90:320-375)
__device__ inline double sd_monod(double c, double k, int deriv) {
  return deriv ? k / (c + k) / (c + k) : c / (c + k);
}
__device__ inline double sd_inhib(int deriv, double c, double k) {
  return deriv ? -k / (c + k) / (c + k) : k / (c + k);
}
__device__ inline double sd_inhib_threshold(int deriv, double c, double k, double k2) {
  if (!deriv) return 0.5 + atan((c - k) * k2) / M_PI;
  const double t = (c - k) * k2;
  return (k2 / (1.0 + t * t)) / M_PI;
}

// GetTemperatureResponse (elm_rspfuncs.F90:61-120)
__device__ inline double sd_temperature(double tc, int itype, double q) {
  if (itype == 2) {
    if (tc > 0.0) return pow(q, (tc - 25.0) / 10.0);
    return pow(q, -25.0 / 10.0) * pow(2.0, tc / 10.0);
  }
  if (itype == 1) {
    const double tk = tc + 273.15;
    return tk > 227.15 ? exp(308.56 * (1.408054069e-2 - 1.0 / (tk - 227.13))) : 0.0;
  }
  if (itype == 3) {
    if (tc < -5.0) return 0.0;
    if (tc >= 30.0) return 1.0;
    return pow(q, (tc - 30.0) / 10.0);
  }
  if (itype == 4) return exp(q / kIdealGas * (1.0 / 298.15 - 1.0 / (tc + 273.15)));
  return 1.0;
}

// GetAerobicCondition, WFPS (elm_rspfuncs.F90:282-316)
__device__ inline double sd_wfps(double sat) {
  const double a = 1.27 - sat, b = sat - 0.0012;
  if (a <= 0.0 || b <= 0.0) return 0.0;
  return fmin(pow(a / 0.67, 3.1777) * pow(b / 0.5988, 2.84), 1.0);
}

// The dtotal entries the sandbox reads (React2's NH4+ / NO3- columns and
// SomDecNemission): rows CO2, NH4+, NO3-, N2O by columns NH4+, NO3-.  With
// aqueous complexes, dtotal_ij = rho/1000 (delta_ij + S_ij / m_j) with
// S_ij = sum over complexes of nu_i nu_j m_k; the kernels accumulate these
// eight S_ij from their complex loops (their Jacobian tiles take other
// terms before the sandbox runs).
__device__ inline int sd_dtot_row(const DevSomdec& sd, int i) {
  if (i < 0) return -1;
  return i == sd.co2 ? 0 : i == sd.nh4 ? 1 : i == sd.no3 ? 2 : i == sd.n2o ? 3 : -1;
}
__device__ inline int sd_dtot_col(const DevSomdec& sd, int j) {
  if (j < 0) return -1;
  return j == sd.nh4 ? 0 : j == sd.no3 ? 1 : -1;
}
// complex k's share nu_i nu_j m_k of the pairs above, added into the
// team's LDS sums S[2 a + b] (rows a, columns b)
template <class P>
__device__ inline void sd_pair_sums(const DevSomdec& sd, const int* ids, const double* st,
                                    double mk, P S) {
  const int ncs = gld(ids);
  for (int a = 0; a < ncs; ++a) {
    const int ra = sd_dtot_row(sd, gld(ids + 1 + a));
    if (ra < 0) continue;
    const double wa = gld(st + a) * mk;
    for (int b = 0; b < ncs; ++b) {
      const int cb = sd_dtot_col(sd, gld(ids + 1 + b));
      if (cb >= 0)
        __hip_atomic_fetch_add(&S[2 * ra + cb], wa * gld(st + b), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
}
// dtotal_ij from the pair sums (i, j among the pairs; else the diagonal)
template <class P>
__device__ inline double sd_dtot_value(const DevSomdec& sd, P S, double dkl, double xj, int i,
                                       int j) {
  const int a = sd_dtot_row(sd, i), b = sd_dtot_col(sd, j);
  const double d = i == j ? 1.0 : 0.0;
  return a >= 0 && b >= 0 ? dkl * (d + S[2 * a + b] / xj) : dkl * d;
}

template <class Env>
__device__ inline double sd_species_conc(Env& e, int id) {
  return id >= 0 ? e.total(id) : e.imm(-id - 1);
}

// species of a 1-based code (aqueous id + 1 > 0, -(immobile id + 1) < 0):
// its concentration (SomDec_SpeciesConc, :3809-3870: total for aqueous,
// immobile concentration) and its residual row (-1 for code 0)
template <class Env>
__device__ inline double sd_code_conc(Env& e, int code) {
  return code > 0 ? e.total(code - 1) : e.imm(-code - 1);
}
template <class Env>
__device__ inline int sd_code_row(Env& e, int code) {
  return code > 0 ? code - 1 : (code < 0 ? e.naq - code - 1 : -1);
}

// Monod oxygen response of reaction r (GetAerobicCondition,
// elm_rspfuncs.F90:282-316, at the oxidizer's concentration; :2129-2147,
// 2764-2784): f = Ox / (K + Ox), and df the derivative in Ox that the
// reference adds to the rate's upstream-pool derivative
template <class Env>
__device__ inline void sd_ox(const DevSomdec& sd, Env& e, int r, double* f, double* df) {
  *f = 1.0;
  *df = 0.0;
  if (gld(sd.Ox_response_function + r) != 1) return;
  const int code = gld(sd.Ox_specid + r);
  if (code == 0) return;
  const double ox = sd_code_conc(e, code), k = gld(sd.Ox_half_saturation + r);
  *f = sd_monod(ox, k, 0);
  *df = sd_monod(ox, k, 1);
}

// general Monod / inhibition factors of reaction r; React2 handles its own
// NH4+ / NO3- Monod terms (skip = 1)
template <class Env>
__device__ void sd_monod_inhib(const DevSomdec& sd, Env& e, int r, int skip, int uc,
                               double c_nh4, double c_no3, double* fmb, double* dfmb,
                               double* fnh4, double* dfnh4, double* fno3, double* dfno3) {
  const int nmo = gld(sd.n_monod + r);
  for (int q = 0; q < nmo; ++q) {
    const int id = gld(sd.monod_specid + r * kSdMonod + q);
    const double k = gld(sd.monod_half_saturation + r * kSdMonod + q);
    const double th = gld(sd.monod_threshold + r * kSdMonod + q);
    const int pn = gld(sd.monod_pool_normalized + r * kSdMonod + q);
    if (skip && id >= 0 && id == sd.nh4) {
      double t = fmax(0.0, c_nh4 - th);
      if (pn) t = t / e.imm(uc);
      *fnh4 = sd_monod(t, k, 0);
      *dfnh4 = sd_monod(t, k, 1);
      continue;
    }
    if (skip && id >= 0 && sd.no3 >= 0 && id == sd.no3) {
      double t = fmax(0.0, c_no3 - th);
      if (pn) t = t / e.imm(uc);
      *fno3 = sd_monod(t, k, 0);
      *dfno3 = sd_monod(t, k, 1);
      continue;
    }
    double t = fmax(0.0, sd_species_conc(e, id) - th);
    if (pn) {
      t = t / e.imm(uc);
      if (id >= 0) t = t * e.theta * 1000.0;
    }
    const double fx = sd_monod(t, k, 0);
    const double dfx = id == -uc - 1 ? sd_monod(t, k, 1) : 0.0;
    *dfmb = *dfmb * fx + *fmb * dfx;
    *fmb = *fmb * fx;
  }
  const int nin = gld(sd.n_inhibition + r);
  for (int q = 0; q < nin; ++q) {
    const int id = gld(sd.inhibition_specid + r * kSdInhib + q);
    const int ty = gld(sd.inhibition_type + r * kSdInhib + q);
    const double k = gld(sd.inhibition_constant + r * kSdInhib + q);
    const double t = sd_species_conc(e, id);
    double fx = 1.0, dfx = 0.0;
    if (ty == 1) {
      const double k2 = gld(sd.inhibition_constant2 + r * kSdInhib + q);
      fx = sd_inhib_threshold(0, t, k, k2);
      dfx = sd_inhib_threshold(1, t, k, k2);
    } else if (ty == 3) {
      fx = sd_inhib(0, t, k);
      dfx = sd_inhib(1, t, k);
    } else if (ty == 4) {
      fx = sd_monod(t, k, 0);
      dfx = sd_monod(t, k, 1);
    }
    if (id != -uc - 1) dfx = 0.0;
    *dfmb = *dfmb * fx + *fmb * dfx;
    *fmb = *fmb * fx;
  }
}

// SomDecReact1 (:1914-2418), N mineralization; returns the mineralized N rate
template <class Env>
__device__ double sd_react1(const DevSomdec& sd, Env& e, int r, double crate_uc, double dcrate,
                            bool deriv) {
  const double n_st = e.mns(r), c_st = e.mcs(r);
  const int uc = gld(sd.upstream_c_id + r);
  const bool vary = gld(sd.upstream_is_varycn + r) != 0;
  const int ruc = e.naq + uc, run = vary ? e.naq + gld(sd.upstream_n_id + r) : -1;
  const int uhr = gld(sd.upstream_hr_id + r), unmin = gld(sd.upstream_nmin_id + r);
  const int nd = gld(sd.n_downstream_pools + r);
  double fmb = 1.0, dfmb = 0.0, d1 = 0, d2 = 0, d3 = 0, d4 = 0;
  sd_monod_inhib(sd, e, r, 0, uc, 0.0, 0.0, &fmb, &dfmb, &d1, &d2, &d3, &d4);
  dcrate = dcrate * fmb + crate_uc * dfmb;
  double crate = crate_uc * fmb;
  int rox = -1;
  if constexpr (Env::kExt) {
    double f_ox, df_ox;
    sd_ox(sd, e, r, &f_ox, &df_ox);
    dcrate = dcrate * f_ox + crate * df_ox;
    crate = crate * f_ox;
    rox = sd_code_row(e, sd.o2);
  }
  e.R(ruc, crate);
  e.R(sd.co2, -c_st * crate);
  if (uhr >= 0) e.R(e.naq + uhr, -c_st * crate);
  if (sd.hr >= 0) e.R(e.naq + sd.hr, -c_st * crate);
  if constexpr (Env::kExt)
    if (rox >= 0) e.R(rox, c_st * crate);  // O2 consumed as CO2 is released (:2176-2186)
  for (int j = 0; j < nd; ++j) {
    const int dc = gld(sd.downstream_c_id + r * kSdDown + j);
    if (dc >= 0) e.R(e.naq + dc, -gld(sd.downstream_stoich + r * kSdDown + j) * crate);
  }
  if (vary) e.R(run, e.unc(r) * crate);
  e.R(sd.nh4, -n_st * crate);
  if (unmin >= 0) e.R(e.naq + unmin, -n_st * crate);
  if (sd.nmin >= 0) e.R(e.naq + sd.nmin, -n_st * crate);
  for (int j = 0; j < nd; ++j) {
    if (!gld(sd.downstream_is_varycn + r * kSdDown + j)) continue;
    const int dn = gld(sd.downstream_n_id + r * kSdDown + j);
    if (dn >= 0)
      e.R(e.naq + dn, -gld(sd.downstream_stoich + r * kSdDown + j) * crate * e.dnc(r, j));
  }
  if (deriv) {
    const double dco2 = dcrate * c_st, duc = -1.0 * dcrate;
    const double dnh4 = dco2 * n_st;  // as the reference writes it (:2273)
    e.J(sd.co2, ruc, -dco2);
    if constexpr (Env::kExt)
      if (rox >= 0) e.J(rox, ruc, dco2);  // (:2300-2308)
    if (uhr >= 0) e.J(e.naq + uhr, ruc, -dco2);
    if (sd.hr >= 0) e.J(e.naq + sd.hr, ruc, -dco2);
    e.J(ruc, ruc, -duc);
    for (int j = 0; j < nd; ++j) {
      const int dc = gld(sd.downstream_c_id + r * kSdDown + j);
      if (dc >= 0) e.J(e.naq + dc, ruc, gld(sd.downstream_stoich + r * kSdDown + j) * duc);
    }
    e.J(sd.nh4, ruc, -dnh4);
    if (unmin >= 0) e.J(e.naq + unmin, ruc, -dnh4);
    if (sd.nmin >= 0) e.J(e.naq + sd.nmin, ruc, -dnh4);
    if (vary) e.J(run, ruc, -e.unc(r) * duc);
    for (int j = 0; j < nd; ++j) {
      if (!gld(sd.downstream_is_varycn + r * kSdDown + j)) continue;
      const int dn = gld(sd.downstream_n_id + r * kSdDown + j);
      if (dn >= 0)
        e.J(e.naq + dn, ruc, gld(sd.downstream_stoich + r * kSdDown + j) * duc * e.dnc(r, j));
    }
  }
  return n_st * crate;
}

// SomDecReact2 (:2423-3472), N immobilization; returns the immobilized N
// rate (<= 0)
template <class Env>
__device__ double sd_react2(const DevSomdec& sd, Env& e, int r, double crate_uc, double dcrate,
                            double dt, bool deriv) {
  const double n_st = e.mns(r), c_st = e.mcs(r);
  const int uc = gld(sd.upstream_c_id + r);
  const bool vary = gld(sd.upstream_is_varycn + r) != 0;
  const int ruc = e.naq + uc, run = vary ? e.naq + gld(sd.upstream_n_id + r) : -1;
  const int uhr = gld(sd.upstream_hr_id + r), unimm = gld(sd.upstream_nimm_id + r);
  const int unimp = gld(sd.upstream_nimp_id + r);
  const int nd = gld(sd.n_downstream_pools + r);
  const int rnh4 = sd.nh4, rno3 = sd.no3;
  const bool has_nh4 = rnh4 >= 0, has_no3 = rno3 >= 0;
  const double c_nh4 = has_nh4 ? e.total(rnh4) * e.theta * 1000.0 : 0.0;
  const double c_no3 = has_no3 ? e.total(rno3) * e.theta * 1000.0 : 0.0;
  double fin = 1.0;
  const double dfin_dnh4 = 0.0, dfin_dno3 = 0.0;  // the split's derivative is off (:2638)
  if (sd.inhib_nh4_no3 > 0.0) {
    if (c_nh4 > sd.x0eps && c_no3 > sd.x0eps) {
      fin = sd_monod(c_nh4 / c_no3, 1.0 / sd.inhib_nh4_no3, 0);
    } else if (c_nh4 > sd.x0eps && c_no3 <= sd.x0eps) {
      fin = 1.0;
    } else if (c_nh4 <= sd.x0eps && c_no3 > sd.x0eps) {
      fin = 0.0;
    } else {
      return 0.0;  // both tiny (:2547)
    }
  }
  double fnh4 = 1.0, dfnh4 = 0.0, fno3 = 1.0, dfno3 = 0.0, fmb = 1.0, dfmb = 0.0;
  sd_monod_inhib(sd, e, r, 1, uc, c_nh4, c_no3, &fmb, &dfmb, &fnh4, &dfnh4, &fno3, &dfno3);
  dcrate = dcrate * fmb + crate_uc * dfmb;
  crate_uc = crate_uc * fmb;
  int rox = -1;
  if constexpr (Env::kExt) {
    double f_ox, df_ox;
    sd_ox(sd, e, r, &f_ox, &df_ox);
    dcrate = dcrate * f_ox + crate_uc * df_ox;
    crate_uc = crate_uc * f_ox;
    rox = sd_code_row(e, sd.o2);
  }
  double h, dh;
  if (has_nh4) {
    if (sd.x0eps > 0.0) {
      sd_hsmooth(c_nh4, sd.x0eps * 10.0, sd.x0eps, &h, &dh);
    } else {
      h = 1.0;
      dh = 0.0;
    }
    dfnh4 = dfnh4 * h + fnh4 * dh;
    fnh4 = fnh4 * h;
  }
  if (has_no3) {
    if (sd.x0eps > 0.0) {
      sd_hsmooth(c_no3, sd.x0eps * 10.0, sd.x0eps, &h, &dh);
    } else {
      h = 1.0;
      dh = 0.0;
    }
    dfno3 = dfno3 * h + fno3 * dh;
    fno3 = fno3 * h;
  }
  // N-rate cap (:2829-2858)
  const double cap = -crate_uc * n_st * dt / 0.45;
  if (has_nh4 && cap * fin > c_nh4 * e.vol) {
    const double f = sd_monod(c_nh4 * e.vol, cap * fin - c_nh4 * e.vol, 0);
    const double df = sd_monod(c_nh4 * e.vol, cap * fin - c_nh4 * e.vol, 1);
    dfnh4 = dfnh4 * f + fnh4 * df;
    fnh4 = fnh4 * f;
  }
  if (has_no3 && cap * (1.0 - fin) > c_no3 * e.vol) {
    const double f = sd_monod(c_no3 * e.vol, cap * (1.0 - fin) - c_no3 * e.vol, 0);
    const double df = sd_monod(c_no3 * e.vol, cap * (1.0 - fin) - c_no3 * e.vol, 1);
    dfno3 = dfno3 * f + fno3 * df;
    fno3 = fno3 * f;
  }
  const double crate_nh4 = crate_uc * fnh4 * fin;
  const double crate_no3 = crate_uc * fno3 * (1.0 - fin);
  const double crate = crate_nh4 + crate_no3;
  e.R(ruc, crate);
  e.R(sd.co2, -c_st * crate);
  if (uhr >= 0) e.R(e.naq + uhr, -c_st * crate);
  if (sd.hr >= 0) e.R(e.naq + sd.hr, -c_st * crate);
  if constexpr (Env::kExt)
    if (rox >= 0) e.R(rox, c_st * crate);  // (:2892-2900)
  for (int j = 0; j < nd; ++j) {
    const int dc = gld(sd.downstream_c_id + r * kSdDown + j);
    if (dc >= 0) e.R(e.naq + dc, -gld(sd.downstream_stoich + r * kSdDown + j) * crate);
  }
  if (vary) e.R(run, e.unc(r) * crate);
  double nimm = 0.0;
  if (has_nh4) {
    e.R(rnh4, -n_st * crate_nh4);
    nimm += n_st * crate_nh4;
  }
  if (has_no3) {
    e.R(rno3, -n_st * crate_no3);
    nimm += n_st * crate_no3;
  }
  if (unimm >= 0) e.R(e.naq + unimm, n_st * crate);
  if (unimp >= 0) e.R(e.naq + unimp, n_st * crate_uc);
  if (sd.nimm >= 0) e.R(e.naq + sd.nimm, n_st * crate);
  if (sd.nimp >= 0) e.R(e.naq + sd.nimp, n_st * crate_uc);
  for (int j = 0; j < nd; ++j) {
    if (!gld(sd.downstream_is_varycn + r * kSdDown + j)) continue;
    const int dn = gld(sd.downstream_n_id + r * kSdDown + j);
    if (dn >= 0)
      e.R(e.naq + dn, -gld(sd.downstream_stoich + r * kSdDown + j) * crate * e.dnc(r, j));
  }
  if (!deriv) return nimm;
  const double unc = e.unc(r);
  double dx = dcrate * (fnh4 * fin + fno3 - fno3 * fin);
  const double dco2_duc = dx * c_st, duc_duc = -1.0 * dx;
  const double dnh4_duc = dcrate * n_st * fnh4 * fin;
  const double dno3_duc = dcrate * n_st * fno3 * (1.0 - fin);
  dx = (dfnh4 * fin + (fnh4 - fno3) * dfin_dnh4) * crate_uc;
  const double dco2_dnh4 = dx * c_st, duc_dnh4 = -1.0 * dx;
  const double dnh4_dnh4 = (fnh4 * dfin_dnh4 + dfnh4 * fin) * crate_uc * n_st;
  const double dno3_dnh4 = (-1.0 * fno3 * dfin_dnh4) * crate_uc * n_st;
  dx = ((fnh4 - fno3) * dfin_dno3 + dfno3 * (1.0 - fnh4 * fin)) * crate_uc;
  const double dco2_dno3 = dx * c_st, duc_dno3 = -1.0 * dx;
  const double dnh4_dno3 = fnh4 * dfin_dno3 * crate_uc * n_st;
  const double dno3_dno3 = (-1.0 * fno3 * dfin_dno3 + dfno3 * (1.0 - fin)) * crate_uc * n_st;
  // column of the upstream pool
  e.J(sd.co2, ruc, -dco2_duc);
  if constexpr (Env::kExt)
    if (rox >= 0) e.J(rox, ruc, dco2_duc);  // (:3104-3111)
  if (uhr >= 0) e.J(e.naq + uhr, ruc, -dco2_duc);
  if (sd.hr >= 0) e.J(e.naq + sd.hr, ruc, -dco2_duc);
  e.J(ruc, ruc, -duc_duc);
  for (int j = 0; j < nd; ++j) {
    const int dc = gld(sd.downstream_c_id + r * kSdDown + j);
    if (dc >= 0) e.J(e.naq + dc, ruc, gld(sd.downstream_stoich + r * kSdDown + j) * duc_duc);
  }
  if (has_nh4) {
    e.J(rnh4, ruc, -dnh4_duc);
    if (unimm >= 0) e.J(e.naq + unimm, ruc, dnh4_duc);
    if (sd.nimm >= 0) e.J(e.naq + sd.nimm, ruc, dnh4_duc);
  }
  if (has_no3) {
    e.J(rno3, ruc, -dno3_duc);
    if (unimm >= 0) e.J(e.naq + unimm, ruc, dno3_duc);
    if (sd.nimm >= 0) e.J(e.naq + sd.nimm, ruc, dno3_duc);
  }
  if (vary) e.J(run, ruc, -unc * duc_duc);
  for (int j = 0; j < nd; ++j) {
    if (!gld(sd.downstream_is_varycn + r * kSdDown + j)) continue;
    const int dn = gld(sd.downstream_n_id + r * kSdDown + j);
    if (dn >= 0)
      e.J(e.naq + dn, ruc, gld(sd.downstream_stoich + r * kSdDown + j) * duc_duc * e.dnc(r, j));
  }
  // columns of NH4+ and NO3- (the factors the reference writes, :3180-3470)
  for (int col = 0; col < 2; ++col) {
    const int rc = col == 0 ? rnh4 : rno3;
    if (rc < 0) continue;
    const double dco2 = col == 0 ? dco2_dnh4 : dco2_dno3;
    const double duc = col == 0 ? duc_dnh4 : duc_dno3;
    e.J(sd.co2, rc, -dco2 * e.dtot(sd.co2, rc));
    if (uhr >= 0) e.J(e.naq + uhr, rc, -dco2);
    if (sd.hr >= 0) e.J(e.naq + sd.hr, rc, -dco2);
    e.J(ruc, rc, -duc);
    for (int j = 0; j < nd; ++j) {
      const int dc = gld(sd.downstream_c_id + r * kSdDown + j);
      if (dc >= 0) e.J(e.naq + dc, rc, gld(sd.downstream_stoich + r * kSdDown + j) * duc);
    }
    const double dnh4 = col == 0 ? dnh4_dnh4 : dnh4_dno3;
    const double dno3 = col == 0 ? dno3_dnh4 : dno3_dno3;
    if (col == 0 || has_nh4) {
      e.J(rnh4, rc, -dnh4 * e.dtot(rnh4, rc));
      if (unimm >= 0) e.J(e.naq + unimm, rc, dnh4);
      if (sd.nimm >= 0) e.J(e.naq + sd.nimm, rc, dnh4);
    }
    if (col == 1 || has_no3) {
      e.J(rno3, rc, -dno3 * e.dtot(rno3, rc));
      if (unimm >= 0) e.J(e.naq + unimm, rc, dno3);
      if (sd.nimm >= 0) e.J(e.naq + sd.nimm, rc, dno3);
    }
    if (vary) e.J(run, rc, -unc * duc);
    for (int j = 0; j < nd; ++j) {
      if (!gld(sd.downstream_is_varycn + r * kSdDown + j)) continue;
      const int dn = gld(sd.downstream_n_id + r * kSdDown + j);
      if (dn >= 0)
        e.J(e.naq + dn, rc, gld(sd.downstream_stoich + r * kSdDown + j) * duc * e.dnc(r, j));
    }
  }
  return nimm;
}

// SomDecReact's per-reaction part (:1620-1880): abiotic factors, rate,
// variable C:N stoichiometry, then React1 or React2; returns the reaction's
// N mineralization (> 0) or immobilization (< 0) rate
template <class Env>
__device__ double sd_reaction(const DevSomdec& sd, Env& e, int r, double dt, bool deriv) {
  double f_w;
  const int mf = gld(sd.moisture_response_function + r);
  if (e.has_f_w) {
    f_w = e.f_w;
  } else if (mf == 3) {
    f_w = e.theta <= 0.08 ? 0.01 : log(e.theta / 0.08) / log(1.0 / 0.08);
  } else {
    f_w = 1.0;
  }
  if (gld(sd.Ox_response_function + r) == 2) f_w = f_w * sd_wfps(e.sat);
  const int tf = gld(sd.temperature_response_function + r);
  double f_t;
  if (tf == 4)
    f_t = sd_temperature(e.temp, 4, gld(sd.Ea + r));
  else if (tf == 1)
    f_t = sd_temperature(e.temp, 1, 0.0);
  else if (tf == 2 || tf == 3)
    f_t = sd_temperature(e.temp, tf, gld(sd.Q10 + r));
  else
    f_t = e.has_f_t ? e.f_t : 1.0;
  if (f_t < 1.0e-20 || f_w < 1.0e-20) return 0.0;
  double k = 0.0;
  const double k0 = gld(sd.rate_constant + r), kd = gld(sd.rate_decomposition + r);
  if (k0 >= 0.0)
    k = k0;
  else if (kd >= 0.0)
    k = (1.0 - exp(-kd * dt)) / dt;
  k = fmin(gld(sd.rate_ad_factor + r) * k, 1.0 / dt);
  const double scaled = k * e.vol * f_t * f_w;
  const int uc = gld(sd.upstream_c_id + r);
  const double c_uc = e.imm(uc);
  double h, dh;
  if (sd.x0eps > 0.0) {
    sd_hsmooth(c_uc, sd.x0eps * 10.0, sd.x0eps, &h, &dh);
  } else {
    h = 1.0;
    dh = 0.0;
    if (c_uc <= sd.x0eps) return 0.0;
  }
  const double crate_uc = scaled * c_uc * h;
  const double dcrate = scaled * (h + c_uc * dh);
  const int nd = gld(sd.n_downstream_pools + r);
  for (int j = 0; j < nd; ++j) {
    if (!gld(sd.downstream_is_varycn + r * kSdDown + j)) continue;
    const int dc = gld(sd.downstream_c_id + r * kSdDown + j);
    const int dn = gld(sd.downstream_n_id + r * kSdDown + j);
    if (dc >= 0 && dn >= 0) {
      const double c_dc = e.imm(dc), c_dn = e.imm(dn);
      if (c_dn >= sd.x0eps && c_dc >= sd.x0eps) e.dnc(r, j) = c_dn / c_dc;
    }
  }
  if (gld(sd.upstream_is_varycn + r)) {
    const int un = gld(sd.upstream_n_id + r);
    if (un >= 0) {
      const double c_un = e.imm(un);
      if (c_un >= sd.x0eps && c_uc >= sd.x0eps) e.unc(r) = c_un / c_uc;
      double st_c = 1.0, st_n = e.unc(r);
      for (int j = 0; j < nd; ++j) {
        const double ds = gld(sd.downstream_stoich + r * kSdDown + j);
        st_c -= ds;
        st_n -= ds * e.dnc(r, j);
      }
      e.mcs(r) = st_c;
      e.mns(r) = st_n;
    }
  }
  if (e.mns(r) >= 0.0) return sd_react1(sd, e, r, crate_uc, dcrate, deriv);
  return sd_react2(sd, e, r, crate_uc, dcrate, dt, deriv);
}

// SomDecNemission (:3477-3640)
template <class Env>
__device__ void sd_nemission(const DevSomdec& sd, Env& e, double net_nmin, double dt,
                             bool deriv) {
  if (sd.n2o < 0 || !(net_nmin > sd.x0eps)) return;
  const double c_nh4 = e.total(sd.nh4) * e.theta * 1000.0;
  double f_t = -0.06 + 0.13 * exp(0.07 * e.temp);
  double f_w = pow((1.27 - e.sat) / 0.67, 3.1777) * pow((e.sat - 0.0012) / 0.5988, 2.84);
  double ph = 6.5;
  if (sd.proton >= 0) ph = -e.lnact(sd.proton) / kLogToLn;
  double f_ph = 0.56 + atan(3.14159265358979323846 * 0.45 * (-5.0 + ph)) /
                           3.14159265358979323846;
  if (!(f_t > sd.x0eps && f_w > sd.x0eps && f_ph > sd.x0eps)) return;
  const double tr = fmin(f_t, 1.0) * fmin(f_w, 1.0) * fmin(f_ph, 1.0);
  double h, dh;
  if (sd.x0eps > 0.0) {
    sd_hsmooth(c_nh4, sd.x0eps * 10.0, sd.x0eps, &h, &dh);
  } else {
    h = 1.0;
    dh = 0.0;
  }
  const double cap = tr * sd.n2o_frac * net_nmin * dt;
  double f = 1.0, df = 0.0;
  if (cap > c_nh4 * e.vol) {
    f = sd_monod(c_nh4 * e.vol, cap - c_nh4 * e.vol, 0);
    df = sd_monod(c_nh4 * e.vol, cap - c_nh4 * e.vol, 1);
  }
  dh = dh * f + h * df;
  h = h * f;
  const double rate = tr * sd.n2o_frac * net_nmin * h;
  e.R(sd.nh4, rate);
  e.R(sd.n2o, -0.5 * rate);
  if (sd.ngasmin >= 0) e.R(e.naq + sd.ngasmin, -rate);
  if (!deriv) return;
  const double d = tr * sd.n2o_frac * net_nmin * dh;
  e.J(sd.nh4, sd.nh4, d * e.dtot(sd.nh4, sd.nh4));
  e.J(sd.n2o, sd.nh4, -0.5 * d * e.dtot(sd.n2o, sd.nh4));
  if (sd.ngasmin >= 0) e.J(e.naq + sd.ngasmin, sd.nh4, -d);
}

}  // namespace pfchem
