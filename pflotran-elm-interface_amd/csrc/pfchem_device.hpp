// pfchem_device.hpp -- device-side network descriptor and common helpers for
// the gfx950 chemistry-step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pfchem.h"

namespace pfchem {

// pflotran_constants.F90:84, 89; reactive_transport_aux.F90:19
constexpr double kLogToLn = 2.30258509299;
constexpr double kIdealGas = 8.31446;
constexpr double kRtMinSaturation = 1.0e-40;

// RReact calls allowed per RStep when pfchem_options.maximum_rreact_calls
// is 0.  The reference's sub-step loop has no bound (its cut counter can go
// negative after runs of good sub-steps, reaction.F90:3721-3725); every
// kernel loop here must end.
constexpr int kMaxRReactAttempts = 10000;

enum FailReason : int { kOk = 0, kMaxIter = 1, kSingular = 2, kActivity = 3 };

// SomDec sandbox descriptor as kernels see it (pfchem_somdec, 0-based ids)
struct DevSomdec {
  int nrxn, co2, nh4, no3, n2o, proton, hr, nmin, nimm, nimp, ngasmin;
  int o2;  // tracked O2: aqueous id + 1, -(immobile id + 1), 0 = none
  int ext;  // the network uses the Monod oxygen response, O2 tracking or complexes
  double x0eps, n2o_frac, inhib_nh4_no3;
  const double* rate_constant;
  const double* rate_decomposition;
  const double* rate_ad_factor;
  const int* upstream_c_id;
  const int* upstream_n_id;
  const int* upstream_hr_id;
  const int* upstream_nmin_id;
  const int* upstream_nimp_id;
  const int* upstream_nimm_id;
  const double* upstream_nc;
  const int* upstream_is_varycn;
  const int* n_downstream_pools;
  const int* downstream_c_id;
  const int* downstream_n_id;
  const double* downstream_stoich;
  const double* downstream_nc;
  const int* downstream_is_varycn;
  const double* mineral_c_stoich;
  const double* mineral_n_stoich;
  const int* temperature_response_function;
  const double* Q10;
  const double* Ea;
  const int* moisture_response_function;
  const int* Ox_response_function;
  const double* Ox_half_saturation;
  const int* n_monod;
  const int* monod_specid;
  const double* monod_half_saturation;
  const double* monod_threshold;
  const int* monod_pool_normalized;
  const int* n_inhibition;
  const int* inhibition_specid;
  const int* inhibition_type;
  const double* inhibition_constant;
  const double* inhibition_constant2;
  const int* Ox_specid;  // [nrxn] oxidizer code (as o2), 0 = none
};

// Network + options as kernels see them: pointers into one device buffer,
// passed by value as a kernel argument.  Field meanings follow
// pfchem_network / pfchem_options (include/pfchem.h).
struct DevNet {
  int naq, nimm, ncomp, nc, nm, ndecay;
  int cplx_ms, mnrl_ms;  // row lengths of the compressed reaction arrays
  int h2o_aq_id;
  // option flags
  int log_form, act_freq, act_alg, act_h2o, total_as_guess, isothermal;
  int max_it, max_cuts, full_geochem;
  int max_calls;  // RReact calls per RStep before the cell fails
  double max_dlnC, tol_abs, tol_rel, tol_change;
  double A, B, Bdot;
  // optional-array presence
  int has_cplx_h2o, has_cplx_logKcoef, has_mnrl_h2o, has_mnrl_logKcoef;
  int has_temkin, has_minscale, has_affpower, has_threshold, has_limiter;
  int has_irrev, has_Ea;
  // primaries
  const double* Zp;
  const double* a0p;
  // complexes
  const int* cplx_ids;  // [nc][cplx_ms+1], [k][0] = count
  const double* cplx_st;
  const double* cplx_h2o;
  const double* cplx_logK;
  const double* cplx_logKcoef;
  const double* cplx_Z;
  const double* cplx_a0;
  // minerals
  const int* mnrl_ids;
  const double* mnrl_st;
  const double* mnrl_h2o;
  const double* mnrl_logK;
  const double* mnrl_logKcoef;
  const double* mnrl_k;
  const double* mnrl_Ea;
  const double* mnrl_Vm;
  const double* mnrl_temkin;
  const double* mnrl_minscale;
  const double* mnrl_affpower;
  const double* mnrl_threshold;
  const double* mnrl_limiter;
  const int* mnrl_irrev;
  // immobile decay
  const int* decay_id;
  const double* decay_k;
  // SomDec sandbox
  int has_somdec;
  DevSomdec sd;
  // per-cell active primaries (options.variable_ncomp): the largest species
  // id of every complex / kinetic mineral; a cell with nact active primaries
  // keeps the reactions whose largest id is below nact
  int variable;
  const int* cplx_maxid;
  const int* mnrl_maxid;
  // kinetic-mineral rate prefactors (kinmnrl_num_prefactors ...): strides
  // mp (prefactors per mineral) and msp (species per prefactor)
  int has_pref, pref_mp, pref_msp;
  const int* pref_n;         // [nm]
  const int* pref_id;        // [nm][mp][msp + 1], [..][0] = count
  const double* pref_alpha;  // [nm][mp][msp]
  const double* pref_beta;
  const double* pref_atten;
  const double* pref_rate;   // [nm][mp]
  const double* pref_Ea;     // [nm][mp] or null
};

// SoA per-cell state, device pointers (pfchem_state)
struct DevState {
  int64_t ld;
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
  const double* somdec_f_t;
  const double* somdec_f_w;
  const int* ncomp_active;
  int32_t* num_iterations;
  int32_t* num_sub_steps;
  int32_t* status;
  // size bucket of a launch over cells with per-cell active counts: the
  // launch works on cells cell_list[0 .. *list_count) (device memory filled
  // by the bucketing kernel); NULL: cells 0 .. ncell
  const int* cell_list;
  const int* list_count;
};

__device__ __forceinline__ double sat_of(const DevState& s, int64_t c) {
  return s.saturation[c];
}

// Loads and stores through the global address space.  Network and state
// pointers reach the kernels inside descriptors, so the compiler sees
// generic pointers and would emit FLAT accesses; those count against the
// same wait counter as LDS operations, so every such read inside a loop of
// LDS atomics would wait for all atomics issued before it.
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)(p);
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)(p) = v;
}

// ReactionInterpolateLogK (reaction_aux.F90:1285-1312) for one entry
__device__ __forceinline__ double interp_logk(const double* b, double temp) {
  double tk = temp + 273.15;
  return gld(b) * log(tk) + gld(b + 1) + gld(b + 2) * tk + gld(b + 3) / tk +
         gld(b + 4) / (tk * tk);
}

// Rate prefactors of mineral m (reaction_mineral.F90:820-860) at ln
// activities lnact: the sum over its prefactors p of
// rate_p Arr_p prod_s a_s^alpha / (1 + atten a_s^beta); 0 prefactors: -1.
__device__ __forceinline__ double pref_arrhenius(const DevNet& n, int m, int p, double temp) {
  const double Ea = n.pref_Ea ? gld(n.pref_Ea + m * n.pref_mp + p) : 0.0;
  return Ea > 0.0 ? exp(Ea / kIdealGas * (1.0 / (25.0 + 273.15) - 1.0 / (temp + 273.15))) : 1.0;
}
__device__ __forceinline__ double pref_ln_term(const DevNet& n, size_t k, double lna) {
  return gld(n.pref_alpha + k) * lna -
         log(1.0 + exp(log(gld(n.pref_atten + k)) + gld(n.pref_beta + k) * lna));
}
template <typename A>
__device__ double prefactor_sum(const DevNet& n, int m, const A* lnact, double temp) {
  const int np = gld(n.pref_n + m);
  double sum = 0.0;
  for (int p = 0; p < np; ++p) {
    const int* pid = n.pref_id + ((size_t)m * n.pref_mp + p) * (n.pref_msp + 1);
    double lp = 0.0;
    const int ns = gld(pid);
    for (int q = 0; q < ns; ++q)
      lp += pref_ln_term(n, ((size_t)m * n.pref_mp + p) * n.pref_msp + q, lnact[gld(pid + 1 + q)]);
    sum += exp(lp) * gld(n.pref_rate + m * n.pref_mp + p) * pref_arrhenius(n, m, p, temp);
  }
  return sum;
}
// Jacobian terms of the prefactors (reaction_mineral.F90:964-1033): for each
// prefactor species j, d(Im V)/d ln m_j = IV / sum * rate_p Arr_p prefactor_p
// * (alpha - atten beta a^beta / (1 + atten a^beta)); add(j, v) receives them
template <typename A, typename F>
__device__ void prefactor_jacobian(const DevNet& n, int m, const A* lnact, double temp,
                                   double IV_over_sum, F&& add) {
  const int np = gld(n.pref_n + m);
  for (int p = 0; p < np; ++p) {
    const size_t base = (size_t)m * n.pref_mp + p;
    const int* pid = n.pref_id + base * (n.pref_msp + 1);
    const int ns = gld(pid);
    double lp = 0.0;
    for (int q = 0; q < ns; ++q) lp += pref_ln_term(n, base * n.pref_msp + q, lnact[gld(pid + 1 + q)]);
    const double c0 = IV_over_sum * exp(lp) * gld(n.pref_rate + base) * pref_arrhenius(n, m, p, temp);
    for (int q = 0; q < ns; ++q) {
      const size_t k = base * n.pref_msp + q;
      const int j = gld(pid + 1 + q);
      const double ab = gld(n.pref_atten + k) * exp(gld(n.pref_beta + k) * lnact[j]);
      add(j, c0 * (gld(n.pref_alpha + k) - ab * gld(n.pref_beta + k) / (1.0 + ab)));
    }
  }
}

// Debye-Hueckel ln(gamma) * 1/ln10 -> gamma (reaction.F90:4586-4592)
__device__ __forceinline__ double dh_gamma(double z, double a0, double sqrtI,
                                           double I, const DevNet& n) {
  return exp((-z * z * sqrtI * n.A / (1.0 + a0 * n.B * sqrtI) + n.Bdot * I) *
             kLogToLn);
}

}  // namespace pfchem
