// pfchem_girt.hip -- the per-cell reaction block of the global implicit
// (GIRT) residual and Jacobian, batched: a team of L lanes (L a power of two
// >= ncomp, one network row per lane) evaluates one cell, a workgroup holds
// 256 / L cells.
//
// Reference: RTResidualNonFlux / RTJacobianNonFlux per cell
// (reactive_transport.F90:2393-2430, 2599-2642, 3172-3183, 3270-3300) in the
// RReact form of one Newton iterate (reaction.F90:3896-3916): RTotalAqueous
// (:4665-4759), RTAccumulation(Derivative) (:5710-5848), RReaction (:4059)
// -> RKineticMineral (reaction_mineral.F90:647-1078) and RImmobileDecay
// (reaction_immobile.F90:244-294), and the SomDec sandbox (SomDecReact,
// reaction_sandbox_somdec.F90:1504-1910, through pfchem_somdec.hpp).
//
// Per cell:  res_i  = (Acc_i(x) - Acc_i(total^k, imm^k)) / dt + R_i(x)
//            jac_ij = d res_i / d x_j   (linear unknowns: free molality for
//            aqueous columns, mol/m^3 bulk for immobile ones)
// Complex k contributes nu_ki m_k to total_i and nu_ki nu_kj m_k / x_j to
// dtotal_ij; the team accumulates the symmetric part nu nu m in an FP64 LDS
// tile with LDS atomics and each row owner applies the column scale 1/x_j.
// Totals, secondary molalities and mineral rates are written back (as
// RTAuxVarCompute updates the auxvar).  The residual and Jacobian leave
// through the workgroup's LDS so that consecutive lanes store consecutive
// cells of one SoA row.
#include "pfchem_team_impl.hpp"

namespace pfchem {
namespace {

// LDS of one team (doubles): lnact, 1/x, x, totals accumulator, reaction
// residual accumulator, final residual, Jacobian tile NM x (NM + 1), then
// nsd doubles of per-cell SomDec stoichiometry
__host__ __device__ constexpr int girt_team_doubles(int NM, int nsd) {
  return 6 * NM + NM * (NM + 1) + nsd;
}
constexpr int kGirtJ = 6;  // offset of the Jacobian tile, in units of NM

// SomDec accumulator of a GIRT team: species values from the team's LDS
// (totals with the complexes' share, the dtotal entries the sandbox reads
// from the team-summed complex pairs S), residual terms into the reaction
// residual, Jacobian terms d res_row / d x_col added as they come (the
// tile is linear here)
template <bool X>
struct GirtSdEnvT {
  static constexpr bool kExt = X;
  typedef __attribute__((address_space(3))) double lds_t;
  const lds_t* xv;
  const lds_t* lna;
  lds_t* racc;
  lds_t* Jt;
  lds_t* st;
  const lds_t* ctot;  // the complexes' share of the totals (molal)
  int naq, nrxn, ld;
  double dkl, theta, vol, sat, temp, f_t, f_w;
  bool has_f_t, has_f_w;
  const DevSomdec* sdp;
  const lds_t* S;  // dtotal pair sums of the cell's complexes
  __device__ double total(int i) const {
    if constexpr (X) return (xv[i] + ctot[i]) * dkl;
    return xv[i] * dkl;
  }
  __device__ double imm(int i) const { return xv[naq + i]; }
  __device__ double lnact(int i) const { return lna[i]; }
  __device__ double dtot(int i, int j) const {
    if constexpr (X) return sd_dtot_value(*sdp, S, dkl, xv[j], i, j);
    return i == j ? dkl : 0.0;
  }
  __device__ void R(int row, double v) {
    __hip_atomic_fetch_add(&racc[row], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ void J(int row, int col, double v) {
    __hip_atomic_fetch_add(&Jt[row * ld + col], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ lds_t& unc(int r) { return st[r]; }
  __device__ lds_t& mcs(int r) { return st[nrxn + r]; }
  __device__ lds_t& mns(int r) { return st[2 * nrxn + r]; }
  __device__ lds_t& dnc(int r, int j) { return st[3 * nrxn + r * kSdDown + j]; }
  static __device__ lds_t* lds(double* p) { return (lds_t*)p; }
};
using GirtSdEnv = GirtSdEnvT<false>;

template <int L, bool SD>
__global__ void __launch_bounds__(256)
girt_team_kernel(const DevNet* np, const DevState* sp, int64_t ncell, double dt,
                 const double* __restrict__ fixed_total, const double* __restrict__ fixed_imm,
                 double* __restrict__ res, double* __restrict__ jac) {
  constexpr int NM = L;
  const DevNet& n = *np;
  const DevState& s = *sp;
  const int nsd = SD ? sd_lds_doubles(n.sd.nrxn) : 0;
  const int TD = girt_team_doubles(NM, nsd);
  extern __shared__ __align__(16) char smem[];
  const int teams = (int)(blockDim.x / L);
  const int team = (int)(threadIdx.x / L);
  const int t = (int)(threadIdx.x % L);
  double* tb = (double*)smem + (size_t)team * TD;
  double* lnact = tb;
  double* xinv = tb + NM;
  double* xv = tb + 2 * NM;
  double* tot = tb + 3 * NM;
  double* racc = tb + 4 * NM;
  double* rfin = tb + 5 * NM;
  double* J = tb + kGirtJ * NM;
  double* sdst = J + NM * (NM + 1);
  double* sdS = SD ? sdst + n.sd.nrxn * (3 + kSdDown) : nullptr;  // SomDec dtotal pair sums
  int* nrow = (int*)((double*)smem + (size_t)teams * TD);  // unknowns per team's cell
  const int64_t c0 = (int64_t)blockIdx.x * teams;
  const int64_t c = c0 + team;
  const int64_t ld = s.ld;
  const int NF = n.ncomp;
  const bool live = c < ncell;  // team-uniform

  if (live) {
    const int naq = n.variable ? min(max(gld(s.ncomp_active + c), 1), n.naq) : n.naq;
    const int N = naq + n.nimm;
    auto cplx_on = [&](int k) { return !n.variable || gld(n.cplx_maxid + k) < naq; };
    auto mnrl_on = [&](int m) { return !n.variable || gld(n.mnrl_maxid + m) < naq; };
    const double por = gld(s.porosity + c), sat = gld(s.saturation + c);
    const double den = gld(s.den_kg + c), vol = gld(s.volume + c);
    const double temp = s.temperature ? gld(s.temperature + c) : 25.0;
    const double ln_aw = s.ln_act_h2o ? gld(s.ln_act_h2o + c) : 0.0;
    const bool wet = !(sat < kRtMinSaturation);
    const double dkl = den * 1.0e-3;
    const double psv = por * sat * 1000.0 * vol;
    const double psvd = psv / dt;
    if (t == 0) nrow[team] = N;

    // (A) row t: unknown, ln activity, column scale; clear accumulators
    const int i = t;
    double x = 1.0;
    if (i < naq) {
      x = gld(s.pri_molal + (i * ld + c));
      lnact[i] = log(x) + log(gld(s.pri_act_coef + (i * ld + c)));
    } else if (i < N) {
      x = gld(s.immobile + ((i - naq) * ld + c));
      lnact[i] = log(x);
    }
    xinv[i] = 1.0 / x;
    xv[i] = x;
    tot[i] = 0.0;
    racc[i] = 0.0;
    for (int j = 0; j <= NM; ++j) J[i * (NM + 1) + j] = 0.0;
    if constexpr (SD) {
      // per-cell SomDec stoichiometry starts from the Setup values
      const int nr = n.sd.nrxn;
      for (int r = t; r < nr; r += L) {
        sdst[r] = gld(n.sd.upstream_nc + r);
        sdst[nr + r] = gld(n.sd.mineral_c_stoich + r);
        sdst[2 * nr + r] = gld(n.sd.mineral_n_stoich + r);
        for (int j = 0; j < kSdDown; ++j)
          sdst[3 * nr + r * kSdDown + j] = gld(n.sd.downstream_nc + r * kSdDown + j);
      }
      for (int q = t; q < kSdPairs; q += L) sdS[q] = 0.0;
    }
    team_sync();

    // (B) complexes k = t, t+L, ...: secondary molality, totals, nu nu m
    const int cms = n.cplx_ms;
    for (int k = t; k < n.nc; k += L) {
      if (!cplx_on(k)) continue;
      const double logK = n.isothermal || !n.has_cplx_logKcoef
                              ? gld(n.cplx_logK + k)
                              : interp_logk(n.cplx_logKcoef + 5 * k, temp);
      double lnQK = -logK * kLogToLn;
      if (n.has_cplx_h2o) lnQK += gld(n.cplx_h2o + k) * ln_aw;
      const int* ids = n.cplx_ids + k * (cms + 1);
      const double* st = n.cplx_st + k * cms;
      const int ncs = gld(ids);
      for (int a = 0; a < ncs; ++a) lnQK += gld(st + a) * lnact[gld(ids + 1 + a)];
      const double mk = exp(lnQK) / gld(s.sec_act_coef + (k * ld + c));
      gst(s.sec_molal + (k * ld + c), mk);
      if constexpr (SD) sd_pair_sums(n.sd, ids, st, mk, GirtSdEnv::lds(sdS));
      for (int a = 0; a < ncs; ++a) {
        const int ia = gld(ids + 1 + a);
        const double wa = gld(st + a) * mk;
        atomicAdd(&tot[ia], wa);
        for (int b = 0; b < ncs; ++b)
          atomicAdd(&J[ia * (NM + 1) + gld(ids + 1 + b)], wa * gld(st + b));
      }
    }
    team_sync();

    // (C) row owner: totals, accumulation, RTAccumulationDerivative, decay
    double r = 0.0;
    if (i < naq) {
      const double total = (x + tot[i]) * dkl;
      gst(s.total + (i * ld + c), total);
      if (wet) {
        r = (psv * total - psv * gld(fixed_total + (i * ld + c))) / dt;
        const double f = dkl * psvd;
        for (int j = 0; j < naq; ++j) {
          const double dt_ij = J[i * (NM + 1) + j] * xinv[j] + (j == i ? 1.0 : 0.0);
          J[i * (NM + 1) + j] = dt_ij * f;
        }
      }
    } else if (i < N && wet) {
      const double imm_k = gld(fixed_imm + ((i - naq) * ld + c));
      r = (x * vol - imm_k * vol) / dt;
      double d = vol / dt;
      for (int q = 0; q < n.ndecay; ++q)
        if (gld(n.decay_id + q) == i - naq) {
          const double kc = gld(n.decay_k + q) * vol;
          r += kc * x;
          d += kc;
        }
      J[i * (NM + 1) + i] = d;
    }
    if (i < N && !wet) {
      for (int j = 0; j < N; ++j) J[i * (NM + 1) + j] = (j == i) ? 1.0 : 0.0;
    }
    team_sync();

    // (D) kinetic minerals m = t, t+L, ... (wet cells only, as RReaction)
    const int mms = n.mnrl_ms;
    for (int m = t; m < n.nm; m += L) {
      if (!mnrl_on(m)) continue;
      double rt = 0.0;
      const double area = gld(s.mnrl_area + (m * ld + c));
      if (wet && area != 0.0) {
        const double logK = n.isothermal || !n.has_mnrl_logKcoef
                                ? gld(n.mnrl_logK + m)
                                : interp_logk(n.mnrl_logKcoef + 5 * m, temp);
        double lnQK = -logK * kLogToLn;
        if (n.has_mnrl_h2o) lnQK += gld(n.mnrl_h2o + m) * ln_aw;
        const int* ids = n.mnrl_ids + m * (mms + 1);
        const double* st = n.mnrl_st + m * mms;
        const int ncs = gld(ids);
        for (int a = 0; a < ncs; ++a) lnQK += gld(st + a) * lnact[gld(ids + 1 + a)];
        const MineralRate mr =
            mineral_rate(n, m, lnQK, gld(s.mnrl_volfrac + (m * ld + c)), area, temp, lnact);
        if (mr.active) {
          rt = mr.Im;
          const double Iv = mr.Im * vol;
          const double g = mr.dIm_dQK * vol * mr.QK * dkl;
          for (int a = 0; a < ncs; ++a) {
            const int ia = gld(ids + 1 + a);
            const double nua = gld(st + a);
            atomicAdd(&racc[ia], nua * Iv);
            for (int b = 0; b < ncs; ++b) {
              const int ib = gld(ids + 1 + b);
              atomicAdd(&J[ia * (NM + 1) + ib], nua * g * gld(st + b) * xinv[ib]);
            }
          }
          // rate prefactors (reaction_mineral.F90:964-1033)
          if (n.has_pref && gld(n.pref_n + m) > 0)
            prefactor_jacobian(n, m, lnact, temp, Iv / mr.kr, [&](int j, double v) {
              for (int a = 0; a < ncs; ++a)
                atomicAdd(&J[gld(ids + 1 + a) * (NM + 1) + j], gld(st + a) * v * xinv[j]);
            });
        }
      }
      if (s.mnrl_rate) gst(s.mnrl_rate + (m * ld + c), rt);
    }
    if (SD && wet) {
      // SomDec sandbox (SomDecReact, reaction_sandbox_somdec.F90:1504-1910):
      // reactions r = t, t+L, ...; the N2O emission (lane 0) needs the net N
      // mineralization of all of them
      const double f_t = s.somdec_f_t ? gld(s.somdec_f_t + c) : 1.0;
      const double f_w = s.somdec_f_w ? gld(s.somdec_f_w + c) : 1.0;
      auto run = [&](auto& e) {
        double nmin = 0.0;
        for (int r = t; r < n.sd.nrxn; r += L) nmin += sd_reaction(n.sd, e, r, dt, true);
        nmin = team_sum<L>(nmin);
        if (t == 0) sd_nemission(n.sd, e, nmin, dt, true);
      };
      if (n.sd.ext) {
        GirtSdEnvT<true> e{GirtSdEnv::lds(xv), GirtSdEnv::lds(lnact), GirtSdEnv::lds(racc),
                           GirtSdEnv::lds(J), GirtSdEnv::lds(sdst), GirtSdEnv::lds(tot), naq,
                           n.sd.nrxn, NM + 1, dkl, por * sat, vol, sat, temp, f_t, f_w,
                           s.somdec_f_t != nullptr, s.somdec_f_w != nullptr, &n.sd,
                           GirtSdEnv::lds(sdS)};
        run(e);
      } else {
        GirtSdEnvT<false> e{GirtSdEnv::lds(xv), GirtSdEnv::lds(lnact), GirtSdEnv::lds(racc),
                            GirtSdEnv::lds(J), GirtSdEnv::lds(sdst), GirtSdEnv::lds(tot), naq,
                            n.sd.nrxn, NM + 1, dkl, por * sat, vol, sat, temp, f_t, f_w,
                            s.somdec_f_t != nullptr, s.somdec_f_w != nullptr, &n.sd,
                            GirtSdEnv::lds(sdS)};
        run(e);
      }
    }
    team_sync();
    if (i < N) rfin[i] = wet ? r + racc[i] : 0.0;
  }

  // (E) residual and Jacobian out through LDS: consecutive lanes store
  // consecutive cells of one SoA row
  __syncthreads();
  const int nent = NF * NF;
  for (int w = threadIdx.x; w < NF * teams; w += blockDim.x) {
    const int b = w % teams, q = w / teams;
    if (c0 + b >= ncell) continue;
    const double* tbb = (const double*)smem + (size_t)b * TD;
    gst(res + ((int64_t)q * ld + c0 + b), q < nrow[b] ? tbb[5 * NM + q] : 0.0);
  }
  if (jac) {
    for (int w = threadIdx.x; w < nent * teams; w += blockDim.x) {
      const int b = w % teams, q = w / teams;
      if (c0 + b >= ncell) continue;
      const int ii = q / NF, jj = q % NF;
      const double* Jb = (const double*)smem + (size_t)b * TD + kGirtJ * NM;
      const int nb = nrow[b];
      gst(jac + ((int64_t)q * ld + c0 + b), (ii < nb && jj < nb) ? Jb[ii * (NM + 1) + jj] : 0.0);
    }
  }
}

template <int L>
hipError_t launch_girt_L(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                         double dt, const double* fixed_total, const double* fixed_imm,
                         double* res, double* jac, hipStream_t stream) {
  const int nsd = n.has_somdec ? sd_lds_doubles(n.sd.nrxn) : 0;
  const int team_bytes = 8 * girt_team_doubles(L, nsd);
  int teams = 256 / L;
  while (teams > 1 && (size_t)teams * (team_bytes + 4) > 65536) teams /= 2;
  const int block = teams * L;
  const size_t lds = (size_t)teams * team_bytes + 4 * (size_t)teams;
  if (lds > 65536) return hipErrorInvalidValue;
  const int64_t grid = (ncell + teams - 1) / teams;
  if (n.has_somdec)
    girt_team_kernel<L, true><<<dim3((unsigned)grid), dim3(block), lds, stream>>>(
        dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac);
  else
    girt_team_kernel<L, false><<<dim3((unsigned)grid), dim3(block), lds, stream>>>(
        dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac);
  return hipGetLastError();
}

}  // namespace

int girt_team_lanes(const DevNet& n) {
  int L = 4;
  while (L < n.ncomp) L *= 2;
  return L <= 64 ? L : 0;
}

hipError_t launch_girt_team(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                            double dt, const double* fixed_total, const double* fixed_imm,
                            double* res, double* jac, hipStream_t stream) {
  switch (girt_team_lanes(n)) {
    case 4: return launch_girt_L<4>(n, dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac, stream);
    case 8: return launch_girt_L<8>(n, dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac, stream);
    case 16: return launch_girt_L<16>(n, dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac, stream);
    case 32: return launch_girt_L<32>(n, dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac, stream);
    case 64: return launch_girt_L<64>(n, dn, ds, ncell, dt, fixed_total, fixed_imm, res, jac, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pfchem
