// pfchem_equil_team.hip -- constraint speciation (ReactionEquilibrateConstraint,
// reaction.F90:1328-2117) batched over teams: a team of L lanes (L a power
// of two >= naqcomp, one primary / constraint row per lane) speciates one
// cell; a workgroup is one wavefront of 64 / L cells.
//
// Per Newton iteration of a cell:
//   * activity coefficients (RActivityCoefficients LAG, reaction.F90:
//     4555-4611): ionic strength as a team sum of the rows' and the
//     complexes' terms; the complexes' coefficients live in the team's LDS;
//   * totals and dtotal (RTotalAqueous, :4665-4759): complexes k = t, t+L,
//     ... evaluate their molality into LDS and scatter nu nu m into the
//     team's FP64 LDS tile with LDS atomics; each row owner forms
//     total_i = (x_i + sum) rho/1000 and the row of dtotal in registers;
//   * constraint rows (:1648-1804): each lane rebuilds its own row (TOTAL
//     keeps dtotal; FREE / LOG / pH on H+ a unit row; charge balance the
//     Z-weighted column sums of dtotal; pH through a complex, mineral and
//     gas rows their mass-action derivatives).  A row that moves its own
//     species (the charge-balance reset, pH on H+) is seen by the rows
//     after it, as in the reference's sequential loop;
//   * RSolve (:5457-5516): row scaling, log-variable columns, then an LU
//     with implicit partial pivoting in registers: the pivot of column p is
//     the unused row with the largest |a_ip| / max_j |a_ij| (the choice of
//     LUDecomposition1, utility.F90:597-688), broadcast to the team by
//     shuffles; rows stay in their lanes;
//   * the log / linear update, positivity, iteration limit and the
//     two-stage convergence test (:1930-2030), all on team reductions, so
//     every lane of a team takes the same decisions.
// Teams of a wavefront run their own iteration counts; a finished team's
// lanes idle until the wavefront's last team finishes.
#include "pfchem_team_impl.hpp"

namespace pfchem {
namespace {

// LDS of one team (doubles): lnact, x before this iteration's row changes,
// x after them, the totals accumulator, then the nu nu m tile NM x (NM+1),
// then the complexes' activity coefficients and molalities
__host__ __device__ constexpr int equil_team_doubles(int NM, int nc) {
  return 4 * NM + NM * (NM + 1) + 2 * nc;
}

template <int L>
__device__ __forceinline__ int team_min_int(int v) {
  return team_reduce<L>(v, [](int a, int b) { return a < b ? a : b; });
}

// ln Q/K of constraint reaction r (mineral :1756-1771, gas :1780-1793) at
// the molalities x_seen(j) and activity coefficients gp(j)
template <typename XS, typename GP>
__device__ double eq_con_lnqk(const DevNet& n, const DevConstraint& k, int r, double temp,
                              double ln_aw, XS&& x_seen, GP&& gp) {
  double logK = gld(k.logK + r);
  if (!n.isothermal && k.has_coef) logK = interp_logk(k.coef + 5 * r, temp);
  double lnQK = -logK * kLogToLn;
  if (k.has_h2o) {
    const double h = gld(k.h2o + r);
    if (h != 0.0) lnQK += h * ln_aw;
  }
  const int* ids = k.ids + r * (k.ms + 1);
  const int ncs = gld(ids);
  for (int j = 0; j < ncs; ++j) {
    const int id = gld(ids + 1 + j);
    lnQK += gld(k.st + r * k.ms + j) * log(x_seen(id) * gp(id));
  }
  return lnQK;
}

template <int L>
__global__ void __launch_bounds__(64)
equil_team_kernel(const DevNet* np, const DevConstraint* kp, DevState s, const double* conc,
                  int64_t ncell) {
  constexpr int NM = L;
  const DevNet& n = *np;
  const DevConstraint& k = *kp;
  const int nc = n.nc;
  const int TD = equil_team_doubles(NM, nc);
  extern __shared__ __align__(16) char smem[];
  const int teams = (int)(blockDim.x / L);
  const int team = (int)(threadIdx.x / L);
  const int t = (int)(threadIdx.x % L);
  // team LDS; the row activity coefficients share the tile's last column
  double* tb = (double*)smem + (size_t)team * TD;
  double* lnact = tb;
  double* xo = tb + NM;      // x before this iteration's row changes
  double* xn = tb + 2 * NM;  // x after them
  double* tot = tb + 3 * NM;
  double* Jt = tb + 4 * NM;  // [NM][NM + 1]
  double* gk = Jt + NM * (NM + 1);
  double* mk = gk + nc;
  double* gpl = Jt;          // gp_j at Jt[j * (NM + 1) + NM]: see gp_at
  auto gp_at = [&](int j) -> double& { return gpl[j * (NM + 1) + NM]; };
  const int64_t c = (int64_t)blockIdx.x * teams + team;
  if (c >= ncell) return;  // team-uniform
  const int64_t ld = s.ld;
  const int naq = n.naq;
  const int i = t;
  const bool row = i < naq;
  const double den = gld(s.den_kg + c);
  const double temp = s.temperature ? gld(s.temperature + c) : 25.0;
  const double dkl = den * 1.0e-3;
  const int type = row ? gld(k.type + i) : PFCHEM_CONSTRAINT_FREE;
  double cv = row ? gld(conc + (i * ld + c)) : 0.0;
  // :1465-1472 (xmass = 1)
  const double to_molar = k.with_molality ? den / 1000.0 : 1.0;
  const double to_molal = k.with_molality ? 1.0 : 1000.0 / den;
  if (!n.full_geochem) {  // :1474-1478
    if (row) {
      gst(s.pri_molal + (i * ld + c), cv * to_molal);
      gst(s.total + (i * ld + c), cv * to_molar);
    }
    if (t == 0) {
      if (s.num_iterations) gst(s.num_iterations + c, 0);
      if (s.status) gst(s.status + c, (int)PFCHEM_CELL_OK);
    }
    return;
  }
  // a fresh auxvar has unit activity coefficients; with the previous
  // solution as guess the state's coefficients are kept (secondary
  // molalities too: the first activity update reads them)
  double gp = 1.0, ln_aw = 0.0;
  if (k.use_prev) {
    if (row) gp = gld(s.pri_act_coef + (i * ld + c));
    ln_aw = s.ln_act_h2o ? gld(s.ln_act_h2o + c) : 0.0;
  }
  for (int q = t; q < nc; q += L) {
    gk[q] = k.use_prev ? gld(s.sec_act_coef + (q * ld + c)) : 1.0;
    mk[q] = k.use_prev ? gld(s.sec_molal + (q * ld + c)) : 0.0;
  }
  // :1486-1599
  double free_conc = 1.0, total_conc = 0.0;
  if (row) {
    free_conc = k.use_prev ? gld(s.pri_molal + (i * ld + c)) : (k.has_guess ? gld(k.guess + i) : 1.0e-9);
    switch (type) {
      case PFCHEM_CONSTRAINT_NULL:
      case PFCHEM_CONSTRAINT_TOTAL: total_conc = cv * to_molar; break;
      case PFCHEM_CONSTRAINT_FREE: free_conc = cv * to_molal; break;
      case PFCHEM_CONSTRAINT_LOG: free_conc = pow(10.0, cv) * to_molal; break;
      case PFCHEM_CONSTRAINT_CHARGE_BAL:
      case PFCHEM_CONSTRAINT_MINERAL:
        if (!k.use_prev) free_conc = cv * to_molal;
        break;
      case PFCHEM_CONSTRAINT_PH: free_conc = pow(10.0, -cv); break;
      case PFCHEM_CONSTRAINT_GAS:
        if (cv <= 0.0) cv = pow(10.0, cv);
        break;
      default: break;
    }
  }
  double x = free_conc;
  const double zi = row ? gld(n.Zp + i) : 0.0;
  const bool has_cb = team_or<L>(type == PFCHEM_CONSTRAINT_CHARGE_BAL ? 1 : 0) != 0;
  int status = PFCHEM_CELL_OK, it = 0, it_act_on = 0;
  bool charge_warned = false, compute_act = k.use_prev != 0;
  double total = 0.0;
  for (;;) {
    team_sync();  // the previous iteration's LDS reads are done
    if (type == PFCHEM_CONSTRAINT_FREE || type == PFCHEM_CONSTRAINT_LOG) x = free_conc;  // :1625-1631
    // ---- activity coefficients, LAG (:1632-1638) ----
    if (n.act_freq != PFCHEM_ACT_COEF_FREQUENCY_OFF && compute_act) {
      double Ip = row ? x * zi * zi : 0.0;
      double sp = (row && i != n.h2o_aq_id) ? x : 0.0;
      double ss = 0.0;
      for (int q = t; q < nc; q += L) {
        const double zq = gld(n.cplx_Z + q);
        Ip += mk[q] * zq * zq;
        ss += mk[q];
      }
      const double I = 0.5 * team_sum<L>(Ip);
      const double sI = sqrt(I);
      if (row) gp = fabs(zi) > 1.0e-10 ? dh_gamma(zi, gld(n.a0p + i), sI, I, n) : 1.0;
      for (int q = t; q < nc; q += L) {
        const double zq = gld(n.cplx_Z + q);
        gk[q] = fabs(zq) > 1.0e-10 ? dh_gamma(zq, gld(n.cplx_a0 + q), sI, I, n) : 1.0;
      }
      if (n.act_h2o) {
        const double a = 1.0 - 0.017 * (team_sum<L>(sp) + team_sum<L>(ss));
        ln_aw = a > 0.0 ? log(a) : 0.0;
      }
    }
    // ---- totals and dtotal (:1639) ----
    lnact[i] = row ? log(x) + log(gp) : 0.0;
    xo[i] = x;
    tot[i] = 0.0;
    for (int j = 0; j < NM; ++j) Jt[i * (NM + 1) + j] = 0.0;
    gp_at(i) = gp;
    team_sync();
    const int cms = n.cplx_ms;
    for (int q = t; q < nc; q += L) {
      const double logK = n.isothermal || !n.has_cplx_logKcoef
                              ? gld(n.cplx_logK + q)
                              : interp_logk(n.cplx_logKcoef + 5 * q, temp);
      double lnQK = -logK * kLogToLn;
      if (n.has_cplx_h2o) lnQK += gld(n.cplx_h2o + q) * ln_aw;
      const int* ids = n.cplx_ids + q * (cms + 1);
      const double* st = n.cplx_st + q * cms;
      const int ncs = gld(ids);
      for (int a = 0; a < ncs; ++a) lnQK += gld(st + a) * lnact[gld(ids + 1 + a)];
      const double m = exp(lnQK) / gk[q];
      mk[q] = m;
      for (int a = 0; a < ncs; ++a) {
        const int ia = gld(ids + 1 + a);
        const double wa = gld(st + a) * m;
        atomicAdd(&tot[ia], wa);
        for (int b = 0; b < ncs; ++b) atomicAdd(&Jt[ia * (NM + 1) + gld(ids + 1 + b)], wa * gld(st + b));
      }
    }
    team_sync();
    double a[NM];  // this lane's row
#pragma unroll
    for (int j = 0; j < NM; ++j)
      a[j] = (row && j < naq) ? (Jt[i * (NM + 1) + j] / xo[j] + (j == i ? 1.0 : 0.0)) * dkl : 0.0;
    total = row ? (x + tot[i]) * dkl : 0.0;
    // charge balance (:1687-1714): residual and row from every dtotal row;
    // the charge-balance lane takes column j's sum in place once every lane
    // has contributed its entry j
    double cb_res = 0.0;
    const bool cb_lane = row && type == PFCHEM_CONSTRAINT_CHARGE_BAL;
    if (has_cb) {
      cb_res = team_sum<L>(zi * total);
#pragma unroll
      for (int j = 0; j < NM; ++j) {
        const double col = team_sum<L>(zi * a[j]);
        if (cb_lane) a[j] = col;
      }
    }
    // ---- constraint rows (:1648-1804) ----
    double res = 0.0;
    if (row) {
      if (type == PFCHEM_CONSTRAINT_NULL || type == PFCHEM_CONSTRAINT_TOTAL) {
        res = total - total_conc;
      } else if (type == PFCHEM_CONSTRAINT_FREE || type == PFCHEM_CONSTRAINT_LOG) {
#pragma unroll
        for (int j = 0; j < NM; ++j) a[j] = j == i ? 1.0 : 0.0;
      } else if (type == PFCHEM_CONSTRAINT_CHARGE_BAL) {
        res = cb_res;  // the row was formed above
        if (x < 1.0e-20 && !charge_warned &&
            ((cb_res > 0.0 && zi > 0.0) || (cb_res < 0.0 && zi < 0.0))) {
          charge_warned = true;
          x = 1.0e-3;
        }
      } else if (type == PFCHEM_CONSTRAINT_PH && k.h_ion_id >= 0) {
        x = pow(10.0, -cv) / gp;
#pragma unroll
        for (int j = 0; j < NM; ++j) a[j] = j == i ? 1.0 : 0.0;
      }
    }
    xn[i] = x;
    team_sync();
    // rows after a moved row see its new molality (the reference's loop
    // order): species j as row i sees it
    auto x_seen = [&](int j) { return j < i ? xn[j] : xo[j]; };
    if (row && type == PFCHEM_CONSTRAINT_PH && k.h_ion_id < 0) {  // :1716-1753
      const int q = -k.h_ion_id - 1, ms = n.cplx_ms;
      double logK = gld(n.cplx_logK + q);
      if (!n.isothermal && n.has_cplx_logKcoef) logK = interp_logk(n.cplx_logKcoef + 5 * q, temp);
      double lnQK = -logK * kLogToLn;
      if (n.has_cplx_h2o) {
        const double h = gld(n.cplx_h2o + q);
        if (h != 0.0) lnQK += h * ln_aw;
      }
      const int* ids = n.cplx_ids + q * (ms + 1);
      const int ncs = gld(ids);
      for (int j = 0; j < ncs; ++j) {
        const int id = gld(ids + 1 + j);
        lnQK += gld(n.cplx_st + q * ms + j) * log(x_seen(id) * gp_at(id));
      }
      lnQK += cv * kLogToLn;
      const double QK = exp(lnQK);
      res = 1.0 - QK;
#pragma unroll
      for (int j = 0; j < NM; ++j) a[j] = 0.0;
      for (int j = 0; j < ncs; ++j) {
        const int id = gld(ids + 1 + j);
        const double v = -QK / x_seen(id) * gld(n.cplx_st + q * ms + j);
#pragma unroll
        for (int jj = 0; jj < NM; ++jj)
          if (jj == id) a[jj] = v;
      }
    } else if (row && (type == PFCHEM_CONSTRAINT_MINERAL || type == PFCHEM_CONSTRAINT_GAS)) {
      const int r = gld(k.rxn + i);  // :1754-1804
      const double lnQK = eq_con_lnqk(n, k, r, temp, ln_aw, x_seen,
                                      [&](int j) -> double { return gp_at(j); });
      res = type == PFCHEM_CONSTRAINT_MINERAL ? lnQK : lnQK - log(cv);
#pragma unroll
      for (int j = 0; j < NM; ++j) a[j] = 0.0;
      const int* ids = k.ids + r * (k.ms + 1);
      const int ncs = gld(ids);
      for (int j = 0; j < ncs; ++j) {
        const int id = gld(ids + 1 + j);
        const double v = gld(k.st + r * k.ms + j) / x_seen(id);
#pragma unroll
        for (int jj = 0; jj < NM; ++jj)
          if (jj == id) a[jj] = v;
      }
    }
    const double max_res = team_max<L>(fabs(res));
    // :1870-1881
    const bool use_log = n.log_form ? ((it > 3 && it < 9) ? (it % 2 == 0) : true) : false;
    // ---- RSolve (:5457-5516): row scaling, log columns, pivoted LU ----
    double nrm = 0.0;
#pragma unroll
    for (int j = 0; j < NM; ++j) nrm = fmax(nrm, fabs(a[j]));
    nrm = 1.0 / fmax(1.0, nrm);
    double rhs = res * nrm;
#pragma unroll
    for (int j = 0; j < NM; ++j) a[j] *= nrm;
    if (use_log) {
#pragma unroll
      for (int j = 0; j < NM; ++j) a[j] *= xn[j < naq ? j : 0];
    }
    // implicit scaling of the pivot search; idle lanes (rows >= naq) never pivot
    double amax = 0.0;
#pragma unroll
    for (int j = 0; j < NM; ++j) amax = fmax(amax, fabs(a[j]));
    const int zero_row = team_or<L>(row && !(amax > 0.0) ? 1 : 0);
    if (zero_row) {
      status = PFCHEM_CELL_CONSTRAINT_SINGULAR;
      break;
    }
    const double vv = row ? 1.0 / amax : 0.0;
    bool used = !row;
    int piv[NM];
    const int base = (int)(threadIdx.x & ~(L - 1));
    double* prow = Jt;  // the pivot row and its right-hand side (the tile is consumed)
    static_for<0, NM>([&](auto P) {
      constexpr int p = decltype(P)::value;
      piv[p] = 0;
      if (p < naq) {
        const double score = used ? -1.0 : vv * fabs(a[p]);
        const double best = team_max<L>(score);
        // the lowest-indexed row of the best score (Crout keeps the last
        // of equal scores; ties only change rounding)
        const int pl = team_min_int<L>((!used && score == best) ? t : L);
        piv[p] = pl;
        if (t == pl) {
          static_for<p, NM>([&](auto Q) {
            constexpr int j = decltype(Q)::value;
            prow[j] = a[j];
          });
          prow[NM] = rhs;
        }
        team_sync();
        double pv = prow[p];
        if (pv == 0.0) pv = 1.0e-20;  // LUDecomposition1's zero-pivot guard
        const double pinv = 1.0 / pv;
        if (t == pl) {
          used = true;
          a[p] = pinv;  // the back substitution reads the reciprocal
        } else if (!used) {
          const double l = a[p] * pinv;
          a[p] = l;
          static_for<p + 1, NM>([&](auto Q) {
            constexpr int j = decltype(Q)::value;
            a[j] -= l * prow[j];
          });
          rhs -= l * prow[NM];
        }
        team_sync();  // prow is rewritten by the next step
      }
    });
    // back substitution: column p's value from its pivot row
    double xs[NM];
    static_for_down<0, NM>([&](auto P) {
      constexpr int p = decltype(P)::value;
      xs[p] = 0.0;
      if (p < naq) {
        double sum = rhs;
        static_for<p + 1, NM>([&](auto Q) {
          constexpr int j = decltype(Q)::value;
          sum -= a[j] * xs[j];
        });
        xs[p] = __shfl(sum * a[p], base + piv[p], 64);
      }
    });
    double upd = 0.0;
    static_for<0, NM>([&](auto P) {
      constexpr int p = decltype(P)::value;
      if (p == i) upd = xs[p];
    });
    // ---- update (:1930-1950) ----
    const double prev = x;
    if (use_log) {
      if (row) {
        const double u = copysign(1.0, upd) * fmin(fabs(upd), k.max_dlnC);
        x *= exp(-u);
      }
    } else {
      double ratio = 1.7976931348623157e308;
      if (row && prev <= upd) ratio = fabs(prev / upd);
      const double min_ratio = team_min<L>(ratio);
      if (min_ratio <= 1.0) upd *= min_ratio * 0.99;
      if (row) x = prev - upd;
    }
    ++it;
    const double min_m = team_min<L>(row ? x : 1.7976931348623157e308);
    if (min_m <= 0.0) {  // :1954-1975
      status = PFCHEM_CELL_CONSTRAINT_NONPOSITIVE;
      break;
    }
    const double max_rel = team_max<L>(row ? fabs((x - prev) / prev) : 0.0);
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
  if (row) {
    gst(s.pri_molal + (i * ld + c), x);
    gst(s.total + (i * ld + c), total);
    gst(s.pri_act_coef + (i * ld + c), gp);
  }
  for (int q = t; q < nc; q += L) {
    gst(s.sec_molal + (q * ld + c), mk[q]);
    gst(s.sec_act_coef + (q * ld + c), gk[q]);
  }
  if (t == 0) {
    if (s.ln_act_h2o) gst(s.ln_act_h2o + c, ln_aw);
    if (s.num_iterations) gst(s.num_iterations + c, it);
    if (s.status) gst(s.status + c, status);
    if (status != PFCHEM_CELL_OK) atomicAdd(k.nfail, 1);
  }
}

template <int L>
hipError_t launch_equil_L(const DevNet& n, const DevNet* dn, const DevConstraint* dk,
                          const DevState& s, const double* conc, int64_t ncell,
                          hipStream_t stream) {
  const int teams = 64 / L;
  const size_t lds = (size_t)teams * equil_team_doubles(L, n.nc) * sizeof(double);
  const dim3 grid((unsigned)((ncell + teams - 1) / teams));
  equil_team_kernel<L><<<grid, 64, lds, stream>>>(dn, dk, s, conc, ncell);
  return hipGetLastError();
}

}  // namespace

int equil_team_lanes(const DevNet& n) {
  if (n.act_freq != PFCHEM_ACT_COEF_FREQUENCY_OFF && n.act_alg == PFCHEM_ACT_COEF_ALGORITHM_NEWTON)
    return 0;  // NEWTON activity: the one-cell-per-thread kernel
  int L = 4;
  while (L < n.naq) L *= 2;
  if (L > 64) return 0;
  const size_t lds = (size_t)(64 / L) * equil_team_doubles(L, n.nc) * sizeof(double);
  return lds <= 64 * 1024 ? L : 0;
}

hipError_t launch_equil_team(const DevNet& n, const DevNet* dn, const DevConstraint* dk,
                             const DevState& s, const double* conc, int64_t ncell,
                             hipStream_t stream) {
  switch (equil_team_lanes(n)) {
    case 4: return launch_equil_L<4>(n, dn, dk, s, conc, ncell, stream);
    case 8: return launch_equil_L<8>(n, dn, dk, s, conc, ncell, stream);
    case 16: return launch_equil_L<16>(n, dn, dk, s, conc, ncell, stream);
    case 32: return launch_equil_L<32>(n, dn, dk, s, conc, ncell, stream);
    case 64: return launch_equil_L<64>(n, dn, dk, s, conc, ncell, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pfchem
