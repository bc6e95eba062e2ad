// pfchem_team.hip -- batched reaction step: persistent, work-stealing
// kernel in which a team of L lanes (L = 4..64, inside one wavefront) owns
// one cell at a time.
//
// Mapping (gfx950, wave64):
//   * lane i of a team owns species / Jacobian row i (i < ncomp <= L) and
//     keeps that row in registers for the LU;
//   * complexes k = t, t+L, ... and minerals m = t, t+L, ... are spread
//     over the team's lanes; their contributions are scattered into the
//     team's LDS tile with LDS atomics (FP64 residual/totals, JT Jacobian);
//   * every wave is an independent persistent worker: when a team's cell
//     finishes (converged, or cuts exhausted) the team takes the next cell
//     from a global cursor, so uneven Newton counts never idle a wave while
//     work remains.  All team communication stays inside the wave (LDS +
//     cross-lane ops), so there is no block barrier in the loop.
// Algorithm: RStep (reaction.F90:3564-3738) / RReact (:3742-4055) with
// RTotalAqueous (:4665-4759), RTAccumulation(Derivative) (:5710-5848),
// RKineticMineral (reaction_mineral.F90:647-1078), RImmobileDecay
// (reaction_immobile.F90:244-294) and RSolve (:5457-5516, back-substituting).
// Deliberate numerical reformulations (same mathematics, different
// rounding):
//   * the Jacobian is assembled directly in the log formulation
//     (d res_i / d ln x_j = J_ij x_j, as RSolve forms it) and each row is
//     divided by its accumulation diagonal; the Newton update is invariant
//     to row scaling, and implicit (scaled) partial pivoting makes the pivot
//     choice scale-invariant too;
//   * the LU is right-looking Gaussian elimination with scaled partial
//     pivoting, which selects the same pivots as the reference Crout form;
//   * JT = float: Jacobian entries, LU and triangular solves in FP32, while
//     the residual, mass-action log K sums, exponentials, the log-space
//     update and every convergence test stay FP64 (the Newton iteration acts
//     as iterative refinement).
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"

#include <type_traits>

namespace pfchem {
namespace {

constexpr int kMaxMnrl = 16;

__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------
// Cross-lane primitives inside a team of L consecutive lanes, built on DPP
// (quad_perm, row_half_mirror, row_mirror, row_newbcast), ds_swizzle and
// readlane, so that reductions and broadcasts stay in the VALU / LDS-swizzle
// path instead of ds_bpermute.  All lanes of a team must be active.
// ---------------------------------------------------------------------
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}
template <int B, int E, typename F>
__device__ __forceinline__ void static_for_down(F&& f) {  // E-1 down to B
  if constexpr (B < E) {
    f(std::integral_constant<int, E - 1>{});
    static_for_down<B, E - 1>(f);
  }
}

// broadcast lane P of every team to the whole team (32-bit)
template <int L, int P>
__device__ __forceinline__ int bcast32(int v) {
  if constexpr (L == 4) {
    return __builtin_amdgcn_update_dpp(0, v, P | (P << 2) | (P << 4) | (P << 6), 0xF, 0xF, false);
  } else if constexpr (L == 8) {
    return __builtin_amdgcn_ds_swizzle(v, 0x18 | (P << 5));  // BROADCAST in groups of 8
  } else if constexpr (L == 16) {
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + P, 0xF, 0xF, false);  // row_newbcast:P
  } else if constexpr (L == 32) {
    return __builtin_amdgcn_ds_swizzle(v, (P << 5));  // BROADCAST in groups of 32
  } else {
    return __builtin_amdgcn_readlane(v, P);
  }
}
template <int L, int P>
__device__ __forceinline__ int bcast(int v) {
  return bcast32<L, P>(v);
}
template <int L, int P>
__device__ __forceinline__ float bcast(float v) {
  return __int_as_float(bcast32<L, P>(__float_as_int(v)));
}
template <int L, int P>
__device__ __forceinline__ double bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = bcast32<L, P>((int)(b & 0xffffffffll));
  const int hi = bcast32<L, P>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// value of the partner lane at butterfly level `lev` (0: xor 1, 1: xor 2,
// 2: mirror within 8, 3: mirror within 16, 4: swap halves of 32)
template <int LEV>
__device__ __forceinline__ int partner32(int v) {
  if constexpr (LEV == 0) return __builtin_amdgcn_update_dpp(0, v, 1 | (0 << 2) | (3 << 4) | (2 << 6), 0xF, 0xF, false);
  else if constexpr (LEV == 1) return __builtin_amdgcn_update_dpp(0, v, 2 | (3 << 2) | (0 << 4) | (1 << 6), 0xF, 0xF, false);
  else if constexpr (LEV == 2) return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
  else if constexpr (LEV == 3) return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
  else return __builtin_amdgcn_ds_swizzle(v, 0x1F | (0x10 << 10));  // SWAP 16
}
template <int LEV>
__device__ __forceinline__ double partner(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = partner32<LEV>((int)(b & 0xffffffffll));
  const int hi = partner32<LEV>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int LEV>
__device__ __forceinline__ float partner(float v) {
  return __int_as_float(partner32<LEV>(__float_as_int(v)));
}
template <int LEV>
__device__ __forceinline__ int partner(int v) {
  return partner32<LEV>(v);
}

// all-reduce over the team; every lane gets the bit-identical result
// (each level combines a lane with a distinct partner, and the combining
// operations used here are commutative)
template <int L, typename T, typename Op>
__device__ __forceinline__ T team_reduce(T v, Op op) {
  static_assert(L == 4 || L == 8 || L == 16 || L == 32 || L == 64, "team width");
  v = op(v, partner<0>(v));
  v = op(v, partner<1>(v));
  if constexpr (L >= 8) v = op(v, partner<2>(v));
  if constexpr (L >= 16) v = op(v, partner<3>(v));
  if constexpr (L >= 32) v = op(v, partner<4>(v));
  if constexpr (L == 64) {
    const T a = bcast<64, 0>(v), b = bcast<64, 32>(v);
    v = op(a, b);
  }
  return v;
}
template <int L>
__device__ __forceinline__ double team_sum(double v) {
  return team_reduce<L>(v, [](double a, double b) { return a + b; });
}
template <int L>
__device__ __forceinline__ double team_max(double v) {
  return team_reduce<L>(v, [](double a, double b) { return fmax(a, b); });
}
template <int L>
__device__ __forceinline__ double team_min(double v) {
  return team_reduce<L>(v, [](double a, double b) { return fmin(a, b); });
}
template <int L>
__device__ __forceinline__ int team_or(int v) {
  return team_reduce<L>(v, [](int a, int b) { return a | b; });
}
// LDS tile of one team (byte offsets); same formula on host and device
struct TeamLds {
  int lnact, rs, tot, racc, xv, mvf, marea, mrate, mk, lng, J, prow, bytes;
};
__host__ __device__ inline TeamLds team_lds_layout(int L, int nc, int jt_size) {
  TeamLds t;
  int o = 0;
  t.lnact = o; o += 8 * L;
  t.rs = o;    o += 8 * L;
  t.tot = o;   o += 8 * L;
  t.racc = o;  o += 8 * L;
  t.xv = o;    o += 8 * L;
  t.mvf = o;   o += 8 * kMaxMnrl;
  t.marea = o; o += 8 * kMaxMnrl;
  t.mrate = o; o += 8 * kMaxMnrl;
  t.mk = o;    o += 8 * (nc > 0 ? nc : 1);
  t.lng = o;   o += 8 * (nc > 0 ? nc : 1);
  t.J = o;     o += jt_size * L * (L + 1);
  o = (o + 7) & ~7;
  t.prow = o;  o += jt_size * (L + 2);
  o = (o + 15) & ~15;
  t.bytes = o;
  return t;
}

// TST rate of one kinetic mineral (RKineticMineral, reaction_mineral.F90:
// 729-908) given ln QK: rate Im (mol/m^3 bulk/s, before the volume factor)
// and dIm/dQK per unit volume, with the affinity options.
struct MineralRate {
  double Im, dIm_dQK, QK;
  bool active;
};

// exponent 1/sigma of QK in the affinity 1 - QK^(1/sigma)
__device__ __forceinline__ double affinity_sigma(const DevNet& n, int m) {
  if (n.has_temkin)
    return n.has_minscale ? n.mnrl_minscale[m] * n.mnrl_temkin[m] : n.mnrl_temkin[m];
  return n.mnrl_minscale[m];
}

__device__ __noinline__ double pow_noinline(double a, double b) { return pow(a, b); }

// Im and dIm/dQK with the optional affinity power, Temkin / scale factor
// and rate limiter (reaction_mineral.F90:839-908)
__device__ __forceinline__ void rate_options(const DevNet& n, int m, double QK, double aff,
                                             double sgn, double kr, double Im_const,
                                             MineralRate* r) {
  const double p = n.has_affpower ? n.mnrl_affpower[m] : 1.0;
  r->Im = n.has_affpower ? Im_const * sgn * pow_noinline(fabs(aff), p) * kr
                         : Im_const * sgn * fabs(aff) * kr;
  double d = n.has_affpower ? -r->Im * p / fabs(aff) : -Im_const * kr;
  if (n.has_temkin || n.has_minscale) d = d * (1.0 / affinity_sigma(n, m)) / QK * (1.0 - aff);
  const double lim = n.has_limiter ? n.mnrl_limiter[m] : 0.0;
  if (lim > 0.0) {
    const double d2 = 1.0 + (1.0 - aff) / lim;
    d *= (1.0 + QK / lim / d2) / d2;
  }
  r->dIm_dQK = d;
}

__device__ __forceinline__ MineralRate mineral_rate(const DevNet& n, int m, double lnQK,
                                                    double volfrac, double area,
                                                    double temp) {
  MineralRate r;
  r.active = false;
  r.Im = r.dIm_dQK = 0.0;
  const double QK = exp(lnQK);
  r.QK = QK;
  const bool plain_aff = !(n.has_temkin || n.has_minscale);
  double aff = plain_aff ? 1.0 - QK : 1.0 - pow_noinline(QK, 1.0 / affinity_sigma(n, m));
  const double sgn = copysign(1.0, aff);
  if (!(volfrac > 0.0 || sgn < 0.0)) return r;
  if (n.has_irrev && n.mnrl_irrev[m] == 1 && sgn < 0.0) return r;
  if (n.has_threshold && n.mnrl_threshold[m] > 0.0 && sgn < 0.0 && QK < n.mnrl_threshold[m])
    return r;
  const double lim = n.has_limiter ? n.mnrl_limiter[m] : 0.0;
  if (lim > 0.0) aff = aff / (1.0 + (1.0 - aff) / lim);
  double arr = 1.0;
  if (n.has_Ea && n.mnrl_Ea[m] > 0.0)
    arr = exp(n.mnrl_Ea[m] / kIdealGas * (1.0 / (25.0 + 273.15) - 1.0 / (temp + 273.15)));
  const double kr = n.mnrl_k[m] * arr;
  double Im_const = -area;
  if (n.has_minscale) Im_const = Im_const / n.mnrl_minscale[m];
  r.active = true;
  if (plain_aff && !n.has_affpower && !(lim > 0.0)) {
    r.Im = Im_const * sgn * fabs(aff) * kr;
    r.dIm_dQK = -Im_const * kr;
  } else {
    rate_options(n, m, QK, aff, sgn, kr, Im_const, &r);
  }
  return r;
}

// relative FP64 residual a mixed-precision iterate must reach before a
// small Newton step may end the iteration
constexpr double kMixedStepExitResidual = 1.0e-11;
// transient per-cell status: the FP32 pass hands the cell to the FP64 pass
constexpr int kEscalate = 16;

enum Outcome : int { kContinue = 0, kConverged = 1, kFailMaxIt = 2, kFailSingular = 3 };

template <int L, typename JT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
team_step_kernel(const DevNet* np, const DevState* sp, int64_t ncell, double dt,
                 unsigned long long* __restrict__ counter, int rescue_only) {
  const DevNet& n = *np;
  const DevState& s = *sp;
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & 63;
  const int t = lane & (L - 1);          // lane within team = species row
  const int team_in_wave = lane / L;
  const int wave_in_block = threadIdx.x >> 6;
  const TeamLds lay = team_lds_layout(L, n.nc, (int)sizeof(JT));
  char* tb = smem + (size_t)(wave_in_block * (64 / L) + team_in_wave) * lay.bytes;
  double* lnact = (double*)(tb + lay.lnact);
  double* rsv = (double*)(tb + lay.rs);
  double* tot = (double*)(tb + lay.tot);
  double* racc = (double*)(tb + lay.racc);
  double* xv = (double*)(tb + lay.xv);
  double* vf = (double*)(tb + lay.mvf);
  double* area = (double*)(tb + lay.marea);
  double* rate = (double*)(tb + lay.mrate);
  double* mk = (double*)(tb + lay.mk);
  double* lng = (double*)(tb + lay.lng);
  JT* Jt = (JT*)(tb + lay.J);
  JT* prow = (JT*)(tb + lay.prow);

  const int naq = n.naq, nimm = n.nimm, N = n.ncomp, nc = n.nc, nm = n.nm;
  const bool mine = t < N;              // this lane owns a species row
  const bool escalate = s.status != nullptr;  // FP32 failures go to the FP64 pass
  const bool is_aq = t < naq;
  const int64_t ld = s.ld;

  // team-uniform cell state
  int64_t cell = -1;
  bool exhausted = false;
  double target = 0.0, cum = 0.0, tdt = 0.0, r0 = 0.0;
  int nconst = 0, ncuts = 0, nsub = 0, nit_total = 0, nit = 0, status = 0;
  int attempts = 0, last_why = 0;
  // per-lane (species t) state
  double x = 1.0, guess = 1.0, fixed = 0.0, init_tot = 0.0, total = 0.0;
  double lngam = 0.0, small_val = 0.0, res = 0.0, delta = 0.0;
  bool small = false;
  double ln_aw = 0.0;
  // cell scalars
  double por = 0.0, sat = 0.0, den = 0.0, vol = 0.0, temp = 25.0;

  // wave-local queue of claimed cells [qnext, qend) (wave-uniform)
  long long qnext = 0, qend = 0;
  const int batch = 2 * (64 / L);

  for (;;) {
    // ---------------- work stealing: hand out cells ----------------
    // One atomic per refill per wave: teams that need a cell take the next
    // entries of the wave's queue; the queue is refilled from the global
    // cursor in batches.
    const int needy = (t == 0 && cell < 0 && !exhausted) ? 1 : 0;
    const unsigned long long needmask = __ballot(needy);
    long long got = -1;
    if (needmask) {
      const int cnt = __popcll(needmask);
      const long long avail = qend - qnext;
      long long base = 0, claim = 0;
      if (cnt > avail) {
        claim = (long long)max((long long)batch, (long long)cnt - avail);
        if (lane == 0) base = (long long)atomicAdd(counter, (unsigned long long)claim);
        base = __shfl(base, 0, 64);
      }
      if (needy) {
        const int r = __popcll(needmask & ((1ull << lane) - 1ull));
        got = (r < avail) ? qnext + r : base + (r - avail);
      }
      if (cnt > avail) {
        qnext = base + (cnt - avail);
        qend = base + claim;
      } else {
        qnext += cnt;
      }
    }
    got = __shfl(got, team_in_wave * L, 64);
    if (cell < 0 && !exhausted && got >= 0) {
      if (got >= ncell) {
        exhausted = true;
      } else if (rescue_only && s.status[got] != kEscalate) {
        // FP64 rescue pass: only cells the FP32 pass handed over
      } else {
        cell = got;
        const int64_t c = cell;
        // load the cell (SoA, coalesced along cells across teams)
        por = s.porosity[c];
        sat = s.saturation[c];
        den = s.den_kg[c];
        vol = s.volume[c];
        temp = s.temperature ? s.temperature[c] : 25.0;
        ln_aw = s.ln_act_h2o ? s.ln_act_h2o[c] : 0.0;
        if (is_aq) {
          x = s.pri_molal[t * ld + c];
          total = s.total[t * ld + c];
          lngam = log(s.pri_act_coef[t * ld + c]);
        } else if (mine) {
          x = s.immobile[(t - naq) * ld + c];
          total = x;
        }
        for (int k = t; k < nc; k += L) {
          lng[k] = log(s.sec_act_coef[k * ld + c]);
          mk[k] = s.sec_molal[k * ld + c];
        }
        for (int m = t; m < nm; m += L) {
          vf[m] = s.mnrl_volfrac[m * ld + c];
          area[m] = s.mnrl_area[m * ld + c];
          rate[m] = 0.0;
        }
        nsub = nit_total = status = 0;
        ncuts = nconst = attempts = last_why = 0;
        if (!n.full_geochem) {
          // RStep early exit (reaction.F90:3601-3605)
          if (is_aq) s.pri_molal[t * ld + c] = total / den * 1.0e3;
          if (t == 0) {
            if (s.num_iterations) s.num_iterations[c] = 0;
            if (s.num_sub_steps) s.num_sub_steps[c] = 0;
            if (s.status) s.status[c] = 0;
          }
          cell = -1;
        } else {
          // RStep set-up: clamp small totals (reaction.F90:3617-3631)
          small = false;
          if (mine && total <= 1.0e-40) {
            small = true;
            small_val = total;
            total = 1.0e-40;
          }
          guess = x;
          if (n.total_as_guess && is_aq) guess = total;
          target = dt;
          cum = 0.0;
          tdt = dt;
          nit = -1;  // request RReact set-up below
        }
      }
    }
    const int active = cell >= 0 ? 1 : 0;
    if (!__any(active)) {
      if (__all((int)exhausted)) break;  // every team of the wave is done
      continue;                          // some team still has to fetch
    }
    if (!active) continue;

    // ---------------- RReact set-up (reaction.F90:3800-3851) ----------------
    const bool wet = !(sat < kRtMinSaturation);
    const double psv = por * sat * 1000.0 * vol;
    const double dkl = den * 1.0e-3;
    if (nit < 0) {
      if (++attempts > kMaxRReactAttempts) {
        // safety bound on RReact calls per RStep: fail the cell
        status = last_why ? last_why : kMaxIter;
        if (is_aq) {
          s.pri_molal[t * ld + cell] = x;
          s.total[t * ld + cell] = total;
        } else if (mine) {
          s.immobile[(t - naq) * ld + cell] = x;
        }
        if (t == 0) {
          if (s.num_iterations) s.num_iterations[cell] = nit_total;
          if (s.num_sub_steps) s.num_sub_steps[cell] = nsub;
          if (s.status) s.status[cell] = status;
        }
        cell = -1;
        continue;
      }
      fixed = 0.0;
      if (wet && mine) fixed = is_aq ? psv * total : total * vol;
      init_tot = total;
      x = guess;
      if (!is_aq && mine) total = x;  // rt_auxvar%immobile = guess
      nit = 0;
      r0 = 0.0;
    }

    // ---------------- one Newton iteration ----------------
    ++nit;
    const double psvd = psv / tdt;
    // row scale: 1 / accumulation diagonal of the log-form Jacobian
    double rs = 1.0;
    if (wet && mine) rs = is_aq ? 1.0 / (psvd * dkl * x) : 1.0 / (vol / tdt * x);
    if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER && naq > 0) {
      // RActivityCoefficients LAG at the current pri / last sec molalities
      double Ip = is_aq ? x * n.Zp[t] * n.Zp[t] : 0.0;
      double sp = (is_aq && t != n.h2o_aq_id) ? x : 0.0;
      double Is = 0.0, ss = 0.0;
      for (int k = t; k < nc; k += L) {
        Is += mk[k] * n.cplx_Z[k] * n.cplx_Z[k];
        ss += mk[k];
      }
      const double I = 0.5 * team_sum<L>(Ip + Is);
      const double sI = sqrt(I);
      if (is_aq)
        lngam = fabs(n.Zp[t]) > 1.0e-10 ? log(dh_gamma(n.Zp[t], n.a0p[t], sI, I, n)) : 0.0;
      for (int k = t; k < nc; k += L)
        lng[k] = fabs(n.cplx_Z[k]) > 1.0e-10
                     ? log(dh_gamma(n.cplx_Z[k], n.cplx_a0[k], sI, I, n)) : 0.0;
      if (n.act_h2o) {
        const double a = 1.0 - 0.017 * team_sum<L>(sp + ss);
        ln_aw = a > 0.0 ? log(a) : 0.0;
      }
    }
    // phase A: per-species quantities into LDS; zero accumulators and J row
    if (mine) {
      const double lx = log(x);
      lnact[t] = is_aq ? lx + lngam : lx;
      rsv[t] = rs;
      xv[t] = x;
      tot[t] = 0.0;
      racc[t] = 0.0;
    }
    if (t < L) {
#pragma unroll
      for (int j = 0; j < L; ++j) Jt[t * (L + 1) + j] = (JT)0;
    }
    team_sync();
    // phase B: complexes (RTotalAqueous) -> sec molality, totals, dtotal
    const int ms = n.cplx_ms;
    for (int k = t; k < nc; k += L) {
      double logK = n.cplx_logK[k];
      if (!n.isothermal && n.has_cplx_logKcoef) logK = interp_logk(n.cplx_logKcoef + 5 * k, temp);
      double lnQK = -logK * kLogToLn;
      if (n.has_cplx_h2o) lnQK += n.cplx_h2o[k] * ln_aw;
      const int* ids = n.cplx_ids + k * (ms + 1);
      const double* st = n.cplx_st + k * ms;
      const int ncs = ids[0];
      for (int q = 0; q < ncs; ++q) lnQK += st[q] * lnact[ids[1 + q]];
      const double m_k = exp(lnQK - lng[k]);
      mk[k] = m_k;
      for (int q = 0; q < ncs; ++q) {
        const int i = ids[1 + q];
        atomicAdd(&tot[i], st[q] * m_k);
        if (wet) {
          const double ri = st[q] * m_k * rsv[i] * psvd * dkl;
          for (int r = 0; r < ncs; ++r) atomicAdd(&Jt[i * (L + 1) + ids[1 + r]], (JT)(ri * st[r]));
        }
      }
    }
    team_sync();
    // phase C: kinetic minerals (RKineticMineral) -> residual and J
    const int mms = n.mnrl_ms;
    for (int m = t; m < nm; m += L) {
      rate[m] = 0.0;
      if (area[m] == 0.0) continue;  // zero-area minerals give Im = 0
      const int* ids = n.mnrl_ids + m * (mms + 1);
      const double* st = n.mnrl_st + m * mms;
      const int ncs = ids[0];
      double logK = n.mnrl_logK[m];
      if (!n.isothermal && n.has_mnrl_logKcoef) logK = interp_logk(n.mnrl_logKcoef + 5 * m, temp);
      double lnQK = -logK * kLogToLn;
      if (n.has_mnrl_h2o) lnQK += n.mnrl_h2o[m] * ln_aw;
      for (int u = 0; u < ncs; ++u) lnQK += st[u] * lnact[ids[1 + u]];
      MineralRate mr = mineral_rate(n, m, lnQK, vf[m], area[m], temp);
      if (!mr.active) continue;
      rate[m] = mr.Im;
      // dry cells: RReaction adds nothing (reaction.F90:4083), but the
      // kinetic-state update still uses this rate
      if (!wet) continue;
      const double Im = mr.Im * vol;
      for (int u = 0; u < ncs; ++u) atomicAdd(&racc[ids[1 + u]], st[u] * Im);
      // log form: d res_i / d ln m_j = nu_i dIm/dQK nu_j QK * rho/1000
      const double g = mr.dIm_dQK * vol * mr.QK * dkl;
      for (int u = 0; u < ncs; ++u) {
        const int i = ids[1 + u];
        const double gi = g * st[u] * rsv[i];
        for (int v = 0; v < ncs; ++v) atomicAdd(&Jt[i * (L + 1) + ids[1 + v]], (JT)(gi * st[v]));
      }
    }
    team_sync();
    // phase D: residual, totals, accumulation diagonal, decay
    // Row t of the scaled log-form Jacobian: the LDS tile plus the
    // accumulation diagonal (1 after scaling) and first-order decay, the
    // identity for a dry cell; columns divided by x_j for the linear form.
    auto build_row = [&](JT* a) {
#pragma unroll
      for (int j = 0; j < L; ++j) a[j] = mine ? Jt[t * (L + 1) + j] : (JT)0;
      if (!mine) return;
      if (wet) {
        double d = 1.0;
        if (!is_aq)
          for (int r = 0; r < n.ndecay; ++r)
            if (n.decay_id[r] == t - naq) d += n.decay_k[r] * vol * x * rs;
#pragma unroll
        for (int j = 0; j < L; ++j)
          if (j == t) a[j] += (JT)d;
      } else {
#pragma unroll
        for (int j = 0; j < L; ++j) a[j] = (JT)(j == t ? 1 : 0);
      }
      if (!n.log_form) {
#pragma unroll
        for (int j = 0; j < L; ++j)
          if (j < N) a[j] = (JT)((double)a[j] / xv[j]);
      }
    };
    double mx = 0.0, nr2 = 0.0;
    JT arow[L];
    build_row(arow);
    if (mine) {
      if (is_aq) {
        total = (x + tot[t]) * dkl;
        res = wet ? (psv * total - fixed) / tdt : 0.0;
      } else {
        total = x;
        res = wet ? (x * vol - fixed) / tdt : 0.0;
        if (wet)
          for (int r = 0; r < n.ndecay; ++r)
            if (n.decay_id[r] == t - naq) res += n.decay_k[r] * vol * x;
      }
      if (wet) res += racc[t];
      mx = fabs(res);
      nr2 = res * res;
    }
    mx = team_max<L>(mx);
    const double nr = sqrt(team_sum<L>(nr2));
    if (nit == 1) r0 = nr;
    int outcome = kContinue;
    if (nit > n.max_it) {
      outcome = kFailMaxIt;
    } else if (mx < n.tol_abs || nr / r0 < n.tol_rel) {
      outcome = kConverged;
    } else {
      // ---------------- linear solve (RSolve) ----------------
      double amax = 0.0;
#pragma unroll
      for (int j = 0; j < L; ++j) amax = fmax(amax, (double)fabs(arow[j]));
      const int zero_row = team_or<L>(mine && !(amax > 0.0) ? 1 : 0);
      if (zero_row) {
        outcome = kFailSingular;
      } else {
        const JT rhs0 = mine ? (JT)(res * rs) : (JT)0;
        JT rhs = rhs0;
        JT xs[L];
        // Fixed-order elimination: the pivot of step p is row p.  The
        // log-form Jacobian of this system (accumulation S^T M S + diag(x),
        // TST minerals c nu nu^T with c >= 0) is symmetric positive definite
        // up to the row scaling, so no pivot search is needed; a pivot that
        // fails the guard sends the team to the pivoting path below.
        constexpr double kPivTol = sizeof(JT) == 4 ? 1.0e-6 : 1.0e-12;
        int bad = 0;
        static_for<0, L>([&](auto P) {
          constexpr int p = decltype(P)::value;
          if (p < N) {
            const JT pv = bcast<L, p>(arow[p]);
            if (t == p && !((double)fabs(arow[p]) > kPivTol * amax)) bad = 1;
            JT pr[L];
            static_for<p + 1, L>([&](auto Q) {
              constexpr int j = decltype(Q)::value;
              pr[j] = bcast<L, p>(arow[j]);
            });
            const JT prhs = bcast<L, p>(rhs);
            if (t > p) {
              const JT l = arow[p] / pv;
              arow[p] = l;
              static_for<p + 1, L>([&](auto Q) {
                constexpr int j = decltype(Q)::value;
                arow[j] -= l * pr[j];
              });
              rhs -= l * prhs;
            }
          }
        });
        bad = team_or<L>(bad);
        if (!bad) {
          // back substitution, rows in natural order
          static_for_down<0, L>([&](auto P) {
            constexpr int p = decltype(P)::value;
            xs[p] = (JT)0;
            if (p < N) {
              JT sum = rhs;
              static_for<p + 1, L>([&](auto Q) {
                constexpr int j = decltype(Q)::value;
                sum -= arow[j] * xs[j];
              });
              xs[p] = bcast<L, p>(sum / arow[p]);
              if (t == p) delta = (double)xs[p];
            }
          });
        } else {
          // scaled partial pivoting (LUDecomposition1, utility.F90:597-688)
          build_row(arow);
          rhs = rhs0;
          const JT vv = mine ? (JT)(1.0 / amax) : (JT)0;
          bool used = !mine;
          int mystep = -1;
#pragma unroll
          for (int p = 0; p < L; ++p) {
            if (p < N) {
              // ties go to the larger row index, as in the Crout loop
              JT cand = used ? (JT)-1 : (JT)fabs(arow[p]) * vv;
              int ci = t;
#pragma unroll
              for (int o = L / 2; o >= 1; o >>= 1) {
                const JT oc = __shfl_xor(cand, o, L);
                const int oi = __shfl_xor(ci, o, L);
                if (oc > cand || (oc == cand && oi > ci)) {
                  cand = oc;
                  ci = oi;
                }
              }
              const int piv = ci;
              if (t == piv) {
                if (arow[p] == (JT)0) arow[p] = (JT)1.0e-20;
#pragma unroll
                for (int j = p; j < L; ++j) prow[j] = arow[j];
                prow[L] = rhs;
                used = true;
                mystep = p;
              }
              team_sync();
              if (!used) {
                const JT l = arow[p] / prow[p];
                arow[p] = l;
#pragma unroll
                for (int j = p + 1; j < L; ++j) arow[j] -= l * prow[j];
                rhs -= l * prow[L];
              }
              team_sync();
            }
          }
          // back substitution: x_p from the row pivoted at step p
#pragma unroll
          for (int p = L - 1; p >= 0; --p) {
            xs[p] = (JT)0;
            if (p < N) {
              if (mystep == p) {
                JT sum = rhs;
#pragma unroll
                for (int j = p + 1; j < L; ++j) sum -= arow[j] * xs[j];
                prow[0] = sum / arow[p];
              }
              team_sync();
              xs[p] = prow[0];
              team_sync();
              if (t == p) delta = (double)xs[p];
            }
          }
        }
        // ---------------- update (reaction.F90:3977-4049) ----------------
        double latest = x, rc = 0.0;
        if (n.log_form) {
          if (mine) {
            const double u = copysign(1.0, delta) * fmin(fabs(delta), n.max_dlnC);
            latest = x * exp(-u);
          }
        } else {
          double ratio = 1.0e308;
          if (mine && x <= delta) ratio = fabs(x / delta);
          const double mr = team_min<L>(ratio);
          if (mine) latest = x - (mr < 1.0 ? delta * mr * 0.99 : delta);
        }
        if (mine) {
          rc = fabs((latest - x) / x);
          if (!(rc == rc)) rc = -1.0;
        }
        const double mrc = team_max<L>(rc);
        // With an FP32 factorization the step is only as accurate as the
        // factorization, so a small-step exit is accepted only when the FP64
        // residual of the iterate confirms it.
        const bool step_exit_ok = sizeof(JT) == 8 || nr <= kMixedStepExitResidual * r0;
        if (mrc >= 0.0 && mrc < n.tol_change && step_exit_ok) {
          outcome = kConverged;
        } else if (mine) {
          x = latest;
        }
      }
    }

    // ---------------- RReact / RStep transitions ----------------
    const int64_t c = cell;
    bool finish = false;
    if (outcome == kConverged) {
      // RUpdateKineticState: mineral volume fractions with the final rates
      for (int m = t; m < nm; m += L) {
        vf[m] += rate[m] * n.mnrl_Vm[m] * tdt;
        if (vf[m] < 0.0) vf[m] = 0.0;
      }
      nit_total += nit;
      cum += tdt;
      ++nsub;
      ++nconst;
      guess = x;
      if (nconst >= 4) {
        --ncuts;
        tdt = fmin(2.0 * tdt, target - cum);
      }
      nit = -1;
      if (cum >= target) {
        finish = true;
        if (small) total = small_val;
      }
    } else if (outcome != kContinue && sizeof(JT) == 4 && escalate) {
      // The FP32 factorization could not bring this cell to convergence:
      // leave its inputs untouched and hand it to the FP64 pass.
      if (t == 0 && s.status) s.status[c] = kEscalate;
      cell = -1;
      continue;
    } else if (outcome != kContinue) {
      nit_total += nit;
      if (outcome == kFailMaxIt) total = init_tot;  // totals / immobile restored
      if (outcome == kFailMaxIt && !is_aq) x = init_tot;
      last_why = outcome == kFailMaxIt ? kMaxIter : kSingular;
      ++ncuts;
      if (ncuts > n.max_cuts) {
        status = last_why;
        finish = true;
      } else {
        tdt *= 0.5;
        nconst = 0;
        nit = -1;
      }
    }
    if (finish) {
      if (is_aq) {
        s.pri_molal[t * ld + c] = x;
        s.total[t * ld + c] = total;
        if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
          s.pri_act_coef[t * ld + c] = exp(lngam);
      } else if (mine) {
        s.immobile[(t - naq) * ld + c] = (outcome == kConverged) ? (small ? small_val : x) : x;
      }
      for (int k = t; k < nc; k += L) {
        s.sec_molal[k * ld + c] = mk[k];
        if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
          s.sec_act_coef[k * ld + c] = exp(lng[k]);
      }
      for (int m = t; m < nm; m += L) {
        s.mnrl_volfrac[m * ld + c] = vf[m];
        if (s.mnrl_rate) s.mnrl_rate[m * ld + c] = rate[m];
      }
      if (t == 0) {
        if (s.ln_act_h2o && n.act_h2o && n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
          s.ln_act_h2o[c] = ln_aw;
        if (s.num_iterations) s.num_iterations[c] = nit_total;
        if (s.num_sub_steps) s.num_sub_steps[c] = nsub;
        if (s.status) s.status[c] = status;
      }
      cell = -1;
    }
  }
}

__global__ void zero_counter(unsigned long long* c) {
  if (threadIdx.x == 0) *c = 0ull;
}

template <int L, typename JT>
hipError_t launch_L(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                    double dt, const LaunchCfg& cfg, unsigned long long* counter,
                    int num_cus, hipStream_t stream, int rescue_only) {
  const TeamLds lay = team_lds_layout(L, n.nc, (int)sizeof(JT));
  const int block = cfg.block_size;
  const size_t lds = (size_t)(block / L) * lay.bytes;
  int64_t blocks = (int64_t)num_cus * cfg.blocks_per_cu;
  const int64_t teams_needed = (ncell + 0);
  const int64_t max_blocks = (teams_needed + (block / L) - 1) / (block / L);
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks < 1) blocks = 1;
  zero_counter<<<1, 64, 0, stream>>>(counter);
  team_step_kernel<L, JT><<<dim3((unsigned)blocks), dim3(block), lds, stream>>>(
      dn, ds, ncell, dt, counter, rescue_only);
  return hipGetLastError();
}

}  // namespace

int team_max_ncomp() { return 64; }

int team_supports(const DevNet& n, const LaunchCfg& cfg, bool mixed) {
  const int L = cfg.lanes_per_cell;
  // L = 64 is switched off: that build gave wrong LU results on the GPU
  // (singular-pivot failures on every cell) that vanish when the code
  // generation changes; it stays off until the cause is shown.
  if (L != 4 && L != 8 && L != 16 && L != 32) return 0;
  if (n.ncomp > L || n.nm > kMaxMnrl) return 0;
  if (cfg.block_size % 64 || cfg.block_size > 256 || cfg.block_size < 64) return 0;
  const TeamLds lay = team_lds_layout(L, n.nc, mixed ? 4 : 8);
  if ((size_t)(cfg.block_size / L) * lay.bytes > 64 * 1024) return 0;
  return 1;
}

hipError_t launch_team_step(const DevNet& n, const DevNet* dn, const DevState* ds,
                            int64_t ncell, double dt, bool mixed, const LaunchCfg& cfg,
                            unsigned long long* counter, int num_cus, hipStream_t stream,
                            bool has_status) {
  switch (cfg.lanes_per_cell) {
#define PFCHEM_CASE(LL)                                                                   \
  case LL:                                                                                \
    if (!mixed) return launch_L<LL, double>(n, dn, ds, ncell, dt, cfg, counter, num_cus, stream, 0); \
    {                                                                                     \
      hipError_t e = launch_L<LL, float>(n, dn, ds, ncell, dt, cfg, counter, num_cus, stream, 0); \
      if (e != hipSuccess || !has_status) return e;                                       \
      LaunchCfg c64 = cfg;                                                                \
      while (c64.block_size > 64 && !team_supports(n, c64, false)) c64.block_size /= 2;   \
      return launch_L<LL, double>(n, dn, ds, ncell, dt, c64, counter, num_cus, stream, 1); \
    }
    PFCHEM_CASE(4)
    PFCHEM_CASE(8)
    PFCHEM_CASE(16)
    PFCHEM_CASE(32)
    PFCHEM_CASE(64)
#undef PFCHEM_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pfchem
