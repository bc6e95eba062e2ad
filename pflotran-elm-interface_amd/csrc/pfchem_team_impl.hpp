// pfchem_team_impl.hpp -- batched reaction step: persistent, work-stealing
// kernel in which a team of L lanes (L = 1..32, inside one wavefront) owns
// one cell at a time and each lane owns R rows of that cell's system.
//
// Mapping (gfx950, wave64):
//   * row / species i of the cell belongs to lane t = i % L, slot r = i / L
//     (NM = L*R rows per team, 64/L cells per wave);
//   * complexes k = t, t+L, ... and minerals m = t, t+L, ... are spread over
//     the team's lanes; their contributions are scattered into the team's
//     LDS tile with LDS atomics (FP64 totals/residual, JT Jacobian);
//   * the LU keeps each lane's rows in registers: the pivot row of step p is
//     row p (owner lane p % L), broadcast with DPP / ds_swizzle / readlane;
//   * every wave is an independent persistent worker.  Teams take cells
//     from a wave-local queue refilled, one atomic per refill, from the
//     cursor of the wave's XCD: each XCD (own L2) works through a
//     contiguous cell range and then steals from the others, so uneven
//     Newton counts never idle a wave while cells remain and neighbouring
//     cells (which share cache lines in every SoA row) stay on one L2.
// Algorithm: RStep (reaction.F90:3564-3738) / RReact (:3742-4055) with
// RTotalAqueous (:4665-4759), RTAccumulation(Derivative) (:5710-5848),
// RKineticMineral (reaction_mineral.F90:647-1078), RImmobileDecay
// (reaction_immobile.F90:244-294) and RSolve (:5457-5516, back-substituting).
// Numerical reformulations (same mathematics, different rounding):
//   * the Jacobian is assembled directly in the log formulation
//     (d res_i / d ln x_j = J_ij x_j, as RSolve forms it) with each row
//     divided by its accumulation diagonal; the Newton update does not
//     depend on row scaling;
//   * that matrix (accumulation S^T diag(m) S + diag(x), TST minerals
//     c nu nu^T with c >= 0, decay on the diagonal) is symmetric positive
//     definite up to the row scaling, so the LU eliminates in natural order;
//     a pivot that fails a size guard sends the team to a Crout
//     factorization with implicit partial pivoting (LUDecomposition1,
//     utility.F90:597-688) run by one lane on the LDS tile;
//   * JT = float: Jacobian entries, LU and triangular solves in FP32, while
//     the residual, mass-action log K sums, exponentials, the log-space
//     update and every convergence test stay FP64 (the Newton iteration is
//     the iterative refinement).  A cell whose FP32 Newton stalls is handed,
//     untouched, to an FP64 pass.
#pragma once
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"
#include "pfchem_somdec.hpp"

#include <type_traits>

namespace pfchem {
namespace {

constexpr int kMaxMnrl = 16;

__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------
// Cross-lane primitives inside a team of L consecutive lanes, built on DPP
// (quad_perm, row_half_mirror, row_mirror, row_newbcast), ds_swizzle and
// readlane.  All lanes of a team must be active.
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
  if constexpr (L == 1) {
    return v;
  } else if constexpr (L == 2) {
    return __builtin_amdgcn_update_dpp(0, v, P | (P << 2) | ((2 + P) << 4) | ((2 + P) << 6),
                                       0xF, 0xF, false);
  } else if constexpr (L == 4) {
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
  if constexpr (L == 1) {
    return v;
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = bcast32<L, P>((int)(b & 0xffffffffll));
    const int hi = bcast32<L, P>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// partner lane at butterfly level LEV (0: xor 1, 1: xor 2, 2: mirror in 8,
// 3: mirror in 16, 4: swap halves of 32, 5: swap halves of the wave)
template <int LEV>
__device__ __forceinline__ int partner32(int v) {
  if constexpr (LEV == 0)
    return __builtin_amdgcn_update_dpp(0, v, 1 | (0 << 2) | (3 << 4) | (2 << 6), 0xF, 0xF, false);
  else if constexpr (LEV == 1)
    return __builtin_amdgcn_update_dpp(0, v, 2 | (3 << 2) | (0 << 4) | (1 << 6), 0xF, 0xF, false);
  else if constexpr (LEV == 2)
    return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
  else if constexpr (LEV == 3)
    return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
  else if constexpr (LEV == 4)
    return __builtin_amdgcn_ds_swizzle(v, 0x1F | (0x10 << 10));  // SWAP 16
  else
    return __shfl_xor(v, 32, 64);
}
template <int LEV>
__device__ __forceinline__ double partner(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = partner32<LEV>((int)(b & 0xffffffffll));
  const int hi = partner32<LEV>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int LEV>
__device__ __forceinline__ int partner(int v) {
  return partner32<LEV>(v);
}

// all-reduce over the team; each level pairs every lane with a distinct
// partner and the combines are commutative, so every lane of the team gets
// the bit-identical result (team-uniform convergence decisions)
template <int L, typename T, typename Op>
__device__ __forceinline__ T team_reduce(T v, Op op) {
  if constexpr (L >= 2) v = op(v, partner<0>(v));
  if constexpr (L >= 4) v = op(v, partner<1>(v));
  if constexpr (L >= 8) v = op(v, partner<2>(v));
  if constexpr (L >= 16) v = op(v, partner<3>(v));
  if constexpr (L >= 32) v = op(v, partner<4>(v));
  if constexpr (L >= 64) v = op(v, partner<5>(v));
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

// ---------------------------------------------------------------------
// LDS tile of one team (byte offsets); the same formula on host and device
// ---------------------------------------------------------------------
struct TeamLds {
  int lnact, rs, tot, racc, xv, mvf, marea, mrate, lng, sol, J, sd, bytes;
};
// nsd: doubles of per-cell SomDec stoichiometry (0 without the sandbox)
__host__ __device__ inline TeamLds team_lds_layout(int NM, int nc, int nm, int jt_size, int nsd) {
  TeamLds t;
  int o = 0;
  t.lnact = o; o += 8 * NM;
  t.rs = o;    o += 8 * NM;
  t.tot = o;   o += 8 * NM;
  t.racc = o;  o += 8 * NM;
  t.xv = o;    o += 8 * NM;
  const int m = nm > 0 ? nm : 1;
  t.mvf = o;   o += 8 * m;
  t.marea = o; o += 8 * m;
  t.mrate = o; o += 8 * m;
  t.lng = o;   o += 8 * (nc > 0 ? nc : 1);
  t.sol = o;   o += 8 * (NM + 2);
  t.J = o;     o += jt_size * NM * (NM + 1);
  t.sd = o;    o += 8 * nsd;
  o = (o + 15) & ~15;
  t.bytes = o;
  return t;
}

// TST rate of one kinetic mineral (RKineticMineral, reaction_mineral.F90:
// 729-908) given ln QK: rate Im (mol/m^3 bulk/s, before the volume factor)
// and dIm/dQK per unit volume, with the affinity options.
struct MineralRate {
  double Im, dIm_dQK, QK;
  double kr;  // rate constant with its Arrhenius factor, or the prefactor sum
  bool active;
};

// exponent 1/sigma of QK in the affinity 1 - QK^(1/sigma)
__device__ __forceinline__ double affinity_sigma(const DevNet& n, int m) {
  if (n.has_temkin)
    return n.has_minscale ? gld(n.mnrl_minscale + m) * gld(n.mnrl_temkin + m)
                          : gld(n.mnrl_temkin + m);
  return gld(n.mnrl_minscale + m);
}

__device__ __noinline__ double pow_noinline(double a, double b) { return pow(a, b); }

// Im and dIm/dQK with the optional affinity power, Temkin / scale factor
// and rate limiter (reaction_mineral.F90:839-908)
__device__ __forceinline__ void rate_options(const DevNet& n, int m, double QK, double aff,
                                             double sgn, double kr, double Im_const,
                                             MineralRate* r) {
  const double p = n.has_affpower ? gld(n.mnrl_affpower + m) : 1.0;
  r->Im = n.has_affpower ? Im_const * sgn * pow_noinline(fabs(aff), p) * kr
                         : Im_const * sgn * fabs(aff) * kr;
  double d = n.has_affpower ? -r->Im * p / fabs(aff) : -Im_const * kr;
  if (n.has_temkin || n.has_minscale) d = d * (1.0 / affinity_sigma(n, m)) / QK * (1.0 - aff);
  const double lim = n.has_limiter ? gld(n.mnrl_limiter + m) : 0.0;
  if (lim > 0.0) {
    const double d2 = 1.0 + (1.0 - aff) / lim;
    d *= (1.0 + QK / lim / d2) / d2;
  }
  r->dIm_dQK = d;
}

template <typename A = double>
__device__ __forceinline__ MineralRate mineral_rate(const DevNet& n, int m, double lnQK,
                                                    double volfrac, double area,
                                                    double temp, const A* lnact = nullptr) {
  MineralRate r;
  r.active = false;
  r.Im = r.dIm_dQK = r.kr = 0.0;
  const double QK = exp(lnQK);
  r.QK = QK;
  const bool plain_aff = !(n.has_temkin || n.has_minscale);
  double aff = plain_aff ? 1.0 - QK : 1.0 - pow_noinline(QK, 1.0 / affinity_sigma(n, m));
  const double sgn = copysign(1.0, aff);
  if (!(volfrac > 0.0 || sgn < 0.0)) return r;
  if (n.has_irrev && gld(n.mnrl_irrev + m) == 1 && sgn < 0.0) return r;
  if (n.has_threshold) {
    const double th = gld(n.mnrl_threshold + m);
    if (th > 0.0 && sgn < 0.0 && QK < th) return r;
  }
  const double lim = n.has_limiter ? gld(n.mnrl_limiter + m) : 0.0;
  if (lim > 0.0) aff = aff / (1.0 + (1.0 - aff) / lim);
  double kr;
  if (n.has_pref && lnact && gld(n.pref_n + m) > 0) {
    kr = prefactor_sum(n, m, lnact, temp);  // reaction_mineral.F90:820-860
  } else {
    double arr = 1.0;
    if (n.has_Ea) {
      const double Ea = gld(n.mnrl_Ea + m);
      if (Ea > 0.0)
        arr = exp(Ea / kIdealGas * (1.0 / (25.0 + 273.15) - 1.0 / (temp + 273.15)));
    }
    kr = gld(n.mnrl_k + m) * arr;
  }
  r.kr = kr;
  double Im_const = -area;
  if (n.has_minscale) Im_const = Im_const / gld(n.mnrl_minscale + m);
  r.active = true;
  if (plain_aff && !n.has_affpower && !(lim > 0.0)) {
    r.Im = Im_const * sgn * fabs(aff) * kr;
    r.dIm_dQK = -Im_const * kr;
  } else {
    rate_options(n, m, QK, aff, sgn, kr, Im_const, &r);
  }
  return r;
}

// mass-action ln QK of complex k from the team's ln-activity vector
__device__ __forceinline__ double complex_lnqk(const DevNet& n, int k, const double* lnact,
                                               double ln_aw, double temp) {
  const int ms = n.cplx_ms;
  double logK = n.isothermal || !n.has_cplx_logKcoef ? gld(n.cplx_logK + k)
                                                     : interp_logk(n.cplx_logKcoef + 5 * k, temp);
  double lnQK = -logK * kLogToLn;
  if (n.has_cplx_h2o) lnQK += gld(n.cplx_h2o + k) * ln_aw;
  const int* ids = n.cplx_ids + k * (ms + 1);
  const double* st = n.cplx_st + k * ms;
  const int ncs = gld(ids);
  for (int q = 0; q < ncs; ++q) lnQK += gld(st + q) * lnact[gld(ids + 1 + q)];
  return lnQK;
}

// Species slots per reaction row held in registers; rows with more species
// take a slower path for the rest.
constexpr int kRowSlots = 4;

// One reaction row (species ids and stoichiometry, the compressed
// eqcplxspecid / eqcplxstoich layout of reaction_database.F90:1704-1772)
// staged in registers together with the row scales of its species.
template <int S>
struct RxnRow {
  int ncs;
  int id[S];
  double nu[S], rs[S];

  // stage the row; returns init + sum nu_a ln a_a (the mass-action sum)
  __device__ __forceinline__ double load(const int* ids, const double* st, const double* lnact,
                                         const double* rsv, double init) {
    ncs = gld(ids);
#pragma unroll
    for (int a = 0; a < S; ++a) {
      id[a] = a < ncs ? gld(ids + 1 + a) : 0;
      nu[a] = a < ncs ? gld(st + a) : 0.0;
    }
    double sum = init;
#pragma unroll
    for (int a = 0; a < S; ++a) {
      rs[a] = 0.0;
      if (a < ncs) {
        sum += nu[a] * lnact[id[a]];
        rs[a] = rsv[id[a]];
      }
    }
    for (int a = S; a < ncs; ++a) sum += gld(st + a) * lnact[gld(ids + 1 + a)];
    return sum;
  }

  // acc_i += nu_i v; J_ij += nu_i nu_j w * (row scale of i)
  template <int NM>
  __device__ __forceinline__ void scatter(const int* ids, const double* st, double v, double w,
                                          bool jac, double* acc, double* Jt,
                                          const double* rsv) const {
#pragma unroll
    for (int a = 0; a < S; ++a) {
      if (a < ncs) {
        atomicAdd(&acc[id[a]], nu[a] * v);
        if (jac) {
          const double wi = nu[a] * w * rs[a];
#pragma unroll
          for (int b = 0; b < S; ++b)
            if (b < ncs) atomicAdd(&Jt[id[a] * (NM + 1) + id[b]], wi * nu[b]);
        }
      }
    }
    if (ncs > S) {
      for (int a = 0; a < ncs; ++a) {
        const int i = gld(ids + 1 + a);
        const double nua = gld(st + a);
        if (a >= S) atomicAdd(&acc[i], nua * v);
        if (!jac) continue;
        const double wi = nua * w * rsv[i];
        for (int b = 0; b < ncs; ++b)
          if (a >= S || b >= S)
            atomicAdd(&Jt[i * (NM + 1) + gld(ids + 1 + b)], wi * gld(st + b));
      }
    }
  }
};

// prefactor Jacobian terms of mineral m into the team's tile (scaled log
// form: d res_i / d ln m_j times the row scale of i)
template <int NM>
__device__ void team_prefactor_jacobian(const DevNet& n, int m, const double* lnact, double temp,
                                        double IV_over_sum, double* Jt, const double* rsv) {
  const int* ids = n.mnrl_ids + m * (n.mnrl_ms + 1);
  const double* st = n.mnrl_st + m * n.mnrl_ms;
  const int ncs = gld(ids);
  prefactor_jacobian(n, m, lnact, temp, IV_over_sum, [&](int j, double v) {
    for (int a = 0; a < ncs; ++a) {
      const int i = gld(ids + 1 + a);
      atomicAdd(&Jt[i * (NM + 1) + j], gld(st + a) * v * rsv[i]);
    }
  });
}

// SomDec accumulator of a team (see pfchem_somdec.hpp): species values from
// the team's LDS tile, residual terms into the reaction residual, Jacobian
// terms d res_row / d c_col into the tile in the scaled log form
// (times c_col and the row scale).  X (the oxygen / complexes build):
// totals are (free molality + the complexes' sum) times density and the
// dtotal entries the sandbox reads come from the complex pair sums S;
// without X, free molality times density and a diagonal dtotal.
template <bool X>
struct TeamSdEnvT {
  static constexpr bool kExt = X;
  // LDS-qualified pointers: sd_reaction is too large to inline, and through
  // generic pointers its loads and FP64 atomic adds would take the FLAT path
  typedef __attribute__((address_space(3))) double lds_t;
  const lds_t* xv;
  const lds_t* lna;
  const lds_t* rsv;
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
    __hip_atomic_fetch_add(&Jt[row * ld + col], v * xv[col] * rsv[row], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ lds_t& unc(int r) { return st[r]; }
  __device__ lds_t& mcs(int r) { return st[nrxn + r]; }
  __device__ lds_t& mns(int r) { return st[2 * nrxn + r]; }
  __device__ lds_t& dnc(int r, int j) { return st[3 * nrxn + r * kSdDown + j]; }
  static __device__ lds_t* lds(double* p) { return (lds_t*)p; }
};
using TeamSdEnv = TeamSdEnvT<false>;

// Crout LU with implicit partial pivoting and back substitution
// (LUDecomposition1 / LUBackSubstitution, utility.F90:597-735) on the
// team's LDS tile, run by one lane; vv / indx are N-entry LDS work arrays.
// Returns 1 on a zero row.
template <typename JT>
__device__ int serial_lu_solve(JT* a, int lda, int N, double* b, double* vv, int* indx) {
  int imax = 0;
  for (int i = 0; i < N; ++i) {
    double amax = 0.0;
    for (int j = 0; j < N; ++j) amax = fmax(amax, fabs((double)a[i * lda + j]));
    if (amax <= 0.0) return 1;
    vv[i] = 1.0 / amax;
  }
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < j; ++i) {
      double sum = a[i * lda + j];
      for (int k = 0; k < i; ++k) sum -= (double)a[i * lda + k] * (double)a[k * lda + j];
      a[i * lda + j] = (JT)sum;
    }
    double amax = 0.0;
    for (int i = j; i < N; ++i) {
      double sum = a[i * lda + j];
      for (int k = 0; k < j; ++k) sum -= (double)a[i * lda + k] * (double)a[k * lda + j];
      a[i * lda + j] = (JT)sum;
      const double dum = vv[i] * fabs(sum);
      if (dum >= amax) {
        imax = i;
        amax = dum;
      }
    }
    if (j != imax) {
      for (int k = 0; k < N; ++k) {
        const JT tt = a[imax * lda + k];
        a[imax * lda + k] = a[j * lda + k];
        a[j * lda + k] = tt;
      }
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (a[j * lda + j] == (JT)0) a[j * lda + j] = (JT)1.0e-20;
    if (j != N - 1) {
      const double dum = 1.0 / (double)a[j * lda + j];
      for (int i = j + 1; i < N; ++i) a[i * lda + j] = (JT)((double)a[i * lda + j] * dum);
    }
  }
  int ii = -1;
  for (int i = 0; i < N; ++i) {
    const int ll = indx[i];
    double sum = b[ll];
    b[ll] = b[i];
    if (ii >= 0) {
      for (int j = ii; j < i; ++j) sum -= (double)a[i * lda + j] * b[j];
    } else if (sum != 0.0) {
      ii = i;
    }
    b[i] = sum;
  }
  for (int i = N - 1; i >= 0; --i) {
    double sum = b[i];
    for (int j = i + 1; j < N; ++j) sum -= (double)a[i * lda + j] * b[j];
    b[i] = sum / (double)a[i * lda + i];
  }
  return 0;
}

// relative FP64 residual a mixed-precision iterate must reach before a
// small Newton step may end the iteration
constexpr double kMixedStepExitResidual = 1.0e-11;
// transient per-cell status: the FP32 pass hands the cell to the FP64 pass
constexpr int kEscalate = 16;
// cells claimed per refill of a wave's queue: a multiple of 16 so that a
// claim covers whole 128-byte lines of every FP64 SoA row
constexpr int kClaimAlign = 16;

enum Outcome : int { kContinue = 0, kConverged = 1, kFailMaxIt = 2, kFailSingular = 3 };

// waves per SIMD the register allocation is sized for
#ifndef PFCHEM_TEAM_WAVES_PER_EU
#define PFCHEM_TEAM_WAVES_PER_EU 2
#endif

// SD: the network carries the SomDec sandbox; VAR: cells carry their own
// active primary counts (options.variable_ncomp).  Both are instantiated for
// a few team shapes only, so the other networks' kernels keep their code.
template <int L, int R, typename JT, int SD, bool VAR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PFCHEM_TEAM_WAVES_PER_EU)))
team_step_kernel(const DevNet* np, const DevState* sp, int64_t ncell, double dt,
                 unsigned long long* __restrict__ cursors, int nxcd, int rescue_only) {
  constexpr int NM = L * R;
  // reaction rows per lane staged in registers: complexes per chunk (more
  // than one spills registers) and kinetic minerals
  constexpr int KC = 1;
  constexpr int KM = (kMaxMnrl + L - 1) / L < 4 ? (kMaxMnrl + L - 1) / L : 4;
  const DevNet& n = *np;
  const DevState& s = *sp;
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & 63;
  const int t = lane & (L - 1);          // lane within team
  const int team_in_wave = lane / L;
  const int wave_in_block = threadIdx.x >> 6;
  const int nsd = sd_lds_doubles(n.sd.nrxn, SD);
  const TeamLds lay = team_lds_layout(NM, n.nc, n.nm, 8, nsd);
  char* tb = smem + (size_t)(wave_in_block * (64 / L) + team_in_wave) * lay.bytes;
  double* lnact = (double*)(tb + lay.lnact);
  double* rsv = (double*)(tb + lay.rs);
  double* tot = (double*)(tb + lay.tot);
  double* racc = (double*)(tb + lay.racc);
  double* xv = (double*)(tb + lay.xv);
  double* vf = (double*)(tb + lay.mvf);
  double* area = (double*)(tb + lay.marea);
  double* rate = (double*)(tb + lay.mrate);
  double* lng = (double*)(tb + lay.lng);
  double* sol = (double*)(tb + lay.sol);
  // Jacobian tile: FP64 for both variants (LDS FP64 atomic adds are cheap
  // on gfx950, FP32 ones are not); JT is the precision of the LU
  double* Jt = (double*)(tb + lay.J);
  double* sdst = (double*)(tb + lay.sd);  // SomDec stoichiometry of the team's cell
  double* sdS = SD >= 2 ? sdst + n.sd.nrxn * (3 + kSdDown) : nullptr;  // its dtotal pair sums

  // aqueous primaries and unknowns of the team's cell: the network's, or
  // with VAR the cell's active count (team-uniform)
  const int naq0 = n.naq, N0 = n.ncomp;
  int naq_var = naq0, N_var = N0;
  const int& naq = VAR ? naq_var : naq0;
  const int& N = VAR ? N_var : N0;
  const int nc = n.nc, nm = n.nm;
  // complex k / mineral m takes part in the cell's system
  auto cplx_on = [&](int k) { return !VAR || gld(n.cplx_maxid + k) < naq; };
  auto mnrl_on = [&](int m) { return !VAR || gld(n.mnrl_maxid + m) < naq; };
  const int64_t ld = s.ld;
  const bool escalate = s.status != nullptr;  // FP32 failures go to the FP64 pass

  // XCD of this workgroup (workgroups are dispatched round robin over XCDs)
  const int xcd = (int)(blockIdx.x % (unsigned)nxcd);
  // cells of this launch: all, or a list (a size bucket's, or the cells an
  // earlier pass handed over)
  const bool listed = s.cell_list != nullptr;
  const int64_t nwork = listed ? (int64_t)gld(s.list_count) : ncell;
  const int64_t chunk = (nwork + nxcd - 1) / nxcd;

  // team-uniform cell state
  int64_t cell = -1;
  bool exhausted = false;
  double target = 0.0, cum = 0.0, tdt = 0.0, r0 = 0.0, ln_aw = 0.0;
  double por = 0.0, sat = 0.0, den = 0.0, vol = 0.0, temp = 25.0;
  double sd_ft = 1.0, sd_fw = 1.0;  // ELM temperature / moisture scalars
  int nconst = 0, ncuts = 0, nsub = 0, nit_total = 0, nit = 0, status = 0;
  int attempts = 0, last_why = 0;
  // per-slot state: row i = r*L + t
  double x[R], guess[R], fixed[R], init_tot[R], total[R], lngam[R], small_val[R];
  bool small[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    x[r] = guess[r] = 1.0;
    fixed[r] = init_tot[r] = total[r] = lngam[r] = small_val[r] = 0.0;
    small[r] = false;
  }
  // LAG activity partial sums over the sec molalities this lane computed
  double sec_I = 0.0, sec_sum = 0.0;

  // wave-local queue of claimed cells [qnext, qend) (wave-uniform)
  long long qnext = 0, qend = 0;
  int steal = 0;  // XCD ranges found empty so far (wave-uniform)
  int batch = ((64 / L) > kClaimAlign ? (64 / L) : kClaimAlign);
  if (listed) {
    // list entries are scattered cells (no lines to keep whole): a short
    // list is claimed one cell per team, so that every team gets a cell
    // rather than a few waves running the whole list several cells deep
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    if (nwork < 4 * (int64_t)batch * waves) batch = 64 / L;
  }

  for (;;) {
    // ---------------- work stealing: hand out cells ----------------
    const int needy = (t == 0 && cell < 0 && !exhausted) ? 1 : 0;
    const unsigned long long needmask = __ballot(needy);
    long long got = -1;
    if (needmask) {
      const int cnt = __popcll(needmask);
      const int rk = needy ? __popcll(needmask & ((1ull << lane) - 1ull)) : 0;
      long long avail = qend - qnext;
      if (needy && rk < avail) got = qnext + rk;
      int taken = (int)(avail < cnt ? avail : cnt);
      qnext += taken;
      // refill from this XCD's range, then steal from the other ranges
      while (taken < cnt && steal < nxcd) {
        const int xr = (xcd + steal) % nxcd;
        const long long lo = (long long)xr * chunk;
        const long long hi = min((long long)nwork, lo + chunk);
        long long base = 0;
        if (lane == 0) base = (long long)atomicAdd(cursors + 16 * xr, (unsigned long long)batch);
        base = __shfl(base, 0, 64);
        if (lo + base >= hi) {
          ++steal;
          continue;
        }
        qnext = lo + base;
        qend = min(hi, qnext + batch);
        avail = qend - qnext;
        const int need = cnt - taken;
        if (needy && rk >= taken && rk < taken + avail) got = qnext + (rk - taken);
        const int take = (int)(avail < need ? avail : need);
        qnext += take;
        taken += take;
      }
      if (needy && got < 0) got = nwork;  // nothing left anywhere
    }
    got = __shfl(got, team_in_wave * L, 64);
    if (cell < 0 && !exhausted && got >= 0) {
      const long long gc = (got < nwork && listed) ? (long long)gld(s.cell_list + got) : got;
      if (got >= nwork) {
        exhausted = true;
      } else if (rescue_only && gld(s.status + (gc)) != kEscalate) {
        // FP64 rescue pass: only cells the FP32 pass handed over
      } else {
        cell = gc;
        const int64_t c = cell;
        if constexpr (VAR) {
          naq_var = min(max(gld(s.ncomp_active + (c)), 1), n.naq);
          N_var = naq_var;
        }
        por = gld(s.porosity + (c));
        sat = gld(s.saturation + (c));
        den = gld(s.den_kg + (c));
        vol = gld(s.volume + (c));
        temp = s.temperature ? gld(s.temperature + (c)) : 25.0;
        ln_aw = s.ln_act_h2o ? gld(s.ln_act_h2o + (c)) : 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = r * L + t;
          if (i < naq) {
            x[r] = gld(s.pri_molal + (i * ld + c));
            total[r] = gld(s.total + (i * ld + c));
            lngam[r] = log(gld(s.pri_act_coef + (i * ld + c)));
          } else if (i < N) {
            x[r] = gld(s.immobile + ((i - naq) * ld + c));
            total[r] = x[r];
          }
        }
        sec_I = sec_sum = 0.0;
        for (int k = t; k < nc; k += L) {
          if (!cplx_on(k)) continue;
          lng[k] = log(gld(s.sec_act_coef + (k * ld + c)));
          const double m_k = gld(s.sec_molal + (k * ld + c));
          sec_I += m_k * gld(n.cplx_Z + k) * gld(n.cplx_Z + k);
          sec_sum += m_k;
        }
        for (int m = t; m < nm; m += L) {
          vf[m] = gld(s.mnrl_volfrac + (m * ld + c));
          // an inactive mineral has no rate (and is not written back)
          area[m] = mnrl_on(m) ? gld(s.mnrl_area + (m * ld + c)) : 0.0;
          rate[m] = 0.0;
        }
        if constexpr (SD != 0) {
          // per-cell SomDec stoichiometry starts from the Setup values
          const int nr = n.sd.nrxn;
          for (int r = t; r < nr; r += L) {
            sdst[r] = gld(n.sd.upstream_nc + r);
            sdst[nr + r] = gld(n.sd.mineral_c_stoich + r);
            sdst[2 * nr + r] = gld(n.sd.mineral_n_stoich + r);
            for (int j = 0; j < kSdDown; ++j)
              sdst[3 * nr + r * kSdDown + j] = gld(n.sd.downstream_nc + r * kSdDown + j);
          }
          sd_ft = s.somdec_f_t ? gld(s.somdec_f_t + c) : 1.0;
          sd_fw = s.somdec_f_w ? gld(s.somdec_f_w + c) : 1.0;
        }
        nsub = nit_total = status = 0;
        ncuts = nconst = attempts = last_why = 0;
        if (!n.full_geochem) {
          // RStep early exit (reaction.F90:3601-3605)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = r * L + t;
            if (i < naq) gst(s.pri_molal + (i * ld + c), total[r] / den * 1.0e3);
          }
          if (t == 0) {
            if (s.num_iterations) gst(s.num_iterations + (c), 0);
            if (s.num_sub_steps) gst(s.num_sub_steps + (c), 0);
            if (s.status) gst(s.status + (c), 0);
          }
          cell = -1;
        } else {
          // RStep set-up: clamp small totals (reaction.F90:3617-3631)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = r * L + t;
            small[r] = false;
            if (i < N && total[r] <= 1.0e-40) {
              small[r] = true;
              small_val[r] = total[r];
              total[r] = 1.0e-40;
            }
            guess[r] = x[r];
            if (n.total_as_guess && i < naq) guess[r] = total[r];
          }
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

    const int64_t c = cell;
    const bool wet = !(sat < kRtMinSaturation);
    const double psv = por * sat * 1000.0 * vol;
    const double dkl = den * 1.0e-3;

    // ---------------- RReact set-up (reaction.F90:3800-3851) ----------------
    if (nit < 0) {
      if (++attempts > n.max_calls) {
        // bound on RReact calls per RStep: fail the cell
        status = last_why ? last_why : kMaxIter;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = r * L + t;
          if (i < naq) {
            gst(s.pri_molal + (i * ld + c), x[r]);
            gst(s.total + (i * ld + c), total[r]);
          } else if (i < N) {
            gst(s.immobile + ((i - naq) * ld + c), x[r]);
          }
        }
        if (t == 0) {
          if (s.num_iterations) gst(s.num_iterations + (c), nit_total);
          if (s.num_sub_steps) gst(s.num_sub_steps + (c), nsub);
          if (s.status) gst(s.status + (c), status);
        }
        cell = -1;
        continue;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = r * L + t;
        fixed[r] = 0.0;
        if (wet && i < N) fixed[r] = i < naq ? psv * total[r] : total[r] * vol;
        init_tot[r] = total[r];
        x[r] = guess[r];
        if (i >= naq && i < N) total[r] = x[r];  // rt_auxvar%immobile = guess
      }
      nit = 0;
      r0 = 0.0;
    }

    // ---------------- one Newton iteration ----------------
    ++nit;
    const double psvd = psv / tdt;
    double rs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = r * L + t;
      rs[r] = 1.0;
      if (wet && i < N) rs[r] = i < naq ? 1.0 / (psvd * dkl * x[r]) : 1.0 / (vol / tdt * x[r]);
    }
    if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER && naq > 0) {
      // RActivityCoefficients LAG at the current pri / last sec molalities
      double Ip = 0.0, sp_ = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = r * L + t;
        if (i < naq) {
          Ip += x[r] * gld(n.Zp + i) * gld(n.Zp + i);
          if (i != n.h2o_aq_id) sp_ += x[r];
        }
      }
      const double I = 0.5 * team_sum<L>(Ip + sec_I);
      const double sI = sqrt(I);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = r * L + t;
        if (i < naq)
          lngam[r] = fabs(gld(n.Zp + i)) > 1.0e-10 ? log(dh_gamma(gld(n.Zp + i), gld(n.a0p + i), sI, I, n)) : 0.0;
      }
      for (int k = t; k < nc; k += L)
        if (cplx_on(k))
          lng[k] = fabs(gld(n.cplx_Z + k)) > 1.0e-10
                     ? log(dh_gamma(gld(n.cplx_Z + k), gld(n.cplx_a0 + k), sI, I, n)) : 0.0;
      if (n.act_h2o) {
        const double a = 1.0 - 0.017 * team_sum<L>(sp_ + sec_sum);
        ln_aw = a > 0.0 ? log(a) : 0.0;
      }
    }
    // phase A: per-species quantities into LDS; zero accumulators and J rows
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = r * L + t;
      if (i < N) {
        const double lx = log(x[r]);
        lnact[i] = i < naq ? lx + lngam[r] : lx;
        rsv[i] = rs[r];
        xv[i] = x[r];
        tot[i] = 0.0;
        racc[i] = 0.0;
      }
#pragma unroll
      for (int j = 0; j < NM; ++j) Jt[i * (NM + 1) + j] = 0.0;
    }
    if constexpr (SD >= 2)
      for (int q = t; q < kSdPairs; q += L) sdS[q] = 0.0;
    team_sync();
    // phases B and C: complexes (RTotalAqueous) and kinetic minerals
    // (RKineticMineral) scattered into the LDS tile.  LDS operations of a
    // wave complete in order, so an LDS read issued after a batch of atomics
    // waits for all of them, and a dependent chain of network loads inside
    // the scatter loops would pay the global-memory latency per element.
    // Each reaction row (ids, stoichiometry) is therefore staged in
    // registers with one batch of global loads, every LDS read (ln
    // activities, row scales, mineral state) is issued before the first
    // atomic of its chunk, and the atomics then run from registers only.
    sec_I = sec_sum = 0.0;
    const double jc = psvd * dkl;
    const int cms = n.cplx_ms, mms = n.mnrl_ms;
    // minerals m = q*L + t (q < KM): rates and scatter weights (reads only)
    RxnRow<kRowSlots> mrow[KM];
    double mv[KM], mg[KM], mkr[KM];
    bool mon[KM];
#pragma unroll
    for (int q = 0; q < KM; ++q) {
      const int m = q * L + t;
      mv[q] = mg[q] = mkr[q] = 0.0;
      mon[q] = false;
      if (m < nm) {
        const double logK = n.isothermal || !n.has_mnrl_logKcoef
                                ? gld(n.mnrl_logK + m)
                                : interp_logk(n.mnrl_logKcoef + 5 * m, temp);
        double lnQK0 = -logK * kLogToLn;
        if (n.has_mnrl_h2o) lnQK0 += gld(n.mnrl_h2o + m) * ln_aw;
        const double lnQK =
            mrow[q].load(n.mnrl_ids + m * (mms + 1), n.mnrl_st + m * mms, lnact, rsv, lnQK0);
        double rt = 0.0;
        if (area[m] != 0.0) {  // zero-area minerals give Im = 0
          const MineralRate mr = mineral_rate(n, m, lnQK, vf[m], area[m], temp, lnact);
          if (mr.active) {
            rt = mr.Im;
            // dry cells: RReaction adds nothing (reaction.F90:4083), but the
            // kinetic-state update still uses this rate
            mon[q] = wet;
            mv[q] = mr.Im * vol;
            mkr[q] = mr.kr;
            // log form: d res_i / d ln m_j = nu_i dIm/dQK nu_j QK * rho/1000
            mg[q] = mr.dIm_dQK * vol * mr.QK * dkl;
          }
        }
        rate[m] = rt;
      }
    }
    // complexes k = k0 + q*L + t, KC per lane per chunk
    for (int k0 = 0; k0 < nc; k0 += KC * L) {
      RxnRow<kRowSlots> crow[KC];
      double mk[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int k = k0 + q * L + t;
        mk[q] = 0.0;
        if (k < nc && cplx_on(k)) {
          const double logK = n.isothermal || !n.has_cplx_logKcoef
                                  ? gld(n.cplx_logK + k)
                                  : interp_logk(n.cplx_logKcoef + 5 * k, temp);
          double lnQK = -logK * kLogToLn;
          if (n.has_cplx_h2o) lnQK += gld(n.cplx_h2o + k) * ln_aw;
          lnQK = crow[q].load(n.cplx_ids + k * (cms + 1), n.cplx_st + k * cms, lnact, rsv, lnQK);
          mk[q] = exp(lnQK - lng[k]);
          const double z = gld(n.cplx_Z + k);
          sec_I += mk[q] * z * z;
          sec_sum += mk[q];
          if constexpr (SD >= 2)
            sd_pair_sums(n.sd, n.cplx_ids + k * (cms + 1), n.cplx_st + k * cms, mk[q],
                         TeamSdEnv::lds(sdS));
        }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int k = k0 + q * L + t;
        if (k < nc && (!VAR || mk[q] != 0.0))
          crow[q].template scatter<NM>(n.cplx_ids + k * (cms + 1), n.cplx_st + k * cms, mk[q],
                                           mk[q] * jc, wet, tot, Jt, rsv);
      }
    }
#pragma unroll
    for (int q = 0; q < KM; ++q) {
      const int m = q * L + t;
      if (m < nm && mon[q]) {
        mrow[q].template scatter<NM>(n.mnrl_ids + m * (mms + 1), n.mnrl_st + m * mms, mv[q],
                                         mg[q], true, racc, Jt, rsv);
        if (n.has_pref && gld(n.pref_n + m) > 0)
          team_prefactor_jacobian<NM>(n, m, lnact, temp, mv[q] / mkr[q], Jt, rsv);
      }
    }
    // minerals beyond the register slots (more than KM*L of them)
    for (int m = KM * L + t; m < nm; m += L) {
      double rt = 0.0;
      if (area[m] != 0.0) {
        RxnRow<kRowSlots> row;
        const double logK = n.isothermal || !n.has_mnrl_logKcoef
                                ? gld(n.mnrl_logK + m)
                                : interp_logk(n.mnrl_logKcoef + 5 * m, temp);
        double lnQK = -logK * kLogToLn;
        if (n.has_mnrl_h2o) lnQK += gld(n.mnrl_h2o + m) * ln_aw;
        lnQK = row.load(n.mnrl_ids + m * (mms + 1), n.mnrl_st + m * mms, lnact, rsv, lnQK);
        const MineralRate mr = mineral_rate(n, m, lnQK, vf[m], area[m], temp, lnact);
        if (mr.active) {
          rt = mr.Im;
          if (wet) {
            row.template scatter<NM>(n.mnrl_ids + m * (mms + 1), n.mnrl_st + m * mms,
                                         mr.Im * vol, mr.dIm_dQK * vol * mr.QK * dkl, true, racc,
                                         Jt, rsv);
            if (n.has_pref && gld(n.pref_n + m) > 0)
              team_prefactor_jacobian<NM>(n, m, lnact, temp, mr.Im * vol / mr.kr, Jt, rsv);
          }
        }
      }
      rate[m] = rt;
    }
    if (SD != 0 && wet) {
      // SomDec sandbox (SomDecReact, reaction_sandbox_somdec.F90:1504-1910):
      // reactions r = t, t+L, ...; the N2O emission needs the net N
      // mineralization of all of them
      auto run = [&](auto& e) {
        double nmin = 0.0;
        for (int r = t; r < n.sd.nrxn; r += L) nmin += sd_reaction(n.sd, e, r, tdt, true);
        nmin = team_sum<L>(nmin);
        if (t == 0) sd_nemission(n.sd, e, nmin, tdt, true);
      };
      if constexpr (SD >= 2) {
        team_sync();  // the complexes' totals and dtotal pairs first: the sandbox reads them
        TeamSdEnvT<true> e{TeamSdEnv::lds(xv), TeamSdEnv::lds(lnact), TeamSdEnv::lds(rsv),
                           TeamSdEnv::lds(racc), TeamSdEnv::lds(Jt), TeamSdEnv::lds(sdst),
                           TeamSdEnv::lds(tot), naq, n.sd.nrxn, NM + 1, dkl, por * sat, vol,
                           sat, temp, sd_ft, sd_fw, s.somdec_f_t != nullptr,
                           s.somdec_f_w != nullptr, &n.sd, TeamSdEnv::lds(sdS)};
        run(e);
      } else {
        TeamSdEnvT<false> e{TeamSdEnv::lds(xv), TeamSdEnv::lds(lnact), TeamSdEnv::lds(rsv),
                            TeamSdEnv::lds(racc), TeamSdEnv::lds(Jt), TeamSdEnv::lds(sdst),
                            TeamSdEnv::lds(tot), naq, n.sd.nrxn, NM + 1, dkl, por * sat, vol,
                            sat, temp, sd_ft, sd_fw, s.somdec_f_t != nullptr,
                            s.somdec_f_w != nullptr, &n.sd, TeamSdEnv::lds(sdS)};
        run(e);
      }
    }
    team_sync();

    // phase D: residual, totals, rows of the scaled log-form Jacobian
    // (LDS tile + accumulation diagonal (1 after scaling) + decay; identity
    // for a dry cell; columns divided by x_j for the linear formulation)
    JT a[R][NM];
    double res[R];
    auto build_rows = [&]() {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = r * L + t;
#pragma unroll
        for (int j = 0; j < NM; ++j) a[r][j] = i < N ? (JT)Jt[i * (NM + 1) + j] : (JT)0;
        if (i >= N) continue;
        if (wet) {
          double d = 1.0;
          if (i >= naq)
            for (int q = 0; q < n.ndecay; ++q)
              if (gld(n.decay_id + q) == i - naq) d += gld(n.decay_k + q) * vol * x[r] * rs[r];
#pragma unroll
          for (int j = 0; j < NM; ++j)
            if (j == i) a[r][j] += (JT)d;
        } else {
#pragma unroll
          for (int j = 0; j < NM; ++j) a[r][j] = (JT)(j == i ? 1 : 0);
        }
        if (!n.log_form) {
#pragma unroll
          for (int j = 0; j < NM; ++j)
            if (j < N) a[r][j] = (JT)((double)a[r][j] / xv[j]);
        }
      }
    };
    build_rows();
    double mx = 0.0, nr2 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = r * L + t;
      res[r] = 0.0;
      if (i >= N) continue;
      if (i < naq) {
        total[r] = (x[r] + tot[i]) * dkl;
        res[r] = wet ? (psv * total[r] - fixed[r]) / tdt : 0.0;
      } else {
        total[r] = x[r];
        res[r] = wet ? (x[r] * vol - fixed[r]) / tdt : 0.0;
        if (wet)
          for (int q = 0; q < n.ndecay; ++q)
            if (gld(n.decay_id + q) == i - naq) res[r] += gld(n.decay_k + q) * vol * x[r];
      }
      if (wet) res[r] += racc[i];
      mx = fmax(mx, fabs(res[r]));
      nr2 += res[r] * res[r];
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
      double amax[R];
      int zero_row = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        amax[r] = 0.0;
#pragma unroll
        for (int j = 0; j < NM; ++j) amax[r] = fmax(amax[r], (double)fabs(a[r][j]));
        if (r * L + t < N && !(amax[r] > 0.0)) zero_row = 1;
      }
      zero_row = team_or<L>(zero_row);
      double delta[R];
      if (zero_row) {
        outcome = kFailSingular;
      } else {
        JT rhs[R];
#pragma unroll
        for (int r = 0; r < R; ++r) rhs[r] = (r * L + t < N) ? (JT)(res[r] * rs[r]) : (JT)0;
        JT xs[NM];
        // natural-order elimination; pivot of step p is row p (lane p % L,
        // slot p / L)
        constexpr double kPivTol = sizeof(JT) == 4 ? 1.0e-6 : 1.0e-12;
        int bad = 0;
        static_for<0, NM>([&](auto P) {
          constexpr int p = decltype(P)::value;
          constexpr int ow = p % L, sl = p / L;
          if (p < N) {
            const JT pv = bcast<L, ow>(a[sl][p]);
            if (t == ow && !((double)fabs(a[sl][p]) > kPivTol * amax[sl])) bad = 1;
            const JT pinv = (JT)1 / pv;
            JT pr[NM];
            static_for<p + 1, NM>([&](auto Q) {
              constexpr int j = decltype(Q)::value;
              pr[j] = bcast<L, ow>(a[sl][j]);
            });
            const JT prhs = bcast<L, ow>(rhs[sl]);
#pragma unroll
            for (int r = 0; r < R; ++r) {
              if (r * L + t > p) {
                const JT l = a[r][p] * pinv;
                a[r][p] = l;
                static_for<p + 1, NM>([&](auto Q) {
                  constexpr int j = decltype(Q)::value;
                  a[r][j] -= l * pr[j];
                });
                rhs[r] -= l * prhs;
              }
            }
            // the diagonal of row p is not read again by the elimination:
            // keep its reciprocal for the back substitution
            if (t == ow) a[sl][p] = pinv;
          }
        });
        bad = team_or<L>(bad);
        if (!bad) {
          // back substitution, rows in natural order
          static_for_down<0, NM>([&](auto P) {
            constexpr int p = decltype(P)::value;
            constexpr int ow = p % L, sl = p / L;
            xs[p] = (JT)0;
            if (p < N) {
              JT sum = rhs[sl];
              static_for<p + 1, NM>([&](auto Q) {
                constexpr int j = decltype(Q)::value;
                sum -= a[sl][j] * xs[j];
              });
              xs[p] = bcast<L, ow>(sum * a[sl][p]);
            }
          });
          // Large systems in mixed precision: one step of iterative
          // refinement.  The residual of the scaled system at the FP32
          // solution is formed in FP64 from the FP64 tile (the matrix the
          // factorization started from, as build_rows assembles it), solved
          // with the same FP32 factors, and the correction is added in FP64,
          // so the Newton step keeps FP64 accuracy on ill-conditioned cells.
          JT xc[NM];
          static_for<0, NM>([&](auto P) { xc[decltype(P)::value] = (JT)0; });
          if constexpr (sizeof(JT) == 4 && NM >= 32) {
            JT r2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int i = r * L + t;
              double acc = 0.0;
              if (i < N) {
                double d = 1.0;
                if (i >= naq)
                  for (int q = 0; q < n.ndecay; ++q)
                    if (gld(n.decay_id + q) == i - naq) d += gld(n.decay_k + q) * vol * x[r] * rs[r];
                acc = res[r] * rs[r];
                static_for<0, NM>([&](auto Q) {
                  constexpr int j = decltype(Q)::value;
                  if (j < N) {
                    double aij = wet ? Jt[i * (NM + 1) + j] + (j == i ? d : 0.0) : (j == i ? 1.0 : 0.0);
                    if (!n.log_form) aij /= xv[j];
                    acc -= aij * (double)xs[j];
                  }
                });
              }
              r2[r] = (JT)acc;
            }
            static_for<0, NM>([&](auto P) {
              constexpr int p = decltype(P)::value;
              constexpr int ow = p % L, sl = p / L;
              if (p < N) {
                const JT pr = bcast<L, ow>(r2[sl]);
#pragma unroll
                for (int r = 0; r < R; ++r)
                  if (r * L + t > p) r2[r] -= a[r][p] * pr;
              }
            });
            static_for_down<0, NM>([&](auto P) {
              constexpr int p = decltype(P)::value;
              constexpr int ow = p % L, sl = p / L;
              if (p < N) {
                JT sum = r2[sl];
                static_for<p + 1, NM>([&](auto Q) {
                  constexpr int j = decltype(Q)::value;
                  sum -= a[sl][j] * xc[j];
                });
                xc[p] = bcast<L, ow>(sum * a[sl][p]);
              }
            });
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            delta[r] = 0.0;
            static_for<0, NM>([&](auto P) {
              constexpr int p = decltype(P)::value;
              if (p == r * L + t) delta[r] = (double)xs[p] + (double)xc[p];
            });
          }
        } else {
          // pivoting fallback: rows and right-hand side to LDS, one lane
          // runs the Crout factorization with implicit pivoting
          build_rows();
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = r * L + t;
            if (i < N) {
#pragma unroll
              for (int j = 0; j < NM; ++j) Jt[i * (NM + 1) + j] = (double)a[r][j];
              sol[i] = res[r] * rs[r];
            }
          }
          team_sync();
          int serr = 0;
          // racc / tot were consumed by the residual: work arrays of the LU
          if (t == 0) serr = serial_lu_solve<double>(Jt, NM + 1, N, sol, racc, (int*)tot);
          serr = bcast<L, 0>(serr);
          team_sync();
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = r * L + t;
            delta[r] = i < N ? sol[i] : 0.0;
          }
          team_sync();
          if (serr) outcome = kFailSingular;
        }
        if (outcome == kContinue) {
          // ---------------- update (reaction.F90:3977-4049) ----------------
          double latest[R], rc = -1.0;
          if (n.log_form) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              latest[r] = x[r];
              if (r * L + t < N) {
                const double u = copysign(1.0, delta[r]) * fmin(fabs(delta[r]), n.max_dlnC);
                latest[r] = x[r] * exp(-u);
              }
            }
          } else {
            double ratio = 1.0e308;
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (r * L + t < N && x[r] <= delta[r]) ratio = fmin(ratio, fabs(x[r] / delta[r]));
            const double mr = team_min<L>(ratio);
#pragma unroll
            for (int r = 0; r < R; ++r)
              latest[r] = (r * L + t < N) ? x[r] - (mr < 1.0 ? delta[r] * mr * 0.99 : delta[r]) : x[r];
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            if (r * L + t < N) {
              const double q = fabs((latest[r] - x[r]) / x[r]);
              if (q == q) rc = fmax(rc, q);
            }
          }
          const double mrc = team_max<L>(rc);
          // With an FP32 factorization the step is only as accurate as the
          // factorization, so a small-step exit needs the FP64 residual of
          // the iterate to confirm it.
          const bool step_exit_ok = sizeof(JT) == 8 || nr <= kMixedStepExitResidual * r0;
          if (mrc >= 0.0 && mrc < n.tol_change && step_exit_ok) {
            outcome = kConverged;
          } else {
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = latest[r];
          }
        }
      }
    }

    // ---------------- RReact / RStep transitions ----------------
    bool finish = false;
    if (outcome == kConverged) {
      // RUpdateKineticState: mineral volume fractions with the final rates
      for (int m = t; m < nm; m += L) {
        vf[m] += rate[m] * gld(n.mnrl_Vm + m) * tdt;
        if (vf[m] < 0.0) vf[m] = 0.0;
      }
      nit_total += nit;
      cum += tdt;
      ++nsub;
      ++nconst;
#pragma unroll
      for (int r = 0; r < R; ++r) guess[r] = x[r];
      if (nconst >= 4) {
        --ncuts;
        tdt = fmin(2.0 * tdt, target - cum);
      }
      nit = -1;
      if (cum >= target) finish = true;
    } else if (outcome != kContinue && sizeof(JT) == 4 && escalate) {
      // The FP32 factorization could not bring this cell to convergence:
      // leave its inputs untouched and hand it to the FP64 pass.
      if (t == 0) gst(s.status + (c), kEscalate);
      cell = -1;
      continue;
    } else if (outcome != kContinue) {
      nit_total += nit;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = r * L + t;
        if (outcome == kFailMaxIt) {
          total[r] = init_tot[r];  // totals / immobile restored
          if (i >= naq) x[r] = init_tot[r];
        }
      }
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
      const bool ok = status == 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = r * L + t;
        if (i < naq) {
          gst(s.pri_molal + (i * ld + c), x[r]);
          gst(s.total + (i * ld + c), (ok && small[r]) ? small_val[r] : total[r]);
          if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
            gst(s.pri_act_coef + (i * ld + c), exp(lngam[r]));
        } else if (i < N) {
          gst(s.immobile + ((i - naq) * ld + c), (ok && small[r]) ? small_val[r] : x[r]);
        }
      }
      // secondary molalities at the final free molalities (the last
      // RTAuxVarCompute, reaction.F90:4052): recomputed here instead of
      // being kept per complex in LDS
      for (int k = t; k < nc; k += L) {
        if (!cplx_on(k)) continue;
        const double lnQK = complex_lnqk(n, k, lnact, ln_aw, temp);
        gst(s.sec_molal + (k * ld + c), exp(lnQK - lng[k]));
        if (n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
          gst(s.sec_act_coef + (k * ld + c), exp(lng[k]));
      }
      for (int m = t; m < nm; m += L) {
        if (!mnrl_on(m)) continue;
        gst(s.mnrl_volfrac + (m * ld + c), vf[m]);
        if (s.mnrl_rate) gst(s.mnrl_rate + (m * ld + c), rate[m]);
      }
      if (t == 0) {
        if (s.ln_act_h2o && n.act_h2o && n.act_freq == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
          gst(s.ln_act_h2o + (c), ln_aw);
        if (s.num_iterations) gst(s.num_iterations + (c), nit_total);
        if (s.num_sub_steps) gst(s.num_sub_steps + (c), nsub);
        if (s.status) gst(s.status + (c), status);
      }
      team_sync();
      cell = -1;
    }
  }
}

__global__ void zero_cursors(unsigned long long* c, int n) {
  if ((int)threadIdx.x < n) c[16 * threadIdx.x] = 0ull;
}

template <int L, int R, typename JT, int SD, bool VAR>
hipError_t launch_LR(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                     double dt, const LaunchCfg& cfg, unsigned long long* cursors,
                     int num_cus, int nxcd, hipStream_t stream, int rescue_only) {
  const TeamLds lay = team_lds_layout(L * R, n.nc, n.nm, 8, sd_lds_doubles(n.sd.nrxn, SD));
  const int block = cfg.block_size;
  const size_t lds = (size_t)(block / L) * lay.bytes;
  int64_t blocks = (int64_t)num_cus * cfg.blocks_per_cu;
  const int64_t max_blocks = (ncell + (block / L) - 1) / (block / L);
  if (blocks > max_blocks) blocks = max_blocks;
  // a multiple of the XCD count keeps the per-XCD load even
  if (blocks > nxcd) blocks -= blocks % nxcd;
  if (blocks < 1) blocks = 1;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&team_step_kernel<L, R, JT, SD, VAR>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  zero_cursors<<<1, 64, 0, stream>>>(cursors, nxcd);
  team_step_kernel<L, R, JT, SD, VAR><<<dim3((unsigned)blocks), dim3(block), lds, stream>>>(
      dn, ds, ncell, dt, cursors, nxcd, rescue_only);
  return hipGetLastError();
}

// mixed = 0: FP64 variant; 1: mixed variant (FP32 pass, then an FP64 pass
// over the cells the FP32 pass handed over); kTeamRescue64: only the FP64
// pass over cells an earlier pass handed over (status kEscalate)
template <int L, int R, int SD, bool VAR>
hipError_t launch_variant(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  if (mixed == kTeamRescue64) {
    LaunchCfg c64 = cfg;
    while (c64.block_size > 64 && !team_supports(n, c64, false)) c64.block_size /= 2;
    return launch_LR<L, R, double, SD, VAR>(n, dn, ds, ncell, dt, c64, cursors, num_cus, nxcd,
                                            stream, 1);
  }
  if (!mixed)
    return launch_LR<L, R, double, SD, VAR>(n, dn, ds, ncell, dt, cfg, cursors, num_cus, nxcd,
                                            stream, 0);
  hipError_t e = launch_LR<L, R, float, SD, VAR>(n, dn, ds, ncell, dt, cfg, cursors, num_cus,
                                                 nxcd, stream, 0);
  if (e != hipSuccess || !has_status) return e;
  LaunchCfg c64 = cfg;
  while (c64.block_size > 64 && !team_supports(n, c64, false)) c64.block_size /= 2;
  return launch_LR<L, R, double, SD, VAR>(n, dn, ds, ncell, dt, c64, cursors, num_cus, nxcd,
                                          stream, 1);
}

template <int L, int R>
hipError_t launch_pair(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                       double dt, int mixed, const LaunchCfg& cfg,
                       unsigned long long* cursors, int num_cus, int nxcd,
                       hipStream_t stream, bool has_status) {
  if (n.variable) {
    if constexpr (team_has_variable(L, R))
      return launch_variant<L, R, 0, true>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus,
                                           nxcd, stream, has_status);
    return hipErrorInvalidValue;
  }
  if (!n.has_somdec)
    return launch_variant<L, R, 0, false>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus,
                                          nxcd, stream, has_status);
  // the sandbox: level 2 with the Monod oxygen response, O2 tracking or
  // complexes, level 1 (a smaller kernel) without them
  if constexpr (team_has_somdec(L, R)) {
    if (n.sd.ext)
      return launch_variant<L, R, 2, false>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus,
                                            nxcd, stream, has_status);
    return launch_variant<L, R, 1, false>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus,
                                          nxcd, stream, has_status);
  }
  return hipErrorInvalidValue;
}

}  // namespace
}  // namespace pfchem
