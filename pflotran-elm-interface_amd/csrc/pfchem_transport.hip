// pfchem_transport.hip -- OSRT transport step on the device: the implicit
// advection-diffusion system of RTCalculateTransportMatrix
// (reactive_transport.F90:1469-1703) with the accumulation and boundary
// right-hand side of RTCalculateRHS_t1 (:1239-1465) and the TFluxCoef
// coefficients (transport.F90:495-565), solved for every aqueous component
// together (one BiCGSTAB per component with a Jacobi preconditioner, the
// components' vectors side by side) where the reference calls KSPSolve once
// per component (pmc_subsurface_osrt.F90:286-333).
//
// Layout.  The off-diagonal pattern is stored as sliced ELL with slices of
// 64 rows, one slice per wavefront: entry (slot j, row 64 s + l) at
// slice_ptr[s] + 64 j + l, so every load of a slot is one coalesced 512-byte
// (values) or 256-byte (int32 columns) access and the slice width is
// uniform in the wave.  Each entry keeps its column and its connection id
// with the row's role in the sign (id + 1: the row is the connection's
// upstream cell, -(id + 1): downstream, 0: padding), so the values of a step
// are formed row by row from the connection coefficients without atomics.
// The solver's vectors hold the components in pairs, [pair][cell] of
// double2 (component k is element k & 1 of pair k >> 1; an odd count has a
// zero partner), so one 16-byte gather serves two components.
//
// Solver.  The Jacobi preconditioner is folded into the matrix as a column
// scaling, A = M D^-1 (unit diagonal), and BiCGSTAB iterates y = D x on
// A y = b: the iterates of right-preconditioned BiCGSTAB on M x = b, without
// the preconditioned vectors.  The shadow residual is a hash of (cell,
// component) recomputed where it is used.  Dot products are summed per
// block (wave shuffles, then LDS) inside the kernels that produce their
// operands; one block per component sums the block partials in a fixed
// order and updates that component's scalars on the device, so the host
// waits once per iteration (for the components' active flags).  The
// gathering kernels place consecutive blocks on one XCD (xcd_block).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "pfchem.h"
#include "pfchem_device.hpp"
#include "pfchem_transport.h"

namespace pfchem {

struct TransportPlan {
  int64_t ncell = 0, nconn = 0, nbc = 0, nslice = 0, nent = 0;  // nent: padded entries
  int naq = 0;
  int wmax = 0;  // widest slice (entries per row)
  // device: slice_ptr[nslice + 1], bc_ptr[ncell + 1] (int64); col and cid
  // [nent + kEntryPad], bc_id[nbc] (int32)
  void* topo = nullptr;
  const int64_t* slice_ptr = nullptr;
  const int64_t* bc_ptr = nullptr;
  const int32_t* col = nullptr;
  const int32_t* cid = nullptr;
  const int32_t* bc_id = nullptr;
  // work vectors, wave partials, scalars and flags (grown on demand)
  double* work = nullptr;
  size_t work_bytes = 0;
  int32_t* host_flags = nullptr;  // pinned: active and breakdown flags per component
};

void transport_plan_free(TransportPlan* p) {
  if (!p) return;
  if (p->topo) (void)hipFree(p->topo);
  if (p->work) (void)hipFree(p->work);
  if (p->host_flags) (void)hipHostFree(p->host_flags);
  delete p;
}

namespace {

constexpr int kTB = 256;      // threads per block: 4 slices
constexpr int kXcd = 8;       // XCDs of the device
// Slots held in registers: the kernels take KW in {2, 4, 6, 8}, the
// smallest that holds the widest slice (wider rows re-read the rest).  The
// KW register slots are gathered without a test on the slice width (a slot
// past the row's width holds the cell itself with value 0), so the gathers
// of a component chunk are issued together instead of one slot at a time.
constexpr int kMaxKW = 8;
constexpr int64_t kEntryPad = (int64_t)kMaxKW * 64;  // entries past the last slice
constexpr int kFinTB = 512;   // threads of the per-component finishing blocks
constexpr int kSc = 8;        // scalars per component
constexpr int kMaxNaq = 512;  // components (LDS of the block partials: 3 x 4 x 8 B each)
constexpr int kRho = 0, kRhoPrev = 1, kAlpha = 2, kOmega = 3, kBeta = 4, kBnorm = 5;

struct Coefs {
  const double *area, *vel, *cod, *bc_area, *bc_vel, *bc_cod, *bc_total;
};

// TFluxCoef (transport.F90:495-565): coefficients of the upstream and
// downstream totals in the flux out of the upstream cell (mol/s per mol/L)
__device__ __forceinline__ void flux_coef(double q, double cod, double area, double& t_up,
                                          double& t_dn) {
  if (q > 0.0) {
    t_up = (cod + q) * area * 1000.0;
    t_dn = -cod * area * 1000.0;
  } else {
    t_up = cod * area * 1000.0;
    t_dn = (-cod + q) * area * 1000.0;
  }
}

// One stored entry of row i: its off-diagonal value, with its part of the
// diagonal added to d.  rr: connection id + 1 with the row's role in the
// sign, 0 for padding; c: the other cell.  Interior connections check both
// saturations (check_upwind_saturation).
__device__ __forceinline__ double entry_value(int rr, int c, double si,
                                              const double* __restrict__ sat, const Coefs& cf,
                                              double& d) {
  if (rr == 0) return 0.0;
  const int64_t q = rr > 0 ? rr - 1 : -rr - 1;
  const double sj = sat[c];
  const double s_up = rr > 0 ? si : sj, s_dn = rr > 0 ? sj : si;
  double t_up = 0.0, t_dn = 0.0;
  if (!(s_dn < 1.0e-40) && !(s_up < 1.0e-40)) flux_coef(cf.vel[q], cf.cod[q], cf.area[q], t_up, t_dn);
  if (rr > 0) {  // this row is upstream: + T_up on the diagonal, + T_dn off it
    d += t_up;
    return t_dn;
  }
  d -= t_dn;  // downstream: - T_dn on the diagonal, - T_up off it
  return -t_up;
}

// the shadow residual: a fixed pseudo-random vector in [0.5, 1.5) (the
// choice rh = r0 breaks down on upwind advection, whose preconditioned
// matrix is the identity plus a nilpotent part: the residual moves
// downstream, orthogonal to r0)
__device__ __forceinline__ double shadow(int64_t i, int k) {
  const unsigned h = (unsigned)((unsigned long long)i * 2654435761ull +
                                (unsigned long long)k * 40503ull);
  return 0.5 + (double)((h >> 8) & 0xffffu) / 65536.0;
}

// Block index with consecutive blocks on one XCD: the hardware hands block
// b to XCD b % 8, so XCD x takes the x-th contiguous run of blocks and the
// neighbours a row gathers (cells +-1, +-nx, +-nx ny) are mostly its own
// blocks' cells, in its own L2.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t g = gridDim.x, b = blockIdx.x;
  const int64_t per = g / kXcd, rem = g % kXcd, x = b % kXcd, j = b / kXcd;
  return x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
}

// Components go through the kernels kKC at a time, so a wave has the
// loads of kKC components in flight at once and interleaves their
// reductions (one component at a time leaves every pass latency-bound).
// on[u]: component k0 + u exists and is active; comp(): the component a
// lane reads for slot u (k0 for an absent one: valid memory, result unused).
constexpr int kKC = 4;

__device__ __forceinline__ bool chunk_on(const int32_t* act, int naq, int k0, bool (&on)[kKC]) {
  bool any = false;
#pragma unroll
  for (int u = 0; u < kKC; ++u) {
    const int k = k0 + u;
    on[u] = k < naq && (act == nullptr || act[k] != 0);
    any = any || on[u];
  }
  return any;
}

__device__ __forceinline__ int comp(const bool (&on)[kKC], int k0, int u) {
  return on[u] ? k0 + u : k0;
}

// wave sums of M values, their shuffle chains interleaved (every lane of
// the wave takes part)
template <int M>
__device__ __forceinline__ void wave_sums(double (&v)[M]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int u = 0; u < M; ++u) v[u] += __shfl_xor(v[u], o, 64);
}

// lane 0's wave sums of NQ products of the chunk's active components into
// the block's LDS slots: pr[q kKC + u] -> lds[((q naq + k0 + u) 4 + wave]
template <int NQ>
__device__ __forceinline__ void chunk_to_lds(double* lds, int naq, int k0, const bool (&on)[kKC],
                                             int lane, const double (&pr)[NQ * kKC]) {
  if (lane != 0) return;
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int u = 0; u < kKC; ++u)
      if (on[u]) lds[((size_t)q * naq + k0 + u) * 4 + (threadIdx.x >> 6)] = pr[q * kKC + u];
}

// the block's partials of products 0..nq-1 of the active components (after
// every wave wrote its slots): part[(q naq + k) nblk + blk]
__device__ __forceinline__ void block_partials(const double* lds, double* part, int nq, int naq,
                                               const int32_t* act, int64_t blk) {
  __syncthreads();
  const int64_t nblk = gridDim.x;
  for (int t = threadIdx.x; t < nq * naq; t += kTB) {
    const int k = t % naq;
    if (act && !act[k]) continue;
    const double* l = lds + (size_t)t * 4;
    part[(size_t)t * nblk + blk] = (l[0] + l[1]) + (l[2] + l[3]);
  }
}

// Diagonal, off-diagonal values (unscaled), b, the initial iterate
// y0 = D x0 (x0: the totals) and r0 = b - M x0, with the block partials of
// (b, b), (r0, r0) and (rh, r0).
template <int KW>
__global__ __launch_bounds__(kTB) void k_assemble(
    int64_t n, int naq, int64_t nslice, const int64_t* __restrict__ sptr,
    const int32_t* __restrict__ col, const int32_t* __restrict__ cid,
    const int64_t* __restrict__ bc_ptr, const int32_t* __restrict__ bc_id, int64_t nbc, Coefs cf,
    const double* __restrict__ por, const double* __restrict__ sat,
    const double* __restrict__ vol, const double* __restrict__ total, int64_t ld, double dt,
    double* __restrict__ diag, double* __restrict__ val, double* __restrict__ y,
    double* __restrict__ r, double* __restrict__ part) {
  extern __shared__ double lds[];
  const int64_t blk = xcd_block();
  const int64_t i = blk * kTB + threadIdx.x;
  const int64_t sl = i >> 6;
  const int lane = threadIdx.x & 63;
  // waves past the last slice (i >= n) take part in the reductions only
  const bool ok = i < n;
  const int64_t e0 = sl < nslice ? sptr[sl] : 0;
  const int w = sl < nslice ? (int)((sptr[sl + 1] - e0) >> 6) : 0;
  double si = 0.0, acc = 0.0, d = 1.0;
  double mv[KW];
  int cv[KW];
#pragma unroll
  for (int j = 0; j < KW; ++j) {
    mv[j] = 0.0;
    cv[j] = ok ? (int)i : 0;
  }
  bool dry = true;
  if (ok) {
    si = sat[i];
    dry = si < 1.0e-40;
    acc = por[i] * si * 1000.0 * vol[i] / dt;
    d = acc;
#pragma unroll
    for (int j = 0; j < KW; ++j)
      if (j < w) {
        const int64_t e = e0 + (int64_t)j * 64 + lane;
        cv[j] = col[e];
        mv[j] = entry_value(cid[e], cv[j], si, sat, cf, d);
        val[e] = mv[j];
      }
    for (int j = KW; j < w; ++j) {
      const int64_t e = e0 + (int64_t)j * 64 + lane;
      val[e] = entry_value(cid[e], col[e], si, sat, cf, d);
    }
    // boundary terms on the diagonal (the cell is the downstream side)
    if (!dry)
      for (int64_t e = bc_ptr[i]; e < bc_ptr[i + 1]; ++e) {
        const int q = bc_id[e];
        double t_up, t_dn;
        flux_coef(cf.bc_vel[q], cf.bc_cod[q], cf.bc_area[q], t_up, t_dn);
        d -= t_dn;
      }
    // a dry cell (no accumulation, no fluxes) has an empty row in the
    // reference's matrix; here it keeps its totals
    if (dry) d = 1.0;
    diag[i] = d;
  }
  for (int k0 = 0; k0 < naq; k0 += kKC) {
    bool on[kKC];
    chunk_on(nullptr, naq, k0, on);
    double pr[3 * kKC];
#pragma unroll
    for (int u = 0; u < 3 * kKC; ++u) pr[u] = 0.0;
    if (ok) {
      const double* tk[kKC];
      double x0[kKC], b[kKC], mx[kKC];
#pragma unroll
      for (int u = 0; u < kKC; ++u) {
        tk[u] = total + (size_t)comp(on, k0, u) * ld;
        x0[u] = tk[u][i];
        b[u] = x0[u];
        mx[u] = d * x0[u];
      }
      if (!dry) {
#pragma unroll
        for (int u = 0; u < kKC; ++u) b[u] = acc * x0[u];
        for (int64_t e = bc_ptr[i]; e < bc_ptr[i + 1]; ++e) {
          const int q = bc_id[e];
          double t_up, t_dn;
          flux_coef(cf.bc_vel[q], cf.bc_cod[q], cf.bc_area[q], t_up, t_dn);
#pragma unroll
          for (int u = 0; u < kKC; ++u) b[u] += t_up * cf.bc_total[(size_t)comp(on, k0, u) * nbc + q];
        }
#pragma unroll
        for (int j = 0; j < KW; ++j)
#pragma unroll
          for (int u = 0; u < kKC; ++u) mx[u] += mv[j] * tk[u][cv[j]];
        for (int j = KW; j < w; ++j) {
          const int64_t e = e0 + (int64_t)j * 64 + lane;
          const double m = val[e];
          const int c = col[e];
#pragma unroll
          for (int u = 0; u < kKC; ++u) mx[u] += m * tk[u][c];
        }
      }
#pragma unroll
      for (int u = 0; u < kKC; ++u) {
        const int k = k0 + u;
        const size_t jv = ((size_t)(k >> 1) * n + i) * 2 + (k & 1);  // pair layout
        if (on[u]) {
          const double res = b[u] - mx[u];
          y[jv] = d * x0[u];
          r[jv] = res;
          pr[u] = b[u] * b[u];
          pr[kKC + u] = res * res;
          pr[2 * kKC + u] = shadow(i, k) * res;
        } else if (k == naq && (naq & 1)) {
          y[jv] = 0.0;  // the zero partner of an odd component count
          r[jv] = 0.0;
        }
      }
    }
    wave_sums<3 * kKC>(pr);
    chunk_to_lds<3>(lds, naq, k0, on, lane, pr);
  }
  block_partials(lds, part, 3, naq, nullptr, blk);
}

// A = M D^-1: every stored value over its column's diagonal
__global__ __launch_bounds__(kTB) void k_scale(int64_t nent, const int32_t* __restrict__ col,
                                               const double* __restrict__ diag,
                                               double* __restrict__ val) {
  const int64_t e = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (e < nent) val[e] = val[e] / diag[col[e]];
}

// p = r (first iteration) or r + beta (p - omega v); values of inactive
// components are computed with zero scalars and never read
__global__ __launch_bounds__(kTB) void k_p(int64_t n, int naq, int first,
                                           const int32_t* __restrict__ act,
                                           const double* __restrict__ sc,
                                           const double* __restrict__ r, double* __restrict__ p,
                                           const double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const int npair = (naq + 1) >> 1;
  const double2* r2 = (const double2*)r;
  const double2* v2 = (const double2*)v;
  double2* p2 = (double2*)p;
  for (int k0 = 0; k0 < naq; k0 += kKC) {
    bool on[kKC];
    if (!chunk_on(act, naq, k0, on)) continue;  // uniform
    const int P0 = k0 >> 1;
    const bool has1 = P0 + 1 < npair;
    const size_t j0 = (size_t)P0 * n + i, j1 = (size_t)(has1 ? P0 + 1 : P0) * n + i;
    const double2 ra = r2[j0], rb = r2[j1];
    if (first) {
      p2[j0] = ra;
      if (has1) p2[j1] = rb;
      continue;
    }
    const double2 pa = p2[j0], pb = p2[j1], va = v2[j0], vb = v2[j1];
    double be[kKC], om[kKC];
#pragma unroll
    for (int u = 0; u < kKC; ++u) {
      be[u] = on[u] ? sc[(k0 + u) * kSc + kBeta] : 0.0;
      om[u] = on[u] ? sc[(k0 + u) * kSc + kOmega] : 0.0;
    }
    p2[j0] = make_double2(ra.x + be[0] * (pa.x - om[0] * va.x), ra.y + be[1] * (pa.y - om[1] * va.y));
    if (has1)
      p2[j1] = make_double2(rb.x + be[2] * (pb.x - om[2] * vb.x), rb.y + be[3] * (pb.y - om[3] * vb.y));
  }
}

// yv = A xv for the active components, with the block partials of
// MODE 0: (rh, yv);  MODE 1: (yv, xv) and (yv, yv).  Each gather loads a
// pair of components (16 bytes).
template <int MODE, int KW>
__global__ __launch_bounds__(kTB) void k_spmv(int64_t n, int naq, int64_t nslice,
                                              const int32_t* __restrict__ act,
                                              const int64_t* __restrict__ sptr,
                                              const int32_t* __restrict__ col,
                                              const double* __restrict__ val,
                                              const double* __restrict__ xv,
                                              double* __restrict__ yv, double* __restrict__ part) {
  extern __shared__ double lds[];
  const int64_t blk = xcd_block();
  const int64_t i = blk * kTB + threadIdx.x;
  const int64_t sl = i >> 6;
  const int lane = threadIdx.x & 63;
  const bool ok = i < n;
  const int64_t e0 = sl < nslice ? sptr[sl] : 0;
  const int w = sl < nslice ? (int)((sptr[sl + 1] - e0) >> 6) : 0;
  const int npair = (naq + 1) >> 1;
  const double2* x2 = (const double2*)xv;
  double2* y2 = (double2*)yv;
  double av[KW];
  int cv[KW];
  // the entry arrays hold kMaxKW x 64 entries past the last slice, so the
  // KW loads stay in bounds; slots past the width are replaced
#pragma unroll
  for (int j = 0; j < KW; ++j) {
    const int64_t e = e0 + (int64_t)j * 64 + lane;
    const double vv = val[e];
    const int cc = col[e];
    av[j] = j < w ? vv : 0.0;
    cv[j] = j < w ? cc : (ok ? (int)i : 0);
  }
  constexpr int NQ = MODE == 0 ? 1 : 2;
  for (int k0 = 0; k0 < naq; k0 += kKC) {
    bool on[kKC];
    if (!chunk_on(act, naq, k0, on)) continue;  // uniform
    double pr[NQ * kKC];
#pragma unroll
    for (int u = 0; u < NQ * kKC; ++u) pr[u] = 0.0;
    if (ok) {
      const int P0 = k0 >> 1;
      const bool has1 = P0 + 1 < npair;
      const double2* xa = x2 + (size_t)P0 * n;
      const double2* xb = x2 + (size_t)(has1 ? P0 + 1 : P0) * n;
      const double2 xia = xa[i], xib = xb[i];
      double a[kKC] = {xia.x, xia.y, xib.x, xib.y};
      const double xi[kKC] = {xia.x, xia.y, xib.x, xib.y};
#pragma unroll
      for (int j = 0; j < KW; ++j) {
        const double2 ga = xa[cv[j]], gb = xb[cv[j]];
        a[0] += av[j] * ga.x;
        a[1] += av[j] * ga.y;
        a[2] += av[j] * gb.x;
        a[3] += av[j] * gb.y;
      }
      for (int j = KW; j < w; ++j) {
        const int64_t e = e0 + (int64_t)j * 64 + lane;
        const double m = val[e];
        const int c = col[e];
        const double2 ga = xa[c], gb = xb[c];
        a[0] += m * ga.x;
        a[1] += m * ga.y;
        a[2] += m * gb.x;
        a[3] += m * gb.y;
      }
      y2[(size_t)P0 * n + i] = make_double2(a[0], a[1]);
      if (has1) y2[(size_t)(P0 + 1) * n + i] = make_double2(a[2], a[3]);
#pragma unroll
      for (int u = 0; u < kKC; ++u)
        if (on[u]) {
          if (MODE == 0) {
            pr[u] = shadow(i, k0 + u) * a[u];
          } else {
            pr[u] = a[u] * xi[u];
            pr[(NQ - 1) * kKC + u] = a[u] * a[u];
          }
        }
    }
    wave_sums<NQ * kKC>(pr);
    chunk_to_lds<NQ>(lds, naq, k0, on, lane, pr);
  }
  block_partials(lds, part, NQ, naq, act, blk);
}

// s = r - alpha v (inactive components: values never read)
__global__ __launch_bounds__(kTB) void k_s(int64_t n, int naq, const int32_t* __restrict__ act,
                                           const double* __restrict__ sc,
                                           const double* __restrict__ r,
                                           const double* __restrict__ v, double* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const int npair = (naq + 1) >> 1;
  const double2* r2 = (const double2*)r;
  const double2* v2 = (const double2*)v;
  double2* s2 = (double2*)s;
  for (int k0 = 0; k0 < naq; k0 += kKC) {
    bool on[kKC];
    if (!chunk_on(act, naq, k0, on)) continue;  // uniform
    const int P0 = k0 >> 1;
    const bool has1 = P0 + 1 < npair;
    const size_t j0 = (size_t)P0 * n + i, j1 = (size_t)(has1 ? P0 + 1 : P0) * n + i;
    const double2 ra = r2[j0], rb = r2[j1], va = v2[j0], vb = v2[j1];
    double al[kKC];
#pragma unroll
    for (int u = 0; u < kKC; ++u) al[u] = on[u] ? sc[(k0 + u) * kSc + kAlpha] : 0.0;
    s2[j0] = make_double2(ra.x - al[0] * va.x, ra.y - al[1] * va.y);
    if (has1) s2[j1] = make_double2(rb.x - al[2] * vb.x, rb.y - al[3] * vb.y);
  }
}

// y += alpha p + omega s (active components; the others keep their
// iterate); r = s - omega t; block partials of (r, r), (rh, r)
__global__ __launch_bounds__(kTB) void k_x(int64_t n, int naq,
                                           const int32_t* __restrict__ act,
                                           const double* __restrict__ sc,
                                           const double* __restrict__ p,
                                           const double* __restrict__ s,
                                           const double* __restrict__ t, double* __restrict__ y,
                                           double* __restrict__ r, double* __restrict__ part) {
  extern __shared__ double lds[];
  const int64_t i = xcd_block() * kTB + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool ok = i < n;
  const int npair = (naq + 1) >> 1;
  const double2* p2 = (const double2*)p;
  const double2* s2 = (const double2*)s;
  const double2* t2 = (const double2*)t;
  double2* y2 = (double2*)y;
  double2* r2 = (double2*)r;
  for (int k0 = 0; k0 < naq; k0 += kKC) {
    bool on[kKC];
    if (!chunk_on(act, naq, k0, on)) continue;  // uniform
    double pr[2 * kKC];
#pragma unroll
    for (int u = 0; u < 2 * kKC; ++u) pr[u] = 0.0;
    if (ok) {
      const int P0 = k0 >> 1;
      const bool has1 = P0 + 1 < npair;
      const size_t j0 = (size_t)P0 * n + i, j1 = (size_t)(has1 ? P0 + 1 : P0) * n + i;
      const double2 pa = p2[j0], pb = p2[j1], sa = s2[j0], sb = s2[j1];
      const double2 ta = t2[j0], tb = t2[j1], ya = y2[j0], yb = y2[j1];
      const double pj[kKC] = {pa.x, pa.y, pb.x, pb.y}, sj[kKC] = {sa.x, sa.y, sb.x, sb.y};
      const double tj[kKC] = {ta.x, ta.y, tb.x, tb.y};
      double yj[kKC] = {ya.x, ya.y, yb.x, yb.y}, rj[kKC];
#pragma unroll
      for (int u = 0; u < kKC; ++u) {
        const int k = k0 + u;
        const double al = on[u] ? sc[k * kSc + kAlpha] : 0.0;
        const double om = on[u] ? sc[k * kSc + kOmega] : 0.0;
        if (on[u]) yj[u] = yj[u] + (al * pj[u] + om * sj[u]);
        rj[u] = sj[u] - om * tj[u];
        if (on[u]) {
          pr[u] = rj[u] * rj[u];
          pr[kKC + u] = shadow(i, k) * rj[u];
        }
      }
      y2[j0] = make_double2(yj[0], yj[1]);
      r2[j0] = make_double2(rj[0], rj[1]);
      if (has1) {
        y2[j1] = make_double2(yj[2], yj[3]);
        r2[j1] = make_double2(rj[2], rj[3]);
      }
    }
    wave_sums<2 * kKC>(pr);
    chunk_to_lds<2>(lds, naq, k0, on, lane, pr);
  }
  block_partials(lds, part, 2, naq, act, xcd_block());
}

// One block per component: the dot products of a stage (the producing
// kernels' block partials summed in a fixed order) and that component's
// scalar update.
// stage 0 (after assembly): (b,b), (r0,r0), (rh,r0); 1: (rh,v);
// 2: (t,s), (t,t); 3: (r,r), (rh,r).  flags[k]: active, flags[naq+k]: a
// breakdown took the component out.
__global__ __launch_bounds__(kFinTB) void k_fin(int stage, int naq, int64_t nblk,
                                                const double* __restrict__ part,
                                                double* __restrict__ sc,
                                                int32_t* __restrict__ flags, double rtol) {
  const int k = blockIdx.x;
  if (stage > 0 && !flags[k]) return;  // uniform in the block
  const int nq = stage == 0 ? 3 : stage == 1 ? 1 : 2;
  __shared__ double red[kFinTB];
  __shared__ double dots[3];
  for (int q = 0; q < nq; ++q) {
    const double* pq = part + ((size_t)q * naq + k) * nblk;
    double s = 0.0;
    for (int64_t j = threadIdx.x; j < nblk; j += kFinTB) s += pq[j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kFinTB / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) dots[q] = red[0];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double* c = sc + k * kSc;
  int32_t* act = flags + k;
  int32_t* brk = flags + naq + k;
  if (stage == 0) {
    c[kBnorm] = sqrt(dots[0]);
    c[kRho] = dots[2];
    c[kRhoPrev] = 1.0;
    c[kAlpha] = 1.0;
    c[kOmega] = 1.0;
    c[kBeta] = 0.0;
    *act = sqrt(dots[1]) > rtol * c[kBnorm] ? 1 : 0;
    *brk = 0;
  } else if (stage == 1) {
    const double rv = dots[0];
    if (rv == 0.0 || !isfinite(rv)) {
      *brk = 1;
      *act = 0;
    } else {
      c[kAlpha] = c[kRho] / rv;
    }
  } else if (stage == 2) {
    c[kOmega] = dots[1] > 0.0 ? dots[0] / dots[1] : 0.0;
  } else {
    c[kRhoPrev] = c[kRho];
    c[kRho] = dots[1];
    const double rn = sqrt(dots[0]);
    if (!(rn > rtol * c[kBnorm])) {
      *act = 0;  // converged
    } else if (c[kOmega] == 0.0 || c[kRho] == 0.0 || !isfinite(rn)) {
      *brk = 1;
      *act = 0;
    } else {
      c[kBeta] = (c[kRho] / c[kRhoPrev]) * (c[kAlpha] / c[kOmega]);
    }
  }
}

// totals of the new time level into the state: x = D^-1 y
__global__ __launch_bounds__(kTB) void k_store(int64_t n, int naq, const double* __restrict__ y,
                                               const double* __restrict__ diag,
                                               double* __restrict__ total, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const double id = 1.0 / diag[i];
  const double2* y2 = (const double2*)y;
  for (int P = 0; 2 * P < naq; ++P) {
    const double2 v = y2[(size_t)P * n + i];
    total[(size_t)(2 * P) * ld + i] = v.x * id;
    if (2 * P + 1 < naq) total[(size_t)(2 * P + 1) * ld + i] = v.y * id;
  }
}

// register slots for a widest slice of wmax entries per row
int slots_for(int wmax) { return wmax <= 2 ? 2 : wmax <= 4 ? 4 : wmax <= 6 ? 6 : 8; }

template <int MODE>
void launch_spmv(int kw, unsigned grid, size_t lds, hipStream_t s, int64_t n, int naq,
                 int64_t ns, const int32_t* act, const int64_t* sptr, const int32_t* col,
                 const double* val, const double* xv, double* yv, double* part) {
  switch (kw) {
    case 2: k_spmv<MODE, 2><<<grid, kTB, lds, s>>>(n, naq, ns, act, sptr, col, val, xv, yv, part); break;
    case 4: k_spmv<MODE, 4><<<grid, kTB, lds, s>>>(n, naq, ns, act, sptr, col, val, xv, yv, part); break;
    case 6: k_spmv<MODE, 6><<<grid, kTB, lds, s>>>(n, naq, ns, act, sptr, col, val, xv, yv, part); break;
    default: k_spmv<MODE, 8><<<grid, kTB, lds, s>>>(n, naq, ns, act, sptr, col, val, xv, yv, part); break;
  }
}

}  // namespace
}  // namespace pfchem

using namespace pfchem;

extern "C" int pfchem_transport_setup(pfchem_ctx* ctx, int64_t ncell,
                                      const pfchem_connections* conn) {
  if (!ctx || !conn || ncell < 0 || conn->nconn < 0 || conn->nbc < 0 ||
      (conn->nconn > 0 && (!conn->id_up || !conn->id_dn)) || (conn->nbc > 0 && !conn->bc_cell))
    return ctx ? ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport setup: bad arguments")
               : PFCHEM_ERR_INVALID;
  if (ncell >= INT32_MAX || conn->nconn >= INT32_MAX - 1 || conn->nbc >= INT32_MAX)
    return ctx_fail(ctx, PFCHEM_ERR_UNSUPPORTED,
                    "transport setup: cells and connections are limited to 2^31 - 2");
  if (ctx_naq(ctx) > kMaxNaq)
    return ctx_fail(ctx, PFCHEM_ERR_UNSUPPORTED, "transport setup: more than 512 components");
  for (int64_t c = 0; c < conn->nconn; ++c)
    if (conn->id_up[c] < 0 || conn->id_up[c] >= ncell || conn->id_dn[c] < 0 ||
        conn->id_dn[c] >= ncell || conn->id_up[c] == conn->id_dn[c])
      return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport setup: connection cell out of range");
  for (int64_t b = 0; b < conn->nbc; ++b)
    if (conn->bc_cell[b] < 0 || conn->bc_cell[b] >= ncell)
      return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport setup: boundary cell out of range");
  // rows: the incident connections of each cell in connection order (both
  // roles), in 64-row slices as wide as their longest row
  std::vector<int64_t> deg(ncell, 0), bcnt(ncell + 1, 0);
  for (int64_t c = 0; c < conn->nconn; ++c) {
    ++deg[conn->id_up[c]];
    ++deg[conn->id_dn[c]];
  }
  for (int64_t b = 0; b < conn->nbc; ++b) ++bcnt[conn->bc_cell[b] + 1];
  for (int64_t i = 0; i < ncell; ++i) bcnt[i + 1] += bcnt[i];
  const int64_t nslice = (ncell + 63) / 64;
  std::vector<int64_t> sptr(nslice + 1, 0);
  int64_t wmax = 0;
  for (int64_t s = 0; s < nslice; ++s) {
    int64_t w = 0;
    for (int64_t i = s * 64; i < std::min<int64_t>(ncell, s * 64 + 64); ++i) w = std::max(w, deg[i]);
    wmax = std::max(wmax, w);
    sptr[s + 1] = sptr[s] + 64 * w;
  }
  const int64_t nent = sptr[nslice];
  std::vector<int32_t> col(nent + kEntryPad, 0), cid(nent + kEntryPad, 0);
  // padding: the row itself (a valid cell; the last cell for rows past the
  // end), connection id 0
  for (int64_t s = 0; s < nslice; ++s) {
    const int64_t w = (sptr[s + 1] - sptr[s]) / 64;
    for (int64_t j = 0; j < w; ++j)
      for (int l = 0; l < 64; ++l)
        col[sptr[s] + j * 64 + l] = (int32_t)std::min<int64_t>(s * 64 + l, ncell - 1);
  }
  std::vector<int64_t> fill(ncell, 0);
  auto put = [&](int64_t row, int64_t other, int32_t id) {
    const int64_t e = sptr[row / 64] + fill[row]++ * 64 + row % 64;
    col[e] = (int32_t)other;
    cid[e] = id;
  };
  for (int64_t c = 0; c < conn->nconn; ++c) {
    put(conn->id_up[c], conn->id_dn[c], (int32_t)(c + 1));
    put(conn->id_dn[c], conn->id_up[c], (int32_t)(-(c + 1)));
  }
  std::vector<int32_t> bid(std::max<int64_t>(conn->nbc, 1), 0);
  std::vector<int64_t> bfill(bcnt.begin(), bcnt.end() - 1);
  for (int64_t b = 0; b < conn->nbc; ++b) bid[bfill[conn->bc_cell[b]]++] = (int32_t)b;
  // one device buffer: slice_ptr, bc_ptr (int64), col, cid, bc_id (int32)
  const size_t n64 = (size_t)(nslice + 1) + (size_t)(ncell + 1);
  const size_t bytes = n64 * 8 + (2 * col.size() + bid.size()) * 4;
  TransportPlan** slot = ctx_transport_plan(ctx);
  if (*slot) {
    transport_plan_free(*slot);
    *slot = nullptr;
  }
  TransportPlan* p = new TransportPlan();
  p->naq = ctx_naq(ctx);
  hipError_t e = hipSetDevice(ctx_device(ctx));
  if (e == hipSuccess) e = hipMalloc(&p->topo, bytes);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&p->host_flags, 2 * (size_t)p->naq * sizeof(int32_t),
                      hipHostMallocDefault);
  if (e != hipSuccess) {
    transport_plan_free(p);
    return ctx_fail(ctx, PFCHEM_ERR_NOMEM, "transport setup: memory");
  }
  std::vector<char> h(bytes);
  char* o = h.data();
  std::memcpy(o, sptr.data(), (size_t)(nslice + 1) * 8);
  std::memcpy(o + (size_t)(nslice + 1) * 8, bcnt.data(), (size_t)(ncell + 1) * 8);
  char* o32 = o + n64 * 8;
  std::memcpy(o32, col.data(), col.size() * 4);
  std::memcpy(o32 + col.size() * 4, cid.data(), cid.size() * 4);
  std::memcpy(o32 + 2 * col.size() * 4, bid.data(), bid.size() * 4);
  e = hipMemcpy(p->topo, h.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    transport_plan_free(p);
    return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport setup: copy");
  }
  char* base = (char*)p->topo;
  p->slice_ptr = (const int64_t*)base;
  p->bc_ptr = (const int64_t*)(base + (size_t)(nslice + 1) * 8);
  const int32_t* b32 = (const int32_t*)(base + n64 * 8);
  p->col = b32;
  p->cid = b32 + col.size();
  p->bc_id = b32 + 2 * col.size();
  p->ncell = ncell;
  p->nconn = conn->nconn;
  p->nbc = conn->nbc;
  p->nslice = nslice;
  p->nent = nent;
  p->wmax = (int)std::min<int64_t>(wmax, INT32_MAX);
  *slot = p;
  return PFCHEM_OK;
}

extern "C" int pfchem_transport_step(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                                     const pfchem_transport_coefs* coefs, double dt, double rtol,
                                     int max_iterations, int32_t* iterations, void* stream) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  TransportPlan* p = *ctx_transport_plan(ctx);
  if (!p) return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport: call pfchem_transport_setup first");
  if (ncell != p->ncell)
    return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport: ncell differs from the setup's");
  if (!state || !coefs || !(dt > 0.0) || state->ld < ncell || !state->total ||
      !state->porosity || !state->saturation || !state->volume ||
      (p->nconn > 0 && (!coefs->area || !coefs->velocity || !coefs->coef_over_dist)) ||
      (p->nbc > 0 && (!coefs->bc_area || !coefs->bc_velocity || !coefs->bc_coef_over_dist ||
                      !coefs->bc_total)))
    return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport: bad arguments");
  if (ncell == 0) {
    if (iterations) *iterations = 0;
    return PFCHEM_OK;
  }
  const int naq = p->naq;
  if (rtol <= 0.0) rtol = 1.0e-12;
  if (max_iterations <= 0) max_iterations = 1000;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(ctx_device(ctx));
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: hipSetDevice");
  // work: diag, values, 6 vectors of naq x n (y, r, p, v, s, t), block
  // partials (3 products), scalars, flags
  const int64_t n = ncell, ns = p->nslice;
  const int64_t nblk = (ns + 3) / 4;  // 4 slices per block
  const size_t vec = (size_t)((naq + 1) & ~1) * n;  // pairs of components
  const size_t need = ((size_t)n + (size_t)(p->nent + kEntryPad) + 6 * vec +
                       (size_t)3 * naq * nblk + (size_t)kSc * naq) * sizeof(double) +
                      2 * (size_t)naq * sizeof(int32_t) + 256;
  if (need > p->work_bytes) {
    if (p->work) (void)hipFree(p->work);
    p->work = nullptr;
    p->work_bytes = 0;
    if (hipMalloc(&p->work, need) != hipSuccess)
      return ctx_fail(ctx, PFCHEM_ERR_NOMEM, "transport: work memory");
    p->work_bytes = need;
  }
  double* diag = p->work;
  double* val = diag + n;
  double* y = val + (p->nent + kEntryPad);
  double* r = y + vec;
  double* pv = r + vec;
  double* v = pv + vec;
  double* sv = v + vec;
  double* t = sv + vec;
  double* part = t + vec;
  double* sc = part + (size_t)3 * naq * nblk;
  int32_t* flags = (int32_t*)(sc + (size_t)kSc * naq);

  const unsigned grid = (unsigned)nblk;
  const size_t lds = (size_t)3 * naq * 4 * sizeof(double);  // wave sums of <= 3 products
  const unsigned grid_n = (unsigned)((n + kTB - 1) / kTB);
  Coefs cf{coefs->area, coefs->velocity, coefs->coef_over_dist, coefs->bc_area,
           coefs->bc_velocity, coefs->bc_coef_over_dist, coefs->bc_total};
  int32_t* hf = p->host_flags;
  // the components' flags on the host (the one wait of an iteration)
  auto poll = [&](int& nactive, bool& broke) {
    hipError_t er = hipMemcpyAsync(hf, flags, 2 * (size_t)naq * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    nactive = 0;
    for (int k = 0; k < naq; ++k) {
      nactive += hf[k] != 0;
      broke = broke || hf[naq + k] != 0;
    }
    return er;
  };

  const int kw = slots_for(p->wmax);
#define PFCHEM_ASSEMBLE(KW)                                                                     \
  k_assemble<KW><<<grid, kTB, lds, s>>>(n, naq, ns, p->slice_ptr, p->col, p->cid, p->bc_ptr,   \
                                        p->bc_id, p->nbc, cf, state->porosity,                 \
                                        state->saturation, state->volume, state->total,        \
                                        state->ld, dt, diag, val, y, r, part)
  switch (kw) {
    case 2: PFCHEM_ASSEMBLE(2); break;
    case 4: PFCHEM_ASSEMBLE(4); break;
    case 6: PFCHEM_ASSEMBLE(6); break;
    default: PFCHEM_ASSEMBLE(8); break;
  }
#undef PFCHEM_ASSEMBLE
  if (p->nent > 0)
    k_scale<<<(unsigned)((p->nent + kTB - 1) / kTB), kTB, 0, s>>>(p->nent, p->col, diag, val);
  k_fin<<<naq, kFinTB, 0, s>>>(0, naq, nblk, part, sc, flags, rtol);
  int nactive = 0;
  bool broke = false;
  e = poll(nactive, broke);
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: set-up kernels");
  int it = 0;
  while (nactive > 0 && it < max_iterations) {
    ++it;
    k_p<<<grid_n, kTB, 0, s>>>(n, naq, it == 1 ? 1 : 0, flags, sc, r, pv, v);
    launch_spmv<0>(kw, grid, lds, s, n, naq, ns, flags, p->slice_ptr, p->col, val, pv, v, part);
    k_fin<<<naq, kFinTB, 0, s>>>(1, naq, nblk, part, sc, flags, rtol);
    k_s<<<grid_n, kTB, 0, s>>>(n, naq, flags, sc, r, v, sv);
    launch_spmv<1>(kw, grid, lds, s, n, naq, ns, flags, p->slice_ptr, p->col, val, sv, t, part);
    k_fin<<<naq, kFinTB, 0, s>>>(2, naq, nblk, part, sc, flags, rtol);
    k_x<<<grid, kTB, lds, s>>>(n, naq, flags, sc, pv, sv, t, y, r, part);
    k_fin<<<naq, kFinTB, 0, s>>>(3, naq, nblk, part, sc, flags, rtol);
    if ((e = poll(nactive, broke)) != hipSuccess) break;
  }
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: solver kernels");
  k_store<<<grid_n, kTB, 0, s>>>(n, naq, y, diag, state->total, state->ld);
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: store");
  if (iterations) *iterations = it;
  if (broke || nactive > 0)
    return ctx_fail(ctx, PFCHEM_ERR_CELLS, "transport: a component did not converge");
  return PFCHEM_OK;
}
