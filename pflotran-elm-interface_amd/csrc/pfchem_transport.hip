// pfchem_transport.hip -- OSRT transport step on the device: the implicit
// advection-diffusion system of RTCalculateTransportMatrix
// (reactive_transport.F90:1469-1703) with the accumulation and boundary
// right-hand side of RTCalculateRHS_t1 (:1239-1465) and the TFluxCoef
// coefficients (transport.F90:495-565), solved for every aqueous component
// together (one BiCGSTAB per component with a Jacobi preconditioner, the
// components' vectors side by side) where the reference calls KSPSolve once
// per component (pmc_subsurface_osrt.F90:286-333).
//
// Layout: the matrix is kept as its diagonal plus, per row, the incident
// connections (column, connection id, role: the row is the connection's
// upstream or downstream cell), so the values of a step are formed row by
// row from the connection coefficients without atomics.  Vectors are
// [component][cell].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pfchem.h"
#include "pfchem_device.hpp"
#include "pfchem_transport.h"

namespace pfchem {

struct TransportPlan {
  int64_t ncell = 0, nconn = 0, nbc = 0, nnz = 0;
  // device: row_ptr[ncell + 1], col[nnz], conn[nnz] (role in the sign: conn
  // id + 1 when the row is the upstream cell, -(id + 1) when downstream),
  // bc_ptr[ncell + 1], bc_id[nbc]
  void* topo = nullptr;
  const int64_t* row_ptr = nullptr;
  const int64_t* col = nullptr;
  const int64_t* conn = nullptr;
  const int64_t* bc_ptr = nullptr;
  const int64_t* bc_id = nullptr;
  // work vectors and reduction partials (grown on demand)
  double* work = nullptr;
  size_t work_bytes = 0;
};

void transport_plan_free(TransportPlan* p) {
  if (!p) return;
  if (p->topo) (void)hipFree(p->topo);
  if (p->work) (void)hipFree(p->work);
  delete p;
}

namespace {

constexpr int kTB = 256;         // threads per block
constexpr int kMaxPartials = 1024;  // reduction blocks

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

// values of the step's matrix, the right-hand side and the initial guess
__global__ __launch_bounds__(kTB) void k_assemble(
    int64_t n, int naq, const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col,
    const int64_t* __restrict__ conn, const int64_t* __restrict__ bc_ptr,
    const int64_t* __restrict__ bc_id, int64_t nbc, Coefs cf, const double* __restrict__ por,
    const double* __restrict__ sat, const double* __restrict__ vol,
    const double* __restrict__ total, int64_t ld, double dt, double* __restrict__ diag,
    double* __restrict__ val, double* __restrict__ b, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const double si = sat[i];
  const double acc = por[i] * si * 1000.0 * vol[i] / dt;
  double d = acc;
  for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
    const int64_t r = conn[e];
    const int64_t c = r > 0 ? r - 1 : -r - 1;
    const int64_t j = col[e];
    // interior connections check both saturations (check_upwind_saturation)
    const double s_up = r > 0 ? si : sat[j], s_dn = r > 0 ? sat[j] : si;
    double t_up = 0.0, t_dn = 0.0;
    if (!(s_dn < 1.0e-40) && !(s_up < 1.0e-40)) flux_coef(cf.vel[c], cf.cod[c], cf.area[c], t_up, t_dn);
    if (r > 0) {  // this row is upstream: + T_up on the diagonal, + T_dn off it
      d += t_up;
      val[e] = t_dn;
    } else {  // downstream: - T_dn on the diagonal, - T_up off it
      d -= t_dn;
      val[e] = -t_up;
    }
  }
  // a dry cell (no accumulation, no fluxes) has an empty row in the
  // reference's matrix; here it keeps its totals
  const bool dry = si < 1.0e-40;
  diag[i] = dry ? 1.0 : d;
  for (int k = 0; k < naq; ++k) {
    if (dry) {
      b[k * n + i] = total[k * ld + i];
      x[k * n + i] = total[k * ld + i];
      continue;
    }
    double rb = acc * total[k * ld + i];
    for (int64_t e = bc_ptr[i]; e < bc_ptr[i + 1]; ++e) {
      const int64_t q = bc_id[e];
      if (si < 1.0e-40) continue;
      double t_up, t_dn;
      flux_coef(cf.bc_vel[q], cf.bc_cod[q], cf.bc_area[q], t_up, t_dn);
      rb += t_up * cf.bc_total[k * nbc + q];
    }
    b[k * n + i] = rb;
    x[k * n + i] = total[k * ld + i];
  }
  // boundary terms on the diagonal (the cell is the downstream side)
  if (!(si < 1.0e-40)) {
    for (int64_t e = bc_ptr[i]; e < bc_ptr[i + 1]; ++e) {
      const int64_t q = bc_id[e];
      double t_up, t_dn;
      flux_coef(cf.bc_vel[q], cf.bc_cod[q], cf.bc_area[q], t_up, t_dn);
      diag[i] -= t_dn;
    }
  }
}

// y_k = M x_k for the active components
__global__ __launch_bounds__(kTB) void k_spmv(int64_t n, int naq, const int* __restrict__ act,
                                              const int64_t* __restrict__ row_ptr,
                                              const int64_t* __restrict__ col,
                                              const double* __restrict__ diag,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const int64_t e0 = row_ptr[i], e1 = row_ptr[i + 1];
  const double d = diag[i];
  for (int k = 0; k < naq; ++k) {
    if (!act[k]) continue;
    const double* xk = x + k * n;
    double s = d * xk[i];
    for (int64_t e = e0; e < e1; ++e) s += val[e] * xk[col[e]];
    y[k * n + i] = s;
  }
}

// block partial sums of (a_k . b_k) and (c_k . d_k) (c null: one product)
__global__ __launch_bounds__(kTB) void k_dot2(int64_t n, int naq, const double* __restrict__ a,
                                              const double* __restrict__ b,
                                              const double* __restrict__ c,
                                              const double* __restrict__ d,
                                              double* __restrict__ part) {
  __shared__ double s1[kTB], s2[kTB];
  const int k = blockIdx.y;
  const double* ak = a + k * n;
  const double* bk = b + k * n;
  double p1 = 0.0, p2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    p1 += ak[i] * bk[i];
    if (c) p2 += c[k * n + i] * d[k * n + i];
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  __syncthreads();
  for (int h = kTB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s1[threadIdx.x] += s1[threadIdx.x + h];
      s2[threadIdx.x] += s2[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[(0 * naq + k) * gridDim.x + blockIdx.x] = s1[0];
    part[(1 * naq + k) * gridDim.x + blockIdx.x] = s2[0];
  }
}

// sums of the partials in a fixed order: out[2 * naq]
__global__ void k_dot_finish(int naq, int nblk, const double* __restrict__ part,
                             double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * naq) return;
  double s = 0.0;
  for (int j = 0; j < nblk; ++j) s += part[q * nblk + j];
  out[q] = s;
}

// r = b - r (r holds M x), p = v = 0, and the shadow residual rh: a fixed
// pseudo-random vector in [0.5, 1.5) (rh = r0 breaks down on upwind
// advection, whose preconditioned matrix is the identity plus a nilpotent
// part: the residual moves downstream, orthogonal to r0)
__global__ __launch_bounds__(kTB) void k_init(int64_t n, int naq, const double* __restrict__ b,
                                              double* __restrict__ r, double* __restrict__ rh,
                                              double* __restrict__ p, double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < naq; ++k) {
    const int64_t j = k * n + i;
    const double rr = b[j] - r[j];
    r[j] = rr;
    const unsigned h = (unsigned)(i * 2654435761ull + (unsigned long long)k * 40503ull);
    rh[j] = 0.5 + (double)((h >> 8) & 0xffffu) / 65536.0;
    p[j] = 0.0;
    v[j] = 0.0;
  }
}

// scalars per component: sc[k*4 + 0] beta, 1 omega, 2 alpha
// p = r + beta (p - omega v); ph = p / diag
__global__ __launch_bounds__(kTB) void k_p(int64_t n, int naq, const int* __restrict__ act,
                                           const double* __restrict__ sc,
                                           const double* __restrict__ diag,
                                           const double* __restrict__ r, double* __restrict__ p,
                                           const double* __restrict__ v,
                                           double* __restrict__ ph) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const double id = 1.0 / diag[i];
  for (int k = 0; k < naq; ++k) {
    if (!act[k]) continue;
    const int64_t j = k * n + i;
    const double pp = r[j] + sc[k * 4 + 0] * (p[j] - sc[k * 4 + 1] * v[j]);
    p[j] = pp;
    ph[j] = pp * id;
  }
}

// s = r - alpha v; sh = s / diag
__global__ __launch_bounds__(kTB) void k_s(int64_t n, int naq, const int* __restrict__ act,
                                           const double* __restrict__ sc,
                                           const double* __restrict__ diag,
                                           const double* __restrict__ r,
                                           const double* __restrict__ v, double* __restrict__ s,
                                           double* __restrict__ sh) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const double id = 1.0 / diag[i];
  for (int k = 0; k < naq; ++k) {
    if (!act[k]) continue;
    const int64_t j = k * n + i;
    const double ss = r[j] - sc[k * 4 + 2] * v[j];
    s[j] = ss;
    sh[j] = ss * id;
  }
}

// x += alpha ph + omega sh; r = s - omega t
__global__ __launch_bounds__(kTB) void k_x(int64_t n, int naq, const int* __restrict__ act,
                                           const double* __restrict__ sc,
                                           const double* __restrict__ ph,
                                           const double* __restrict__ sh,
                                           const double* __restrict__ s,
                                           const double* __restrict__ t, double* __restrict__ x,
                                           double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < naq; ++k) {
    if (!act[k]) continue;
    const int64_t j = k * n + i;
    x[j] += sc[k * 4 + 2] * ph[j] + sc[k * 4 + 1] * sh[j];
    r[j] = s[j] - sc[k * 4 + 1] * t[j];
  }
}

// totals of the new time level into the state
__global__ __launch_bounds__(kTB) void k_store(int64_t n, int naq, const double* __restrict__ x,
                                               double* __restrict__ total, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < naq; ++k) total[k * ld + i] = x[k * n + i];
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
  for (int64_t c = 0; c < conn->nconn; ++c)
    if (conn->id_up[c] < 0 || conn->id_up[c] >= ncell || conn->id_dn[c] < 0 ||
        conn->id_dn[c] >= ncell || conn->id_up[c] == conn->id_dn[c])
      return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport setup: connection cell out of range");
  for (int64_t b = 0; b < conn->nbc; ++b)
    if (conn->bc_cell[b] < 0 || conn->bc_cell[b] >= ncell)
      return ctx_fail(ctx, PFCHEM_ERR_INVALID, "transport setup: boundary cell out of range");
  // rows: the incident connections of each cell (upstream role first in
  // connection order, then downstream), the boundary connections
  std::vector<int64_t> cnt(ncell + 1, 0), bcnt(ncell + 1, 0);
  for (int64_t c = 0; c < conn->nconn; ++c) {
    ++cnt[conn->id_up[c] + 1];
    ++cnt[conn->id_dn[c] + 1];
  }
  for (int64_t b = 0; b < conn->nbc; ++b) ++bcnt[conn->bc_cell[b] + 1];
  for (int64_t i = 0; i < ncell; ++i) {
    cnt[i + 1] += cnt[i];
    bcnt[i + 1] += bcnt[i];
  }
  const int64_t nnz = cnt[ncell];
  std::vector<int64_t> col(nnz > 0 ? nnz : 1), cid(nnz > 0 ? nnz : 1), fill(cnt.begin(), cnt.end() - 1);
  for (int64_t c = 0; c < conn->nconn; ++c) {
    const int64_t u = conn->id_up[c], d = conn->id_dn[c];
    col[fill[u]] = d;
    cid[fill[u]++] = c + 1;
    col[fill[d]] = u;
    cid[fill[d]++] = -(c + 1);
  }
  std::vector<int64_t> bid(conn->nbc > 0 ? conn->nbc : 1), bfill(bcnt.begin(), bcnt.end() - 1);
  for (int64_t b = 0; b < conn->nbc; ++b) bid[bfill[conn->bc_cell[b]]++] = b;
  // one device buffer: row_ptr, col, conn, bc_ptr, bc_id
  const size_t n1 = (size_t)ncell + 1;
  const size_t words = n1 + 2 * (size_t)std::max<int64_t>(nnz, 1) + n1 +
                       (size_t)std::max<int64_t>(conn->nbc, 1);
  TransportPlan** slot = ctx_transport_plan(ctx);
  if (*slot) {
    transport_plan_free(*slot);
    *slot = nullptr;
  }
  TransportPlan* p = new TransportPlan();
  hipError_t e = hipSetDevice(ctx_device(ctx));
  if (e == hipSuccess) e = hipMalloc(&p->topo, words * sizeof(int64_t));
  if (e != hipSuccess) {
    delete p;
    return ctx_fail(ctx, PFCHEM_ERR_NOMEM, "transport setup: device memory");
  }
  std::vector<int64_t> h(words, 0);
  size_t o = 0;
  std::memcpy(&h[o], cnt.data(), n1 * sizeof(int64_t));
  o += n1;
  std::memcpy(&h[o], col.data(), col.size() * sizeof(int64_t));
  const size_t ocol = n1, ocid = n1 + std::max<int64_t>(nnz, 1);
  std::memcpy(&h[ocid], cid.data(), cid.size() * sizeof(int64_t));
  const size_t obp = ocid + std::max<int64_t>(nnz, 1);
  std::memcpy(&h[obp], bcnt.data(), n1 * sizeof(int64_t));
  const size_t obi = obp + n1;
  std::memcpy(&h[obi], bid.data(), bid.size() * sizeof(int64_t));
  e = hipMemcpy(p->topo, h.data(), words * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    transport_plan_free(p);
    return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport setup: copy");
  }
  const int64_t* base = (const int64_t*)p->topo;
  p->row_ptr = base;
  p->col = base + ocol;
  p->conn = base + ocid;
  p->bc_ptr = base + obp;
  p->bc_id = base + obi;
  p->ncell = ncell;
  p->nconn = conn->nconn;
  p->nbc = conn->nbc;
  p->nnz = nnz;
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
  const int naq = ctx_naq(ctx);
  if (rtol <= 0.0) rtol = 1.0e-12;
  if (max_iterations <= 0) max_iterations = 1000;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(ctx_device(ctx));
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: hipSetDevice");
  // work: diag, val, 10 vectors of naq x n, partials, dot results,
  // scalars, active flags
  const int64_t n = ncell;
  const int nblk = (int)std::min<int64_t>(kMaxPartials, (n + kTB - 1) / kTB);
  const size_t vec = (size_t)naq * n;
  const size_t need = ((size_t)n + (size_t)std::max<int64_t>(p->nnz, 1) + 10 * vec +
                       (size_t)2 * naq * nblk + 2 * naq + 4 * naq) * sizeof(double) +
                      naq * sizeof(int) + 256;
  if (need > p->work_bytes) {
    if (p->work) (void)hipFree(p->work);
    p->work = nullptr;
    p->work_bytes = 0;
    if (hipMalloc(&p->work, need) != hipSuccess)
      return ctx_fail(ctx, PFCHEM_ERR_NOMEM, "transport: work memory");
    p->work_bytes = need;
  }
  double* w = p->work;
  double* diag = w;
  double* val = diag + n;
  double* b = val + std::max<int64_t>(p->nnz, 1);
  double* x = b + vec;
  double* r = x + vec;
  double* rh = r + vec;
  double* pv = rh + vec;
  double* v = pv + vec;
  double* sv = v + vec;
  double* t = sv + vec;
  double* ph = t + vec;
  double* sh = ph + vec;
  double* part = sh + vec;
  double* dots = part + (size_t)2 * naq * nblk;
  double* sc = dots + 2 * naq;
  int* act = (int*)(sc + 4 * naq);

  const unsigned grid = (unsigned)((n + kTB - 1) / kTB);
  Coefs cf{coefs->area, coefs->velocity, coefs->coef_over_dist, coefs->bc_area,
           coefs->bc_velocity, coefs->bc_coef_over_dist, coefs->bc_total};
  std::vector<int> hact(naq, 1);
  std::vector<double> hsc(4 * naq, 0.0), hd(2 * naq), rho0_tmp(naq, 0.0);
  auto upload = [&]() {
    hipError_t er = hipMemcpyAsync(sc, hsc.data(), 4 * naq * sizeof(double),
                                   hipMemcpyHostToDevice, s);
    if (er == hipSuccess)
      er = hipMemcpyAsync(act, hact.data(), naq * sizeof(int), hipMemcpyHostToDevice, s);
    return er;
  };
  // the two dot products of every component, on the host
  auto dot2 = [&](const double* a1, const double* b1, const double* c1, const double* d1) {
    k_dot2<<<dim3(nblk, naq), kTB, 0, s>>>(n, naq, a1, b1, c1, d1, part);
    k_dot_finish<<<(2 * naq + 63) / 64, 64, 0, s>>>(naq, nblk, part, dots);
    hipError_t er = hipMemcpyAsync(hd.data(), dots, 2 * naq * sizeof(double),
                                   hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    return er;
  };

  k_assemble<<<grid, kTB, 0, s>>>(n, naq, p->row_ptr, p->col, p->conn, p->bc_ptr, p->bc_id,
                                  p->nbc, cf, state->porosity, state->saturation, state->volume,
                                  state->total, state->ld, dt, diag, val, b, x);
  e = upload();
  if (e == hipSuccess) {
    k_spmv<<<grid, kTB, 0, s>>>(n, naq, act, p->row_ptr, p->col, diag, val, x, r);
    k_init<<<grid, kTB, 0, s>>>(n, naq, b, r, rh, pv, v);
    e = dot2(b, b, r, r);  // ||b||^2, ||r0||^2
    if (e == hipSuccess) {
      const std::vector<double> nb(hd.begin(), hd.end());
      e = dot2(rh, r, nullptr, nullptr);  // rho_0 = (rh, r0)
      for (int k = 0; k < naq; ++k) {
        rho0_tmp[k] = hd[k];
        hd[k] = nb[k];
        hd[naq + k] = nb[naq + k];
      }
    }
  }
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: set-up kernels");
  std::vector<double> bnorm(naq), rho(naq), rho_prev(naq, 1.0), alpha(naq, 1.0), omega(naq, 1.0);
  int nactive = 0;
  for (int k = 0; k < naq; ++k) {
    bnorm[k] = std::sqrt(hd[k]);
    const double rn = std::sqrt(hd[naq + k]);
    hact[k] = rn > rtol * bnorm[k] ? 1 : 0;
    nactive += hact[k];
    rho[k] = rho0_tmp[k];
  }
  int it = 0;
  bool broke = false;
  while (nactive > 0 && it < max_iterations) {
    ++it;
    for (int k = 0; k < naq; ++k) {
      if (!hact[k]) continue;
      const double beta = it == 1 ? 0.0 : (rho[k] / rho_prev[k]) * (alpha[k] / omega[k]);
      hsc[k * 4 + 0] = beta;
      hsc[k * 4 + 1] = omega[k];
    }
    if ((e = upload()) != hipSuccess) break;
    k_p<<<grid, kTB, 0, s>>>(n, naq, act, sc, diag, r, pv, v, ph);
    k_spmv<<<grid, kTB, 0, s>>>(n, naq, act, p->row_ptr, p->col, diag, val, ph, v);
    if ((e = dot2(rh, v, nullptr, nullptr)) != hipSuccess) break;
    for (int k = 0; k < naq; ++k) {
      if (!hact[k]) continue;
      if (hd[k] == 0.0 || !std::isfinite(hd[k])) {
        broke = true;
        hact[k] = 0;
        continue;
      }
      alpha[k] = rho[k] / hd[k];
      hsc[k * 4 + 2] = alpha[k];
    }
    if ((e = upload()) != hipSuccess) break;
    k_s<<<grid, kTB, 0, s>>>(n, naq, act, sc, diag, r, v, sv, sh);
    k_spmv<<<grid, kTB, 0, s>>>(n, naq, act, p->row_ptr, p->col, diag, val, sh, t);
    if ((e = dot2(t, sv, t, t)) != hipSuccess) break;
    for (int k = 0; k < naq; ++k) {
      if (!hact[k]) continue;
      omega[k] = hd[naq + k] > 0.0 ? hd[k] / hd[naq + k] : 0.0;
      hsc[k * 4 + 1] = omega[k];
    }
    if ((e = upload()) != hipSuccess) break;
    k_x<<<grid, kTB, 0, s>>>(n, naq, act, sc, ph, sh, sv, t, x, r);
    if ((e = dot2(r, r, rh, r)) != hipSuccess) break;
    nactive = 0;
    for (int k = 0; k < naq; ++k) {
      if (!hact[k]) continue;
      rho_prev[k] = rho[k];
      rho[k] = hd[naq + k];
      const double rn = std::sqrt(hd[k]);
      if (!(rn > rtol * bnorm[k])) {
        hact[k] = 0;  // converged
      } else if (omega[k] == 0.0 || rho[k] == 0.0 || !std::isfinite(rn)) {
        broke = true;
        hact[k] = 0;
      } else {
        ++nactive;
      }
    }
  }
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: solver kernels");
  k_store<<<grid, kTB, 0, s>>>(n, naq, x, state->total, state->ld);
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return ctx_fail(ctx, PFCHEM_ERR_HIP, "transport: store");
  if (iterations) *iterations = it;
  if (broke || nactive > 0)
    return ctx_fail(ctx, PFCHEM_ERR_CELLS, "transport: a component did not converge");
  return PFCHEM_OK;
}
