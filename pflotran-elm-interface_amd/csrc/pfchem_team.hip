// pfchem_team.hip -- shape table and dispatch of the batched kernel.
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"
#include "pfchem_somdec.hpp"

namespace pfchem {

hipError_t launch_team_4_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_2_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_2_8(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_4_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_8_2(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_16_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_8_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_32_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_64_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);

namespace {
constexpr int kMaxMnrl = 16;
// (lanes per cell, rows per lane) shapes compiled into the library
// One-lane teams (1x4, 1x8) and 8x1 are not offered: on MI355X their
// instantiations returned wrong answers and faulted (1x4) or did not finish
// (8x1) with no cause found in the kernel (a host build that runs one lane
// per wavefront matches the oracle).
const int kShapes[][2] = {{4, 1}, {2, 4}, {2, 8}, {4, 4}, {8, 2}, {16, 1}, {8, 4}, {32, 1},
                          {64, 1}};

// per-team LDS bytes: the same layout as team_lds_layout in the kernel
int team_lds_bytes(int NM, int nc, int nm, int jt_size, int nsd) {
  int o = 8 * NM * 5;
  const int m = nm > 0 ? nm : 1;
  o += 8 * m * 3 + 8 * (nc > 0 ? nc : 1) + 8 * (NM + 2) + jt_size * NM * (NM + 1) + 8 * nsd;
  return (o + 15) & ~15;
}
}  // namespace

int team_max_ncomp() { return 64; }

int team_supports(const DevNet& n, const LaunchCfg& cfg, bool mixed) {
  (void)mixed;  // both variants accumulate the Jacobian tile in FP64
  const int L = cfg.lanes_per_cell, R = cfg.rows_per_lane;
  bool known = false;
  for (const auto& sh : kShapes)
    if (sh[0] == L && sh[1] == R) known = true;
  if (!known) return 0;
  // with per-cell active counts the size buckets bring every cell within
  // its launch's team (nact <= L * R)
  if ((n.ncomp > L * R && !n.variable) || n.nm > kMaxMnrl) return 0;
  if (n.has_somdec && !team_has_somdec(L, R)) return 0;
  if (n.variable && !team_has_variable(L, R)) return 0;
  if (cfg.block_size % 64 || cfg.block_size > 256 || cfg.block_size < 64) return 0;
  if ((size_t)(cfg.block_size / L) * team_lds_bytes(L * R, n.nc, n.nm, 8,
                     n.has_somdec ? sd_lds_doubles(n.sd.nrxn, n.sd.ext ? 2 : 1) : 0) >
      (size_t)cfg.lds_limit)
    return 0;
  return 1;
}

namespace {
__global__ void clear_bucket_counts(int32_t* counts) {
  if (threadIdx.x < kBuckets) counts[threadIdx.x] = 0;
}
__global__ void __launch_bounds__(256)
bucket_cells(const int32_t* nact, int64_t ncell, int naq, int32_t* lists, int32_t* counts) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int a = min(max(nact[c], 1), naq);
  const int b = a <= kBucketBound[0] ? 0 : (a <= kBucketBound[1] ? 1 : 2);
  const int pos = atomicAdd(&counts[b], 1);
  lists[(int64_t)b * ncell + pos] = (int32_t)c;
}
__global__ void clear_count(int32_t* count) {
  if (threadIdx.x < 1) count[threadIdx.x] = 0;
}
// one atomic per workgroup of kListBlock threads: its handed-over cells
// take consecutive entries, in cell order within the workgroup
constexpr int kListBlock = 1024;
__global__ void __launch_bounds__(kListBlock)
escalated_cells(const int32_t* status, int64_t ncell, int32_t* list, int32_t* count) {
  __shared__ int wave_off[kListBlock / 64];
  __shared__ int block_base;
  const int64_t c = (int64_t)blockIdx.x * kListBlock + threadIdx.x;
  const bool esc = c < ncell && status[c] == kTeamEscalate;
  const unsigned long long mask = __ballot(esc);
  const int lane = (int)__lane_id();
  const int wave = (int)(threadIdx.x >> 6);
  if (lane == 0) wave_off[wave] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int w = 0; w < kListBlock / 64; ++w) {
      const int k = wave_off[w];
      wave_off[w] = total;
      total += k;
    }
    block_base = total > 0 ? atomicAdd(count, total) : 0;
  }
  __syncthreads();
  if (esc) list[block_base + wave_off[wave] + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)c;
}
}  // namespace

hipError_t launch_escalated_list(const int32_t* status, int64_t ncell, int32_t* list,
                                 int32_t* count, hipStream_t stream) {
  clear_count<<<1, 64, 0, stream>>>(count);
  if (ncell > 0)
    escalated_cells<<<dim3((unsigned)((ncell + kListBlock - 1) / kListBlock)), dim3(kListBlock), 0,
                      stream>>>(status, ncell, list, count);
  return hipGetLastError();
}

hipError_t launch_buckets(const int32_t* nact, int64_t ncell, int naq, int32_t* lists,
                          int32_t* counts, hipStream_t stream) {
  clear_bucket_counts<<<1, 64, 0, stream>>>(counts);
  const int block = 256;
  bucket_cells<<<dim3((unsigned)((ncell + block - 1) / block)), dim3(block), 0, stream>>>(
      nact, ncell, naq, lists, counts);
  return hipGetLastError();
}

hipError_t launch_team_step(const DevNet& n, const DevNet* dn, const DevState* ds,
                            int64_t ncell, double dt, int mixed, const LaunchCfg& cfg,
                            unsigned long long* cursors, int num_cus, int nxcd,
                            hipStream_t stream, bool has_status) {
#define PFCHEM_SHAPE(LL, RR)                                                               \
  if (cfg.lanes_per_cell == LL && cfg.rows_per_lane == RR)                                 \
    return launch_team_##LL##_##RR(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, \
                                   stream, has_status);
  PFCHEM_SHAPE(4, 1)
  PFCHEM_SHAPE(2, 4)
  PFCHEM_SHAPE(2, 8)
  PFCHEM_SHAPE(4, 4)
  PFCHEM_SHAPE(8, 2)
  PFCHEM_SHAPE(16, 1)
  PFCHEM_SHAPE(8, 4)
  PFCHEM_SHAPE(32, 1)
  PFCHEM_SHAPE(64, 1)
#undef PFCHEM_SHAPE
  return hipErrorInvalidValue;
}

}  // namespace pfchem
