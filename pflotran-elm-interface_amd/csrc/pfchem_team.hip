// pfchem_team.hip -- shape table and dispatch of the batched kernel.
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"

namespace pfchem {

hipError_t launch_team_4_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_2_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_2_8(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_4_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_8_2(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_16_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_8_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_32_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status);
hipError_t launch_team_64_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, bool mixed, const LaunchCfg& cfg,
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
  if (n.ncomp > L * R || n.nm > kMaxMnrl) return 0;
  if (n.has_somdec && !team_has_somdec(L, R)) return 0;
  if (n.variable && !team_has_variable(L, R)) return 0;
  if (cfg.block_size % 64 || cfg.block_size > 256 || cfg.block_size < 64) return 0;
  if ((size_t)(cfg.block_size / L) * team_lds_bytes(L * R, n.nc, n.nm, 8,
                     n.has_somdec ? n.sd.nrxn * (3 + PFCHEM_SOMDEC_MAX_DOWN) : 0) >
      (size_t)cfg.lds_limit)
    return 0;
  return 1;
}

hipError_t launch_team_step(const DevNet& n, const DevNet* dn, const DevState* ds,
                            int64_t ncell, double dt, bool mixed, const LaunchCfg& cfg,
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
