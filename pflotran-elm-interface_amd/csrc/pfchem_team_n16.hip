// pfchem_team_n16.hip -- instantiations of the batched kernel for teams of
// 16 rows (lanes per cell x rows per lane).
#include "pfchem_team_impl.hpp"

namespace pfchem {

hipError_t launch_team_2_8(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  return launch_pair<2, 8>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                           has_status);
}

hipError_t launch_team_4_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  return launch_pair<4, 4>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                           has_status);
}

hipError_t launch_team_8_2(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  return launch_pair<8, 2>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                           has_status);
}

hipError_t launch_team_16_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  return launch_pair<16, 1>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                           has_status);
}

}  // namespace pfchem
