// pfchem_team_n32.hip -- instantiations of the batched kernel for teams of
// 32 rows (lanes per cell x rows per lane).
#include "pfchem_team_impl.hpp"

namespace pfchem {

hipError_t launch_team_8_4(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  return launch_pair<8, 4>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                           has_status);
}

hipError_t launch_team_32_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                          double dt, int mixed, const LaunchCfg& cfg,
                          unsigned long long* cursors, int num_cus, int nxcd,
                          hipStream_t stream, bool has_status) {
  return launch_pair<32, 1>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                           has_status);
}

}  // namespace pfchem
