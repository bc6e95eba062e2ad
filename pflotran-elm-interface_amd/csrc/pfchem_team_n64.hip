// pfchem_team_n64.hip -- instantiation of the batched kernel for teams of
// 64 rows: one cell per wavefront (networks of 33..64 unknowns).
#include "pfchem_team_impl.hpp"

namespace pfchem {

hipError_t launch_team_64_1(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                            double dt, int mixed, const LaunchCfg& cfg,
                            unsigned long long* cursors, int num_cus, int nxcd,
                            hipStream_t stream, bool has_status) {
  return launch_pair<64, 1>(n, dn, ds, ncell, dt, mixed, cfg, cursors, num_cus, nxcd, stream,
                            has_status);
}

}  // namespace pfchem
