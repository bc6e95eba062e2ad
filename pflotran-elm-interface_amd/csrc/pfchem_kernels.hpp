// pfchem_kernels.hpp -- host-side launchers of the chemistry-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "pfchem_device.hpp"

namespace pfchem {

// persistent-kernel launch shape (see pfchem_launch in include/pfchem.h)
struct LaunchCfg {
  int lanes_per_cell;
  int rows_per_lane;
  int block_size;
  int blocks_per_cu;
  int lds_limit = 65536;  // LDS bytes one workgroup may allocate on the device
};

int plain_max_ncomp();

// One aqueous constraint as the constraint-speciation kernel reads it
// (pfchem_constraint with 0-based ids); lives in one device buffer followed
// by the reaction table its pointers address.
constexpr int kConMaxComp = 64;
struct DevConstraint {
  int type[kConMaxComp];
  int rxn[kConMaxComp];
  double guess[kConMaxComp];
  int nrxn, ms, h_ion_id, with_molality, use_prev, max_it, has_guess, has_h2o, has_coef;
  double max_dlnC;
  const int* ids;      // [nrxn][ms+1], [r][0] = count
  const double* st;    // [nrxn][ms]
  const double* h2o;   // [nrxn] (has_h2o)
  const double* logK;  // [nrxn]
  const double* coef;  // [nrxn][5] (has_coef)
  int* nfail;          // cells that failed (zero on entry)
};
// ReactionEquilibrateConstraint, one cell per thread, FP64;
// conc[i * s.ld + c] the constraint values of cell c
hipError_t launch_equilibrate(const DevNet& n, const DevNet* dn, const DevConstraint* dk,
                              const DevState& s, const double* conc, int64_t ncell,
                              hipStream_t stream);
hipError_t launch_plain_step(const DevNet& n, const DevNet* dn, const DevState& s, int64_t ncell,
                             double dt, hipStream_t stream);
// ReactionEquilibrateConstraint over teams of L lanes (one row per lane,
// pfchem_equil_team.hip); equil_team_lanes returns L, 0 when the network
// takes the one-cell-per-thread kernel (NEWTON activity, more than 64
// primaries, team LDS above 64 KiB)
int equil_team_lanes(const DevNet& n);
hipError_t launch_equil_team(const DevNet& n, const DevNet* dn, const DevConstraint* dk,
                             const DevState& s, const double* conc, int64_t ncell,
                             hipStream_t stream);
hipError_t launch_activity_lag(const DevNet& n, const DevNet* dn, const DevState& s, int64_t ncell,
                               hipStream_t stream);
// GIRT reaction block of one iterate, FP64: res[i*ld+c] and, if jac !=
// nullptr, jac[(i*ncomp+j)*ld+c]; teams of L >= ncomp lanes (one row per
// lane, 256 / L cells per workgroup); ds: device copy of the state
// descriptor.  girt_team_lanes returns L (0: ncomp > 64)
int girt_team_lanes(const DevNet& n);
hipError_t launch_girt_team(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                            double dt, const double* fixed_total, const double* fixed_imm,
                            double* res, double* jac, hipStream_t stream);

// batched (team) kernel: `cursors` holds one 64-bit work cursor per XCD
// (128-byte stride) that the launcher zeroes with a fill kernel (no memset
// node) before each launch.
int team_max_ncomp();
// team shapes (lanes per cell, rows per lane) with a SomDec instantiation
constexpr bool team_has_somdec(int L, int R) {
  return (L == 16 && R == 1) || (L == 8 && R == 2) || (L == 32 && R == 1);
}
// team shapes with an instantiation for per-cell active primary counts
constexpr bool team_has_variable(int L, int R) {
  return R == 1 && (L == 16 || L == 32 || L == 64);
}
int team_supports(const DevNet& n, const LaunchCfg& cfg, bool mixed);
// dn / ds: device copies of the network / state descriptors.  mixed = 0:
// FP64; 1: the FP32 pass and then an FP64 pass over the cells the FP32
// pass could not converge (needs the per-cell status array); kTeamRescue64:
// only the FP64 pass over cells whose status is kTeamEscalate.
constexpr int kTeamRescue64 = 2;
constexpr int kTeamEscalate = 16;
hipError_t launch_team_step(const DevNet& n, const DevNet* dn, const DevState* ds,
                            int64_t ncell, double dt, int mixed, const LaunchCfg& cfg,
                            unsigned long long* cursors, int num_cus, int nxcd,
                            hipStream_t stream, bool has_status);

// size buckets of cells with per-cell active counts: bucket 0 nact <= 16,
// 1: <= 32, 2: <= 64.  counts[0..2] and lists[b * ncell + i] (device); the
// counts are cleared by a fill kernel on `stream` first.
constexpr int kBuckets = 3;
constexpr int kBucketBound[kBuckets] = {16, 32, 64};
// the cells whose status is kTeamEscalate into list[0 .. *count) (any
// order); ncell < 2^31
hipError_t launch_escalated_list(const int32_t* status, int64_t ncell, int32_t* list,
                                 int32_t* count, hipStream_t stream);
hipError_t launch_buckets(const int32_t* nact, int64_t ncell, int naq, int32_t* lists,
                          int32_t* counts, hipStream_t stream);

// sensitivities (tangent: mode 0, out = df per cell; adjoint: mode 1, out =
// per-cell gradients [q * ldo + cell], sum = their sum over converged cells)
int sens_supported(const DevNet& n, const pfchem_options& opt);
hipError_t launch_sens(const DevNet& n, const DevNet* dn, const DevState* ds, int64_t ncell,
                       double dt, const double* wv, const double* ws, const double* dp, int mode,
                       int jacobian, double* fcell, double* out, int64_t ldo, int32_t* status,
                       double* sum, hipStream_t stream, int rescue = 0);

}  // namespace pfchem
