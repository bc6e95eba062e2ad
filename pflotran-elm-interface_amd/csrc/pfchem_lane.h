// pfchem_lane.h -- library-internal interface of the network-specialized
// lane kernel (device source pfchem_lane.hpp, host side pfchem_lane.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pfchem.h"

namespace pfchem {

// networks the lane kernel takes (the team kernel takes the others)
constexpr int kLaneMaxAq = 16;
constexpr int kLaneMaxCplx = 256;
constexpr int kLaneMaxMnrl = 16;
// per-lane LDS doubles of one cell's state (x 256 lanes x 8 B <= 160 KB)
constexpr int kLaneMaxLdsFields = 80;

// kernel arguments; the same layout as pflane::Args in pfchem_lane.hpp
struct LaneArgs {
  long long ld, ncell;
  double* pri_molal;
  double* total;
  double* immobile;
  double* sec_molal;
  double* pri_act_coef;
  double* sec_act_coef;
  double* ln_act_h2o;
  double* mnrl_volfrac;
  const double* mnrl_area;
  double* mnrl_rate;
  const double* porosity;
  const double* saturation;
  const double* den_kg;
  const double* volume;
  const double* temperature;
  int* num_iterations;
  int* num_sub_steps;
  int* status;
  unsigned long long* cursors;
  double dt;
  int nxcd;
  int rescue;
  int escalate;
  int list_min;
  const int* list;
  const int* list_count;
};
static_assert(sizeof(LaneArgs) == 208, "LaneArgs layout");

// adjoint kernel arguments; the same layout as pflane::SensArgs
struct LaneSensArgs {
  long long ld, ncell;
  const double* pri_molal;
  const double* total;
  const double* pri_act_coef;
  const double* sec_act_coef;
  const double* ln_act_h2o;
  const double* mnrl_volfrac;
  const double* mnrl_area;
  const double* porosity;
  const double* saturation;
  const double* den_kg;
  const double* volume;
  const double* temperature;
  const double* wv;
  const double* ws;
  double* grad;
  long long ldo;
  double* fcell;
  int* status;
  double* log;
  long long ldl;
  double dt;
  int jacobian;
  int pad;
};
static_assert(sizeof(LaneSensArgs) == 192, "LaneSensArgs layout");
constexpr int kLaneSensSub = 8;  // sub-steps the adjoint kernel records per cell
// first line of an adjoint module's source (followed by lane_source)
constexpr char kLaneAdjointDefine[] = "#define PFLANE_ADJOINT 1\n";

// 1 if the lane kernel handles this network and these options (else `why`)
int lane_supported(const pfchem_network* net, const pfchem_options* opt, std::string& why);
// full device source: the network as constants followed by the template
std::string lane_source(const pfchem_network* net, const pfchem_options* opt);
// compile `src` for gfx950 and load it on `device` (cached per process)
int lane_compile(const std::string& src, int device, std::string& err);
// compile only (no device needed); code object in *code when non-null
int lane_compile_only(const std::string& src, std::vector<char>* code, std::string& err);
// launch the FP32 (f32 = 1) or FP64 kernel of a compiled source over
// args.ncell cells; persistent grid of num_cus x min(blocks_per_cu,
// occupancy) workgroups of 256 lanes (blocks_per_cu <= 0: occupancy)
hipError_t lane_launch(const std::string& src, int device, int f32, const LaneArgs& args,
                       int num_cus, int blocks_per_cu, unsigned long long* cursors,
                       hipStream_t stream);
// resident 256-lane workgroups per CU of a compiled kernel (0: not compiled)
int lane_occupancy(const std::string& src, int device, int f32);
// launch the adjoint kernel of a compiled source, one cell per lane
hipError_t lane_adjoint_launch(const std::string& src, int device, const LaneSensArgs& args,
                               hipStream_t stream);

}  // namespace pfchem
