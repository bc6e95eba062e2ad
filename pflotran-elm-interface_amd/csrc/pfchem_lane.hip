// pfchem_lane.hip -- host side of the network-specialized lane kernel
// (pfchem_lane.hpp): source generation from one reaction network and its
// solver options, run-time compilation for gfx950 (hiprtc), a process-wide
// cache of compiled modules, and the launch.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pfchem.h"
#include "pfchem_kernels.hpp"
#include "pfchem_lane.h"

namespace pfchem {

namespace {

// device source of the kernel template (pfchem_lane.hpp), embedded at build
const char kLaneTemplate[] =
#include "pfchem_lane_src.inc"
    ;

// exact decimal-free spelling of a double (C++17 hexadecimal literal)
std::string hexd(double v) {
  char b[64];
  if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
  std::snprintf(b, sizeof b, "%a", v);
  return b;
}

struct Gen {
  std::string s;
  void line(const std::string& l) { s += l + "\n"; }
  void i(const char* name, long v) { line("  static constexpr int " + std::string(name) + " = " + std::to_string(v) + ";"); }
  void d(const char* name, double v) { line("  static constexpr double " + std::string(name) + " = " + hexd(v) + ";"); }
  void ivec(const char* name, const std::vector<int>& v) {
    std::string l = "  static constexpr int " + std::string(name) + "[" +
                    std::to_string(v.empty() ? 1 : v.size()) + "] = {";
    if (v.empty()) l += "0";
    for (size_t k = 0; k < v.size(); ++k) l += (k ? "," : "") + std::to_string(v[k]);
    line(l + "};");
  }
  void dvec(const char* name, const std::vector<double>& v) {
    std::string l = "  static constexpr double " + std::string(name) + "[" +
                    std::to_string(v.empty() ? 1 : v.size()) + "] = {";
    if (v.empty()) l += "0.0";
    for (size_t k = 0; k < v.size(); ++k) l += (k ? "," : "") + hexd(v[k]);
    line(l + "};");
  }
  // rows x w tables (at least one row and one column)
  void imat(const char* name, const std::vector<std::vector<int>>& v, int w) {
    if (w < 1) w = 1;
    std::string l = "  static constexpr int " + std::string(name) + "[" +
                    std::to_string(v.empty() ? 1 : v.size()) + "][" + std::to_string(w) + "] = {";
    if (v.empty()) l += "{0}";
    for (size_t k = 0; k < v.size(); ++k) {
      l += k ? ",{" : "{";
      for (int q = 0; q < w; ++q) l += (q ? "," : "") + std::to_string(q < (int)v[k].size() ? v[k][q] : 0);
      l += "}";
    }
    line(l + "};");
  }
  void dmat(const char* name, const std::vector<std::vector<double>>& v, int w) {
    if (w < 1) w = 1;
    std::string l = "  static constexpr double " + std::string(name) + "[" +
                    std::to_string(v.empty() ? 1 : v.size()) + "][" + std::to_string(w) + "] = {";
    if (v.empty()) l += "{0.0}";
    for (size_t k = 0; k < v.size(); ++k) {
      l += k ? ",{" : "{";
      for (int q = 0; q < w; ++q) l += (q ? "," : "") + hexd(q < (int)v[k].size() ? v[k][q] : 0.0);
      l += "}";
    }
    line(l + "};");
  }
};

bool finite_all(const double* p, size_t n) {
  if (!p) return true;
  for (size_t k = 0; k < n; ++k)
    if (!std::isfinite(p[k])) return false;
  return true;
}

// reaction rows of the compressed layout (count in entry 0)
void rows(const int32_t* ids, const double* st, int n, int ms, std::vector<int>& cnt,
          std::vector<std::vector<int>>& id, std::vector<std::vector<double>>& nu) {
  for (int k = 0; k < n; ++k) {
    const int32_t* r = ids + (size_t)k * (ms + 1);
    cnt.push_back(r[0]);
    std::vector<int> a;
    std::vector<double> b;
    for (int q = 0; q < r[0]; ++q) {
      a.push_back(r[1 + q]);
      b.push_back(st[(size_t)k * ms + q]);
    }
    id.push_back(a);
    nu.push_back(b);
  }
}

struct Module {
  hipModule_t mod = nullptr;
  hipFunction_t f32 = nullptr, f64 = nullptr, adj = nullptr;
  int blocks_per_cu32 = 1, blocks_per_cu64 = 1;
};
std::mutex g_mu;
std::map<std::pair<int, std::string>, Module> g_mods;  // (device, source) -> module

__global__ void lane_zero_cursors(unsigned long long* c, int n) {
  if ((int)threadIdx.x < n) c[16 * threadIdx.x] = 0ull;
}

}  // namespace

int lane_supported(const pfchem_network* net, const pfchem_options* opt, std::string& why) {
  const int naq = net->naqcomp, nimm = net->nimmobile;
  if (net->somdec) { why = "SomDec sandbox"; return 0; }
  if (opt->variable_ncomp) { why = "per-cell active counts"; return 0; }
  if (naq < 1 || naq > kLaneMaxAq) { why = "aqueous primaries outside 1.." + std::to_string(kLaneMaxAq); return 0; }
  if (naq + nimm > 32) { why = "more than 32 unknowns"; return 0; }
  if (net->neqcplx > kLaneMaxCplx) { why = "too many complexes"; return 0; }
  if (net->nkinmnrl > kLaneMaxMnrl) { why = "too many kinetic minerals"; return 0; }
  // per-lane LDS slots: 2 (naq + nimm) + naq + 3 nkinmnrl doubles plus one
  // per activity-coefficient class of the complexes, x 256 lanes
  int ncls = 0;
  {
    std::vector<std::pair<double, double>> cls;
    for (int k = 0; k < net->neqcplx; ++k) {
      const double z = net->eqcplx_Z ? net->eqcplx_Z[k] : 0.0;
      if (!(std::fabs(z) > 1.0e-10)) continue;
      const std::pair<double, double> key(z * z, net->eqcplx_a0 ? net->eqcplx_a0[k] : 0.0);
      bool seen = false;
      for (const auto& c : cls) seen = seen || c == key;
      if (!seen) cls.push_back(key);
    }
    ncls = (int)cls.size();
  }
  if (2 * (naq + nimm) + naq + 3 * net->nkinmnrl + ncls > kLaneMaxLdsFields) {
    why = "per-cell state exceeds the LDS slots";
    return 0;
  }
  if (net->neqcplx > 0 && net->eqcplx_maxsp > 8) { why = "complex rows longer than 8"; return 0; }
  if (net->nkinmnrl > 0 && net->kinmnrl_maxsp > 8) { why = "mineral rows longer than 8"; return 0; }
  if (net->kinmnrl_num_prefactors)
    for (int m = 0; m < net->nkinmnrl; ++m)
      if (net->kinmnrl_num_prefactors[m] > 0) {
        // their Jacobian terms are not symmetric
        why = "mineral rate prefactors";
        return 0;
      }
  if (opt->act_coef_update_algorithm == PFCHEM_ACT_COEF_ALGORITHM_NEWTON &&
      opt->act_coef_update_frequency == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER) {
    why = "NEWTON activity algorithm inside the Newton iteration";
    return 0;
  }
  const int nc = net->neqcplx, nm = net->nkinmnrl;
  if (!finite_all(net->eqcplx_logK, nc) || !finite_all(net->eqcplxstoich, (size_t)nc * net->eqcplx_maxsp) ||
      !finite_all(net->kinmnrl_logK, nm) || !finite_all(net->kinmnrl_rate_constant, nm)) {
    why = "non-finite network constants";
    return 0;
  }
  return 1;
}

std::string lane_source(const pfchem_network* net, const pfchem_options* opt) {
  const int naq = net->naqcomp, nimm = net->nimmobile, nc = net->neqcplx, nm = net->nkinmnrl;
  Gen g;
  // (hiprtc provides the HIP device runtime without an include)
  g.line("// generated: one reaction network and its solver options as constants");
  g.line("struct Net {");
  g.i("NAQ", naq);
  g.i("NIMM", nimm);
  g.i("NC", nc);
  g.i("NM", nm);
  g.i("NDECAY", net->ndecay_rxn);
  g.i("H2O_ID", net->h2o_aq_id);
  g.i("FULL_GEOCHEM", opt->use_full_geochemistry ? 1 : 0);
  g.i("LOG_FORM", opt->use_log_formulation ? 1 : 0);
  g.i("ACT_FREQ", opt->act_coef_update_frequency);
  g.i("ACT_H2O", opt->use_activity_h2o ? 1 : 0);
  g.i("TOTAL_AS_GUESS", opt->use_total_as_guess ? 1 : 0);
  g.i("ISOTHERMAL", opt->isothermal ? 1 : 0);
  g.i("MAX_IT", opt->maximum_reaction_iterations);
  g.i("MAX_CUTS", opt->maximum_reaction_cuts);
  g.i("MAX_CALLS", opt->maximum_rreact_calls > 0 ? opt->maximum_rreact_calls : 10000);
  g.d("MAX_DLNC", opt->max_dlnC_rreact);
  g.d("TOL_ABS", opt->max_residual_tolerance);
  g.d("TOL_REL", opt->max_rel_residual_tolerance);
  g.d("TOL_CHANGE", opt->max_relative_change_tolerance);
  g.d("DH_A", net->debyeA);
  g.d("DH_B", net->debyeB);
  g.d("DH_BDOT", net->debyeBdot);
  g.dvec("ZP", std::vector<double>(net->primary_spec_Z, net->primary_spec_Z + naq));
  g.dvec("A0P", std::vector<double>(net->primary_spec_a0, net->primary_spec_a0 + naq));
  // complexes
  std::vector<int> cn;
  std::vector<std::vector<int>> cid;
  std::vector<std::vector<double>> cnu;
  if (nc) rows(net->eqcplxspecid, net->eqcplxstoich, nc, net->eqcplx_maxsp, cn, cid, cnu);
  auto vec = [](const double* p, int n, double dflt) {
    std::vector<double> v(n, dflt);
    if (p)
      for (int k = 0; k < n; ++k) v[k] = p[k];
    return v;
  };
  auto coef = [](const double* p, int n) {
    std::vector<std::vector<double>> v(n, std::vector<double>(5, 0.0));
    if (p)
      for (int k = 0; k < n; ++k)
        for (int q = 0; q < 5; ++q) v[k][q] = p[5 * k + q];
    return v;
  };
  g.ivec("CPLX_N", cn);
  g.imat("CPLX_ID", cid, net->eqcplx_maxsp);
  g.dmat("CPLX_NU", cnu, net->eqcplx_maxsp);
  g.dvec("CPLX_H2O", vec(net->eqcplxh2ostoich, nc, 0.0));
  g.dvec("CPLX_LOGK", vec(net->eqcplx_logK, nc, 0.0));
  g.i("HAS_CPLX_COEF", net->eqcplx_logKcoef ? 1 : 0);
  g.dmat("CPLX_COEF", coef(net->eqcplx_logKcoef, nc), 5);
  g.dvec("CPLX_Z", vec(net->eqcplx_Z, nc, 0.0));
  g.dvec("CPLX_A0", vec(net->eqcplx_a0, nc, 0.0));
  // activity-coefficient classes of the complexes: one per distinct
  // (Z^2, a0) of the charged ones (Debye-Hueckel gives them equal
  // coefficients), -1 for neutral ones (coefficient 1)
  {
    std::vector<std::pair<double, double>> cls;
    std::vector<int> id(nc, -1), first(nc, 0);
    for (int k = 0; k < nc; ++k) {
      const double z = net->eqcplx_Z[k];
      if (!(std::fabs(z) > 1.0e-10)) continue;
      const std::pair<double, double> key(z * z, net->eqcplx_a0[k]);
      int q = -1;
      for (size_t j = 0; j < cls.size(); ++j)
        if (cls[j] == key) q = (int)j;
      if (q < 0) {
        q = (int)cls.size();
        cls.push_back(key);
        first[k] = 1;
      }
      id[k] = q;
    }
    g.i("NGCLS", (long)cls.size());
    g.ivec("CPLX_GCLS", id);
    g.ivec("CPLX_GFIRST", first);
  }
  // kinetic minerals
  std::vector<int> mn;
  std::vector<std::vector<int>> mid;
  std::vector<std::vector<double>> mnu;
  if (nm) rows(net->kinmnrlspecid, net->kinmnrlstoich, nm, net->kinmnrl_maxsp, mn, mid, mnu);
  g.ivec("MNRL_N", mn);
  g.imat("MNRL_ID", mid, net->kinmnrl_maxsp);
  g.dmat("MNRL_NU", mnu, net->kinmnrl_maxsp);
  g.dvec("MNRL_H2O", vec(net->kinmnrlh2ostoich, nm, 0.0));
  g.dvec("MNRL_LOGK", vec(net->kinmnrl_logK, nm, 0.0));
  g.i("HAS_MNRL_COEF", net->kinmnrl_logKcoef ? 1 : 0);
  g.dmat("MNRL_COEF", coef(net->kinmnrl_logKcoef, nm), 5);
  g.dvec("MNRL_K", vec(net->kinmnrl_rate_constant, nm, 0.0));
  g.dvec("MNRL_VM", vec(net->kinmnrl_molar_vol, nm, 0.0));
  g.i("HAS_EA", net->kinmnrl_activation_energy ? 1 : 0);
  g.dvec("MNRL_EA", vec(net->kinmnrl_activation_energy, nm, 0.0));
  g.i("HAS_TEMKIN", net->kinmnrl_Temkin_const ? 1 : 0);
  g.dvec("MNRL_TEMKIN", vec(net->kinmnrl_Temkin_const, nm, 1.0));
  g.i("HAS_MINSCALE", net->kinmnrl_min_scale_factor ? 1 : 0);
  g.dvec("MNRL_MINSCALE", vec(net->kinmnrl_min_scale_factor, nm, 1.0));
  g.i("HAS_AFFPOWER", net->kinmnrl_affinity_power ? 1 : 0);
  g.dvec("MNRL_AFFPOWER", vec(net->kinmnrl_affinity_power, nm, 1.0));
  g.i("HAS_THRESHOLD", net->kinmnrl_affinity_threshold ? 1 : 0);
  g.dvec("MNRL_THRESHOLD", vec(net->kinmnrl_affinity_threshold, nm, 0.0));
  g.i("HAS_LIMITER", net->kinmnrl_rate_limiter ? 1 : 0);
  g.dvec("MNRL_LIMITER", vec(net->kinmnrl_rate_limiter, nm, 0.0));
  g.i("HAS_IRREV", net->kinmnrl_irreversible ? 1 : 0);
  std::vector<int> irr(nm, 0);
  if (net->kinmnrl_irreversible)
    for (int m = 0; m < nm; ++m) irr[m] = net->kinmnrl_irreversible[m];
  g.ivec("MNRL_IRREV", irr);
  // immobile decay
  std::vector<int> did;
  std::vector<double> dk;
  for (int r = 0; r < net->ndecay_rxn; ++r) {
    did.push_back(net->decayspecid[r]);
    dk.push_back(net->decay_rate_constant[r]);
  }
  g.ivec("DECAY_ID", did);
  g.dvec("DECAY_K", dk);
  g.line("};");
  g.s += kLaneTemplate;
  return g.s;
}

int lane_compile(const std::string& src, int device, std::string& err) {
  std::lock_guard<std::mutex> lk(g_mu);
  const auto key = std::make_pair(device, src);
  if (g_mods.count(key)) return PFCHEM_OK;
  std::vector<char> code;
  const int rc = lane_compile_only(src, &code, err);
  if (rc) return rc;
  Module m;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipModuleLoadData(&m.mod, code.data());
  // a step module holds the two step kernels, an adjoint module the adjoint
  const bool adjoint = src.compare(0, sizeof(kLaneAdjointDefine) - 1, kLaneAdjointDefine) == 0;
  if (e == hipSuccess && adjoint) e = hipModuleGetFunction(&m.adj, m.mod, "pfchem_lane_adjoint");
  if (e == hipSuccess && !adjoint) e = hipModuleGetFunction(&m.f32, m.mod, "pfchem_lane_step_f32");
  if (e == hipSuccess && !adjoint) e = hipModuleGetFunction(&m.f64, m.mod, "pfchem_lane_step_f64");
  if (e == hipSuccess && !adjoint)
    e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&m.blocks_per_cu32, m.f32, 256, 0);
  if (e == hipSuccess && !adjoint)
    e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&m.blocks_per_cu64, m.f64, 256, 0);
  if (e != hipSuccess) {
    err = std::string("lane kernel load: ") + hipGetErrorString(e);
    return PFCHEM_ERR_HIP;
  }
  if (m.blocks_per_cu32 < 1) m.blocks_per_cu32 = 1;
  if (m.blocks_per_cu64 < 1) m.blocks_per_cu64 = 1;
  g_mods[key] = m;
  return PFCHEM_OK;
}

int lane_compile_only(const std::string& src, std::vector<char>* code, std::string& err) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "pfchem_lane.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return PFCHEM_ERR_HIP;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    err = "lane kernel compilation failed: " + log.substr(0, 4000);
    hiprtcDestroyProgram(&prog);
    return PFCHEM_ERR_HIP;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  if (code) {
    code->resize(cs);
    hiprtcGetCode(prog, code->data());
  }
  hiprtcDestroyProgram(&prog);
  return PFCHEM_OK;
}

hipError_t lane_launch(const std::string& src, int device, int f32, const LaneArgs& args,
                       int num_cus, int blocks_per_cu, unsigned long long* cursors,
                       hipStream_t stream) {
  Module m;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mods.find(std::make_pair(device, src));
    if (it == g_mods.end()) return hipErrorInvalidHandle;
    m = it->second;
  }
  const int occ = f32 ? m.blocks_per_cu32 : m.blocks_per_cu64;
  const int bpc = blocks_per_cu > 0 && blocks_per_cu < occ ? blocks_per_cu : occ;
  int64_t blocks = (int64_t)num_cus * bpc;
  // no more waves than cells
  const int64_t max_blocks = (args.ncell + 255) / 256;
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks > args.nxcd) blocks -= blocks % args.nxcd;
  if (blocks < 1) blocks = 1;
  lane_zero_cursors<<<1, 64, 0, stream>>>(cursors, args.nxcd);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  LaneArgs a = args;
  a.cursors = cursors;
  void* params[] = {&a};
  return hipModuleLaunchKernel(f32 ? m.f32 : m.f64, (unsigned)blocks, 1, 1, 256, 1, 1, 0, stream,
                               params, nullptr);
}

hipError_t lane_adjoint_launch(const std::string& src, int device, const LaneSensArgs& args,
                               hipStream_t stream) {
  Module m;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mods.find(std::make_pair(device, src));
    if (it == g_mods.end()) return hipErrorInvalidHandle;
    m = it->second;
  }
  const int64_t blocks = (args.ncell + 255) / 256;
  if (blocks < 1) return hipSuccess;
  LaneSensArgs a = args;
  void* params[] = {&a};
  return hipModuleLaunchKernel(m.adj, (unsigned)blocks, 1, 1, 256, 1, 1, 0, stream, params,
                               nullptr);
}

int lane_occupancy(const std::string& src, int device, int f32) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_mods.find(std::make_pair(device, src));
  if (it == g_mods.end()) return 0;
  return f32 ? it->second.blocks_per_cu32 : it->second.blocks_per_cu64;
}

}  // namespace pfchem
