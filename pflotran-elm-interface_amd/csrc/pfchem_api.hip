// pfchem_api.hip -- the C ABI (include/pfchem.h): context, network upload,
// argument checks and kernel dispatch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pfchem.h"
#include "pfchem_device.hpp"
#include "pfchem_kernels.hpp"
#include "pfchem_lane.h"
#include "pfchem_transport.h"
#include "pfchem_tune.h"

using namespace pfchem;

struct pfchem_ctx {
  int device = 0;
  int num_cus = 256;
  DevNet dnet{};
  void* dbuf = nullptr;  // device copy of the network arrays
  // per-XCD work cursors (128-byte stride), one block of kCursorWords per
  // descriptor slot: a launch uses its slot's block, so launches on
  // different streams never share cursors
  unsigned long long* counter = nullptr;
  int nxcd = 8;                           // XCDs of the device
  int max_lds = 65536;                    // LDS bytes per workgroup
  DevNet* dnet_dev = nullptr;             // device copy of dnet
  DevState* dstate_ring = nullptr;        // device slots for per-launch state descriptors
  DevState* hstate_ring = nullptr;        // pinned host mirror the slots are copied from
  hipEvent_t ring_done[256] = {};         // per slot: recorded after the launch reading it
  bool ring_used[256] = {};               // the slot's event has been recorded
  int ring_next = 0;
  LaunchCfg launch[3] = {{1, 1, 128, 0}, {16, 1, 256, 4}, {16, 1, 256, 4}};
  // per-cell active counts: team shape per size bucket, per batched variant
  LaunchCfg vlaunch[2][kBuckets];
  int32_t* buckets = nullptr;  // kBuckets lists of ncell entries, then the counts
  int64_t bucket_cap = 0;
  int32_t* esc_list = nullptr;  // cells the one-lane pass handed over, then their count
  int64_t esc_cap = 0;
  // steps that use the scratch lists held here (escalated cells, status
  // scratch, size buckets) are ordered on the device: each records this
  // event after its last launch and the next one's stream waits for it
  hipEvent_t scratch_done = nullptr;
  bool scratch_done_used = false;
  // pfchem_set_pass_timing: events around the one-lane pass of a step
  bool pass_timing = false;
  bool pass_recorded = false;
  hipEvent_t pass_t0 = nullptr, pass_t1 = nullptr;
  pfchem_network net{};   // sizes only (host pointers are not kept)
  pfchem_options opt{};
  bool has_somdec = false;
  // host staging (pfchem_step_host)
  void* stage = nullptr;
  size_t stage_bytes = 0;
  // sensitivities: weights / direction (device + pinned mirror), per-cell
  // gradient and status scratch, gradient sum
  double* sens_in = nullptr;     // device: w_vf[nm], w_sec[nc], dp[nm+nc]
  double* sens_in_h = nullptr;   // pinned host mirror
  double* sens_sum = nullptr;    // device: sum of per-cell gradients
  double* sens_sum_h = nullptr;  // pinned host mirror
  hipEvent_t sens_copied = nullptr;
  void* sens_work = nullptr;     // per-cell gradients + status when not given
  size_t sens_work_bytes = 0;
  // constraint speciation: DevConstraint + reaction table + failure count
  void* con_buf = nullptr;
  size_t con_bytes = 0;
  // network-specialized lane kernel (pfchem_lane.hpp): generated source,
  // whether the network qualifies, whether it is compiled on the device
  std::string lane_src;
  bool lane_ok = false;
  bool lane_ready = false;
  std::string lane_why;
  std::string lane_adj_src;  // the adjoint module's source (compiled at first use)
  bool lane_adj_ready = false;
  LaunchCfg team_default{16, 1, 256, 4};  // team shape of the FP64 rescue pass
  int32_t* status_scratch = nullptr;      // per-cell status when the caller gives none
  int64_t status_cap = 0;
  // launch-shape cache identity: device architecture and a hash of the
  // network arrays and options
  std::string arch = "gfx";
  uint64_t net_hash = 0;
  // OSRT transport: connection topology (pfchem_transport_setup)
  TransportPlan* tplan = nullptr;
  std::string err;
};

namespace {

int fail(pfchem_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(pfchem_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, PFCHEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// byte buffer with 256-B aligned sub-arrays; offsets resolved after upload
struct Packer {
  std::vector<char> bytes;
  size_t add(const void* p, size_t nbytes) {
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + (nbytes ? nbytes : 1));
    if (p && nbytes) std::memcpy(bytes.data() + off, p, nbytes);
    return off;
  }
};

const size_t kNone = ~size_t(0);

// Fixed device array for pfchem_equilibrate_static: the constraint
// descriptor, its reaction table and the failure count.  Each GPU holds its
// own copy for the life of the process; host code reaches it only through
// the address hipGetSymbolAddress returns on that device.
constexpr size_t kStaticScratchBytes = 32768;
static __device__ __align__(256) char scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// resolve the array's address on `device` (current); once per device
hipError_t resolve_static_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[device] = p;
  return e;
}

void* static_scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}
// Launches take their state descriptor from a ring of device slots, filled
// from a pinned host mirror by a copy on the launch's stream.  An event is
// recorded after every launch that reads a slot; before the slot (and its
// host mirror) is rewritten, that event must have completed, whatever
// stream the earlier launch used.
const int kStateRing = 256;
constexpr size_t kCursorWords = 1024;  // 8 KiB of cursors per descriptor slot
// handed-over cells above which the FP64 one-lane kernel runs before the
// team pass (the team pass alone takes a short list faster: its time is one
// cell's latency)
constexpr int kLaneListMin = 4096;

unsigned long long* slot_cursors(pfchem_ctx* ctx, int k) {
  return ctx->counter + (size_t)k * kCursorWords;
}

// before a scratch list is freed / regrown: every step that used it is done
void scratch_quiesce(pfchem_ctx* ctx) {
  if (ctx->scratch_done_used) (void)hipEventSynchronize(ctx->scratch_done);
}
// a step that uses the scratch lists starts after the previous such step
hipError_t scratch_order(pfchem_ctx* ctx, hipStream_t s) {
  return ctx->scratch_done_used ? hipStreamWaitEvent(s, ctx->scratch_done, 0) : hipSuccess;
}
hipError_t scratch_release(pfchem_ctx* ctx, hipStream_t s) {
  hipError_t e = hipEventRecord(ctx->scratch_done, s);
  if (e == hipSuccess) ctx->scratch_done_used = true;
  return e;
}

DevState* acquire_slot(pfchem_ctx* ctx, const DevState& d, hipStream_t s, int* idx,
                       hipError_t* e) {
  const int k = ctx->ring_next;
  ctx->ring_next = (k + 1) % kStateRing;
  *idx = k;
  if (ctx->ring_used[k]) {
    *e = hipEventSynchronize(ctx->ring_done[k]);
    if (*e != hipSuccess) return nullptr;
    ctx->ring_used[k] = false;
  }
  ctx->hstate_ring[k] = d;
  *e = hipMemcpyAsync(ctx->dstate_ring + k, ctx->hstate_ring + k, sizeof(DevState),
                      hipMemcpyHostToDevice, s);
  return ctx->dstate_ring + k;
}

hipError_t release_slot(pfchem_ctx* ctx, int k, hipStream_t s) {
  const hipError_t e = hipEventRecord(ctx->ring_done[k], s);
  ctx->ring_used[k] = e == hipSuccess;
  return e;
}

// SomDec descriptor: ids in range
int check_somdec(const pfchem_network* net, std::string& msg) {
  const pfchem_somdec* sd = net->somdec;
  const int naq = net->naqcomp, nimm = net->nimmobile, nr = sd->nrxn;
  auto aq = [&](int id, bool req) { return req ? (id >= 0 && id < naq) : (id >= -1 && id < naq); };
  auto im = [&](int id, bool req) { return req ? (id >= 0 && id < nimm) : (id >= -1 && id < nimm); };
  if (nr < 0 || nr > 32) {
    msg = "somdec: 0..32 reactions";
    return PFCHEM_ERR_UNSUPPORTED;
  }
  if (naq + nimm > 32) {
    msg = "somdec: at most 32 primary + immobile species";
    return PFCHEM_ERR_UNSUPPORTED;
  }
  if (!aq(sd->species_id_co2, true) || !aq(sd->species_id_nh4, true) ||
      !aq(sd->species_id_no3, false) || !aq(sd->species_id_n2o, false) ||
      !aq(sd->species_id_proton, false) || !im(sd->species_id_hr, false) ||
      !im(sd->species_id_nmin, false) || !im(sd->species_id_nimm, false) ||
      !im(sd->species_id_nimp, false) || !im(sd->species_id_ngasmin, false)) {
    msg = "somdec: species id out of range";
    return PFCHEM_ERR_INVALID;
  }
  if (nr > 0 && (!sd->rate_constant || !sd->rate_decomposition || !sd->rate_ad_factor ||
                 !sd->upstream_c_id || !sd->upstream_n_id || !sd->upstream_hr_id ||
                 !sd->upstream_nmin_id || !sd->upstream_nimp_id || !sd->upstream_nimm_id ||
                 !sd->upstream_nc || !sd->upstream_is_varycn || !sd->n_downstream_pools ||
                 !sd->downstream_c_id || !sd->downstream_n_id || !sd->downstream_stoich ||
                 !sd->downstream_nc || !sd->downstream_is_varycn || !sd->mineral_c_stoich ||
                 !sd->mineral_n_stoich || !sd->temperature_response_function || !sd->Q10 ||
                 !sd->Ea || !sd->moisture_response_function || !sd->Ox_response_function ||
                 !sd->Ox_half_saturation || !sd->n_monod || !sd->monod_specid ||
                 !sd->monod_half_saturation || !sd->monod_threshold ||
                 !sd->monod_pool_normalized || !sd->n_inhibition || !sd->inhibition_specid ||
                 !sd->inhibition_type || !sd->inhibition_constant || !sd->inhibition_constant2)) {
    msg = "somdec: missing per-reaction array";
    return PFCHEM_ERR_INVALID;
  }
  for (int r = 0; r < nr; ++r) {
    if (!im(sd->upstream_c_id[r], true) || !im(sd->upstream_n_id[r], false) ||
        !im(sd->upstream_hr_id[r], false) || !im(sd->upstream_nmin_id[r], false) ||
        !im(sd->upstream_nimp_id[r], false) || !im(sd->upstream_nimm_id[r], false) ||
        (sd->upstream_is_varycn[r] && sd->upstream_n_id[r] < 0) ||
        sd->n_downstream_pools[r] < 0 || sd->n_downstream_pools[r] > PFCHEM_SOMDEC_MAX_DOWN ||
        sd->n_monod[r] < 0 || sd->n_monod[r] > PFCHEM_SOMDEC_MAX_MONOD ||
        sd->n_inhibition[r] < 0 || sd->n_inhibition[r] > PFCHEM_SOMDEC_MAX_INHIB) {
      msg = "somdec: bad reaction " + std::to_string(r);
      return PFCHEM_ERR_INVALID;
    }
    for (int j = 0; j < sd->n_downstream_pools[r]; ++j) {
      const int k = r * PFCHEM_SOMDEC_MAX_DOWN + j;
      if (!im(sd->downstream_c_id[k], false) || !im(sd->downstream_n_id[k], false) ||
          (sd->downstream_is_varycn[k] && sd->downstream_n_id[k] < 0)) {
        msg = "somdec: bad downstream pool in reaction " + std::to_string(r);
        return PFCHEM_ERR_INVALID;
      }
    }
    for (int q = 0; q < sd->n_monod[r]; ++q) {
      const int id = sd->monod_specid[r * PFCHEM_SOMDEC_MAX_MONOD + q];
      if (id >= naq || id < -nimm) {
        msg = "somdec: bad Monod species in reaction " + std::to_string(r);
        return PFCHEM_ERR_INVALID;
      }
    }
    for (int q = 0; q < sd->n_inhibition[r]; ++q) {
      const int id = sd->inhibition_specid[r * PFCHEM_SOMDEC_MAX_INHIB + q];
      if (id >= naq || id < -nimm) {
        msg = "somdec: bad inhibition species in reaction " + std::to_string(r);
        return PFCHEM_ERR_INVALID;
      }
    }
    // oxygen codes: aqueous id + 1 or -(immobile id + 1), 0 = none
    if (sd->Ox_specid) {
      const int o = sd->Ox_specid[r];
      if (o > naq || o < -nimm) {
        msg = "somdec: bad oxidizer species in reaction " + std::to_string(r);
        return PFCHEM_ERR_INVALID;
      }
    }
  }
  if (sd->species_o2 > naq || sd->species_o2 < -nimm) {
    msg = "somdec: bad O2 species";
    return PFCHEM_ERR_INVALID;
  }
  return 0;
}

int check_reactions(const int32_t* ids, int n, int maxsp, int naq, const char* what,
                    std::string& msg) {
  for (int k = 0; k < n; ++k) {
    const int32_t* row = ids + (size_t)k * (maxsp + 1);
    if (row[0] < 0 || row[0] > maxsp) {
      msg = std::string(what) + ": bad species count in reaction " + std::to_string(k);
      return 1;
    }
    for (int i = 0; i < row[0]; ++i)
      if (row[1 + i] < 0 || row[1 + i] >= naq) {
        msg = std::string(what) + ": species id out of range in reaction " + std::to_string(k);
        return 1;
      }
  }
  return 0;
}

}  // namespace

extern "C" {

const char* pfchem_strerror(int code) {
  switch (code) {
    case PFCHEM_OK: return "ok";
    case PFCHEM_ERR_INVALID: return "invalid argument";
    case PFCHEM_ERR_HIP: return "HIP runtime error";
    case PFCHEM_ERR_UNSUPPORTED: return "unsupported network, option or size";
    case PFCHEM_ERR_NOMEM: return "device allocation failed";
    case PFCHEM_ERR_CELLS: return "one or more cells failed (see per-cell status)";
    default: return "unknown pfchem error";
  }
}

const char* pfchem_last_error(const pfchem_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int pfchem_create(const pfchem_network* net, const pfchem_options* opt, int device,
                  pfchem_ctx** out) {
  if (!out) return PFCHEM_ERR_INVALID;
  *out = nullptr;
  if (!net || !opt) return PFCHEM_ERR_INVALID;
  pfchem_ctx* ctx = new pfchem_ctx();
  std::string msg;
  const int naq = net->naqcomp, nimm = net->nimmobile;
  if (naq < 0 || nimm < 0 || naq + nimm < 1) {
    delete ctx;
    return PFCHEM_ERR_INVALID;
  }
  if (naq + nimm > 64 || net->nkinmnrl > 16 || net->neqcplx > 4096) {
    delete ctx;
    return PFCHEM_ERR_UNSUPPORTED;
  }
  if ((naq && (!net->primary_spec_Z || !net->primary_spec_a0)) ||
      (net->neqcplx > 0 && (!net->eqcplxspecid || !net->eqcplxstoich || !net->eqcplx_logK ||
                            !net->eqcplx_Z || !net->eqcplx_a0)) ||
      (net->nkinmnrl > 0 && (!net->kinmnrlspecid || !net->kinmnrlstoich || !net->kinmnrl_logK ||
                             !net->kinmnrl_rate_constant || !net->kinmnrl_molar_vol)) ||
      (net->ndecay_rxn > 0 && (!net->decayspecid || !net->decay_rate_constant))) {
    delete ctx;
    return PFCHEM_ERR_INVALID;
  }
  if (check_reactions(net->eqcplxspecid, net->neqcplx, net->eqcplx_maxsp, naq, "complex", msg) ||
      check_reactions(net->kinmnrlspecid, net->nkinmnrl, net->kinmnrl_maxsp, naq, "mineral", msg)) {
    delete ctx;
    return PFCHEM_ERR_INVALID;
  }
  for (int r = 0; r < net->ndecay_rxn; ++r)
    if (net->decayspecid[r] < 0 || net->decayspecid[r] >= nimm) {
      delete ctx;
      return PFCHEM_ERR_INVALID;
    }
  if (net->kinmnrl_num_prefactors && net->nkinmnrl > 0) {
    // rate prefactors: strides, counts and primary-species ids
    const int mp = net->kinmnrl_max_pref, msp = net->kinmnrl_max_pref_spec;
    if (mp < 1 || mp > 8 || msp < 1 || msp > 8 || !net->kinmnrl_prefactor_id ||
        !net->kinmnrl_pref_alpha || !net->kinmnrl_pref_beta || !net->kinmnrl_pref_atten_coef ||
        !net->kinmnrl_pref_rate) {
      delete ctx;
      return PFCHEM_ERR_INVALID;
    }
    for (int m = 0; m < net->nkinmnrl; ++m) {
      const int np = net->kinmnrl_num_prefactors[m];
      if (np < 0 || np > mp) {
        delete ctx;
        return PFCHEM_ERR_INVALID;
      }
      for (int p = 0; p < np; ++p) {
        const int32_t* pid = net->kinmnrl_prefactor_id + ((size_t)m * mp + p) * (msp + 1);
        if (pid[0] < 0 || pid[0] > msp) {
          delete ctx;
          return PFCHEM_ERR_INVALID;
        }
        for (int q = 0; q < pid[0]; ++q) {
          if (pid[1 + q] < 0) {  // a complex (the reference's negative ids)
            delete ctx;
            return PFCHEM_ERR_UNSUPPORTED;
          }
          if (pid[1 + q] >= naq) {
            delete ctx;
            return PFCHEM_ERR_INVALID;
          }
        }
      }
    }
  }
  if (net->somdec) {
    const int rc = check_somdec(net, msg);
    if (rc) {
      delete ctx;
      return rc;
    }
  }
  // per-cell active primaries: aqueous networks without the sandbox
  if (opt->variable_ncomp && (nimm > 0 || net->somdec)) {
    delete ctx;
    return PFCHEM_ERR_UNSUPPORTED;
  }
  if (opt->max_dlnC_rreact <= 0.0 || opt->maximum_reaction_iterations < 1 ||
      opt->maximum_reaction_cuts < 0 || opt->maximum_rreact_calls < 0) {
    delete ctx;
    return PFCHEM_ERR_INVALID;
  }

  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete ctx;
    return PFCHEM_ERR_HIP;
  }
  ctx->device = device;
  if (resolve_static_scratch(device) != hipSuccess) {
    delete ctx;
    return PFCHEM_ERR_HIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    ctx->num_cus = cus;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
          hipSuccess && lds >= 65536)
    ctx->max_lds = lds;
  // MI355X / MI300-class parts have 32 CUs per XCD
  ctx->nxcd = ctx->num_cus >= 64 && ctx->num_cus % 32 == 0 ? ctx->num_cus / 32 : 1;

  const int nc = net->neqcplx, nm = net->nkinmnrl;
  Packer P;
  auto addd = [&](const double* p, size_t n) { return p ? P.add(p, n * sizeof(double)) : kNone; };
  auto addi = [&](const int32_t* p, size_t n) { return p ? P.add(p, n * sizeof(int32_t)) : kNone; };
  size_t oZ = addd(net->primary_spec_Z, naq), oa0 = addd(net->primary_spec_a0, naq);
  size_t oci = addi(net->eqcplxspecid, (size_t)nc * (net->eqcplx_maxsp + 1));
  size_t ocs = addd(net->eqcplxstoich, (size_t)nc * net->eqcplx_maxsp);
  size_t och = addd(net->eqcplxh2ostoich, nc), ock = addd(net->eqcplx_logK, nc);
  size_t ockc = addd(net->eqcplx_logKcoef, (size_t)nc * 5);
  size_t ocZ = addd(net->eqcplx_Z, nc), oca = addd(net->eqcplx_a0, nc);
  size_t omi = addi(net->kinmnrlspecid, (size_t)nm * (net->kinmnrl_maxsp + 1));
  size_t oms = addd(net->kinmnrlstoich, (size_t)nm * net->kinmnrl_maxsp);
  size_t omh = addd(net->kinmnrlh2ostoich, nm), omk = addd(net->kinmnrl_logK, nm);
  size_t omkc = addd(net->kinmnrl_logKcoef, (size_t)nm * 5);
  size_t omr = addd(net->kinmnrl_rate_constant, nm), omE = addd(net->kinmnrl_activation_energy, nm);
  size_t omV = addd(net->kinmnrl_molar_vol, nm), omT = addd(net->kinmnrl_Temkin_const, nm);
  size_t omS = addd(net->kinmnrl_min_scale_factor, nm), omP = addd(net->kinmnrl_affinity_power, nm);
  size_t omt = addd(net->kinmnrl_affinity_threshold, nm), oml = addd(net->kinmnrl_rate_limiter, nm);
  size_t omI = addi(net->kinmnrl_irreversible, nm);
  // rate prefactors
  const bool pref = net->kinmnrl_num_prefactors && nm > 0;
  const size_t pmp = pref ? (size_t)net->kinmnrl_max_pref : 0;
  const size_t pms = pref ? (size_t)net->kinmnrl_max_pref_spec : 0;
  size_t opn = pref ? addi(net->kinmnrl_num_prefactors, nm) : kNone;
  size_t opi = pref ? addi(net->kinmnrl_prefactor_id, nm * pmp * (pms + 1)) : kNone;
  size_t opa = pref ? addd(net->kinmnrl_pref_alpha, nm * pmp * pms) : kNone;
  size_t opb = pref ? addd(net->kinmnrl_pref_beta, nm * pmp * pms) : kNone;
  size_t opt_ = pref ? addd(net->kinmnrl_pref_atten_coef, nm * pmp * pms) : kNone;
  size_t opr = pref ? addd(net->kinmnrl_pref_rate, nm * pmp) : kNone;
  size_t ope = pref ? addd(net->kinmnrl_pref_activation_energy, nm * pmp) : kNone;
  size_t odi = addi(net->decayspecid, net->ndecay_rxn), odk = addd(net->decay_rate_constant, net->ndecay_rxn);
  // largest species id per complex / mineral (per-cell active primaries)
  size_t ocmx = kNone, ommx = kNone;
  if (opt->variable_ncomp) {
    auto maxids = [](const int32_t* ids, int nr, int ms) {
      std::vector<int32_t> v(nr > 0 ? nr : 1, -1);
      for (int r = 0; r < nr; ++r) {
        const int32_t* row = ids + (size_t)r * (ms + 1);
        for (int a = 1; a <= row[0]; ++a) v[r] = row[a] > v[r] ? row[a] : v[r];
      }
      return v;
    };
    const std::vector<int32_t> cm = maxids(net->eqcplxspecid, nc, net->eqcplx_maxsp);
    const std::vector<int32_t> mm = maxids(net->kinmnrlspecid, nm, net->kinmnrl_maxsp);
    ocmx = P.add(cm.data(), cm.size() * sizeof(int32_t));
    ommx = P.add(mm.data(), mm.size() * sizeof(int32_t));
  }
  // SomDec per-reaction arrays
  const pfchem_somdec* sdh = net->somdec;
  const int nr = sdh ? sdh->nrxn : 0;
  const size_t nrd = (size_t)nr * PFCHEM_SOMDEC_MAX_DOWN, nrm = (size_t)nr * PFCHEM_SOMDEC_MAX_MONOD;
  const size_t nri = (size_t)nr * PFCHEM_SOMDEC_MAX_INHIB;
  size_t so[38];
  std::vector<int32_t> ox_none;  // Ox_specid when the descriptor has none
  if (sdh) {
    int q = 0;
    so[q++] = addd(sdh->rate_constant, nr);
    so[q++] = addd(sdh->rate_decomposition, nr);
    so[q++] = addd(sdh->rate_ad_factor, nr);
    so[q++] = addi(sdh->upstream_c_id, nr);
    so[q++] = addi(sdh->upstream_n_id, nr);
    so[q++] = addi(sdh->upstream_hr_id, nr);
    so[q++] = addi(sdh->upstream_nmin_id, nr);
    so[q++] = addi(sdh->upstream_nimp_id, nr);
    so[q++] = addi(sdh->upstream_nimm_id, nr);
    so[q++] = addd(sdh->upstream_nc, nr);
    so[q++] = addi(sdh->upstream_is_varycn, nr);
    so[q++] = addi(sdh->n_downstream_pools, nr);
    so[q++] = addi(sdh->downstream_c_id, nrd);
    so[q++] = addi(sdh->downstream_n_id, nrd);
    so[q++] = addd(sdh->downstream_stoich, nrd);
    so[q++] = addd(sdh->downstream_nc, nrd);
    so[q++] = addi(sdh->downstream_is_varycn, nrd);
    so[q++] = addd(sdh->mineral_c_stoich, nr);
    so[q++] = addd(sdh->mineral_n_stoich, nr);
    so[q++] = addi(sdh->temperature_response_function, nr);
    so[q++] = addd(sdh->Q10, nr);
    so[q++] = addd(sdh->Ea, nr);
    so[q++] = addi(sdh->moisture_response_function, nr);
    so[q++] = addi(sdh->Ox_response_function, nr);
    so[q++] = addd(sdh->Ox_half_saturation, nr);
    so[q++] = addi(sdh->n_monod, nr);
    so[q++] = addi(sdh->monod_specid, nrm);
    so[q++] = addd(sdh->monod_half_saturation, nrm);
    so[q++] = addd(sdh->monod_threshold, nrm);
    so[q++] = addi(sdh->monod_pool_normalized, nrm);
    so[q++] = addi(sdh->n_inhibition, nr);
    so[q++] = addi(sdh->inhibition_specid, nri);
    so[q++] = addi(sdh->inhibition_type, nri);
    so[q++] = addd(sdh->inhibition_constant, nri);
    so[q++] = addd(sdh->inhibition_constant2, nri);
    if (!sdh->Ox_specid) ox_none.assign(nr > 0 ? nr : 1, 0);
    so[q++] = addi(sdh->Ox_specid ? sdh->Ox_specid : ox_none.data(), nr);
  }

  {
    // FNV-1a over the packed network, its sizes and the options
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void* p, size_t n) {
      const unsigned char* b = static_cast<const unsigned char*>(p);
      for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(P.bytes.data(), P.bytes.size());
    const int32_t sizes[8] = {naq, nimm, nc, nm, net->ndecay_rxn, net->eqcplx_maxsp,
                              net->kinmnrl_maxsp, sdh ? sdh->nrxn : -1};
    mix(sizes, sizeof(sizes));
    mix(opt, sizeof(*opt));
    if (sdh) {
      const double sd_scalars[3] = {sdh->x0eps, sdh->n2o_frac_mineralization,
                                    sdh->inhibition_nh4_no3};
      const int32_t sd_ids[11] = {sdh->species_id_co2,    sdh->species_id_nh4,
                                  sdh->species_id_no3,    sdh->species_id_n2o,
                                  sdh->species_id_proton, sdh->species_id_hr,
                                  sdh->species_id_nmin,   sdh->species_id_nimm,
                                  sdh->species_id_nimp,   sdh->species_id_ngasmin,
                                  sdh->species_o2};
      mix(sd_scalars, sizeof(sd_scalars));
      mix(sd_ids, sizeof(sd_ids));
    }
    ctx->net_hash = h;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      std::string a = prop.gcnArchName;
      ctx->arch = a.substr(0, a.find(':'));
    }
  }
  e = hipMalloc(&ctx->dbuf, P.bytes.size());
  if (e != hipSuccess) {
    delete ctx;
    return PFCHEM_ERR_NOMEM;
  }
  e = hipMemcpy(ctx->dbuf, P.bytes.data(), P.bytes.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&ctx->counter, sizeof(unsigned long long) * kCursorWords * kStateRing);
  if (e != hipSuccess) {
    (void)hipFree(ctx->dbuf);
    if (ctx->counter) (void)hipFree(ctx->counter);
    delete ctx;
    return PFCHEM_ERR_HIP;
  }
  char* base = (char*)ctx->dbuf;
  auto pd = [&](size_t o) { return o == kNone ? (const double*)nullptr : (const double*)(base + o); };
  auto pi = [&](size_t o) { return o == kNone ? (const int*)nullptr : (const int*)(base + o); };
  DevNet& d = ctx->dnet;
  d.naq = naq;
  d.nimm = nimm;
  d.ncomp = naq + nimm;
  d.nc = nc;
  d.nm = nm;
  d.ndecay = net->ndecay_rxn;
  d.cplx_ms = net->eqcplx_maxsp;
  d.mnrl_ms = net->kinmnrl_maxsp;
  d.h2o_aq_id = net->h2o_aq_id;
  d.log_form = opt->use_log_formulation;
  d.act_freq = opt->act_coef_update_frequency;
  d.act_alg = opt->act_coef_update_algorithm;
  d.act_h2o = opt->use_activity_h2o;
  d.total_as_guess = opt->use_total_as_guess;
  d.isothermal = opt->isothermal;
  d.max_it = opt->maximum_reaction_iterations;
  d.max_cuts = opt->maximum_reaction_cuts;
  d.max_calls = opt->maximum_rreact_calls > 0 ? opt->maximum_rreact_calls : kMaxRReactAttempts;
  d.full_geochem = opt->use_full_geochemistry;
  d.max_dlnC = opt->max_dlnC_rreact;
  d.tol_abs = opt->max_residual_tolerance;
  d.tol_rel = opt->max_rel_residual_tolerance;
  d.tol_change = opt->max_relative_change_tolerance;
  d.A = net->debyeA;
  d.B = net->debyeB;
  d.Bdot = net->debyeBdot;
  d.has_cplx_h2o = och != kNone;
  d.has_cplx_logKcoef = ockc != kNone;
  d.has_mnrl_h2o = omh != kNone;
  d.has_mnrl_logKcoef = omkc != kNone;
  d.has_temkin = omT != kNone;
  d.has_minscale = omS != kNone;
  d.has_affpower = omP != kNone;
  d.has_threshold = omt != kNone;
  d.has_limiter = oml != kNone;
  d.has_irrev = omI != kNone;
  d.has_Ea = omE != kNone;
  d.Zp = pd(oZ);
  d.a0p = pd(oa0);
  d.cplx_ids = pi(oci);
  d.cplx_st = pd(ocs);
  d.cplx_h2o = pd(och);
  d.cplx_logK = pd(ock);
  d.cplx_logKcoef = pd(ockc);
  d.cplx_Z = pd(ocZ);
  d.cplx_a0 = pd(oca);
  d.mnrl_ids = pi(omi);
  d.mnrl_st = pd(oms);
  d.mnrl_h2o = pd(omh);
  d.mnrl_logK = pd(omk);
  d.mnrl_logKcoef = pd(omkc);
  d.mnrl_k = pd(omr);
  d.mnrl_Ea = pd(omE);
  d.mnrl_Vm = pd(omV);
  d.mnrl_temkin = pd(omT);
  d.mnrl_minscale = pd(omS);
  d.mnrl_affpower = pd(omP);
  d.mnrl_threshold = pd(omt);
  d.mnrl_limiter = pd(oml);
  d.mnrl_irrev = pi(omI);
  d.decay_id = pi(odi);
  d.decay_k = pd(odk);
  d.has_pref = pref ? 1 : 0;
  d.pref_mp = (int)pmp;
  d.pref_msp = (int)pms;
  d.pref_n = pi(opn);
  d.pref_id = pi(opi);
  d.pref_alpha = pd(opa);
  d.pref_beta = pd(opb);
  d.pref_atten = pd(opt_);
  d.pref_rate = pd(opr);
  d.pref_Ea = pd(ope);
  d.variable = opt->variable_ncomp != 0;
  d.cplx_maxid = pi(ocmx);
  d.mnrl_maxid = pi(ommx);
  d.has_somdec = sdh != nullptr;
  if (sdh) {
    DevSomdec& sd = d.sd;
    sd.nrxn = nr;
    sd.co2 = sdh->species_id_co2;
    sd.nh4 = sdh->species_id_nh4;
    sd.no3 = sdh->species_id_no3;
    sd.n2o = sdh->species_id_n2o;
    sd.proton = sdh->species_id_proton;
    sd.hr = sdh->species_id_hr;
    sd.nmin = sdh->species_id_nmin;
    sd.nimm = sdh->species_id_nimm;
    sd.nimp = sdh->species_id_nimp;
    sd.ngasmin = sdh->species_id_ngasmin;
    sd.x0eps = sdh->x0eps;
    sd.n2o_frac = sdh->n2o_frac_mineralization;
    sd.inhib_nh4_no3 = sdh->inhibition_nh4_no3;
    sd.o2 = sdh->species_o2;
    int q = 0;
    sd.rate_constant = pd(so[q++]);
    sd.rate_decomposition = pd(so[q++]);
    sd.rate_ad_factor = pd(so[q++]);
    sd.upstream_c_id = pi(so[q++]);
    sd.upstream_n_id = pi(so[q++]);
    sd.upstream_hr_id = pi(so[q++]);
    sd.upstream_nmin_id = pi(so[q++]);
    sd.upstream_nimp_id = pi(so[q++]);
    sd.upstream_nimm_id = pi(so[q++]);
    sd.upstream_nc = pd(so[q++]);
    sd.upstream_is_varycn = pi(so[q++]);
    sd.n_downstream_pools = pi(so[q++]);
    sd.downstream_c_id = pi(so[q++]);
    sd.downstream_n_id = pi(so[q++]);
    sd.downstream_stoich = pd(so[q++]);
    sd.downstream_nc = pd(so[q++]);
    sd.downstream_is_varycn = pi(so[q++]);
    sd.mineral_c_stoich = pd(so[q++]);
    sd.mineral_n_stoich = pd(so[q++]);
    sd.temperature_response_function = pi(so[q++]);
    sd.Q10 = pd(so[q++]);
    sd.Ea = pd(so[q++]);
    sd.moisture_response_function = pi(so[q++]);
    sd.Ox_response_function = pi(so[q++]);
    sd.Ox_half_saturation = pd(so[q++]);
    sd.n_monod = pi(so[q++]);
    sd.monod_specid = pi(so[q++]);
    sd.monod_half_saturation = pd(so[q++]);
    sd.monod_threshold = pd(so[q++]);
    sd.monod_pool_normalized = pi(so[q++]);
    sd.n_inhibition = pi(so[q++]);
    sd.inhibition_specid = pi(so[q++]);
    sd.inhibition_type = pi(so[q++]);
    sd.inhibition_constant = pd(so[q++]);
    sd.inhibition_constant2 = pd(so[q++]);
    sd.Ox_specid = pi(so[q++]);
    bool ox = sdh->species_o2 != 0;
    for (int r = 0; r < nr; ++r)
      ox = ox || (sdh->Ox_response_function[r] == 1 && sdh->Ox_specid && sdh->Ox_specid[r] != 0);
    sd.ext = (ox || net->neqcplx > 0) ? 1 : 0;
  }
  ctx->net = *net;
  ctx->opt = *opt;
  ctx->has_somdec = net->somdec != nullptr;
  // default team shape (lanes per cell x rows per lane) by system size;
  // pfchem_autotune measures the alternatives, and a cache of its choices
  // is applied at the end of set-up
  int L = 32, R = 1;
  if (d.ncomp > 32) {
    L = 64;  // one cell per wavefront
  } else if (d.variable) {
    if (d.ncomp <= 16) L = 16;  // per-cell active counts: 16x1, 32x1, 64x1
  } else if (d.has_somdec) {
    if (d.ncomp <= 16) L = 16;  // SomDec kernels exist for 16x1, 8x2, 32x1
  } else if (d.ncomp <= 4) {
    L = 4;
    R = 1;
  } else if (d.ncomp <= 8) {
    L = 16;
    R = 1;
  } else if (d.ncomp <= 16) {
    L = 16;
    R = 1;
  }
  for (int v = PFCHEM_KERNEL_BATCHED_FP64; v <= PFCHEM_KERNEL_BATCHED_MIXED; ++v) {
    LaunchCfg c{L, R, 256, 4, ctx->max_lds};
    while (c.block_size > 64 && !team_supports(d, c, v == PFCHEM_KERNEL_BATCHED_MIXED))
      c.block_size /= 2;
    ctx->launch[v] = c;
    // size buckets of per-cell active counts: 16x1, 32x1, 64x1 teams
    for (int b = 0; b < kBuckets; ++b) {
      LaunchCfg cb{kBucketBound[b], 1, 256, 4, ctx->max_lds};
      while (cb.block_size > 64 && !team_supports(d, cb, v == PFCHEM_KERNEL_BATCHED_MIXED))
        cb.block_size /= 2;
      ctx->vlaunch[v - 1][b] = cb;
    }
  }
  for (int k = 0; k < kStateRing; ++k)
    if (hipEventCreateWithFlags(&ctx->ring_done[k], hipEventDisableTiming) != hipSuccess) {
      pfchem_destroy(ctx);
      return PFCHEM_ERR_HIP;
    }
  if (hipEventCreateWithFlags(&ctx->scratch_done, hipEventDisableTiming) != hipSuccess) {
    pfchem_destroy(ctx);
    return PFCHEM_ERR_HIP;
  }
  // the FP64 team shape also takes the cells the lane kernel hands over
  ctx->team_default = ctx->launch[PFCHEM_KERNEL_BATCHED_FP64];
  // network-specialized one-cell-per-lane kernel: the default batched
  // kernel of every network it takes (compiled at the first step)
  ctx->lane_ok = lane_supported(net, opt, ctx->lane_why) != 0;
  if (ctx->lane_ok) {
    ctx->lane_src = lane_source(net, opt);
    for (int v = PFCHEM_KERNEL_BATCHED_FP64; v <= PFCHEM_KERNEL_BATCHED_MIXED; ++v)
      ctx->launch[v] = LaunchCfg{1, 1, 256, 0, ctx->max_lds};
  }
  if (hipMalloc(&ctx->dnet_dev, sizeof(DevNet)) != hipSuccess ||
      hipMalloc(&ctx->dstate_ring, sizeof(DevState) * kStateRing) != hipSuccess ||
      hipHostMalloc(&ctx->hstate_ring, sizeof(DevState) * kStateRing) != hipSuccess ||
      hipMemcpy(ctx->dnet_dev, &ctx->dnet, sizeof(DevNet), hipMemcpyHostToDevice) != hipSuccess) {
    pfchem_destroy(ctx);
    return PFCHEM_ERR_NOMEM;
  }
  // measured launch shapes (pfchem_autotune) from the cache file the
  // environment names, where they match this device and network
  if (const char* cache = std::getenv("PFCHEM_TUNE_CACHE"))
    if (*cache) (void)tune_apply_cache(ctx, cache);
  *out = ctx;
  return PFCHEM_OK;
}

}  // extern "C"

namespace pfchem {

int ctx_device(const pfchem_ctx* ctx) { return ctx->device; }
int ctx_naq(const pfchem_ctx* ctx) { return ctx->net.naqcomp; }
TransportPlan** ctx_transport_plan(pfchem_ctx* ctx) { return &ctx->tplan; }
int ctx_fail(pfchem_ctx* ctx, int code, const char* msg) { return fail(ctx, code, msg); }

TuneInfo tune_info(const pfchem_ctx* ctx) {
  TuneInfo t;
  const DevNet& d = ctx->dnet;
  t.naq = ctx->net.naqcomp;
  t.nimm = ctx->net.nimmobile;
  t.neqcplx = ctx->net.neqcplx;
  t.nkinmnrl = ctx->net.nkinmnrl;
  t.ncomp = d.ncomp;
  t.variable = d.variable;
  t.lane_ok = ctx->lane_ok;
  int p2 = 1;
  while (p2 < d.ncomp) p2 *= 2;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s/%dcu/aq%d/im%d/cx%d/mn%d/sd%d/var%d/%016llx",
                ctx->arch.c_str(), ctx->num_cus, t.naq, t.nimm, t.neqcplx, t.nkinmnrl,
                ctx->has_somdec ? 1 : 0, t.variable ? 1 : 0, (unsigned long long)ctx->net_hash);
  t.key = buf;
  std::snprintf(buf, sizeof(buf), "%s/%dcu/n%d/sd%d/var%d/lane%d", ctx->arch.c_str(),
                ctx->num_cus, p2, ctx->has_somdec ? 1 : 0, t.variable ? 1 : 0,
                ctx->lane_ok ? 1 : 0);
  t.bucket = buf;
  if (t.variable) {
    for (int b = 0; b < kBuckets && b < kTuneSizeBuckets; ++b) {
      if (b > 0 && kBucketBound[b - 1] >= d.naq) break;  // no cell that large
      t.bucket_bound[b] = kBucketBound[b];
      for (int v = 0; v < 2; ++v) {
        const LaunchCfg& c = ctx->vlaunch[v][b];
        t.bucket_launch[v][b] = pfchem_launch{c.lanes_per_cell, c.block_size, c.blocks_per_cu,
                                              c.rows_per_lane};
      }
      t.size_buckets = b + 1;
    }
  }
  return t;
}

}  // namespace pfchem

extern "C" {

int pfchem_destroy(pfchem_ctx* ctx) {
  if (!ctx) return PFCHEM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->dbuf) (void)hipFree(ctx->dbuf);
  if (ctx->counter) (void)hipFree(ctx->counter);
  if (ctx->stage) (void)hipFree(ctx->stage);
  if (ctx->buckets) (void)hipFree(ctx->buckets);
  if (ctx->esc_list) (void)hipFree(ctx->esc_list);
  if (ctx->dnet_dev) (void)hipFree(ctx->dnet_dev);
  if (ctx->dstate_ring) (void)hipFree(ctx->dstate_ring);
  if (ctx->hstate_ring) (void)hipHostFree(ctx->hstate_ring);
  if (ctx->sens_in) (void)hipFree(ctx->sens_in);
  if (ctx->sens_in_h) (void)hipHostFree(ctx->sens_in_h);
  if (ctx->sens_sum) (void)hipFree(ctx->sens_sum);
  if (ctx->sens_sum_h) (void)hipHostFree(ctx->sens_sum_h);
  if (ctx->sens_copied) (void)hipEventDestroy(ctx->sens_copied);
  if (ctx->sens_work) (void)hipFree(ctx->sens_work);
  if (ctx->con_buf) (void)hipFree(ctx->con_buf);
  if (ctx->status_scratch) (void)hipFree(ctx->status_scratch);
  if (ctx->tplan) transport_plan_free(ctx->tplan);
  for (int k = 0; k < kStateRing; ++k)
    if (ctx->ring_done[k]) (void)hipEventDestroy(ctx->ring_done[k]);
  if (ctx->scratch_done) (void)hipEventDestroy(ctx->scratch_done);
  if (ctx->pass_t0) (void)hipEventDestroy(ctx->pass_t0);
  if (ctx->pass_t1) (void)hipEventDestroy(ctx->pass_t1);
  delete ctx;
  return PFCHEM_OK;
}

static int to_dev_state(pfchem_ctx* ctx, const pfchem_state* st, int64_t ncell, DevState* d) {
  if (!st) return fail(ctx, PFCHEM_ERR_INVALID, "null state");
  if (st->ld < ncell) return fail(ctx, PFCHEM_ERR_INVALID, "state ld < ncell");
  const DevNet& n = ctx->dnet;
  if (!st->pri_molal || !st->total || !st->pri_act_coef || !st->porosity || !st->saturation ||
      !st->den_kg || !st->volume)
    return fail(ctx, PFCHEM_ERR_INVALID, "missing required state array");
  if (n.nimm > 0 && !st->immobile) return fail(ctx, PFCHEM_ERR_INVALID, "missing immobile");
  if (n.nc > 0 && (!st->sec_molal || !st->sec_act_coef))
    return fail(ctx, PFCHEM_ERR_INVALID, "missing sec_molal / sec_act_coef");
  if (n.nm > 0 && (!st->mnrl_volfrac || !st->mnrl_area))
    return fail(ctx, PFCHEM_ERR_INVALID, "missing mineral arrays");
  if (n.variable && !st->ncomp_active)
    return fail(ctx, PFCHEM_ERR_INVALID, "variable_ncomp needs ncomp_active");
  d->ld = st->ld;
  d->pri_molal = st->pri_molal;
  d->total = st->total;
  d->immobile = st->immobile;
  d->sec_molal = st->sec_molal;
  d->pri_act_coef = st->pri_act_coef;
  d->sec_act_coef = st->sec_act_coef;
  d->ln_act_h2o = st->ln_act_h2o;
  d->mnrl_volfrac = st->mnrl_volfrac;
  d->mnrl_area = st->mnrl_area;
  d->mnrl_rate = st->mnrl_rate;
  d->porosity = st->porosity;
  d->saturation = st->saturation;
  d->den_kg = st->den_kg;
  d->volume = st->volume;
  d->temperature = st->temperature;
  d->somdec_f_t = st->somdec_f_t;
  d->somdec_f_w = st->somdec_f_w;
  d->ncomp_active = n.variable ? st->ncomp_active : nullptr;
  d->cell_list = nullptr;
  d->list_count = nullptr;
  d->num_iterations = st->num_iterations;
  d->num_sub_steps = st->num_sub_steps;
  d->status = st->status;
  return PFCHEM_OK;
}

// Batched step on the network-specialized lane kernel: FP32 (mixed) or FP64
// pass over every cell, then the FP64 team kernel (pivoting) over the cells
// handed over.  The hand-over pass runs few cells, so its time is the
// latency of its slowest cell: a team of lanes runs one Newton iteration
// several times faster than one lane does.
static int lane_step(pfchem_ctx* ctx, int64_t ncell, const DevState& d, double dt, bool mixed,
                     hipStream_t s) {
  if (!ctx->lane_ok)
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "one-lane kernel: " + ctx->lane_why);
  if (!ctx->lane_ready) {
    std::string msg;
    const int rc = lane_compile(ctx->lane_src, ctx->device, msg);
    if (rc) return fail(ctx, rc, msg);
    ctx->lane_ready = true;
  }
  int32_t* st = d.status;
  if (!st) {
    if (ncell > ctx->status_cap) {
      scratch_quiesce(ctx);
      if (ctx->status_scratch) (void)hipFree(ctx->status_scratch);
      ctx->status_scratch = nullptr;
      ctx->status_cap = 0;
      if (hipMalloc(&ctx->status_scratch, sizeof(int32_t) * (size_t)ncell) != hipSuccess)
        return fail(ctx, PFCHEM_ERR_NOMEM, "status scratch");
      ctx->status_cap = ncell;
    }
    st = ctx->status_scratch;
  }
  LaneArgs a{};
  a.ld = d.ld;
  a.ncell = ncell;
  a.pri_molal = d.pri_molal;
  a.total = d.total;
  a.immobile = d.immobile;
  a.sec_molal = d.sec_molal;
  a.pri_act_coef = d.pri_act_coef;
  a.sec_act_coef = d.sec_act_coef;
  a.ln_act_h2o = d.ln_act_h2o;
  a.mnrl_volfrac = d.mnrl_volfrac;
  a.mnrl_area = d.mnrl_area;
  a.mnrl_rate = d.mnrl_rate;
  a.porosity = d.porosity;
  a.saturation = d.saturation;
  a.den_kg = d.den_kg;
  a.volume = d.volume;
  a.temperature = d.temperature;
  a.num_iterations = d.num_iterations;
  a.num_sub_steps = d.num_sub_steps;
  a.status = st;
  a.dt = dt;
  a.nxcd = ctx->nxcd;
  a.escalate = 1;
  const int bpc = ctx->launch[mixed ? PFCHEM_KERNEL_BATCHED_MIXED : PFCHEM_KERNEL_BATCHED_FP64]
                      .blocks_per_cu;
  hipError_t e = hipSuccess;
  a.rescue = 0;
  // the rescue pass walks only the cells the one-lane pass handed over
  // (listed by a compaction kernel; larger grids walk all cells)
  DevState dr = d;
  dr.status = st;
  const bool listed = ncell <= 0x7fffffff;
  if (listed) {
    if (ncell > ctx->esc_cap) {
      scratch_quiesce(ctx);
      if (ctx->esc_list) (void)hipFree(ctx->esc_list);
      ctx->esc_list = nullptr;
      ctx->esc_cap = 0;
      if (hipMalloc(&ctx->esc_list, sizeof(int32_t) * ((size_t)ncell + 64)) != hipSuccess)
        return fail(ctx, PFCHEM_ERR_NOMEM, "escalated-cell list");
      ctx->esc_cap = ncell;
    }
    dr.cell_list = ctx->esc_list;
    dr.list_count = ctx->esc_list + ncell;
  }
  // one descriptor slot (and its cursors) for both launches of this step
  int k = 0;
  DevState* slot = acquire_slot(ctx, dr, s, &k, &e);
  if (e != hipSuccess) return hip_fail(ctx, e, "state descriptor copy");
  unsigned long long* cursors = slot_cursors(ctx, k);
  // a step on another stream may still use the list / status scratch
  e = scratch_order(ctx, s);
  if (e == hipSuccess && ctx->pass_timing) e = hipEventRecord(ctx->pass_t0, s);
  if (e == hipSuccess)
    e = lane_launch(ctx->lane_src, ctx->device, mixed ? 1 : 0, a, ctx->num_cus, bpc, cursors, s);
  if (e == hipSuccess && ctx->pass_timing) {
    e = hipEventRecord(ctx->pass_t1, s);
    if (e == hipSuccess) ctx->pass_recorded = true;
  }
  if (e != hipSuccess)
    return hip_fail(ctx, e, mixed ? "lane kernel launch (FP32)" : "lane kernel launch (FP64)");
  if (listed) {
    e = launch_escalated_list(st, ncell, ctx->esc_list, ctx->esc_list + ncell, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "escalated-cell list");
    if (mixed) {
      // a long list goes through the FP64 one-lane kernel first (per cell
      // several times cheaper than a team, at a longer latency for its
      // slowest cell); the team pass takes what that pass hands over.  A
      // short list (the kernel reads its length) is left to the team pass.
      LaneArgs b = a;
      b.rescue = 1;
      b.list = ctx->esc_list;
      b.list_count = ctx->esc_list + ncell;
      b.list_min = kLaneListMin;
      e = lane_launch(ctx->lane_src, ctx->device, 0, b, ctx->num_cus,
                      ctx->launch[PFCHEM_KERNEL_BATCHED_FP64].blocks_per_cu, cursors, s);
      if (e != hipSuccess) return hip_fail(ctx, e, "lane kernel launch (FP64 list)");
    }
  }
  e = launch_team_step(ctx->dnet, ctx->dnet_dev, slot, ncell, dt, kTeamRescue64,
                       ctx->team_default, cursors, ctx->num_cus, ctx->nxcd, s, true);
  if (e == hipSuccess) e = release_slot(ctx, k, s);
  if (e == hipSuccess) e = scratch_release(ctx, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "team rescue launch");
  return PFCHEM_OK;
}

int pfchem_set_pass_timing(pfchem_ctx* ctx, int on) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess && on && !ctx->pass_t0) e = hipEventCreate(&ctx->pass_t0);
  if (e == hipSuccess && on && !ctx->pass_t1) e = hipEventCreate(&ctx->pass_t1);
  if (e != hipSuccess) return hip_fail(ctx, e, "pass timing events");
  ctx->pass_timing = on != 0;
  ctx->pass_recorded = false;
  return PFCHEM_OK;
}

int pfchem_pass_time(pfchem_ctx* ctx, float* ms) {
  if (!ctx || !ms) return PFCHEM_ERR_INVALID;
  if (!ctx->pass_recorded) return fail(ctx, PFCHEM_ERR_INVALID, "no timed one-lane pass recorded");
  hipError_t e = hipEventSynchronize(ctx->pass_t1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->pass_t0, ctx->pass_t1);
  if (e != hipSuccess) return hip_fail(ctx, e, "pass timing");
  return PFCHEM_OK;
}

int pfchem_step(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state, double dt,
                int variant, void* stream) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  if (ncell < 0) return fail(ctx, PFCHEM_ERR_INVALID, "ncell < 0");
  if (!(dt > 0.0)) return fail(ctx, PFCHEM_ERR_INVALID, "dt must be > 0");
  if (variant < PFCHEM_KERNEL_PLAIN_FP64 || variant > PFCHEM_KERNEL_BATCHED_MIXED)
    return fail(ctx, PFCHEM_ERR_INVALID, "unknown kernel variant");
  DevState d;
  int rc = to_dev_state(ctx, state, ncell, &d);
  if (rc) return rc;
  if (ncell == 0) return PFCHEM_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  // RActivityCoefficients NEWTON inside every Newton iteration: only the
  // one-cell-per-thread kernel carries that fixed point
  if (variant != PFCHEM_KERNEL_PLAIN_FP64 &&
      ctx->opt.act_coef_update_algorithm == PFCHEM_ACT_COEF_ALGORITHM_NEWTON &&
      ctx->opt.act_coef_update_frequency == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER)
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED,
                "the NEWTON activity algorithm per Newton iteration runs on the plain kernel");
  if (variant == PFCHEM_KERNEL_PLAIN_FP64) {
    // the one-cell-per-thread SomDec kernel faulted on MI355X without a
    // cause found in its code (its index checks pass on a host build); the
    // batched variants run the sandbox
    if (ctx->dnet.has_somdec)
      return fail(ctx, PFCHEM_ERR_UNSUPPORTED,
                  "the plain kernel does not run the SomDec sandbox; use a batched variant");
    e = launch_plain_step(ctx->dnet, ctx->dnet_dev, d, ncell, dt, s);
  } else {
    const bool mixed = variant == PFCHEM_KERNEL_BATCHED_MIXED;
    if (ctx->dnet.variable) {
      // per-cell active counts: cells sorted into size buckets on the
      // device, one launch per bucket with the bucket's team shape
      if (ncell > 0x7fffffff) return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "variable_ncomp: ncell >= 2^31");
      if (ncell > ctx->bucket_cap) {
        scratch_quiesce(ctx);
        if (ctx->buckets) (void)hipFree(ctx->buckets);
        ctx->buckets = nullptr;
        ctx->bucket_cap = 0;
        if (hipMalloc(&ctx->buckets, sizeof(int32_t) * ((size_t)kBuckets * ncell + 64)) != hipSuccess)
          return fail(ctx, PFCHEM_ERR_NOMEM, "bucket lists");
        ctx->bucket_cap = ncell;
      }
      int32_t* counts = ctx->buckets + (size_t)kBuckets * ncell;
      e = scratch_order(ctx, s);
      if (e == hipSuccess)
        e = launch_buckets(d.ncomp_active, ncell, ctx->dnet.naq, ctx->buckets, counts, s);
      if (e != hipSuccess) return hip_fail(ctx, e, "bucket launch");
      for (int b = 0; b < kBuckets; ++b) {
        if (b > 0 && kBucketBound[b - 1] >= ctx->dnet.naq) break;  // no cell that large
        const LaunchCfg& cb = ctx->vlaunch[variant - 1][b];
        if (!team_supports(ctx->dnet, cb, mixed))
          return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "bucket launch configuration does not fit");
        DevState db = d;
        db.cell_list = ctx->buckets + (size_t)b * ncell;
        db.list_count = counts + b;
        int k = 0;
        DevState* slot = acquire_slot(ctx, db, s, &k, &e);
        if (e != hipSuccess) return hip_fail(ctx, e, "state descriptor copy");
        e = launch_team_step(ctx->dnet, ctx->dnet_dev, slot, ncell, dt, mixed, cb, slot_cursors(ctx, k),
                             ctx->num_cus, ctx->nxcd, s, d.status != nullptr);
        if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
        e = release_slot(ctx, k, s);
        if (e != hipSuccess) return hip_fail(ctx, e, "descriptor event");
      }
      e = scratch_release(ctx, s);
      if (e != hipSuccess) return hip_fail(ctx, e, "scratch event");
      return PFCHEM_OK;
    }
    const LaunchCfg& cfg = ctx->launch[variant];
    if (cfg.lanes_per_cell == 1) return lane_step(ctx, ncell, d, dt, mixed, s);
    if (!team_supports(ctx->dnet, cfg, mixed))
      return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "launch configuration does not fit this network");
    int k = 0;
    DevState* slot = acquire_slot(ctx, d, s, &k, &e);
    if (e != hipSuccess) return hip_fail(ctx, e, "state descriptor copy");
    e = launch_team_step(ctx->dnet, ctx->dnet_dev, slot, ncell, dt, mixed, cfg, slot_cursors(ctx, k),
                         ctx->num_cus, ctx->nxcd, s, d.status != nullptr);
    if (e == hipSuccess) e = release_slot(ctx, k, s);
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
  return PFCHEM_OK;
}

int pfchem_update_activity(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                           void* stream) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  DevState d;
  int rc = to_dev_state(ctx, state, ncell, &d);
  if (rc) return rc;
  if (ncell == 0) return PFCHEM_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = launch_activity_lag(ctx->dnet, ctx->dnet_dev, d, ncell, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "activity launch");
  return PFCHEM_OK;
}

int pfchem_residual_jacobian(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                             const double* fixed_total, const double* fixed_immobile, double dt,
                             double* res, double* jac, void* stream) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  if (ncell < 0) return fail(ctx, PFCHEM_ERR_INVALID, "ncell < 0");
  if (!(dt > 0.0)) return fail(ctx, PFCHEM_ERR_INVALID, "dt must be > 0");
  if (!res) return fail(ctx, PFCHEM_ERR_INVALID, "null res");
  if (ctx->dnet.naq > 0 && !fixed_total) return fail(ctx, PFCHEM_ERR_INVALID, "null fixed_total");
  if (ctx->dnet.nimm > 0 && !fixed_immobile)
    return fail(ctx, PFCHEM_ERR_INVALID, "null fixed_immobile");
  DevState d;
  int rc = to_dev_state(ctx, state, ncell, &d);
  if (rc) return rc;
  if (ncell == 0) return PFCHEM_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  int k = 0;
  DevState* slot = acquire_slot(ctx, d, s, &k, &e);
  if (e != hipSuccess) return hip_fail(ctx, e, "state descriptor copy");
  e = launch_girt_team(ctx->dnet, ctx->dnet_dev, slot, ncell, dt, fixed_total, fixed_immobile, res,
                       jac, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "residual launch");
  e = release_slot(ctx, k, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "descriptor event");
  return PFCHEM_OK;
}

int pfchem_residual(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                    const double* fixed_total, const double* fixed_immobile, double dt,
                    double* res, void* stream) {
  return pfchem_residual_jacobian(ctx, ncell, state, fixed_total, fixed_immobile, dt, res,
                                  nullptr, stream);
}

int pfchem_set_launch(pfchem_ctx* ctx, int variant, const pfchem_launch* cfg) {
  if (!ctx || !cfg) return PFCHEM_ERR_INVALID;
  if (variant < PFCHEM_KERNEL_PLAIN_FP64 || variant > PFCHEM_KERNEL_BATCHED_MIXED)
    return fail(ctx, PFCHEM_ERR_INVALID, "unknown kernel variant");
  if (cfg->block_size <= 0 || cfg->block_size % 64 || cfg->block_size > 1024 ||
      cfg->blocks_per_cu < 0 || cfg->blocks_per_cu > 64)
    return fail(ctx, PFCHEM_ERR_INVALID, "bad launch shape");
  // blocks_per_cu 0: the default (team kernels 4 per CU; the one-lane kernel
  // as many as its occupancy allows)
  LaunchCfg c{cfg->lanes_per_cell, cfg->rows_per_lane > 0 ? cfg->rows_per_lane : 1,
              cfg->block_size, cfg->blocks_per_cu ? cfg->blocks_per_cu
                                                  : (cfg->lanes_per_cell == 1 ? 0 : 4),
              ctx->max_lds};
  if (variant != PFCHEM_KERNEL_PLAIN_FP64 && c.lanes_per_cell == 1) {
    // the network-specialized lane kernel: 256-lane workgroups, at most as
    // many per CU as its occupancy allows
    if (!ctx->lane_ok)
      return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "one-lane kernel: " + ctx->lane_why);
    if (c.rows_per_lane != 1 || c.block_size != 256)
      return fail(ctx, PFCHEM_ERR_INVALID, "the one-lane kernel runs 256-lane workgroups, 1 row");
    ctx->launch[variant] = c;
    return PFCHEM_OK;
  }
  if (variant != PFCHEM_KERNEL_PLAIN_FP64 &&
      !team_supports(ctx->dnet, c, variant == PFCHEM_KERNEL_BATCHED_MIXED))
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "launch configuration does not fit this network");
  ctx->launch[variant] = c;
  if (ctx->dnet.variable && variant != PFCHEM_KERNEL_PLAIN_FP64 && c.rows_per_lane == 1)
    for (int b = 0; b < kBuckets; ++b)
      if (kBucketBound[b] == c.lanes_per_cell) ctx->vlaunch[variant - 1][b] = c;
  return PFCHEM_OK;
}

int pfchem_prepare_kernels(const pfchem_network* net, const pfchem_options* opt,
                           const char* source_path) {
  if (!net || !opt) return PFCHEM_ERR_INVALID;
  if (net->naqcomp < 1 || net->nimmobile < 0 ||
      (net->neqcplx > 0 && (!net->eqcplxspecid || !net->eqcplxstoich || !net->eqcplx_logK)) ||
      (net->nkinmnrl > 0 && (!net->kinmnrlspecid || !net->kinmnrlstoich ||
                             !net->kinmnrl_logK || !net->kinmnrl_rate_constant ||
                             !net->kinmnrl_molar_vol)) ||
      !net->primary_spec_Z || !net->primary_spec_a0)
    return PFCHEM_ERR_INVALID;
  std::string why;
  if (!lane_supported(net, opt, why)) return PFCHEM_ERR_UNSUPPORTED;
  const std::string src = lane_source(net, opt);
  if (source_path) {
    FILE* f = std::fopen(source_path, "w");
    if (!f) return PFCHEM_ERR_INVALID;
    std::fwrite(src.data(), 1, src.size(), f);
    std::fclose(f);
  }
  std::string err;
  const int rc = lane_compile_only(src, nullptr, err);
  if (rc) std::fprintf(stderr, "pfchem_prepare_kernels: %s\n", err.c_str());
  return rc;
}

int pfchem_get_launch(const pfchem_ctx* ctx, int variant, pfchem_launch* cfg) {
  if (!ctx || !cfg || variant < 0 || variant > 2) return PFCHEM_ERR_INVALID;
  cfg->lanes_per_cell = ctx->launch[variant].lanes_per_cell;
  cfg->block_size = ctx->launch[variant].block_size;
  cfg->blocks_per_cu = ctx->launch[variant].blocks_per_cu;
  cfg->rows_per_lane = ctx->launch[variant].rows_per_lane;
  return PFCHEM_OK;
}

int pfchem_step_host(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* h, double dt,
                     int variant) {
  if (!ctx || !h) return PFCHEM_ERR_INVALID;
  if (ncell <= 0) return ncell == 0 ? PFCHEM_OK : PFCHEM_ERR_INVALID;
  if (h->ld < ncell) return fail(ctx, PFCHEM_ERR_INVALID, "state ld < ncell");
  const DevNet& n = ctx->dnet;
  struct F { const void* src; int rows; int esz; bool out; size_t off; };
  F f[] = {
      {h->pri_molal, n.naq, 8, true, 0},       {h->total, n.naq, 8, true, 0},
      {h->immobile, n.nimm, 8, true, 0},       {h->sec_molal, n.nc, 8, true, 0},
      {h->pri_act_coef, n.naq, 8, true, 0},    {h->sec_act_coef, n.nc, 8, true, 0},
      {h->ln_act_h2o, 1, 8, true, 0},          {h->mnrl_volfrac, n.nm, 8, true, 0},
      {h->mnrl_area, n.nm, 8, false, 0},       {h->mnrl_rate, n.nm, 8, true, 0},
      {h->porosity, 1, 8, false, 0},           {h->saturation, 1, 8, false, 0},
      {h->den_kg, 1, 8, false, 0},             {h->volume, 1, 8, false, 0},
      {h->temperature, 1, 8, false, 0},        {h->somdec_f_t, 1, 8, false, 0},
      {h->somdec_f_w, 1, 8, false, 0},         {h->num_iterations, 1, 4, true, 0},
      {h->num_sub_steps, 1, 4, true, 0},       {h->status, 1, 4, true, 0},
      {h->ncomp_active, 1, 4, false, 0}};
  const int nf = sizeof(f) / sizeof(f[0]);
  size_t bytes = 0;
  for (int i = 0; i < nf; ++i) {
    f[i].off = bytes;
    if (f[i].src && f[i].rows > 0)
      bytes += (((size_t)f[i].rows * ncell * f[i].esz) + 255) & ~size_t(255);
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  if (bytes > ctx->stage_bytes) {
    if (ctx->stage) (void)hipFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_bytes = 0;
    if (hipMalloc(&ctx->stage, bytes) != hipSuccess) return fail(ctx, PFCHEM_ERR_NOMEM, "staging");
    ctx->stage_bytes = bytes;
  }
  char* base = (char*)ctx->stage;
  pfchem_state d{};
  d.ld = ncell;
  void** dp[] = {(void**)&d.pri_molal, (void**)&d.total, (void**)&d.immobile, (void**)&d.sec_molal,
                 (void**)&d.pri_act_coef, (void**)&d.sec_act_coef, (void**)&d.ln_act_h2o,
                 (void**)&d.mnrl_volfrac, (void**)&d.mnrl_area, (void**)&d.mnrl_rate,
                 (void**)&d.porosity, (void**)&d.saturation, (void**)&d.den_kg,
                 (void**)&d.volume, (void**)&d.temperature, (void**)&d.somdec_f_t,
                 (void**)&d.somdec_f_w, (void**)&d.num_iterations, (void**)&d.num_sub_steps,
                 (void**)&d.status, (void**)&d.ncomp_active};
  for (int i = 0; i < nf; ++i) {
    if (!f[i].src || f[i].rows <= 0) {
      *dp[i] = nullptr;
      continue;
    }
    *dp[i] = base + f[i].off;
    // rows are ld-strided on the host, packed (ld = ncell) on the device
    e = hipMemcpy2D(base + f[i].off, (size_t)ncell * f[i].esz, f[i].src, (size_t)h->ld * f[i].esz,
                    (size_t)ncell * f[i].esz, f[i].rows, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, "stage in");
  }
  int rc = pfchem_step(ctx, ncell, &d, dt, variant, nullptr);
  if (rc) return rc;
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(ctx, e, "step");
  for (int i = 0; i < nf; ++i) {
    if (!f[i].src || f[i].rows <= 0 || !f[i].out) continue;
    e = hipMemcpy2D((void*)f[i].src, (size_t)h->ld * f[i].esz, base + f[i].off,
                    (size_t)ncell * f[i].esz, (size_t)ncell * f[i].esz, f[i].rows,
                    hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "stage out");
  }
  if (h->status)
    for (int64_t c = 0; c < ncell; ++c)
      if (h->status[c]) return fail(ctx, PFCHEM_ERR_CELLS, "cells failed");
  return PFCHEM_OK;
}

}  // extern "C"

namespace {

// weights / direction to the device through the pinned mirror; the mirror
// is rewritten only after the previous call's copy has completed
int sens_stage(pfchem_ctx* ctx, const double* wv, const double* ws, const double* dp,
               hipStream_t s) {
  const int nm = ctx->dnet.nm, nc = ctx->dnet.nc, np = nm + nc;
  const size_t nin = (size_t)2 * np + 1;
  hipError_t e = hipSuccess;
  if (!ctx->sens_in) {
    if (hipMalloc(&ctx->sens_in, nin * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&ctx->sens_in_h, nin * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->sens_sum, (np + 1) * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&ctx->sens_sum_h, (np + 1) * sizeof(double)) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->sens_copied, hipEventDisableTiming) != hipSuccess)
      return fail(ctx, PFCHEM_ERR_NOMEM, "sensitivity buffers");
  } else {
    e = hipEventSynchronize(ctx->sens_copied);
    if (e != hipSuccess) return hip_fail(ctx, e, "sensitivity staging");
  }
  double* h = ctx->sens_in_h;
  for (int q = 0; q < nm; ++q) h[q] = wv ? wv[q] : 0.0;
  for (int k = 0; k < nc; ++k) h[nm + k] = ws ? ws[k] : 0.0;
  for (int q = 0; q < np; ++q) h[np + q] = dp ? dp[q] : 0.0;
  e = hipMemcpyAsync(ctx->sens_in, h, nin * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipEventRecord(ctx->sens_copied, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "sensitivity staging");
  return PFCHEM_OK;
}

int sens_check(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state, double dt,
               DevState* d) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  if (ncell < 0) return fail(ctx, PFCHEM_ERR_INVALID, "ncell < 0");
  if (!(dt > 0.0)) return fail(ctx, PFCHEM_ERR_INVALID, "dt must be > 0");
  if (!sens_supported(ctx->dnet, ctx->opt))
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED,
                "sensitivities need an aqueous network with plain TST minerals and "
                "activity coefficients fixed during the step");
  int rc = to_dev_state(ctx, state, ncell, d);
  if (rc) return rc;
  if (ncell == 0) return PFCHEM_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  return PFCHEM_OK;
}


}  // namespace

extern "C" {

int pfchem_tangent(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state, double dt,
                   const double* w_vf, const double* w_sec, const double* dp, int jacobian,
                   double* df_cell, double* f_cell, int32_t* status, void* stream) {
  DevState d;
  int rc = sens_check(ctx, ncell, state, dt, &d);
  if (rc || ncell == 0) return rc;
  if (!df_cell || !dp) return fail(ctx, PFCHEM_ERR_INVALID, "tangent needs dp and df_cell");
  hipStream_t s = (hipStream_t)stream;
  rc = sens_stage(ctx, w_vf, w_sec, dp, s);
  if (rc) return rc;
  hipError_t e;
  int k = 0;
  DevState* slot = acquire_slot(ctx, d, s, &k, &e);
  if (e != hipSuccess) return hip_fail(ctx, e, "state descriptor copy");
  const int np = ctx->dnet.nm + ctx->dnet.nc;
  e = launch_sens(ctx->dnet, ctx->dnet_dev, slot, ncell, dt, ctx->sens_in,
                  ctx->sens_in + ctx->dnet.nm, ctx->sens_in + np, 0, jacobian ? 1 : 0, f_cell,
                  df_cell, ncell, status, nullptr, s);
  if (e == hipSuccess) e = release_slot(ctx, k, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "tangent launch");
  return PFCHEM_OK;
}

int pfchem_adjoint(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state, double dt,
                   const double* w_vf, const double* w_sec, int jacobian, double* grad,
                   double* grad_cell, double* f_cell, int32_t* status, void* stream) {
  DevState d;
  int rc = sens_check(ctx, ncell, state, dt, &d);
  const int np = ctx ? ctx->dnet.nm + ctx->dnet.nc : 0;
  if (rc) return rc;
  if (ncell == 0) {
    if (grad)
      for (int q = 0; q < np; ++q) grad[q] = 0.0;
    return PFCHEM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  rc = sens_stage(ctx, w_vf, w_sec, nullptr, s);
  if (rc) return rc;
  // the network-specialized adjoint (one cell per lane) where the one-lane
  // kernel takes the network; the cells it hands over, and every cell of
  // other networks, go to the general kernel
  const bool lane = ctx->lane_ok;
  if (lane && !ctx->lane_adj_ready) {
    ctx->lane_adj_src = std::string(kLaneAdjointDefine) + ctx->lane_src;
    std::string msg;
    const int crc = lane_compile(ctx->lane_adj_src, ctx->device, msg);
    if (crc) return fail(ctx, crc, msg);
    ctx->lane_adj_ready = true;
  }
  const size_t nlog =
      lane ? (size_t)kLaneSensSub * (ctx->dnet.naq + ctx->dnet.nm + 1) * (size_t)ncell : 0;
  // per-cell gradients and status: caller's or the context's scratch; the
  // sub-step log of the one-lane adjoint
  size_t need = nlog * sizeof(double);
  if (!grad_cell) need += (size_t)np * ncell * sizeof(double);
  if (!status) need += (size_t)ncell * sizeof(int32_t);
  if (need > ctx->sens_work_bytes) {
    if (ctx->sens_work) (void)hipFree(ctx->sens_work);
    ctx->sens_work = nullptr;
    ctx->sens_work_bytes = 0;
    if (hipMalloc(&ctx->sens_work, need) != hipSuccess)
      return fail(ctx, PFCHEM_ERR_NOMEM, "sensitivity scratch");
    ctx->sens_work_bytes = need;
  }
  char* wp = (char*)ctx->sens_work;
  double* lg = nullptr;
  if (nlog) {
    lg = (double*)wp;
    wp += nlog * sizeof(double);
  }
  double* gc = grad_cell;
  if (!gc) {
    gc = (double*)wp;
    wp += (size_t)np * ncell * sizeof(double);
  }
  int32_t* st = status ? status : (int32_t*)wp;
  hipError_t e;
  if (lane) {
    LaneSensArgs a{};
    a.ld = d.ld;
    a.ncell = ncell;
    a.pri_molal = d.pri_molal;
    a.total = d.total;
    a.pri_act_coef = d.pri_act_coef;
    a.sec_act_coef = d.sec_act_coef;
    a.ln_act_h2o = d.ln_act_h2o;
    a.mnrl_volfrac = d.mnrl_volfrac;
    a.mnrl_area = d.mnrl_area;
    a.porosity = d.porosity;
    a.saturation = d.saturation;
    a.den_kg = d.den_kg;
    a.volume = d.volume;
    a.temperature = d.temperature;
    a.wv = ctx->sens_in;
    a.ws = ctx->sens_in + ctx->dnet.nm;
    a.grad = gc;
    a.ldo = ncell;
    a.fcell = f_cell;
    a.status = st;
    a.log = lg;
    a.ldl = ncell;
    a.dt = dt;
    a.jacobian = jacobian ? 1 : 0;
    e = lane_adjoint_launch(ctx->lane_adj_src, ctx->device, a, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "adjoint launch");
  }
  int k = 0;
  DevState* slot = acquire_slot(ctx, d, s, &k, &e);
  if (e != hipSuccess) return hip_fail(ctx, e, "state descriptor copy");
  e = launch_sens(ctx->dnet, ctx->dnet_dev, slot, ncell, dt, ctx->sens_in,
                  ctx->sens_in + ctx->dnet.nm, nullptr, 1, jacobian ? 1 : 0, f_cell, gc, ncell,
                  st, ctx->sens_sum, s, lane ? 1 : 0);
  if (e == hipSuccess) e = release_slot(ctx, k, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->sens_sum_h, ctx->sens_sum, np * sizeof(double),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(ctx, e, "adjoint");
  if (grad)
    for (int q = 0; q < np; ++q) grad[q] = ctx->sens_sum_h[q];
  return PFCHEM_OK;
}

}  // extern "C"

extern "C" {

// ReactionEquilibrateConstraint over ncell cells (reaction.F90:1328-2117)
// Argument checks of pfchem_equilibrate / pfchem_equilibrate_static; on
// success *d holds the state descriptor
static int equilibrate_check(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                             const pfchem_constraint* con, const double* conc, DevState* dout) {
  if (!ctx) return PFCHEM_ERR_INVALID;
  if (ncell < 0) return fail(ctx, PFCHEM_ERR_INVALID, "ncell < 0");
  if (!con || !con->constraint_type || !con->constraint_rxn)
    return fail(ctx, PFCHEM_ERR_INVALID, "null constraint");
  if (ncell > 0 && !conc) return fail(ctx, PFCHEM_ERR_INVALID, "null conc");
  const DevNet& n = ctx->dnet;
  if (n.variable)
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "constraint speciation with per-cell active counts");
  if (n.naq > kConMaxComp)
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED, "constraint speciation above 64 primaries");
  if (con->max_iterations < 1 || !(con->max_dlnC > 0.0))
    return fail(ctx, PFCHEM_ERR_INVALID, "max_iterations < 1 or max_dlnC <= 0");
  const int nr = con->nrxn, ms = con->rxn_maxsp;
  if (nr < 0 || (nr > 0 && (ms < 1 || !con->rxnspecid || !con->rxnstoich || !con->rxn_logK)))
    return fail(ctx, PFCHEM_ERR_INVALID, "constraint reaction table");
  for (int r = 0; r < nr; ++r) {
    const int32_t* ids = con->rxnspecid + (size_t)r * (ms + 1);
    if (ids[0] < 0 || ids[0] > ms)
      return fail(ctx, PFCHEM_ERR_INVALID, "constraint reaction species count");
    for (int j = 0; j < ids[0]; ++j)
      if (ids[1 + j] < 0 || ids[1 + j] >= n.naq)
        return fail(ctx, PFCHEM_ERR_INVALID, "constraint reaction species id");
  }
  for (int i = 0; i < n.naq; ++i) {
    const int t = con->constraint_type[i];
    switch (t) {
      case PFCHEM_CONSTRAINT_NULL:
      case PFCHEM_CONSTRAINT_FREE:
      case PFCHEM_CONSTRAINT_TOTAL:
      case PFCHEM_CONSTRAINT_LOG:
      case PFCHEM_CONSTRAINT_CHARGE_BAL:
        break;
      case PFCHEM_CONSTRAINT_PH:
        if (!(con->h_ion_id >= 0 ? con->h_ion_id < n.naq
                                 : (-con->h_ion_id - 1) < n.nc && con->h_ion_id > -(1 << 30)))
          return fail(ctx, PFCHEM_ERR_INVALID, "pH constraint, but H+ is not among the species");
        break;
      case PFCHEM_CONSTRAINT_MINERAL:
      case PFCHEM_CONSTRAINT_GAS:
        if (con->constraint_rxn[i] < 0 || con->constraint_rxn[i] >= nr)
          return fail(ctx, PFCHEM_ERR_INVALID, "mineral / gas constraint without a reaction row");
        break;
      case PFCHEM_CONSTRAINT_PE:
      case PFCHEM_CONSTRAINT_EH:
      case PFCHEM_CONSTRAINT_TOTAL_SORB:
      case PFCHEM_CONSTRAINT_SUPERCRIT_CO2:
      case PFCHEM_CONSTRAINT_TOTAL_AQ_PLUS_SORB:
        return fail(ctx, PFCHEM_ERR_UNSUPPORTED,
                    "redox, sorption and supercritical-CO2 constraints are not supported");
      default:
        return fail(ctx, PFCHEM_ERR_INVALID, "unknown constraint type");
    }
  }
  return to_dev_state(ctx, state, ncell, dout);
}

// Byte layout of the constraint staging buffer: DevConstraint | ids | stoich
// | h2o | logK | coef | failure count (256-byte aligned parts)
struct ConLayout {
  size_t o_ids, o_st, o_h2o, o_lk, o_cf, o_nf, bytes;
};
static ConLayout con_layout(int nr, int ms) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int rows = nr > 0 ? nr : 1, rms = nr > 0 ? ms : 1;
  ConLayout L;
  L.o_ids = up(sizeof(DevConstraint));
  L.o_st = L.o_ids + up(sizeof(int) * rows * (rms + 1));
  L.o_h2o = L.o_st + up(sizeof(double) * rows * rms);
  L.o_lk = L.o_h2o + up(sizeof(double) * rows);
  L.o_cf = L.o_lk + up(sizeof(double) * rows);
  L.o_nf = L.o_cf + up(sizeof(double) * rows * 5);
  L.bytes = L.o_nf + 256;
  return L;
}

// Fill the host image h (L.bytes, zeroed) of the staging buffer that will
// live at device address dbase; the failure count starts at zero
static void con_image(const DevNet& n, const pfchem_constraint* con, const ConLayout& L,
                      char* dbase, char* h) {
  const int nr = con->nrxn, ms = con->rxn_maxsp;
  DevConstraint k{};
  for (int i = 0; i < n.naq; ++i) {
    k.type[i] = con->constraint_type[i];
    k.rxn[i] = con->constraint_rxn[i];
    k.guess[i] = con->free_ion_guess ? con->free_ion_guess[i] : 1.0e-9;
  }
  k.nrxn = nr;
  k.ms = nr > 0 ? ms : 1;
  k.h_ion_id = con->h_ion_id;
  k.with_molality = con->initialize_with_molality != 0;
  k.use_prev = con->use_prev_soln_as_guess != 0;
  k.max_it = con->max_iterations;
  k.has_guess = con->free_ion_guess != nullptr;
  k.has_h2o = nr > 0 && con->rxnh2ostoich != nullptr;
  k.has_coef = nr > 0 && con->rxn_logKcoef != nullptr;
  k.max_dlnC = con->max_dlnC;
  k.ids = (const int*)(dbase + L.o_ids);
  k.st = (const double*)(dbase + L.o_st);
  k.h2o = (const double*)(dbase + L.o_h2o);
  k.logK = (const double*)(dbase + L.o_lk);
  k.coef = (const double*)(dbase + L.o_cf);
  k.nfail = (int*)(dbase + L.o_nf);
  std::memcpy(h, &k, sizeof(k));
  if (nr > 0) {
    std::memcpy(h + L.o_ids, con->rxnspecid, sizeof(int) * nr * (ms + 1));
    std::memcpy(h + L.o_st, con->rxnstoich, sizeof(double) * nr * ms);
    if (k.has_h2o) std::memcpy(h + L.o_h2o, con->rxnh2ostoich, sizeof(double) * nr);
    std::memcpy(h + L.o_lk, con->rxn_logK, sizeof(double) * nr);
    if (k.has_coef) std::memcpy(h + L.o_cf, con->rxn_logKcoef, sizeof(double) * nr * 5);
  }
}

// Upload the image, launch, read the failure count back and synchronize
static int con_run(pfchem_ctx* ctx, int64_t ncell, const DevState& d, const double* conc,
                   const ConLayout& L, char* dbase, const char* h, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(dbase, h, L.bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "constraint copy");
  e = launch_equilibrate(ctx->dnet, ctx->dnet_dev, (const DevConstraint*)dbase, d, conc, ncell, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "equilibrate launch");
  int nfail = 0;
  e = hipMemcpyAsync(&nfail, dbase + L.o_nf, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "failure count copy");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(ctx, e, "equilibrate");
  if (nfail > 0)
    return fail(ctx, PFCHEM_ERR_CELLS, std::to_string(nfail) + " cell(s) failed to equilibrate");
  return PFCHEM_OK;
}

int pfchem_equilibrate(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                       const pfchem_constraint* con, const double* conc, void* stream) {
  DevState d;
  int rc = equilibrate_check(ctx, ncell, state, con, conc, &d);
  if (rc) return rc;
  if (ncell == 0) return PFCHEM_OK;
  const ConLayout L = con_layout(con->nrxn, con->rxn_maxsp);
  const size_t bytes = L.bytes;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  if (bytes > ctx->con_bytes) {
    if (ctx->con_buf) {
      e = hipFree(ctx->con_buf);  // the previous call synchronized before returning
      ctx->con_buf = nullptr;
      ctx->con_bytes = 0;
      if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    }
    e = hipMalloc(&ctx->con_buf, bytes);
    if (e != hipSuccess) {
      ctx->con_buf = nullptr;
      return fail(ctx, PFCHEM_ERR_NOMEM, "constraint buffer");
    }
    ctx->con_bytes = bytes;
  }
  std::vector<char> h(bytes, 0);
  con_image(ctx->dnet, con, L, (char*)ctx->con_buf, h.data());
  return con_run(ctx, ncell, d, conc, L, (char*)ctx->con_buf, h.data(), s);
}

// The same speciation with its staging buffer in the library's fixed device
// array (no allocation in the call; see pfchem.h)
int pfchem_equilibrate_static(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                              const pfchem_constraint* con, const double* conc, void* stream) {
  DevState d;
  int rc = equilibrate_check(ctx, ncell, state, con, conc, &d);
  if (rc) return rc;
  const ConLayout L = con_layout(con->nrxn, con->rxn_maxsp);
  if (L.bytes > kStaticScratchBytes)
    return fail(ctx, PFCHEM_ERR_UNSUPPORTED,
                "constraint reaction table too large for the fixed staging array (" +
                    std::to_string(L.bytes) + " > " + std::to_string(kStaticScratchBytes) +
                    " bytes; use pfchem_equilibrate)");
  if (ncell == 0) return PFCHEM_OK;
  char* dbase = (char*)static_scratch_ptr(ctx->device);
  if (!dbase) return fail(ctx, PFCHEM_ERR_HIP, "fixed staging array not resolved on this device");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  // host image on the stack: no heap allocation either
  alignas(256) char h[kStaticScratchBytes];
  std::memset(h, 0, L.bytes);
  con_image(ctx->dnet, con, L, dbase, h);
  return con_run(ctx, ncell, d, conc, L, dbase, h, (hipStream_t)stream);
}

}  // extern "C"
