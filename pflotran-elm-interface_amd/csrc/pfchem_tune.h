// pfchem_tune.h -- internal interface between the context (pfchem_api.hip)
// and the launch-shape autotuner with its JSON cache (pfchem_tune.hip).
#pragma once

#include <cstdint>
#include <string>

#include "pfchem.h"

namespace pfchem {

constexpr int kTuneSizeBuckets = 3;  // size buckets of per-cell active counts

// What the tuner needs to know of a context's network and device.
struct TuneInfo {
  std::string key;     // device + network content (exact match)
  std::string bucket;  // device + system-size bucket (power of two) + kernel family
  int naq = 0, nimm = 0, neqcplx = 0, nkinmnrl = 0, ncomp = 0;
  bool variable = false;  // per-cell active counts (size buckets of teams)
  bool lane_ok = false;   // the one-lane kernel takes the network
  // per-cell active counts: the size buckets that can hold cells, the team
  // size of each and its current launch per batched variant
  int size_buckets = 0;
  int bucket_bound[kTuneSizeBuckets] = {};
  pfchem_launch bucket_launch[2][kTuneSizeBuckets] = {};
};

TuneInfo tune_info(const pfchem_ctx* ctx);  // pfchem_api.hip

// Apply the cache entries of `path` that match the context (both batched
// variants); entries the context refuses are skipped.  Returns the number
// of variants set.
int tune_apply_cache(pfchem_ctx* ctx, const char* path);  // pfchem_tune.hip

}  // namespace pfchem
