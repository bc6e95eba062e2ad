// lane_host_shim.h -- host build of the generated one-lane kernel source
// (pfchem_lane.hpp + its generated network constants) for the CPU tests:
// one thread plays a wavefront of one lane, so ballots, shuffles and the
// work cursor reduce to their one-lane meaning.  Test infrastructure only.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(x)
#define __shared__ static

struct PfDim3 {
  unsigned x, y, z;
};
static PfDim3 threadIdx = {0, 0, 0};
static PfDim3 blockIdx = {0, 0, 0};
static PfDim3 blockDim = {1, 1, 1};

using std::copysign;
using std::exp;
using std::fabs;
using std::fmax;
using std::fmin;
using std::log;
using std::pow;
using std::sqrt;
using std::fma;

static inline unsigned __lane_id() { return 0u; }
static inline unsigned long long __ballot(int p) { return p ? 1ull : 0ull; }
static inline int __all(int p) { return p != 0; }
static inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
template <typename T>
static inline T __shfl(T v, int, int) {
  return v;
}
static inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  const unsigned long long o = *p;
  *p += v;
  return o;
}
static inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
static inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
