// pfchem_lu.hpp -- per-thread dense FP64 solve used by the one-cell-per-
// thread kernels: Crout LU with implicit partial pivoting, zero pivots
// replaced by 1e-20, then forward / back substitution.
#pragma once
#include <hip/hip_runtime.h>

namespace pfchem {

// LUDecomposition1 (utility.F90:597-688) + LUBackSubstitution (:692-735)
template <int MAXN>
__device__ inline int crout_lu_solve(double* a, int N, double* b) {
  double vv[MAXN];
  int indx[MAXN];
  int imax = 0;
  for (int i = 0; i < N; ++i) {
    double amax = 0.0;
    for (int j = 0; j < N; ++j) amax = fmax(amax, fabs(a[i * N + j]));
    if (amax <= 0.0) return 1;
    vv[i] = 1.0 / amax;
  }
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < j; ++i) {
      double sum = a[i * N + j];
      for (int k = 0; k < i; ++k) sum -= a[i * N + k] * a[k * N + j];
      a[i * N + j] = sum;
    }
    double amax = 0.0;
    for (int i = j; i < N; ++i) {
      double sum = a[i * N + j];
      for (int k = 0; k < j; ++k) sum -= a[i * N + k] * a[k * N + j];
      a[i * N + j] = sum;
      const double dum = vv[i] * fabs(sum);
      if (dum >= amax) {
        imax = i;
        amax = dum;
      }
    }
    if (j != imax) {
      for (int k = 0; k < N; ++k) {
        const double t = a[imax * N + k];
        a[imax * N + k] = a[j * N + k];
        a[j * N + k] = t;
      }
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (a[j * N + j] == 0.0) a[j * N + j] = 1.0e-20;
    if (j != N - 1) {
      const double dum = 1.0 / a[j * N + j];
      for (int i = j + 1; i < N; ++i) a[i * N + j] *= dum;
    }
  }
  int ii = -1;
  for (int i = 0; i < N; ++i) {
    const int ll = indx[i];
    double sum = b[ll];
    b[ll] = b[i];
    if (ii >= 0) {
      for (int j = ii; j < i; ++j) sum -= a[i * N + j] * b[j];
    } else if (sum != 0.0) {
      ii = i;
    }
    b[i] = sum;
  }
  for (int i = N - 1; i >= 0; --i) {
    double sum = b[i];
    for (int j = i + 1; j < N; ++j) sum -= a[i * N + j] * b[j];
    b[i] = sum / a[i * N + i];
  }
  return 0;
}

}  // namespace pfchem
