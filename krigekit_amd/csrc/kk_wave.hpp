// Lane-group reductions and the reciprocal used by the row-per-lane solvers
// (G lanes per system, 64 / G systems per wave).
#pragma once
#include <hip/hip_runtime.h>

namespace kk {

template <int G>
__device__ inline double gsum(double v) {
#pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
  return v;
}
template <int G>
__device__ inline int gmax(int v) {
#pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, G));
  return v;
}
// 1/a to within an ulp or so: hardware estimate and two Newton steps (the
// factorisation needs no correctly rounded division; a > 0 here or the
// system is flagged)
__device__ inline double recip(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = fma(-a, r, 1.0);
  r = fma(r, e, r);
  e = fma(-a, r, 1.0);
  return fma(r, e, r);
}
__device__ inline int wmax(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}

}  // namespace kk
