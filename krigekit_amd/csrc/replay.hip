// Estimates from stored kriging weights: the reference's block loop with
// use_old_weight, load_block_weights + estimate_block_base
// (kriging_base.F90:2125-2126, 2792-2812, 3407-3503).  No search and no
// solve: per block and target variable q,
//   est_q = sum_g [ sum_k w(k, g, q) z_g(inear(k, g))
//                   + (m_q - m_g) sum_k w(k, g, q)   (Isaaks-Srivastava, g != q) ]
//           + (1 - sum_g sum_k w(k, g, q)) sk_mean_q  (simple kriging, sk_mean != 0)
// clipped to the bounds, m_g the mean of group g's neighbour values; the
// variance is the stored one.  A block without neighbours gets NaN for both
// (kriging_base.F90:3424-3428).  A stored id outside the variable's
// observations (a group the reference left with a stale id, or a store
// built elsewhere) contributes NaN: the gather never leaves the array.
//
// G lanes share one block (G = 16, 32 or 64, the smallest that covers the
// store's nmax), so the store is read in coalesced runs; each lane sums its
// strided share and a fixed xor-tree combines the lanes, so the result does
// not depend on scheduling.  Memory bound: 4 + 8 nvar bytes of store per
// neighbour slot plus the gathered value.
#include <hip/hip_runtime.h>

#include "kk_device.hpp"
#include "kk_types.hpp"

namespace kk {

constexpr int RPT = 256;  // threads per workgroup

template <int G>
__device__ inline double lanes_sum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

template <int G>
__global__ __launch_bounds__(RPT) void k_replay(ReplayArgs a) {
  const int lane = threadIdx.x % G;
  const int b = blockIdx.x * (RPT / G) + threadIdx.x / G;
  if (b >= a.nb) return;  // the whole lane group leaves together
  const int ng = a.ng, nvar = a.nvar, nm = a.nm;
  int nn[MAXVAR];
  int ntot = 0;
#pragma unroll
  for (int g = 0; g < MAXVAR; ++g) {
    nn[g] = g < ng ? a.nnear[g + (size_t)ng * b] : 0;
    ntot += nn[g];
  }
  const double qnan = __builtin_nan("");
  if (ntot == 0) {
    if (lane < nvar) a.est[b + (size_t)a.nb * lane] = qnan;
    if (lane < nvar * nvar) a.var[b + (size_t)a.nb * lane] = qnan;
    return;
  }
  // per group g and target q: sum w z and sum w; per group: sum z
  double dot[MAXVAR][MAXVAR], tw[MAXVAR][MAXVAR], zs[MAXVAR];
#pragma unroll
  for (int g = 0; g < MAXVAR; ++g) {
    zs[g] = 0.0;
#pragma unroll
    for (int q = 0; q < MAXVAR; ++q) dot[g][q] = tw[g][q] = 0.0;
  }
#pragma unroll
  for (int g = 0; g < MAXVAR; ++g) {
    if (g >= ng) break;
    const int* in = a.inear + (size_t)nm * (g + (size_t)ng * b);
    const double* z = a.ovalue[g];
    const int nobs = a.nobs[g];
    for (int k = lane; k < nn[g]; k += G) {
      const int id = in[k];
      const double zk = (id >= 1 && id <= nobs) ? z[id - 1] : qnan;
      zs[g] += zk;
#pragma unroll
      for (int q = 0; q < MAXVAR; ++q) {
        if (q >= nvar) break;
        const double w = a.weight[k + (size_t)nm * (g + (size_t)ng * (q + (size_t)nvar * b))];
        dot[g][q] = fma(w, zk, dot[g][q]);
        tw[g][q] += w;
      }
    }
  }
#pragma unroll
  for (int g = 0; g < MAXVAR; ++g) {
    if (g >= ng) break;
    zs[g] = lanes_sum<G>(zs[g]);
#pragma unroll
    for (int q = 0; q < MAXVAR; ++q) {
      if (q >= nvar) break;
      dot[g][q] = lanes_sum<G>(dot[g][q]);
      tw[g][q] = lanes_sum<G>(tw[g][q]);
    }
  }
  const bool isa = !a.std_ck && a.unbias && nvar > 1;
  if (lane < nvar) {
    const int q = lane;
    double avg = 0.0, twsum = 0.0, mq = 0.0;
#pragma unroll
    for (int g = 0; g < MAXVAR; ++g)
      if (g == q && nn[g] > 0) mq = zs[g] / nn[g];
#pragma unroll
    for (int g = 0; g < MAXVAR; ++g) {
      if (g >= ng) break;
      // the lane's own target: select, no runtime index into the arrays
      double d = 0.0, t = 0.0;
#pragma unroll
      for (int u = 0; u < MAXVAR; ++u)
        if (u == q) {
          d = dot[g][u];
          t = tw[g][u];
        }
      avg += d;
      if (isa && g != q && nn[g] > 0) avg += (mq - zs[g] / nn[g]) * t;
      twsum += t;
    }
    double skm = 0.0;
#pragma unroll
    for (int u = 0; u < MAXVAR; ++u)
      if (u == q) skm = a.sk_mean[u];
    if (!a.unbias && skm != 0.0) avg += (1.0 - twsum) * skm;
    avg = avg < a.bound_lo ? a.bound_lo : avg;
    avg = avg > a.bound_hi ? a.bound_hi : avg;
    a.est[b + (size_t)a.nb * q] = avg;
  }
  if (lane < nvar * nvar) {
    const int i = lane % nvar, j = lane / nvar;
    a.var[b + (size_t)a.nb * lane] = a.wvar[i + (size_t)nvar * (j + (size_t)nvar * b)];
  }
}

hipError_t launch_replay(const ReplayArgs& a, hipStream_t st) {
  if (a.nb <= 0) return hipSuccess;
  if (a.nvar < 1 || a.nvar > MAXVAR || a.ng != a.nvar || a.nm < 1) return hipErrorInvalidValue;
  const int G = a.nm <= 16 ? 16 : a.nm <= 32 ? 32 : 64;
  dim3 blk(RPT), grid((unsigned)((a.nb + RPT / G - 1) / (RPT / G)));
  if (G == 16) hipLaunchKernelGGL(k_replay<16>, grid, blk, 0, st, a);
  else if (G == 32) hipLaunchKernelGGL(k_replay<32>, grid, blk, 0, st, a);
  else hipLaunchKernelGGL(k_replay<64>, grid, blk, 0, st, a);
  return hipGetLastError();
}

}  // namespace kk
