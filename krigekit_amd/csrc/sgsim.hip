// Sequential Gaussian simulation (nsim > 0, one variable) on the GPU.
//
// The reference visits the blocks along a random path and, at each, kriges
// from the observations and the blocks simulated before it
// (kriging_base.F90:2055-2181, kriging.F90:535-612, kriging_base.F90:3405-
// 3505).  The kriging weights depend only on where the data are, not on the
// simulated values, so the library solves every block's system in one
// batched pass over the observations followed by the path-ordered blocks
// (each search limited to the points before its block), and the realisations
// then reduce to a recurrence in path order:
//   val_s(p) = sum_obs w z + sum_{q < p} w_q val_s(q) + sqrt(var_p) u_s(p)
// run here one dependency level at a time (a block's level is one above its
// simulated neighbours' highest), every (block, realisation) of a level in
// parallel.
#include <hip/hip_runtime.h>

#include "kk_device.hpp"

namespace kk {

// Post-search lists of the simulation pass: inear rows hold combined indices
// (observations 0..nobs-1, then blocks in path order), ascending, so each row
// is its observation group followed by its simulated group.  A simulated
// block that lies on one of the row's observations (every coordinate within
// 1e-6 (1 + |x_obs|)) is dropped: the datum already conditions that location
// and the two rows would be equal (kriging.F90:588-612).
__global__ __launch_bounds__(256) void k_sgsim_lists(int* nnear, int* inear, int mmax, int nb, int nobs, int ndim,
                                                      const double* coord) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nb) return;
  const int n = nnear[p];
  int* row = inear + (size_t)p * mmax;
  int no = 0;
  while (no < n && row[no] < nobs) ++no;
  int keep = no;
  for (int j = no; j < n; ++j) {
    const int b = row[j];
    const double* xb = coord + (size_t)ndim * b;
    bool dup = false;
    for (int i = 0; i < no && !dup; ++i) {
      const double* xo = coord + (size_t)ndim * row[i];
      bool same = true;
      for (int d = 0; d < ndim; ++d)
        if (fabs(xb[d] - xo[d]) > 1.0e-6 * (1.0 + fabs(xo[d]))) same = false;
      dup = same;
    }
    if (!dup) row[keep++] = b;
  }
  nnear[p] = keep;
}

hipError_t launch_sgsim_lists(int* nnear, int* inear, int mmax, int nb, int nobs, int ndim, const double* coord,
                              hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sgsim_lists, dim3((nb + 255) / 256), dim3(256), 0, st, nnear, inear, mmax, nb, nobs, ndim,
                     coord);
  return hipGetLastError();
}

// One dependency level: thread (i, s) simulates block order[i] of the level
// for realisation s (estimate_block_base, kriging_base.F90:3444-3503):
//   the observation part sum_obs w z (simple kriging with a mean adds
//   (1 - sum_obs w) mean), then the simulated neighbours' values of this
//   realisation, then sqrt(var) u (the 1x1 Cholesky factor of the
//   conditional variance, 0 when that is not positive), clipped to the
//   bounds.  Blocks without a solved system (no neighbours, or one) are NaN.
__global__ __launch_bounds__(256) void k_sgsim_level(SgsimArgs a, int l0, int l1) {
  const int i = l0 + (int)(blockIdx.x * blockDim.x + threadIdx.x) / a.nsim;
  const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x) % a.nsim;
  if (i >= l1) return;
  const int p = a.order[i];
  const int n = a.nnear[p];
  const int* row = a.inear + (size_t)p * a.mmax;
  const double* w = a.weights + (size_t)p * a.mmax;
  double v;
  if (n < 2) {
    v = __builtin_nan("");
  } else {
    double avg = 0.0, tw = 0.0;
    int k = 0;
    for (; k < n && row[k] < a.nobs; ++k) {
      avg = avg + a.ovalue[row[k]] * w[k];
      tw = tw + w[k];
    }
    if (a.unbias == 0 && a.sk_mean != 0.0) avg = avg + (1.0 - tw) * a.sk_mean;
    double sim = 0.0;
    for (; k < n; ++k) sim = sim + a.val[(size_t)s + (size_t)a.nsim * (row[k] - a.nobs)] * w[k];
    const double var = a.var[p];
    const double L = var > 0.0 ? sqrt(var) : 0.0;
    v = (avg + sim) + L * a.sample[(size_t)s + (size_t)a.nsim * p];
    v = v < a.bound_lo ? a.bound_lo : v;
    v = v > a.bound_hi ? a.bound_hi : v;
  }
  a.val[(size_t)s + (size_t)a.nsim * p] = v;
}

hipError_t launch_sgsim_level(const SgsimArgs& a, int l0, int l1, hipStream_t st) {
  const long work = (long)(l1 - l0) * a.nsim;
  if (work <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sgsim_level, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, a, l0, l1);
  return hipGetLastError();
}

}  // namespace kk
