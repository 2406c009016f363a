// Morton (Z-order) permutation of the query points on the GPU, so that the
// 64 queries of one search wave lie close together.  Per query b the
// coordinates (the fourth, the time, scaled by time_at) are mapped to
// bits-wide integers over the bounding cube, (x - lo_k) * ((2^bits - 1) /
// extent), the bits interleaved from the top, and the points ordered by code
// and, for equal codes, by index: a stable LSD radix sort (rocPRIM) of the
// codes with the indices 0..nb-1 as values.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstdint>

#include "kk_device.hpp"

namespace kk {

namespace {

constexpr int OT = 256;    // threads per workgroup
constexpr int OP = 1024;   // min/max partials at most

__device__ inline double ord_coord(const double* bc, int nd, int k, int b, double tsc) {
  const double v = bc[k + (size_t)nd * b];
  return k == 3 ? v * tsc : v;
}

// per-workgroup minima and maxima of each coordinate: part[g * 8 + k] = lo_k,
// part[g * 8 + 4 + k] = hi_k
__global__ __launch_bounds__(OT) void k_order_minmax(const double* bc, int nb, int nd, double tsc, double* part) {
  __shared__ double slo[4][OT], shi[4][OT];
  double lo[4], hi[4];
  for (int k = 0; k < 4; ++k) {
    lo[k] = __builtin_inf();
    hi[k] = -__builtin_inf();
  }
  for (int b = blockIdx.x * OT + threadIdx.x; b < nb; b += gridDim.x * OT)
    for (int k = 0; k < nd; ++k) {
      const double v = ord_coord(bc, nd, k, b, tsc);
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  for (int k = 0; k < 4; ++k) {
    slo[k][threadIdx.x] = lo[k];
    shi[k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = OT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 4; ++k) {
        slo[k][threadIdx.x] = fmin(slo[k][threadIdx.x], slo[k][threadIdx.x + s]);
        shi[k][threadIdx.x] = fmax(shi[k][threadIdx.x], shi[k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 4) {
    part[(size_t)blockIdx.x * 8 + threadIdx.x] = slo[threadIdx.x][0];
    part[(size_t)blockIdx.x * 8 + 4 + threadIdx.x] = shi[threadIdx.x][0];
  }
}

// the bounds over the partials, then lohs[0..3] = lo_k, lohs[4] = scale
__global__ void k_order_bounds(const double* part, int np, int nd, int bits, double* lohs) {
  if (threadIdx.x != 0) return;
  double ext = 0.0;
  for (int k = 0; k < nd; ++k) {
    double lo = part[k], hi = part[4 + k];
    for (int g = 1; g < np; ++g) {
      lo = fmin(lo, part[(size_t)g * 8 + k]);
      hi = fmax(hi, part[(size_t)g * 8 + 4 + k]);
    }
    lohs[k] = lo;
    ext = fmax(ext, hi - lo);
  }
  if (!(ext > 0.0)) ext = 1.0;
  lohs[4] = (double)((1u << bits) - 1) / ext;
}

__global__ __launch_bounds__(OT) void k_order_codes(const double* bc, int nb, int nd, double tsc, int bits,
                                                    const double* lohs, unsigned long long* key, int* val) {
  const int b = blockIdx.x * OT + threadIdx.x;
  if (b >= nb) return;
  const double scale = lohs[4];
  unsigned int c[4] = {0u, 0u, 0u, 0u};
  for (int k = 0; k < nd; ++k) c[k] = (unsigned int)((ord_coord(bc, nd, k, b, tsc) - lohs[k]) * scale);
  unsigned long long code = 0ull;
  for (int bit = bits - 1; bit >= 0; --bit)
    for (int k = 0; k < nd; ++k) code = (code << 1) | ((c[k] >> bit) & 1u);
  key[b] = code;
  val[b] = b;
}

int order_bits(int nd) { return nd == 1 ? 30 : nd == 2 ? 21 : nd == 3 ? 20 : 15; }

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t sort_temp_bytes(int nb, int nd) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                  (int*)nullptr, (int*)nullptr, (size_t)nb, 0, order_bits(nd) * nd);
  return bytes;
}

}  // namespace

size_t order_work_bytes(int nb, int nd) {
  return 2 * align256(sizeof(unsigned long long) * (size_t)nb) + align256(sizeof(int) * (size_t)nb) +
         align256(sizeof(double) * 8 * OP) + align256(sizeof(double) * 8) + align256(sort_temp_bytes(nb, nd));
}

hipError_t launch_order(const double* bc, int nb, int nd, double tsc, int* perm, void* work, size_t work_bytes,
                        hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  if (nd < 1 || nd > 4 || work_bytes < order_work_bytes(nb, nd)) return hipErrorInvalidValue;
  char* w = static_cast<char*>(work);
  auto take = [&](size_t bytes) {
    char* p = w;
    w += align256(bytes);
    return p;
  };
  auto* kin = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * (size_t)nb));
  auto* kout = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * (size_t)nb));
  auto* vin = reinterpret_cast<int*>(take(sizeof(int) * (size_t)nb));
  auto* part = reinterpret_cast<double*>(take(sizeof(double) * 8 * OP));
  auto* lohs = reinterpret_cast<double*>(take(sizeof(double) * 8));
  size_t tbytes = sort_temp_bytes(nb, nd);
  void* tmp = take(tbytes);
  const int bits = order_bits(nd);
  const int np = std::max(1, std::min(OP, (nb + OT - 1) / OT));
  hipLaunchKernelGGL(k_order_minmax, dim3(np), dim3(OT), 0, st, bc, nb, nd, tsc, part);
  hipLaunchKernelGGL(k_order_bounds, dim3(1), dim3(64), 0, st, part, np, nd, bits, lohs);
  hipLaunchKernelGGL(k_order_codes, dim3((nb + OT - 1) / OT), dim3(OT), 0, st, bc, nb, nd, tsc, bits, lohs, kin,
                     vin);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return rocprim::radix_sort_pairs(tmp, tbytes, kin, kout, vin, perm, (size_t)nb, 0, bits * nd, st);
}

}  // namespace kk
