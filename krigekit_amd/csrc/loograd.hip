// Gradient of the leave-one-out estimates with respect to the variogram
// parameters, for the ordinary-kriging system the reference solves per block
// (kriging.F90:714-939, solver.f90:224-300):
//
//   A = [K 1; 1' 0],  x = A^-1 [k; 1] (weights w, multiplier),  zhat = w'z.
//
// With the adjoint y = A^-1 [z; 0] (A is symmetric), for any parameter t
//
//   d zhat / dt = y_w' dk/dt - y_w' (dK/dt) w,
//
// because only K and k depend on the variogram; the neighbour set does not
// (the search metric is fixed).  K is recomputed here rather than stored.
//
// Covariance derivatives of one structure with a nugget:
//   analytic: C = sill * rho(h), h = |M lag| > 0;  C(0) = sill + nugget
//   table:    C = t_i + f (t_{i+1} - t_i) on the reference's lookup table
//             (t_0 = sill + nugget, t_i = sill * rho(h_i)), so dC/dh is the
//             slope of the table segment, dC/dsill interpolates rho(h_i) and
//             dC/dnugget = 1 - f in the first segment;
//   dh/dt = (M lag) . (dM/dt lag) / h for the ranges and angles.
// One 64-lane workgroup per block, K in LDS, FP64 throughout.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kk_cov.hpp"
#include "kk_device.hpp"
#include "kk_types.hpp"
#include "kk_wave.hpp"

namespace kk {

// d rho / dr for the correlation shapes of vgmfunc.f90:48-87
__device__ inline double corefunc_d(int t, double r) {
  const double pi = 3.141592653589793238462643383279502884;
  switch (t) {
    case SPH: return (r < 1.0) ? -1.5 + 1.5 * (r * r) : 0.0;
    case EXP: return -3.0 * exp(-3.0 * r);
    case GAU: return -6.125 * r * exp(-3.0625 * (r * r));
    case HOL: return -pi * sin(pi * r);
    case POW: return (r < 1.0) ? -1.5 * sqrt(r) : 0.0;
    case BSQ: return (r < 1.0) ? -4.0 * r * (1.0 - r * r) : 0.0;
    case CIR: return (r < 1.0) ? -4.0 * sqrt(1.0 - r * r) / pi : 0.0;
    case LIN: return (r < 1.0) ? -1.0 : 0.0;
    case CYC: { double s = sin(pi * r); return -2.0 * pi * sin(2.0 * pi * r) * exp(-2.0 * (s * s)); }
    case DCO: return -exp(-3.0 * r) * (3.0 * cos(pi * r) + pi * sin(pi * r));
    default: return 0.0;
  }
}

template <int DIM>
__device__ inline void mat_apply(const double* m, const double* l, double* u) {
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < DIM; ++c) s += m[r + 3 * c] * l[c];
    u[r] = s;
  }
}

// covariance of one structure at lag l and, when D, its derivatives dc[NGRAD]
template <int DIM, bool D = true>
__device__ inline double cov_grad(const Model& M, const double (*dmat)[9], const double* l, double* dc) {
  const int vt = M.vtype[0];
  const double sill = M.sill[0];
  double u[DIM];
  mat_apply<DIM>(M.mat[0], l, u);
  double hh = 0.0;
#pragma unroll
  for (int r = 0; r < DIM; ++r) hh += u[r] * u[r];
  const double h = sqrt(hh);
  if constexpr (D) {
#pragma unroll
    for (int p = 0; p < NGRAD; ++p) dc[p] = 0.0;
  }
  const double eps = 2.2250738585072014e-308 * 1.0e3;
  double c, dcdh = 0.0;
  if (M.mode == COV_TABLE_B && h > 0.0 && h < M.zb[3]) {
    const int k = (h >= M.zb[1]) + (h >= M.zb[2]);
    double fi = (h - M.zb[k]) * M.zinv[k];
    const int ifi = (int)fi;
    const int il = min(ifi, M.zn[k] - 1);
    const int i = M.zoff[k] + il;
    fi = fi - (double)ifi;
    const double t0 = M.tab[i], t1 = M.tab[i + 1];
    c = t0 + fi * (t1 - t0);
    if constexpr (D) {
      dcdh = (t1 - t0) * M.zinv[k];
      const double step = 1.0 / M.zinv[k];
      const double r0 = i == 0 ? 1.0 : corefunc(vt, M.zb[k] + (double)il * step);
      const double r1 = corefunc(vt, M.zb[k] + (double)(il + 1) * step);
      dc[1] = r0 + fi * (r1 - r0);
      if (i == 0) dc[0] = 1.0 - fi;
    }
  } else if (M.mode == COV_TABLE_B && h <= 0.0) {
    c = M.tab[0];
    if constexpr (D) {
      dc[0] = 1.0;
      dc[1] = 1.0;
    }
  } else if (h > eps && (M.mode != COV_TABLE_B || M.b_tail)) {
    const double rho = corefunc(vt, h);
    c = sill * rho;
    if constexpr (D) {
      dc[1] = rho;
      dcdh = sill * corefunc_d(vt, h);
    }
  } else if (h > eps) {
    c = 0.0;  // beyond the table of a shape without a tail
  } else {
    c = sill + M.nugget[0];
    if constexpr (D) {
      dc[0] = 1.0;
      dc[1] = 1.0;
    }
  }
  if (D && dcdh != 0.0 && h > 0.0) {
#pragma unroll
    for (int g = 0; g < 6; ++g) {
      double du[DIM];
      mat_apply<DIM>(dmat[g], l, du);
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) s += u[r] * du[r];
      dc[2 + g] = dcdh * s / h;
    }
  }
  return c;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int GMAX = 64;

// One block per 64-lane workgroup, K in LDS: systems of 33-64 rows.  Their
// K takes 9-33 KB of LDS, so LDS caps the waves per SIMD (at most 4), and
// no register bound is set (one for five waves only added spills).
template <int DIM>
__global__ __launch_bounds__(64) void k_loo_grad(GradArgs a) {
  extern __shared__ double Ks[];  // mmax x (mmax + 1)
  __shared__ double sx[DIM][GMAX];
  __shared__ double rh[3][GMAX];
  const int b = blockIdx.x;
  const int j = threadIdx.x;
  const Model& M = *a.model;
  const int n = a.nnear[b];
  const double qnan = __builtin_nan("");
  if (n < 2) {
    // no neighbour: NaN; one neighbour: the reference skips the solve
    // (kriging_base.F90:2127), NaN here as in the forward path
    if (j == 0) a.zhat[b] = qnan;
    if (j < NGRAD) a.dz[(size_t)b * NGRAD + j] = 0.0;
    return;
  }
  const int ld = a.mmax + 1;
  const int* in = a.inear + (size_t)b * a.mmax;
  const bool row = j < n;
  const int idx = row ? in[j] : 0;
  double x0[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) x0[k] = a.bcoord[k + DIM * (size_t)b];
  // neighbour coordinates relative to the target (node-local origin)
#pragma unroll
  for (int k = 0; k < DIM; ++k) sx[k][j] = row ? a.ocoord[k + DIM * (size_t)idx] - x0[k] : 0.0;
  __syncthreads();
  double dc[NGRAD];
  if (row) {
    for (int m = 0; m <= j; ++m) {
      double l[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) l[k] = sx[k][j] - sx[k][m];
      Ks[j * ld + m] = m == j ? M.cov0 + a.ovar[idx] : cov_grad<DIM>(M, a.dmat, l, dc);
    }
    double l[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) l[k] = sx[k][j];
    rh[0][j] = cov_grad<DIM>(M, a.dmat, l, dc);
    rh[1][j] = 1.0;
    rh[2][j] = a.ovalue[idx];
  }
  __syncthreads();
  // Cholesky K = L L' (lower triangle, in place), lanes own rows
  bool ok = true;
  for (int c = 0; c < n; ++c) {
    const double piv = Ks[c * ld + c];
    if (!(piv > 0.0)) {
      ok = false;
      break;
    }
    const double d = sqrt(piv);
    __syncthreads();
    if (j == c) Ks[c * ld + c] = d;
    if (j > c && row) Ks[j * ld + c] = Ks[j * ld + c] / d;
    __syncthreads();
    if (j > c && row) {
      const double ljc = Ks[j * ld + c];
      for (int m = c + 1; m <= j; ++m) Ks[j * ld + m] -= ljc * Ks[m * ld + c];
    }
    __syncthreads();
  }
  if (!ok) {
    if (j == 0) a.zhat[b] = qnan;
    if (j < NGRAD) a.dz[(size_t)b * NGRAD + j] = qnan;
    return;
  }
  // K^-1 k, K^-1 1, K^-1 z: forward then backward substitution
  for (int c = 0; c < n; ++c) {
    if (j < 3) rh[j][c] = rh[j][c] / Ks[c * ld + c];
    __syncthreads();
    if (j > c && row) {
      const double l = Ks[j * ld + c];
#pragma unroll
      for (int r = 0; r < 3; ++r) rh[r][j] -= l * rh[r][c];
    }
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; --c) {
    if (j < 3) rh[j][c] = rh[j][c] / Ks[c * ld + c];
    __syncthreads();
    if (j < c) {
      const double l = Ks[c * ld + j];
#pragma unroll
      for (int r = 0; r < 3; ++r) rh[r][j] -= l * rh[r][c];
    }
    __syncthreads();
  }
  const double kk = row ? rh[0][j] : 0.0, k1 = row ? rh[1][j] : 0.0, kz = row ? rh[2][j] : 0.0;
  const double uu = wave_sum(k1), uk = wave_sum(kk), uz = wave_sum(kz);
  const double mu = (uk - 1.0) / uu;
  const double w = kk - mu * k1;      // kriging weight of row j
  const double y = kz - (uz / uu) * k1;  // adjoint weight of row j
  __shared__ double sw[GMAX];
  sw[j] = row ? w : 0.0;
  const double zj = row ? a.ovalue[idx] : 0.0;
  const double zh = wave_sum(row ? w * zj : 0.0);
  __syncthreads();
  // row j of  y' dk - y' dK w
  double g[NGRAD];
#pragma unroll
  for (int p = 0; p < NGRAD; ++p) g[p] = 0.0;
  if (row) {
    double l[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) l[k] = sx[k][j];
    (void)cov_grad<DIM>(M, a.dmat, l, dc);
#pragma unroll
    for (int p = 0; p < NGRAD; ++p) g[p] = y * dc[p];
    for (int m = 0; m < n; ++m) {
      const double wm = sw[m];
      if (m == j) {
        g[0] -= y * wm;  // diagonal: d cov0 / d nugget = d cov0 / d sill = 1
        g[1] -= y * wm;
        continue;
      }
#pragma unroll
      for (int k = 0; k < DIM; ++k) l[k] = sx[k][j] - sx[k][m];
      (void)cov_grad<DIM>(M, a.dmat, l, dc);
#pragma unroll
      for (int p = 0; p < NGRAD; ++p) g[p] -= y * wm * dc[p];
    }
  }
#pragma unroll
  for (int p = 0; p < NGRAD; ++p) g[p] = wave_sum(g[p]);
  if (j == 0) a.zhat[b] = zh;
#pragma unroll
  for (int p = 0; p < NGRAD; ++p)
    if (j == p) a.dz[(size_t)b * NGRAD + p] = g[p];
}

// The same quantities with G lanes per block (systems of at most G rows), 64 / G
// blocks per 64-lane workgroup.  Lane r of a group owns row r of K in
// registers; one elimination serves the estimate and the adjoint:
//   1. row r of K (cov0 + variance on the diagonal) and k_r = C(x_r - x0);
//   2. Gaussian elimination of [K | k 1 z] without pivoting (K is symmetric
//      positive definite; a non-positive pivot makes the block NaN), the
//      pivot column broadcast through LDS;
//   3. back substitution for K^-1 k, K^-1 1, K^-1 z, then w, the adjoint y
//      and zhat = w'z as above;
//   4. y' dk - y' dK w over the unordered pairs of the block: lane r takes
//      the pairs (r, r + d mod n), d = 1..n/2 (for even n the pair at
//      d = n/2 once), so each pair's covariance derivatives are evaluated
//      once and weighted by y_r w_m + y_m w_r.
// Padding rows (r >= n) are rows of the identity with a zero right-hand side.
template <int DIM, int G>
__global__ __launch_bounds__(64) void k_loo_grad_g(GradArgs a) {
  constexpr int S = 64 / G;
  __shared__ double sx[S][DIM][G];
  __shared__ double Ks[S][G][G + 1];  // each lane's assembled row (its own scratch)
  __shared__ double col[S][G];
  __shared__ double ext[S][3];
  __shared__ double sw[S][G], sy[S][G];
  const int t = threadIdx.x, s = t / G, r = t % G;
  const int b = blockIdx.x * S + s;
  const bool act = b < a.nb;
  const int n0 = act ? a.nnear[b] : 0;
  // no neighbour: NaN; one neighbour: the reference skips the solve
  // (kriging_base.F90:2127), NaN here as in the forward path.  (The host
  // picks G from the largest count, so n0 <= G; a larger count would be
  // left unsolved, never read past the group.)
  const int n = (n0 >= 2 && n0 <= G) ? n0 : 0;
  const bool live = r < n;
  const int idx = live ? a.inear[(size_t)b * a.mmax + r] : 0;
  const Model& M = *a.model;
  // neighbour coordinates relative to the target (node-local origin)
  double x[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    x[k] = live ? a.ocoord[k + DIM * (size_t)idx] - a.bcoord[k + DIM * (size_t)b] : 0.0;
    sx[s][k][r] = x[k];
  }
  const double z = live ? a.ovalue[idx] : 0.0;
  const double ov = live ? a.ovar[idx] : 0.0;
  const int nw = __builtin_amdgcn_readfirstlane(wmax(n));
  __syncthreads();
  for (int m = 0; m < nw; ++m) {
    double v = m == r ? 1.0 : 0.0;
    if (live && m < n) {
      if (m == r) {
        v = M.cov0 + ov;
      } else {
        double l[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) l[k] = x[k] - sx[s][k][m];
        v = cov_grad<DIM, false>(M, a.dmat, l, nullptr);
      }
    }
    Ks[s][r][m] = v;
  }
  double A[G];
#pragma unroll
  for (int m = 0; m < G; ++m) A[m] = m < nw ? Ks[s][r][m] : (m == r ? 1.0 : 0.0);
  double cb = live ? cov_grad<DIM, false>(M, a.dmat, x, nullptr) : 0.0;
  double c1 = live ? 1.0 : 0.0, cz = z;
  int fail = 0;
#pragma unroll
  for (int k = 0; k < G; ++k) {
    if (k < nw) {
      col[s][r] = A[k];
      if (r == k) {
        ext[s][0] = cb;
        ext[s][1] = c1;
        ext[s][2] = cz;
      }
      __syncthreads();
      const double akk = col[s][k];
      fail |= (k < n) && !(akk > 0.0);
      const double f = r > k ? A[k] * recip(akk) : 0.0;
#pragma unroll
      for (int j = k + 1; j < G; ++j) A[j] = fma(-f, col[s][j], A[j]);
      cb = fma(-f, ext[s][0], cb);
      c1 = fma(-f, ext[s][1], c1);
      cz = fma(-f, ext[s][2], cz);
      __syncthreads();
    }
  }
  double xb = 0.0, x1 = 0.0, xz = 0.0;
#pragma unroll
  for (int k = G - 1; k >= 0; --k) {
    if (k < nw) {
      if (r == k) {
        const double id = recip(A[k]);
        xb = cb * id;
        x1 = c1 * id;
        xz = cz * id;
        ext[s][0] = xb;
        ext[s][1] = x1;
        ext[s][2] = xz;
      }
      __syncthreads();
      if (r < k) {
        cb = fma(-A[k], ext[s][0], cb);
        c1 = fma(-A[k], ext[s][1], c1);
        cz = fma(-A[k], ext[s][2], cz);
      }
      __syncthreads();
    }
  }
  fail = gmax<G>(fail);
  const double uu = gsum<G>(x1), uk = gsum<G>(xb), uz = gsum<G>(xz);
  const double mu = (uk - 1.0) / uu;
  const double w = live ? xb - mu * x1 : 0.0;     // kriging weight of row r
  const double y = live ? xz - (uz / uu) * x1 : 0.0;  // adjoint weight of row r
  const double zh = gsum<G>(w * z);
  sw[s][r] = w;
  sy[s][r] = y;
  __syncthreads();
  double g[NGRAD], dc[NGRAD];
#pragma unroll
  for (int p = 0; p < NGRAD; ++p) g[p] = 0.0;
  if (live) {
    (void)cov_grad<DIM, true>(M, a.dmat, x, dc);
#pragma unroll
    for (int p = 0; p < NGRAD; ++p) g[p] = y * dc[p];
    g[0] -= y * w;  // diagonal: d cov0 / d nugget = d cov0 / d sill = 1
    g[1] -= y * w;
  }
  const int half = n / 2;
  for (int d = 1; d <= nw / 2; ++d) {
    if (live && d <= half && !(2 * d == n && r >= d)) {
      int m = r + d;
      if (m >= n) m -= n;
      double l[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) l[k] = x[k] - sx[s][k][m];
      (void)cov_grad<DIM, true>(M, a.dmat, l, dc);
      const double cf = y * sw[s][m] + sy[s][m] * w;
#pragma unroll
      for (int p = 0; p < NGRAD; ++p) g[p] = fma(-cf, dc[p], g[p]);
    }
  }
  double gr = 0.0;
#pragma unroll
  for (int p = 0; p < NGRAD; ++p) {
    const double v = gsum<G>(g[p]);
    if (r == p) gr = v;
  }
  if (!act) return;
  const double qnan = __builtin_nan("");
  if (r == 0) a.zhat[b] = (n == 0 || fail) ? qnan : zh;
  if (r < NGRAD) a.dz[(size_t)b * NGRAD + r] = n == 0 ? 0.0 : (fail ? qnan : gr);
}

// LOO squared error and gradient sums reduced on the device in a fixed order
// (no atomics), so repeated evaluations give identical sums: RED_BLOCKS
// workgroups each sum a strided share of the blocks, then one workgroup adds
// the partials in workgroup order.  out[0] = sum r^2, out[1 + p] = sum r dz_p,
// out[NGRAD + 1] = number of finite estimates, r = zhat - z.
constexpr int RED_BLOCKS = 256, RED_T = 256, RED_N = NGRAD + 2;

__global__ __launch_bounds__(RED_T) void k_loo_reduce(const double* zhat, const double* dz, const double* val,
                                                      int nb, double* part) {
  __shared__ double sh[RED_N][RED_T];
  double acc[RED_N];
#pragma unroll
  for (int i = 0; i < RED_N; ++i) acc[i] = 0.0;
  for (int b = blockIdx.x * RED_T + threadIdx.x; b < nb; b += RED_BLOCKS * RED_T) {
    const double z = zhat[b];
    if (!isfinite(z)) continue;
    const double r = z - val[b];
    acc[0] += r * r;
#pragma unroll
    for (int p = 0; p < NGRAD; ++p) acc[1 + p] += r * dz[(size_t)b * NGRAD + p];
    acc[NGRAD + 1] += 1.0;
  }
#pragma unroll
  for (int i = 0; i < RED_N; ++i) sh[i][threadIdx.x] = acc[i];
  __syncthreads();
  for (int s = RED_T / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int i = 0; i < RED_N; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + s];
    }
    __syncthreads();
  }
  if ((int)threadIdx.x < RED_N) part[(size_t)blockIdx.x * RED_N + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_loo_reduce_final(const double* part, double* out) {
  const int i = threadIdx.x;
  if (i >= RED_N) return;
  double s = 0.0;
  for (int k = 0; k < RED_BLOCKS; ++k) s += part[(size_t)k * RED_N + i];
  out[i] = s;
}

int loo_reduce_words() { return (RED_BLOCKS + 1) * RED_N; }

hipError_t launch_loo_reduce(const double* zhat, const double* dz, const double* val, int nb, double* work,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_loo_reduce, dim3(RED_BLOCKS), dim3(RED_T), 0, st, zhat, dz, val, nb, work);
  hipLaunchKernelGGL(k_loo_reduce_final, dim3(1), dim3(64), 0, st, work, work + (size_t)RED_BLOCKS * RED_N);
  return hipGetLastError();
}

template <int G>
void launch_loo_grad_g(const GradArgs& a, hipStream_t st) {
  const dim3 grid((a.nb + 64 / G - 1) / (64 / G));
  if (a.ndim == 1) hipLaunchKernelGGL((k_loo_grad_g<1, G>), grid, dim3(64), 0, st, a);
  else if (a.ndim == 2) hipLaunchKernelGGL((k_loo_grad_g<2, G>), grid, dim3(64), 0, st, a);
  else hipLaunchKernelGGL((k_loo_grad_g<3, G>), grid, dim3(64), 0, st, a);
}

// largest neighbour count over the blocks, into *out (cleared first)
__global__ __launch_bounds__(256) void k_max_count(const int* nnear, int nb, int* out) {
  int m = 0;
  for (int b = blockIdx.x * 256 + threadIdx.x; b < nb; b += gridDim.x * 256) m = max(m, nnear[b]);
  m = wmax(m);
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_zero_int(int* p) { *p = 0; }

hipError_t launch_max_count(const int* nnear, int nb, int* out, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, out);
  const int grid = std::max(1, std::min((nb + 255) / 256, 1024));
  hipLaunchKernelGGL(k_max_count, dim3(grid), dim3(256), 0, st, nnear, nb, out);
  return hipGetLastError();
}

hipError_t launch_loo_grad(const GradArgs& a, hipStream_t st) {
  if (a.nb <= 0) return hipSuccess;
  if (a.mmax > GMAX || a.rows > a.mmax) return hipErrorInvalidValue;
  // systems of at most 32 rows: several blocks per wave
  if (a.rows <= 8) {
    launch_loo_grad_g<8>(a, st);
    return hipGetLastError();
  }
  if (a.rows <= 16) {
    launch_loo_grad_g<16>(a, st);
    return hipGetLastError();
  }
  if (a.rows <= 32) {
    launch_loo_grad_g<32>(a, st);
    return hipGetLastError();
  }
  const size_t lds = sizeof(double) * (size_t)a.mmax * (a.mmax + 1);
  if (a.ndim == 1) hipLaunchKernelGGL(k_loo_grad<1>, dim3(a.nb), dim3(64), lds, st, a);
  else if (a.ndim == 2) hipLaunchKernelGGL(k_loo_grad<2>, dim3(a.nb), dim3(64), lds, st, a);
  else hipLaunchKernelGGL(k_loo_grad<3>, dim3(a.nb), dim3(64), lds, st, a);
  return hipGetLastError();
}

}  // namespace kk
