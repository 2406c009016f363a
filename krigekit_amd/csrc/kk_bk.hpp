// Bunch-Kaufman factorisation and solve of a symmetric indefinite system, the
// reference's fallback when the Cholesky path fails (ssytrf_setup /
// ssytrs_solve, solver.f90:337-406, on the full augmented system).
//
// Restated from the vendored LAPACK (src/libkriging/lapack.f): SSYTF2 with
// UPLO = 'U' (:6213-6445; SSYTRF runs it for every system of up to 64 rows,
// :6811-6900), ISAMAX (:669-720), the reference loop orders of SSYR, SGER,
// SGEMV and SSCAL, and SSYTRS with UPLO = 'U' (:7478-7600).  The pivot
// decisions, the order of every update and the exact-zero test that sets INFO
// are the reference's, and the arithmetic is not contracted into fused
// multiply-adds (the reference's x86-64 build has none), so a singular system
// is detected as it is there.
//
// Column-major with leading dimension ld, 0-based indices; ipiv holds
// LAPACK's 1-based codes (k+1 for a 1x1 pivot, -(kp+1) on both columns of a
// 2x2 pivot).  One thread runs one system; the matrices live in scratch memory.
#pragma once
#include <hip/hip_runtime.h>

namespace kk {

// first index of the largest magnitude among x[0], x[inc], ... (ISAMAX)
__device__ inline int bk_iamax(int n, const double* x, int inc) {
  int im = 0;
  double mx = fabs(x[0]);
  for (int i = 1; i < n; ++i) {
    const double v = fabs(x[(size_t)i * inc]);
    if (v > mx) {
      mx = v;
      im = i;
    }
  }
  return im;
}

// INFO: 0, or the 1-based column whose pivot block is exactly zero
__device__ inline int bk_factor(double* A, int n, int ld, int* ipiv) {
#pragma clang fp contract(off)
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  int info = 0;
  int k = n - 1;
  while (k >= 0) {
    int kstep = 1, kp, imax = 0;
    const double absakk = fabs(A[k + ld * k]);
    double colmax = 0.0;
    if (k > 0) {
      imax = bk_iamax(k, A + (size_t)ld * k, 1);
      colmax = fabs(A[imax + ld * k]);
    }
    const double mx = absakk > colmax ? absakk : colmax;
    if (mx == 0.0 || isnan(absakk)) {
      if (info == 0) info = k + 1;
      kp = k;
    } else {
      if (absakk >= alpha * colmax) {
        kp = k;
      } else {
        int jmax = imax + 1 + bk_iamax(k - imax, A + imax + (size_t)ld * (imax + 1), ld);
        double rowmax = fabs(A[imax + ld * jmax]);
        if (imax > 0) {
          jmax = bk_iamax(imax, A + (size_t)ld * imax, 1);
          const double v = fabs(A[jmax + ld * imax]);
          rowmax = rowmax > v ? rowmax : v;
        }
        if (absakk >= alpha * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (fabs(A[imax + ld * imax]) >= alpha * rowmax) {
          kp = imax;
        } else {
          kp = imax;
          kstep = 2;
        }
      }
      const int kk = k - kstep + 1;
      if (kp != kk) {
        // interchange rows and columns kk and kp of the leading submatrix
        for (int i = 0; i < kp; ++i) {
          const double t = A[i + ld * kk];
          A[i + ld * kk] = A[i + ld * kp];
          A[i + ld * kp] = t;
        }
        for (int i = kp + 1; i < kk; ++i) {
          const double t = A[i + ld * kk];
          A[i + ld * kk] = A[kp + ld * i];
          A[kp + ld * i] = t;
        }
        double t = A[kk + ld * kk];
        A[kk + ld * kk] = A[kp + ld * kp];
        A[kp + ld * kp] = t;
        if (kstep == 2) {
          t = A[(k - 1) + ld * k];
          A[(k - 1) + ld * k] = A[kp + ld * k];
          A[kp + ld * k] = t;
        }
      }
      if (kstep == 1) {
        // rank-1 update of the leading block (upper triangle), then scale
        const double r1 = 1.0 / A[k + ld * k];
        for (int j = 0; j < k; ++j) {
          const double xj = A[j + ld * k];
          if (xj != 0.0) {
            const double tmp = -r1 * xj;
            for (int i = 0; i <= j; ++i) A[i + ld * j] = A[i + ld * j] + A[i + ld * k] * tmp;
          }
        }
        for (int i = 0; i < k; ++i) A[i + ld * k] = r1 * A[i + ld * k];
      } else if (k > 1) {
        // rank-2 update with the inverse of the 2x2 pivot block
        double d12 = A[(k - 1) + ld * k];
        const double d22 = A[(k - 1) + ld * (k - 1)] / d12;
        const double d11 = A[k + ld * k] / d12;
        const double t = 1.0 / (d11 * d22 - 1.0);
        d12 = t / d12;
        for (int j = k - 2; j >= 0; --j) {
          const double wkm1 = d12 * (d11 * A[j + ld * (k - 1)] - A[j + ld * k]);
          const double wk = d12 * (d22 * A[j + ld * k] - A[j + ld * (k - 1)]);
          for (int i = j; i >= 0; --i) A[i + ld * j] = A[i + ld * j] - A[i + ld * k] * wk - A[i + ld * (k - 1)] * wkm1;
          A[j + ld * k] = wk;
          A[j + ld * (k - 1)] = wkm1;
        }
      }
    }
    if (kstep == 1) {
      ipiv[k] = kp + 1;
    } else {
      ipiv[k] = -(kp + 1);
      ipiv[k - 1] = -(kp + 1);
    }
    k -= kstep;
  }
  return info;
}

// solve with the factors of bk_factor for nrhs right-hand sides B (ldb)
__device__ inline void bk_solve(const double* A, int n, int ld, const int* ipiv, double* B, int nrhs, int ldb) {
#pragma clang fp contract(off)
  auto swaprow = [&](int r1, int r2) {
    for (int j = 0; j < nrhs; ++j) {
      const double t = B[r1 + ldb * j];
      B[r1 + ldb * j] = B[r2 + ldb * j];
      B[r2 + ldb * j] = t;
    }
  };
  // backward sweep: U D y = P b
  int k = n - 1;
  while (k >= 0) {
    if (ipiv[k] > 0) {
      const int kp = ipiv[k] - 1;
      if (kp != k) swaprow(k, kp);
      for (int j = 0; j < nrhs; ++j) {
        const double y = B[k + ldb * j];
        if (y != 0.0) {
          const double tmp = -y;
          for (int i = 0; i < k; ++i) B[i + ldb * j] = B[i + ldb * j] + A[i + ld * k] * tmp;
        }
      }
      const double r = 1.0 / A[k + ld * k];
      for (int j = 0; j < nrhs; ++j) B[k + ldb * j] = r * B[k + ldb * j];
      k -= 1;
    } else {
      const int kp = -ipiv[k] - 1;
      if (kp != k - 1) swaprow(k - 1, kp);
      for (int c = 0; c < 2; ++c) {
        const int col = k - c;
        for (int j = 0; j < nrhs; ++j) {
          const double y = B[col + ldb * j];
          if (y != 0.0) {
            const double tmp = -y;
            for (int i = 0; i < k - 1; ++i) B[i + ldb * j] = B[i + ldb * j] + A[i + ld * col] * tmp;
          }
        }
      }
      const double akm1k = A[(k - 1) + ld * k];
      const double akm1 = A[(k - 1) + ld * (k - 1)] / akm1k;
      const double ak = A[k + ld * k] / akm1k;
      const double denom = akm1 * ak - 1.0;
      for (int j = 0; j < nrhs; ++j) {
        const double bkm1 = B[(k - 1) + ldb * j] / akm1k;
        const double bk = B[k + ldb * j] / akm1k;
        B[(k - 1) + ldb * j] = (ak * bkm1 - bk) / denom;
        B[k + ldb * j] = (akm1 * bk - bkm1) / denom;
      }
      k -= 2;
    }
  }
  // forward sweep: U' x = y
  k = 0;
  while (k < n) {
    const bool two = ipiv[k] < 0;
    for (int c = 0; c <= (two ? 1 : 0); ++c) {
      const int col = k + c;
      if (k == 0) continue;
      for (int j = 0; j < nrhs; ++j) {
        double tmp = 0.0;
        for (int i = 0; i < k; ++i) tmp = tmp + B[i + ldb * j] * A[i + ld * col];
        B[col + ldb * j] = B[col + ldb * j] + (-1.0) * tmp;
      }
    }
    const int kp = (two ? -ipiv[k] : ipiv[k]) - 1;
    if (kp != k) swaprow(k, kp);
    k += two ? 2 : 1;
  }
}

// The reference's fallback for one augmented system of m rows, n of them data
// rows (kriging_base.F90:3299-3345): factor a copy A of K; when the
// factorisation is exactly singular, add 1e-6 * mean(data diagonal) to the
// data diagonal of K (as the reference modifies its matA) and factor once
// more.  On success B (nrhs right-hand sides, ldb = ld) holds the solutions.
// Returns INFO (0 = solved).
__device__ inline int bk_fallback(double* K, double* A, int* ipiv, int n, int m, int ld, double* B, int nrhs) {
#pragma clang fp contract(off)
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) A[i + ld * j] = K[i + ld * j];
  int info = bk_factor(A, m, ld, ipiv);
  if (info != 0 && n > 0) {
    double ds = 0.0;
    for (int i = 0; i < n; ++i) ds = ds + K[i + ld * i];
    if (ds > 0.0) {
      ds = 1.0e-6 * (ds / (double)n);
      for (int i = 0; i < n; ++i) K[i + ld * i] = K[i + ld * i] + ds;
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) A[i + ld * j] = K[i + ld * j];
      info = bk_factor(A, m, ld, ipiv);
    }
  }
  if (info == 0) bk_solve(A, m, ld, ipiv, B, nrhs, ld);
  return info;
}

}  // namespace kk
