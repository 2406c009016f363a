// Multi-variable (co)kriging per block, as the reference assembles and solves
// it for nvar > 1 (kriging_base.F90:3159-3235, kriging.F90:714-1028,
// solver.f90:224-300, kriging_base.F90:3407-3503):
//
//   rows       the variables' neighbour groups, in variable order;
//   K(r, c)    C from vgm(max(v_r, v_c), min(v_r, v_c)) at (x_r - x_c)/rs,
//              diagonal cov0 + observation variance + local nugget;
//   F          constraint rows: std_ck -> F(r, v_r) = 1 (one per variable),
//              otherwise a single row of ones; none for simple kriging;
//   b_q        per target variable q: sum_k w_k C_{vgm(v_r, q)}((x_r - g_k)/rs)
//              over the block's integration nodes; d_q = e_q or 1;
//   solution   Cholesky K = L L', S = F' K^-1 F, mu_q = S^-1 (F' K^-1 b_q - d_q),
//              w_q = K^-1 b_q - K^-1 F mu_q (the bordered system's solution);
//   estimate   w_q' z, plus the Isaaks-Srivastava shift (non-standard
//              cokriging) and the simple-kriging mean term, clipped to bounds;
//   variance   base_ij - (w_i' b_j + mu_i' d_j), symmetric, diagonal and
//              off-diagonal clamped at 0; base = cov0 of vgm(i, j) for points,
//              or sum_k sum_l w_k w_l C_ij(g_k - g_l) (lag not range-scaled).
// The model array carries, in table mode, lookup tables only for vgm(a, b)
// with a <= b -- the pairs the reference builds tables for -- and analytic
// models for the others, so every evaluation matches the reference's.
//
// One 128-thread workgroup per block: thread r owns row r (rows <= 128), K
// and the solve columns live in LDS, FP64 throughout.
#include <hip/hip_runtime.h>

#include "kk_cov.hpp"
#include "kk_device.hpp"
#include "kk_types.hpp"

namespace kk {

constexpr int MVT = 128;           // threads per workgroup = max rows
constexpr int MAXCOL = 2 * MAXVAR; // right-hand sides + constraint columns
constexpr int MAXSUB = 64;

// sum over the workgroup (two waves); every thread gets the total
__device__ inline double block_sum(double v, double* part) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) part[w] = v;
  __syncthreads();
  return part[0] + part[1];
}

template <int DIM>
__device__ inline double mv_cov(const Model* models, int a, int b, int nvar, const double* l) {
  return cov_general<DIM>(models + a + nvar * b, l[0], DIM > 1 ? l[1] : 0.0, DIM > 2 ? l[2] : 0.0);
}

template <int DIM, int NV>
__global__ __launch_bounds__(MVT) void k_solve_mv(MVArgs a) {
  extern __shared__ double lds[];
  __shared__ double sx[DIM][MVT];
  __shared__ double sz[MVT];
  __shared__ int svr[MVT];
  __shared__ double sg[DIM][MAXSUB];
  __shared__ double sgw[MAXSUB];
  __shared__ double part[2];
  const int b = blockIdx.x, t = threadIdx.x;
  constexpr int nvar = NV;
  int nv[NV], st[NV], n = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    nv[v] = v < nvar ? a.nnear[v][b] : 0;
    st[v] = n;
    n += nv[v];
  }
  const double qnan = __builtin_nan("");
  if (n < 2) {
    // no neighbour: NaN; one neighbour: the reference skips the solve
    // (kriging_base.F90:2127), NaN here
    if (t < nvar) a.est[b + (size_t)a.nb * t] = qnan;
    if (t < nvar * nvar) a.var[b + (size_t)a.nb * t] = qnan;
    if (t == 0) a.flag[b] = 0;
    return;
  }
  const int p = a.unbias ? (a.std_ck ? nvar : 1) : 0;
  const int ncol = nvar + p;
  const int ld = a.ld;
  double* Kd = lds;                         // rows x ld, lower triangle by rows
  double* cols = lds + (size_t)ld * ld;     // MAXCOL x ld
  const double rs = a.rangescale ? a.rangescale[b] : 1.0;
  const double ln = a.localnugget ? a.localnugget[b] : 0.0;
  double x0[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) x0[k] = a.bcoord[k + DIM * (size_t)b];
  // rows: variable, coordinates relative to the block centre, value
  const bool row = t < n;
  int vr = 0, idx = 0;
  double ovar = 0.0;
  if (row) {
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (v < nvar && t >= st[v] && t < st[v] + nv[v]) vr = v;
    idx = a.inear[vr][(size_t)b * a.mmax[vr] + (t - st[vr])];
#pragma unroll
    for (int k = 0; k < DIM; ++k) sx[k][t] = a.ocoord[vr][k + DIM * (size_t)idx] - x0[k];
    sz[t] = a.ovalue[vr][idx];
    ovar = a.ovar[vr][idx];
    svr[t] = vr;
  }
  const int ns = a.nsub;
  if (t < ns) {
    const size_t s = (size_t)b * ns + t;
#pragma unroll
    for (int k = 0; k < DIM; ++k) sg[k][t] = ns > 1 ? a.scoord[k + DIM * s] - x0[k] : 0.0;
    sgw[t] = ns > 1 ? a.sweight[s] : 1.0;
  }
  __syncthreads();
  // ---- assembly ---------------------------------------------------------------
  double bq[NV];
  if (row) {
    for (int m = 0; m < t; ++m) {
      const int vm = svr[m];
      double l[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) l[k] = (sx[k][t] - sx[k][m]) / rs;
      Kd[t * ld + m] = mv_cov<DIM>(a.models, vr > vm ? vr : vm, vr > vm ? vm : vr, nvar, l);
    }
    Kd[t * ld + t] = a.models[vr + nvar * vr].cov0 + ovar + ln;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      if (q < nvar) {
        for (int k = 0; k < ns; ++k) {
          double l[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) l[d] = (sx[d][t] - sg[d][k]) / rs;
          s += mv_cov<DIM>(a.models, vr, q, nvar, l) * sgw[k];
        }
        cols[q * ld + t] = s;
      }
      bq[q] = s;
    }
    for (int c = 0; c < p; ++c) cols[(nvar + c) * ld + t] = (a.std_ck ? (vr == c) : true) ? 1.0 : 0.0;
  }
  __syncthreads();
  // ---- Cholesky K = L L' ------------------------------------------------------
  bool ok = true;
  for (int c = 0; c < n; ++c) {
    const double piv = Kd[c * ld + c];
    if (!(piv > 0.0)) {
      ok = false;
      break;
    }
    const double d = sqrt(piv);
    __syncthreads();
    if (t == c) Kd[c * ld + c] = d;
    if (t > c && row) Kd[t * ld + c] = Kd[t * ld + c] / d;
    __syncthreads();
    if (t > c && row) {
      const double ltc = Kd[t * ld + c];
      for (int m = c + 1; m <= t; ++m) Kd[t * ld + m] -= ltc * Kd[m * ld + c];
    }
    __syncthreads();
  }
  if (!ok) {
    if (t < nvar) a.est[b + (size_t)a.nb * t] = qnan;
    if (t < nvar * nvar) a.var[b + (size_t)a.nb * t] = qnan;
    if (t == 0) a.flag[b] = 3;
    return;
  }
  // ---- K^-1 [b_0 .. b_{nvar-1}, F_0 .. F_{p-1}] -----------------------------------
  for (int c = 0; c < n; ++c) {
    if (t < ncol) cols[t * ld + c] = cols[t * ld + c] / Kd[c * ld + c];
    __syncthreads();
    if (t > c && row) {
      const double l = Kd[t * ld + c];
      for (int j = 0; j < ncol; ++j) cols[j * ld + t] -= l * cols[j * ld + c];
    }
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; --c) {
    if (t < ncol) cols[t * ld + c] = cols[t * ld + c] / Kd[c * ld + c];
    __syncthreads();
    if (t < c) {
      const double l = Kd[c * ld + t];
      for (int j = 0; j < ncol; ++j) cols[j * ld + t] -= l * cols[j * ld + c];
    }
    __syncthreads();
  }
  double kb[NV], kf[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    kb[q] = (row && q < nvar) ? cols[q * ld + t] : 0.0;
    kf[q] = (row && q < p) ? cols[(nvar + q) * ld + t] : 0.0;
  }
  // ---- Schur complement S = F' K^-1 F, G = F' K^-1 b, multipliers ---------------
  double S[NV][NV], G[NV][NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double fi = row && i < p ? ((a.std_ck ? (vr == i) : true) ? 1.0 : 0.0) : 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      S[i][j] = (i < p && j < p) ? block_sum(fi * kf[j], part) : 0.0;
      G[i][j] = (i < p && j < nvar) ? block_sum(fi * kb[j], part) : 0.0;
    }
  }
  // mu_q = S^-1 (G_q - d_q): Gaussian elimination with partial pivoting on the
  // small p x p system, every thread the same
  double mu[NV][NV];  // mu[q][a]
#pragma unroll
  for (int q = 0; q < NV; ++q)
#pragma unroll
    for (int i = 0; i < NV; ++i) mu[q][i] = 0.0;
  if (p > 0) {
    double A[NV][NV], R[NV][NV];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        A[i][j] = S[i][j];
        const double d = a.std_ck ? (i == j ? 1.0 : 0.0) : 1.0;
        R[i][j] = G[i][j] - d;  // row i (constraint), column j (target)
      }
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      if (c >= p) break;
      int pr = c;
#pragma unroll
      for (int r = 0; r < NV; ++r)
        if (r > c && r < p && fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        double x = A[c][j];
        A[c][j] = A[pr][j];
        A[pr][j] = x;
        x = R[c][j];
        R[c][j] = R[pr][j];
        R[pr][j] = x;
      }
#pragma unroll
      for (int r = 0; r < NV; ++r) {
        if (r <= c || r >= p) continue;
        const double f = A[r][c] / A[c][c];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          A[r][j] -= f * A[c][j];
          R[r][j] -= f * R[c][j];
        }
      }
    }
#pragma unroll
    for (int c = NV - 1; c >= 0; --c) {
      if (c >= p) continue;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        double s = R[c][q];
#pragma unroll
        for (int j = 0; j < NV; ++j)
          if (j > c && j < p) s -= A[c][j] * mu[q][j];
        mu[q][c] = s / A[c][c];
      }
    }
  }
  // ---- weights, estimates -------------------------------------------------------
  double w[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = kb[q];
#pragma unroll
    for (int c = 0; c < NV; ++c) s -= mu[q][c] * kf[c];
    w[q] = (row && q < nvar) ? s : 0.0;
  }
  const double zr = row ? sz[t] : 0.0;
  const bool isa = !a.std_ck && a.unbias && nvar > 1;
  double gm[NV];  // group means of the neighbour values (Isaaks-Srivastava)
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    gm[v] = 0.0;
    if (isa && v < nvar && nv[v] > 0) gm[v] = block_sum((row && vr == v) ? zr : 0.0, part) / nv[v];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if (q >= nvar) break;
    double e = block_sum(w[q] * zr, part);
    if (isa) {
#pragma unroll
      for (int v = 0; v < NV; ++v)
        if (v < nvar && v != q && nv[v] > 0) e += (gm[q] - gm[v]) * block_sum((row && vr == v) ? w[q] : 0.0, part);
    }
    if (!a.unbias && a.sk_mean[q] != 0.0) e += (1.0 - block_sum(w[q], part)) * a.sk_mean[q];
    e = e < a.bound_lo ? a.bound_lo : e;
    e = e > a.bound_hi ? a.bound_hi : e;
    if (t == 0) a.est[b + (size_t)a.nb * q] = e;
  }
  // ---- variance -------------------------------------------------------------------
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i >= nvar) break;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if (j < i || j >= nvar) continue;
      const Model& M = a.models[i + nvar * j];
      double base;
      if (ns == 1) {
        base = M.cov0;
      } else {
        double s = 0.0;
        for (int e = t; e < ns * ns; e += MVT) {
          const int k = e / ns, l = e % ns;
          if (l < k) continue;
          if (l == k) {
            s += M.cov0 * sgw[k] * sgw[k];
          } else {
            double lag[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) lag[d] = sg[d][k] - sg[d][l];
            s += mv_cov<DIM>(a.models, i, j, nvar, lag) * sgw[k] * sgw[l] * 2.0;
          }
        }
        base = block_sum(s, part);
      }
      double mud = 0.0;
#pragma unroll
      for (int c = 0; c < NV; ++c)
        if (c < p) mud += mu[i][c] * (a.std_ck ? (c == j ? 1.0 : 0.0) : 1.0);
      double v = base - (block_sum(w[i] * bq[j], part) + mud);
      v = v > 0.0 ? v : 0.0;  // diagonal and off-diagonal clamped at 0
      if (t == 0) {
        a.var[b + (size_t)a.nb * (i + nvar * j)] = v;
        a.var[b + (size_t)a.nb * (j + nvar * i)] = v;
      }
    }
  }
  if (t == 0) a.flag[b] = 0;
}

size_t mv_lds_bytes(int ld, int ncol) { return sizeof(double) * ((size_t)ld * ld + (size_t)MAXCOL * ld); }

hipError_t launch_solve_mv(const MVArgs& a, hipStream_t st) {
  if (a.nb <= 0) return hipSuccess;
  const size_t lds = mv_lds_bytes(a.ld, MAXCOL);
  if (a.ld > MVT + 1 || a.nsub > MAXSUB || a.nvar < 2 || a.nvar > MAXVAR) return hipErrorInvalidValue;
  dim3 g(a.nb), blk(MVT);
#define KK_MV(D)                                                                      \
  do {                                                                               \
    if (a.nvar == 2) hipLaunchKernelGGL((k_solve_mv<D, 2>), g, blk, lds, st, a);     \
    else if (a.nvar == 3) hipLaunchKernelGGL((k_solve_mv<D, 3>), g, blk, lds, st, a); \
    else hipLaunchKernelGGL((k_solve_mv<D, 4>), g, blk, lds, st, a);                 \
  } while (0)
  if (a.ndim == 1) KK_MV(1);
  else if (a.ndim == 2) KK_MV(2);
  else KK_MV(3);
#undef KK_MV
  return hipGetLastError();
}

}  // namespace kk
