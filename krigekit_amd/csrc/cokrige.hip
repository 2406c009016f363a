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
//   solution   K = L L', S = F' K^-1 F, mu_q = S^-1 (F' K^-1 b_q - d_q),
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
// One 256-thread workgroup per block, a 16 x 16 thread grid.  The symmetric
// bordered matrix [K X'; X 0], X = (b_0.., F_0.., z, (I_0.. | 1)) as extra
// rows, is held in registers, lower triangle, 2-D block-cyclic: thread
// (tr, tc) owns rows tr + 16a and columns tc + 16b.  Right-looking Cholesky
// over the n pivots of K leaves -X K^-1 X' in the trailing block, i.e. minus
// the Gram matrix of Y = L^-1 X' (c'K^-1 d = y_c' y_d), so no substitution is
// needed: it gives S, F'K^-1 b, z'K^-1 b, ..., and one thread finishes the
// small algebra.  The factorisation takes four pivots per barrier (a panel
// factored by the one wave that holds its columns).  The variable pairs'
// models are staged in LDS.
//
// MF (FP32 factorisation, MFMA trailing update, FP64 refinement; up to 80
// columns): the FP64 bordered matrix stays in registers for the residuals;
// an FP32 copy in LDS (column-major) is factored by right-looking blocked
// Cholesky over 16-column panels: one wave factors the diagonal 16 x 16 tile,
// the rows below are solved against it (one row per thread), and the
// trailing tiles take the rank-16 update A -= L L' on the matrix cores
// (v_mfma_f32_16x16x4f32, four per tile, tiles spread over the waves).  The
// extra rows are eliminated alongside, as in the FP64 path.  Each extra
// column x_f = K^-1 X_f' is then obtained in FP32 by one backward
// substitution, and the Gram matrix is refined with FP64 residuals:
// X_e K^-1 X_f' = X_e x_f + x_e' (X_f' - K x_f) + O(e^2), e the FP32 solve
// error.  A block whose FP32 pivot is not positive or whose pivot ratio
// falls below pivot_tol is flagged (1) and solved again by the FP64 path.
#include <hip/hip_runtime.h>

#include "kk_bk.hpp"
#include "kk_cov.hpp"
#include "kk_device.hpp"
#include "kk_types.hpp"

// k_solve_mv instances are compiled in units of their own, one per
// coordinate count (KK_MV_DIM 1-3) and weights mode (KK_MV_W 0/1), so they
// compile in parallel; the unit without KK_MV_DIM holds the fallback and the
// dispatch
#ifndef KK_MV_DIM
#define KK_MV_DIM 0
#endif
#ifndef KK_MV_W
#define KK_MV_W 0
#endif

namespace kk {

hipError_t launch_solve_mv_d1_w0(const MVArgs& a, int rb, hipStream_t st);
hipError_t launch_solve_mv_d2_w0(const MVArgs& a, int rb, hipStream_t st);
hipError_t launch_solve_mv_d3_w0(const MVArgs& a, int rb, hipStream_t st);
hipError_t launch_solve_mv_d1_w1(const MVArgs& a, int rb, hipStream_t st);
hipError_t launch_solve_mv_d2_w1(const MVArgs& a, int rb, hipStream_t st);
hipError_t launch_solve_mv_d3_w1(const MVArgs& a, int rb, hipStream_t st);

constexpr int MVT = 256;            // threads per workgroup (16 x 16)
constexpr int MVR = 128;            // max rows + extra rows
constexpr int MAXE = 3 * MAXVAR + 1 + MAXDRIFT;  // extra columns: b, F (drift + unbiasedness), z, indicators or ones
constexpr int MAXSUB = 64;

template <int DIM, bool SIMPLE>
__device__ inline double mv_cov(const Model* models, int a, int b, int nvar, double dx, double dy, double dz) {
  const Model* M = models + a + nvar * b;
  if (SIMPLE) return cov_simple<DIM>(*M, dx, dy, dz);
  return cov_general<DIM>(M, dx, dy, dz);
}

// Increment (gradient-pair) rows, out of line so that the kernels' register
// allocation does not pay for them (kriging.F90:796-840, 904-920).
// Element (i, j), j <= i, of increment row i >= no: against a data row
// C(x1 - x) - C(x2 - x) with vgm(increment variable, data variable);
// between increments of one variable c11 - c12 - c21 + c22 of the pair (j, i);
// zero across variables; diagonal 2 C(0) - 2 C(x1 - x2) + variance, without
// the local nugget.  sx: first points (and data rows), stride ldx; sx2: second
// points of the increment rows (index i - no), stride MAXGRAD; all relative
// to the block centre.
template <int DIM, bool SIMPLE>
__device__ __noinline__ double mv_inc_cov(const Model* models, int nvar, const double* sx, int ldx, const double* sx2,
                                          const int* svr, const double* sov, int i, int j, int no, bool scaled,
                                          double rs) {
  auto c = [&](int va, int vb, const double* u, int su, int iu, const double* w, int sw_, int iw) {
    double d[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < DIM; ++k) {
      d[k] = u[k * su + iu] - w[k * sw_ + iw];
      if (scaled) d[k] = d[k] / rs;
    }
    return mv_cov<DIM, SIMPLE>(models, va, vb, nvar, d[0], d[1], d[2]);
  };
  const int vi = svr[i], vj = svr[j];
  if (j < no) return c(vi, vj, sx, ldx, i, sx, ldx, j) - c(vi, vj, sx2, MAXGRAD, i - no, sx, ldx, j);
  if (j == i)
    return (2.0 * models[vi + nvar * vi].cov0 - 2.0 * c(vi, vi, sx, ldx, i, sx2, MAXGRAD, i - no)) + sov[i];
  if (vj != vi) return 0.0;
  const double c11 = c(vi, vi, sx, ldx, j, sx, ldx, i);
  const double c12 = c(vi, vi, sx, ldx, j, sx2, MAXGRAD, i - no);
  const double c21 = c(vi, vi, sx2, MAXGRAD, j - no, sx, ldx, i);
  const double c22 = c(vi, vi, sx2, MAXGRAD, j - no, sx2, MAXGRAD, i - no);
  return ((c11 - c12) - c21) + c22;
}

// sum_k w_k C_{vgm(v, q)}(x2 - g_k) over nodes k0..k1 of increment row gi
template <int DIM, bool SIMPLE>
__device__ __noinline__ double mv_inc_rhs(const Model* models, int nvar, int v, int q, const double* sx2, int gi,
                                          const double* sg, const double* sgw, int k0, int k1, bool scaled,
                                          double rs) {
  double s = 0.0;
  for (int k = k0; k < k1; ++k) {
    double d[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < DIM; ++m) {
      d[m] = sx2[m * MAXGRAD + gi] - sg[m * MAXSUB + k];
      if (scaled) d[m] = d[m] / rs;
    }
    s += mv_cov<DIM, SIMPLE>(models, v, q, nvar, d[0], d[1], d[2]) * sgw[k];
  }
  return s;
}

__device__ inline double readlane_d(double v, int lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// 1 / sqrt(x) for x > 0: the hardware estimate and two Newton steps
// (relative error e -> 1.5 e^2 per step), a few ulp from the rounded value
__device__ inline double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// orders one wave's LDS writes before its later LDS reads (and the reverse):
// a wave's LDS operations complete in issue order, the fence keeps the
// compiler from moving them across
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// sum over the workgroup; every thread gets the total
__device__ inline double wg_sum(double v, double* part) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  return (part[0] + part[1]) + (part[2] + part[3]);
}

// The second launch bound is the minimum number of waves per SIMD the
// registers must allow: five for the 80-column instance (its 31 KB of LDS at
// two variables admits five workgroups per CU; measured faster than four
// despite a few spills), unconstrained for the 128-column one.
//
// WEIGHTS: the instance also writes the kriging weights (the weight store,
// kriging_base.F90:2428-2439, 2760-2782).  The factor is kept in LDS
// (packed lower triangle of the bordered matrix), and after the multipliers
// the weights w_q = K^-1 (b_q - F mu_q) are obtained by one backward
// substitution with the unit factor, all targets together.
constexpr int cmax3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
typedef float mf_f4 __attribute__((ext_vector_type(4)));

// DRIFT instances hold up to MAXDRIFT external-drift constraint rows; the
// others none, so that their one-thread Schur step stays NV x NV (sized for
// the drifts it took the 2-variable solve from 73 to 82 ms).
template <int DIM, int NV, int RB, bool SIMPLE, bool WEIGHTS, bool MF = false, bool GRAD = false, bool DRIFT = GRAD>
__global__ __launch_bounds__(MVT, MF ? 4 : (RB == 5 && !WEIGHTS) ? 5 : 1) void k_solve_mv(MVArgs a) {
  static_assert(!MF || (RB == 5 && !WEIGHTS && !DRIFT),
                "the FP32 MFMA instance covers up to 80 columns, no weight store, no drift");
  static_assert(!GRAD || (!SIMPLE && !WEIGHTS && !MF && DRIFT), "increment rows: general-shape FP64 instances only");
  constexpr int ND = DRIFT ? MAXDRIFT : 0;  // drift constraint rows at most
  // rows + extra rows fit 16 RB (the instance is chosen by the column count)
  constexpr int ROWS = 16 * RB;
  __shared__ double sx[DIM][ROWS];
  // increment rows (GRAD instances only: their code in every general
  // instance took those from 2 to 1 wave per SIMD): the second point
  __shared__ double sx2[GRAD ? DIM : 1][GRAD ? MAXGRAD : 1];
  __shared__ double sz[ROWS];
  __shared__ double sov[ROWS];
  __shared__ int svr[ROWS];
  __shared__ double sg[DIM][MAXSUB];
  __shared__ double sgw[MAXSUB];
  // right-hand sides as NPART partial sums over the integration nodes
  constexpr int NPART = NV == 2 ? 8 : 2;
  // assembly: each thread's RB values of one row slot; then, for the
  // factorisation, the panel broadcast L(:, c0..c0+3), double-buffered.
  // MF: after the assembly the same area holds the packed lower triangle
  // (mirror exchange), then the FP32 matrix (column stride LD32 floats).
  // panel rows padded to PLD doubles (48 bytes): the 16 rows one wave's
  // 128-bit reads take in the trailing update then start in distinct banks
  constexpr int PLD = 6;
  static_assert(RB * MVT >= 2 * ROWS * PLD, "panel buffer fits the assembly slots");
  constexpr int LD32 = ROWS + 4;
  constexpr int SBW = NPART * NV * ROWS;
  constexpr int UW = MF ? cmax3(SBW + RB * MVT, ROWS * LD32 / 2, ROWS * (ROWS + 1) / 2) : SBW + RB * MVT;
  // (MF only: the FP64 instances keep sb and sbuf as arrays of their own,
  // whose accesses the compiler can tell apart; carved out of one array they
  // cost the 2-variable FP64 solve 5%)
  __shared__ __align__(16) double uarea[MF ? UW : 1];
  __shared__ double sb_own[MF ? 1 : NPART][NV][ROWS];
  __shared__ __align__(16) double sbuf_own[MF ? 1 : RB * MVT];
  double(*const sb)[NV][ROWS] = MF ? reinterpret_cast<double(*)[NV][ROWS]>(uarea) : sb_own;
  double* const sbuf = MF ? uarea + SBW : sbuf_own;
  double(*const vpan)[ROWS][PLD] = reinterpret_cast<double(*)[ROWS][PLD]>(sbuf);
  // MF: the refined solutions x_f (columns of K^-1 X', FP64), the matrix-vector
  // products M [x_f; -e_f], 1 / L(k, k), K's diagonal
  constexpr int MXE = 2 * NV + 2 + ND;  // extra columns at most (non-standard cokriging, drift)
  __shared__ double sxd[MF ? MXE : 1][MF ? ROWS : 1];
  __shared__ double red[MF ? ROWS : 1][MF ? MXE : 1];
  __shared__ float sli[MF ? ROWS : 1];
  __shared__ float sdg[MF ? ROWS : 1];
  __shared__ double gram[MAXE][MAXE];
  __shared__ double sdr[DRIFT ? MAXDRIFT : 1][DRIFT ? ROWS : 1];  // external drift values of the rows
  __shared__ double sbase[NV * NV];
  __shared__ double part[4];
  __shared__ double partb[4][NV * NV];

  __shared__ Model smodels[NV * NV];
  __shared__ int sfail;
  // WEIGHTS: A(i, j), j <= i, j < n, at sl[i (i + 1) / 2 + j]; multipliers;
  // the weights finished so far
  __shared__ double sl[WEIGHTS ? ROWS * (ROWS + 1) / 2 : 1];
  __shared__ double smu[NV][NV + ND];
  __shared__ double sw[NV][WEIGHTS ? ROWS : 1];
  const int b = a.list ? a.list[blockIdx.x] : (int)blockIdx.x, t = threadIdx.x;
  // column class tc = t / 16: the 16 threads of a column class (all row
  // classes) sit in one wave, so a column's pivot and scaling never leave it
  const int tc = t >> 4, tr = t & 15;
  const int wave = t >> 6;
  constexpr int nvar = NV;
  int nv[NV], st[NV], n = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    nv[v] = a.nnear[v][b];
    st[v] = n;
    n += nv[v];
  }
  // increment rows after the neighbour groups (gst[v]: first of variable v)
  const int no = n;
  int gst[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    gst[v] = n;
    if constexpr (GRAD) n += a.ngrad[v];
  }
  const double qnan = __builtin_nan("");
  if (no == 0 || n < 2) {
    // no neighbour (increments alone do not make a system,
    // kriging_base.F90:3198-3207): NaN; one row: the reference skips the solve
    // (kriging_base.F90:2127), NaN here
    if (!a.weights_only) {
      if (t < nvar) a.est[b + (size_t)a.nb * t] = qnan;
      if (t < nvar * nvar) a.var[b + (size_t)a.nb * t] = qnan;
    }
    if (t == 0) a.flag[b] = 0;
    return;
  }
  // constraint rows: the external drifts, then unbiasedness (one per variable
  // under std_ck, else one row of ones)
  const int nd = DRIFT ? a.ndrift : 0;
  const int p = nd + (a.unbias ? (a.std_ck ? nvar : 1) : 0);
  const bool isa = !a.std_ck && a.unbias && nvar > 1;  // Isaaks-Srivastava shift
  // right-hand side of constraint row c for target q
  auto dval = [&](int c, int q) -> double {
    if (c < nd) return a.bdrift[c + nd * ((size_t)q + nvar * (size_t)b)];
    return a.std_ck ? (c - nd == q ? 1.0 : 0.0) : 1.0;
  };
  // extra column layout (offsets from n)
  const int eb = 0, ef = nvar, ez = nvar + p, ei = ez + 1;
  const int ne = ez + 1 + (isa ? nvar : 0) + (a.unbias ? 0 : 1);
  const int nx = n + ne;
  const double rs = a.rangescale ? a.rangescale[b] : 1.0;
  const double ln = a.localnugget ? a.localnugget[b] : 0.0;
  const bool scaled = a.rangescale != nullptr;
  double x0[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) x0[k] = a.bcoord[k + DIM * (size_t)b];

  // ---- stage rows (coordinates relative to the block centre) and nodes -------
  if (t < no) {
    // the row's variable, by selects (a runtime index into the argument
    // arrays would copy them to scratch)
    int vr = 0, s0 = 0, mm = a.mmax[0];
    const int* inr = a.inear[0];
    const double *ocr = a.ocoord[0], *ovl = a.ovalue[0], *ovr = a.ovar[0];
#pragma unroll
    for (int v = 1; v < NV; ++v)
      if (t >= st[v]) {
        vr = v;
        s0 = st[v];
        mm = a.mmax[v];
        inr = a.inear[v];
        ocr = a.ocoord[v];
        ovl = a.ovalue[v];
        ovr = a.ovar[v];
      }
    const int idx = inr[(size_t)b * mm + (t - s0)];
#pragma unroll
    for (int k = 0; k < DIM; ++k) sx[k][t] = ocr[k + DIM * (size_t)idx] - x0[k];
    sz[t] = ovl[idx];
    sov[t] = ovr[idx];
    svr[t] = vr;
    if (nd > 0) {
      const double* odr = a.odrift[0];
#pragma unroll
      for (int v = 1; v < NV; ++v)
        if (v == vr) odr = a.odrift[v];
      for (int c = 0; c < nd; ++c) sdr[c][t] = odr[c + nd * (size_t)idx];
    }
  } else if (GRAD && t < n) {
    // increment row: both points, value, variance, drift values
    int vr = 0, g0 = gst[0];
    const double *p1 = a.gx1[0], *p2 = a.gx2[0], *pv = a.gval[0], *pe = a.gvar[0], *pd = a.gdrift[0];
#pragma unroll
    for (int v = 1; v < NV; ++v)
      if (t >= gst[v]) {
        vr = v;
        g0 = gst[v];
        p1 = a.gx1[v];
        p2 = a.gx2[v];
        pv = a.gval[v];
        pe = a.gvar[v];
        pd = a.gdrift[v];
      }
    const int gi = t - g0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      sx[k][t] = p1[k + DIM * (size_t)gi] - x0[k];
      if constexpr (GRAD) sx2[k][t - no] = p2[k + DIM * (size_t)gi] - x0[k];
    }
    sz[t] = pv[gi];
    sov[t] = pe[gi];
    svr[t] = vr;
    for (int c = 0; c < nd; ++c) sdr[c][t] = pd[c + nd * (size_t)gi];
  }
  const int ns = a.nsub;
  if (t < ns) {
    const size_t s = (size_t)b * ns + t;
#pragma unroll
    for (int k = 0; k < DIM; ++k) sg[k][t] = ns > 1 ? a.scoord[k + DIM * s] - x0[k] : 0.0;
    sgw[t] = ns > 1 ? a.sweight[s] : 1.0;
  }
  if (t == 0) sfail = 0;
  {
    // the variable pairs' models, read by every covariance evaluation
    static_assert(sizeof(Model) % 8 == 0, "Model copied as 8-byte words");
    const long long* src = reinterpret_cast<const long long*>(a.models);
    long long* dst = reinterpret_cast<long long*>(smodels);
    for (int w = t; w < (int)(NV * NV * sizeof(Model) / 8); w += MVT) dst[w] = src[w];
  }
  __syncthreads();

  // ---- right-hand sides: items (row, target, part of the nodes) ---------------
  {
    for (int it = t; it < NPART * n * nvar; it += MVT) {
      const int h = it / (n * nvar), rq = it - h * n * nvar;
      const int q = rq / n, r = rq - q * n;
      const int k0 = h * ns / NPART, k1 = (h + 1) * ns / NPART;
      const int vr = svr[r];
      double s = 0.0, s2 = 0.0;
      for (int k = k0; k < k1; ++k) {
        double dx = sx[0][r] - sg[0][k];
        double dy = DIM > 1 ? sx[1][r] - sg[1][k] : 0.0;
        double dz = DIM > 2 ? sx[2][r] - sg[2][k] : 0.0;
        if (scaled) {
          dx = dx / rs;
          dy = dy / rs;
          dz = dz / rs;
        }
        s += mv_cov<DIM, SIMPLE>(smodels, vr, q, nvar, dx, dy, dz) * sgw[k];
      }
      // increment row: sum_k w_k C(x1 - g_k) - sum_k w_k C(x2 - g_k)
      if constexpr (GRAD) {
        if (r >= no)
          s2 = mv_inc_rhs<DIM, SIMPLE>(smodels, nvar, vr, q, &sx2[0][0], r - no, &sg[0][0], sgw, k0, k1, scaled, rs);
      }
      sb[h][q][r] = r >= no ? s - s2 : s;
    }
  }
  // ---- block base covariances sum_k sum_l w_k w_l C_ij(g_k - g_l) -------------
  // Each thread takes node pairs (k, l), k <= l, for every variable pair;
  // the sums are reduced together: one barrier for all variable pairs.
  if (ns > 1) {
    const int npair = ns * (ns + 1) / 2;
    double bs[NV][NV];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < NV; ++j) bs[i][j] = 0.0;
    for (int e = t; e < npair; e += MVT) {
      // e -> (k, l), k <= l, row-major upper triangle
      int k = 0, rem = e;
      while (rem >= ns - k) {
        rem -= ns - k;
        ++k;
      }
      const int l = k + rem;
      const double dx = sg[0][k] - sg[0][l];
      const double dy = DIM > 1 ? sg[1][k] - sg[1][l] : 0.0;
      const double dz = DIM > 2 ? sg[2][k] - sg[2][l] : 0.0;
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          if (j < i) continue;
          if (l == k) bs[i][j] += smodels[i + nvar * j].cov0 * sgw[k] * sgw[k];
          else bs[i][j] += mv_cov<DIM, SIMPLE>(smodels, i, j, nvar, dx, dy, dz) * sgw[k] * sgw[l] * 2.0;
        }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if (j < i) continue;
        double v = bs[i][j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((t & 63) == 0) partb[wave][i + NV * j] = v;
      }
    __syncthreads();
    if (t < NV * NV) sbase[t] = (partb[0][t] + partb[1][t]) + (partb[2][t] + partb[3][t]);
  }
  __syncthreads();

  // ---- assemble the owned elements of the bordered matrix ------------------------
  // M = [K X'; X 0] (lower triangle), X = the extra columns as rows n..nx-1:
  // eliminating the n pivots of K leaves -X K^-1 X' (minus the Gram matrix of
  // the eliminated extra columns) in the trailing block.  One row slot at a
  // time, the values go through the thread's own LDS slots (a loop that is
  // not unrolled evaluates them), then into registers with static indices.
  double A[RB][RB];
#pragma unroll
  for (int ra = 0; ra < RB; ++ra) {
    const int i = tr + 16 * ra;
    const bool ri = i < n;
    const int vi = ri ? svr[i] : 0;
    double xi[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) xi[k] = ri ? sx[k][i] : 0.0;
#pragma unroll 1
    for (int cb = 0; cb < RB; ++cb) {
      const int j = tc + 16 * cb;
      double v = 0.0;
      if (GRAD && ri && j <= i && i >= no) {
        if constexpr (GRAD)
          v = mv_inc_cov<DIM, SIMPLE>(smodels, nvar, &sx[0][0], ROWS, &sx2[0][0], svr, sov, i, j, no, scaled, rs);
      } else if (ri && j <= i) {
        if (j == i) {
          v = smodels[vi + nvar * vi].cov0 + sov[i] + ln;
        } else {
          const int vj = svr[j];
          double dx = xi[0] - sx[0][j];
          double dy = DIM > 1 ? xi[1] - sx[1][j] : 0.0;
          double dz = DIM > 2 ? xi[2] - sx[2][j] : 0.0;
          if (scaled) {
            dx = dx / rs;
            dy = dy / rs;
            dz = dz / rs;
          }
          v = mv_cov<DIM, SIMPLE>(smodels, vi > vj ? vi : vj, vi > vj ? vj : vi, nvar, dx, dy, dz);
        }
      } else if (i >= n && i < nx && j < n) {
        // extra row e = i - n, data column j
        const int e = i - n, vj = svr[j];
        if (e < ef) {
          // pairwise sums of the partial sums, in a fixed order
          double w[NPART];
#pragma unroll
          for (int h = 0; h < NPART; ++h) w[h] = sb[h][e][j];
#pragma unroll
          for (int step = 1; step < NPART; step *= 2)
#pragma unroll
            for (int h = 0; h + step < NPART; h += 2 * step) w[h] += w[h + step];
          v = w[0];
        }
        else if (e < ez) {
          // increment rows: their drift values, 0 in the unbiasedness rows
          const int c = e - ef;
          v = c < nd ? sdr[c][j] : (j < no && (a.std_ck ? (vj == c - nd) : true) ? 1.0 : 0.0);
        }
        else if (e == ez) v = sz[j];
        else if (isa) v = (j < no && vj == e - ei) ? 1.0 : 0.0;
        else v = j < no ? 1.0 : 0.0;  // simple kriging: the ones column for the mean term (data rows)
      }
      sbuf[cb * MVT + t] = v;
    }
#pragma unroll
    for (int cb = 0; cb < RB; ++cb) A[ra][cb] = sbuf[cb * MVT + t];
  }
  if constexpr (MF) {
    // ---- FP32 blocked Cholesky with an MFMA trailing update, FP64 refinement -----
    float* const M32 = reinterpret_cast<float*>(uarea);  // M32[j * LD32 + i] = M(i, j)
    double* const L64 = uarea;                             // packed lower triangle
    const int NT = (nx + 15) >> 4;                         // 16 x 16 tiles per side
    const int lane = t & 63;
    // the full symmetric matrix in registers (the residuals need both
    // triangles): the upper entries from the lower ones, through LDS
    __syncthreads();  // every thread has read its assembly slots
#pragma unroll
    for (int ra = 0; ra < RB; ++ra)
#pragma unroll
      for (int cb = 0; cb <= ra; ++cb) {
        const int i = tr + 16 * ra, j = tc + 16 * cb;
        if (i < nx && j <= i) L64[i * (i + 1) / 2 + j] = A[ra][cb];
      }
    __syncthreads();
#pragma unroll
    for (int ra = 0; ra < RB; ++ra)
#pragma unroll
      for (int cb = ra; cb < RB; ++cb) {
        const int i = tr + 16 * ra, j = tc + 16 * cb;
        if (j > i && j < nx) A[ra][cb] = L64[j * (j + 1) / 2 + i];
      }
    __syncthreads();
    // FP32 copy, zero past nx; K's diagonal for the pivot ratios
#pragma unroll
    for (int ra = 0; ra < RB; ++ra)
#pragma unroll
      for (int cb = 0; cb < RB; ++cb) {
        const int i = tr + 16 * ra, j = tc + 16 * cb;
        if (i < 16 * NT && j < 16 * NT) M32[j * LD32 + i] = (i < nx && j < nx) ? (float)A[ra][cb] : 0.0f;
      }
    if (tr == tc) {
#pragma unroll
      for (int ra = 0; ra < RB; ++ra)
        if (tr + 16 * ra < n) sdg[tr + 16 * ra] = (float)A[ra][ra];
    }
    if (t == 0) sfail = 0;
    __syncthreads();
    const int npan = (n + 15) >> 4;
    for (int P = 0; P < npan; ++P) {
      const int c0 = 16 * P, npv = n - c0 < 16 ? n - c0 : 16;
      // diagonal tile: lane t holds row c0 + t; pivot k from lane k, the
      // scaled column broadcast by readlane, rank-1 updates of the later
      // columns (extra columns inside the tile are updated, not pivoted)
      if (t < 16) {
        float r[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) r[j] = M32[(c0 + j) * LD32 + c0 + t];
        float pm = 1.0f;
        bool bad = false;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (k >= npv) break;
          const float piv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r[k]), k));
          bad = bad || !(piv > 0.0f);
          pm = fminf(pm, piv / sdg[c0 + k]);
          float d = __builtin_amdgcn_rsqf(piv);
          d = d * fmaf(-0.5f * piv * d, d, 1.5f);
          r[k] = t == k ? piv * d : (t > k ? r[k] * d : r[k]);
#pragma unroll
          for (int j = k + 1; j < 16; ++j) {
            const float ljk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r[k]), j));
            if (t >= j) r[j] = fmaf(-r[k], ljk, r[j]);
          }
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) M32[(c0 + j) * LD32 + c0 + t] = r[j];
        if (t < npv) sli[c0 + t] = 1.0f / r[t];
        if (t == 0 && (bad || pm < (float)a.pivot_tol)) sfail = 1;
      }
      __syncthreads();
      // rows below the tile (extra rows included): solved against it, one
      // row per thread
      const int nrow = nx - (c0 + 16);
      if (t < nrow) {
        const int i = c0 + 16 + t;
        float r[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) r[j] = M32[(c0 + j) * LD32 + i];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (k >= npv) break;
          r[k] *= sli[c0 + k];
#pragma unroll
          for (int j = k + 1; j < 16; ++j) r[j] = fmaf(-r[k], M32[(c0 + k) * LD32 + c0 + j], r[j]);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) M32[(c0 + j) * LD32 + i] = r[j];
      }
      __syncthreads();
      // trailing tiles (I, J), P < J <= I < NT: C -= L_I L_J' over this
      // panel's pivots, four 16x16x4 MFMAs per tile, tiles round-robin over
      // the waves.  A(m, k) = -L(16 I + m, c0 + k) at lane m + 16 k,
      // B(k, n) = L(16 J + n, c0 + k) at lane n + 16 k, C(m, n) in register
      // m % 4 of lane n + 16 (m / 4).
      const int T0 = P + 1, ntr = NT - T0;
      const int npair = ntr > 0 ? ntr * (ntr + 1) / 2 : 0;
      const int col = lane & 15, kq = lane >> 4;
      for (int q = wave; q < npair; q += 4) {
        int J = 0, rem = q;
        while (rem >= ntr - J) {
          rem -= ntr - J;
          ++J;
        }
        const int Jt = T0 + J, It = Jt + rem;
        mf_f4* cp = reinterpret_cast<mf_f4*>(&M32[(16 * Jt + col) * LD32 + 16 * It + 4 * kq]);
        mf_f4 acc = *cp;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (4 * c >= npv) break;
          const int kc = c0 + 4 * c + kq;
          const bool on = 4 * c + kq < npv;
          const float av = on ? -M32[kc * LD32 + 16 * It + col] : 0.0f;
          const float bv = on ? M32[kc * LD32 + 16 * Jt + col] : 0.0f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
        }
        *cp = acc;
      }
      __syncthreads();
    }
    if (sfail) {
      // FP32 pivot not positive or below pivot_tol x K(k, k): FP64 re-solve
      if (t == 0) a.flag[b] = 1;
      return;
    }
    // x_e = K~^-1 X_e' for every extra column, refined twice.  First the
    // backward substitution with L' from y_e = L^-1 X_e' (row n + e of the
    // factor); then the residual r_e = X_e' - K x_e in FP64 and the FP32
    // correction K~ d_e = r_e (forward and backward substitution), x_e += d_e;
    // then the residual of the corrected x_e.  32 threads per extra column,
    // rows l32 + 32 m.
    constexpr int NM = (ROWS + 31) / 32;
    const int g = t >> 5, l32 = t & 31;
    // lane src of this thread's half wave (readlane: no LDS round trip)
    auto bcast = [&](float v, int src) {
      const int lo = __builtin_amdgcn_readlane(__float_as_int(v), src);
      const int hi = __builtin_amdgcn_readlane(__float_as_int(v), src + 32);
      return __int_as_float(lane < 32 ? lo : hi);
    };
    // y <- L'^-1 y (rows i < n)
    auto backward = [&](float (&y)[NM]) {
      for (int k = n - 1; k >= 0; --k) {
        const int mk = k >> 5;
        float yk = y[0];
#pragma unroll
        for (int m = 1; m < NM; ++m) yk = m == mk ? y[m] : yk;
        const float wk = bcast(yk, k & 31) * sli[k];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
          const int i = l32 + 32 * m;
          if (i < k) y[m] = fmaf(-M32[i * LD32 + k], wk, y[m]);
          else if (i == k) y[m] = wk;
        }
      }
    };
    for (int e = g; e < ne; e += MVT / 32) {
      float y[NM];
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const int i = l32 + 32 * m;
        y[m] = i < n ? M32[i * LD32 + n + e] : 0.0f;
      }
      backward(y);
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const int i = l32 + 32 * m;
        if (i < n) sxd[e][i] = (double)y[m];
      }
    }
    __syncthreads();
    for (int step = 0; step < 2; ++step) {
      // red[j][f] = (M v_f)_j with v_f = [x_f; -e_f]: -(X_f' - K x_f)_j for
      // j < n, X_e x_f at j = n + e.  FP64, from the registers: M is
      // symmetric, so thread (tr, tc) forms sum_a M(tr + 16 a, tc + 16 b)
      // v(tr + 16 a) for output row tc + 16 b, and the 16 lanes of one column
      // class (one wave) complete the sum
      for (int f = 0; f < ne; ++f) {
        double v[RB], pr[RB];
#pragma unroll
        for (int ra = 0; ra < RB; ++ra) {
          const int i = tr + 16 * ra;
          v[ra] = i < n ? sxd[f][i] : (i == n + f ? -1.0 : 0.0);
        }
#pragma unroll
        for (int cb = 0; cb < RB; ++cb) {
          double acc = 0.0;
#pragma unroll
          for (int ra = 0; ra < RB; ++ra) acc = fma(A[ra][cb], v[ra], acc);
#pragma unroll
          for (int o = 8; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 16);
          pr[cb] = acc;
        }
        if (tr == 0) {
#pragma unroll
          for (int cb = 0; cb < RB; ++cb)
            if (tc + 16 * cb < nx) red[tc + 16 * cb][f] = pr[cb];
        }
      }
      __syncthreads();
      if (step == 0) {
        // the correction d_e: L u = r_e, L' d_e = u, in FP32
        for (int e = g; e < ne; e += MVT / 32) {
          float y[NM];
#pragma unroll
          for (int m = 0; m < NM; ++m) {
            const int i = l32 + 32 * m;
            y[m] = i < n ? (float)-red[i][e] : 0.0f;
          }
          for (int k = 0; k < n; ++k) {
            const int mk = k >> 5;
            float yk = y[0];
#pragma unroll
            for (int m = 1; m < NM; ++m) yk = m == mk ? y[m] : yk;
            const float uk = bcast(yk, k & 31) * sli[k];
#pragma unroll
            for (int m = 0; m < NM; ++m) {
              const int i = l32 + 32 * m;
              if (i > k && i < n) y[m] = fmaf(-M32[k * LD32 + i], uk, y[m]);
              else if (i == k) y[m] = uk;
            }
          }
          backward(y);
#pragma unroll
          for (int m = 0; m < NM; ++m) {
            const int i = l32 + 32 * m;
            if (i < n) sxd[e][i] += (double)y[m];
          }
        }
      } else {
        // H_ef = X_e x_f + x_e' (X_f' - K x_f), e <= f, mirrored
        for (int w = t; w < ne * ne; w += MVT) {
          const int e = w / ne, f = w - e * ne;
          if (e > f) continue;
          double h = red[n + e][f];
          for (int i = 0; i < n; ++i) h = fma(-sxd[e][i], red[i][f], h);
          gram[e][f] = h;
          gram[f][e] = h;
        }
      }
      __syncthreads();
    }
  } else {
  // the panel buffer takes the assembly slots: cleared once every thread has
  // read its own (rows past nx keep zero factors throughout)
  __syncthreads();
  for (int w = t; w < 2 * ROWS * PLD; w += MVT) sbuf[w] = 0.0;
  __syncthreads();

  // ---- Cholesky of the n pivots, four columns per barrier ------------------------
  // Columns c0..c0+3 (c0 a multiple of 4) are column classes c0 % 16 .. +3,
  // which all sit in wave (c0 % 16) / 4.  That wave factors the panel alone:
  // for each column it reads the pivot from its owner lane, scales the column
  // by 1 / sqrt(pivot) (as LAPACK's unblocked Cholesky does, DSCAL by
  // ONE / AJJ) into the panel buffer, and updates the panel's later columns
  // from it, exchanging through LDS within the wave.  It then clears the
  // panel's own rows in the buffer, and after one barrier every thread
  // applies the rank-4 update A(i, j) -= sum_k L(i, c0+k) L(j, c0+k) to the
  // trailing entries, in the order of the unblocked algorithm.  The buffer is
  // double-buffered, so the next panel's writer cannot overtake a reader.
  // A pivot that is not positive sets the failure flag, which is read once
  // after the loop: the steps after it compute values that are discarded.
#pragma unroll
  for (int cb = 0; cb < RB; ++cb) {
    for (int pp = 0; pp < 4; ++pp) {
      const int c0 = 16 * cb + 4 * pp;
      if (c0 >= n) break;
      const int np = n - c0 < 4 ? n - c0 : 4;  // pivots in this panel
      const int last = c0 + np - 1;
      double(*vp)[PLD] = vpan[pp & 1];
      if (wave == pp) {
        const int kl = (t & 63) >> 4;  // this lane's panel column: tc = 4 pp + kl
        for (int k = 0; k < np; ++k) {
          const int c = c0 + k;
          const double piv = readlane_d(A[cb][cb], 16 * k + (c & 15));
          if (!(piv > 0.0) && (t & 63) == 0) sfail = 1;
          const double rd = rsqrt_nr(piv);
          if (kl == k) {
#pragma unroll
            for (int ra = 0; ra < RB; ++ra) {
              const int i = tr + 16 * ra;
              if (16 * ra < nx && i < nx) vp[i][k] = i > c ? A[ra][cb] * rd : 0.0;
            }
          }
          wave_lds_sync();
          // the panel's later pivot columns; columns past the last pivot
          // (extra columns) take the whole panel in the trailing update
          if (kl > k && kl < np) {
            const double f = vp[c0 + kl][k];  // L(c0 + kl, c)
#pragma unroll
            for (int ra = 0; ra < RB; ++ra) {
              if (16 * ra + 15 <= c || 16 * ra >= nx) continue;
              A[ra][cb] = fma(-vp[tr + 16 * ra][k], f, A[ra][cb]);
            }
          }
        }
        wave_lds_sync();
        // columns past the last pivot: zero factors
        if (kl >= np) {
#pragma unroll
          for (int ra = 0; ra < RB; ++ra) {
            const int i = tr + 16 * ra;
            if (16 * ra < nx && i < nx) vp[i][kl] = 0.0;
          }
        }
        // the panel's pivot rows take no part in the trailing update
        if ((t & 63) < 16 && (t & 3) < np) vp[c0 + (t & 3)][(t & 15) >> 2] = 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int ra = 0; ra < RB; ++ra) {
        if (16 * ra + 15 <= last || 16 * ra >= nx) continue;
        const double* li = vp[tr + 16 * ra];
        const double2 l01 = *reinterpret_cast<const double2*>(li);
        const double2 l23 = *reinterpret_cast<const double2*>(li + 2);
#pragma unroll
        for (int kk = 0; kk <= ra; ++kk) {
          if (16 * kk + 15 <= last) continue;
          const double* lj = vp[tc + 16 * kk];
          const double2 m01 = *reinterpret_cast<const double2*>(lj);
          const double2 m23 = *reinterpret_cast<const double2*>(lj + 2);
          double v = A[ra][kk];
          v = fma(-l01.x, m01.x, v);
          v = fma(-l01.y, m01.y, v);
          v = fma(-l23.x, m23.x, v);
          v = fma(-l23.y, m23.y, v);
          A[ra][kk] = v;
        }
      }
    }
  }
  __syncthreads();
  if (sfail) {
    // K not positive definite: the Bunch-Kaufman fallback solves the block
    if (t == 0) {
      a.flag[b] = 2;
      if (a.ctr) atomicAdd(a.ctr, 1);
    }
    return;
  }

  if (WEIGHTS) {
    // the factor's columns, the extra rows included, for the substitution
#pragma unroll
    for (int ra = 0; ra < RB; ++ra) {
      const int i = tr + 16 * ra;
#pragma unroll
      for (int k = 0; k <= ra; ++k) {
        const int j = tc + 16 * k;
        if (i < nx && j < n && j <= i) sl[i * (i + 1) / 2 + j] = A[ra][k];
      }
    }
  }

  // ---- Gram matrix of the eliminated extra columns: minus the trailing block ----
#pragma unroll
  for (int ra = 0; ra < RB; ++ra) {
    const int i = tr + 16 * ra;
#pragma unroll
    for (int k = 0; k <= ra; ++k) {
      const int j = tc + 16 * k;
      if (i >= n && i < nx && j >= n && j <= i) gram[i - n][j - n] = gram[j - n][i - n] = -A[ra][k];
    }
  }
  }
  // group means of the neighbour values (Isaaks-Srivastava)
  double gm[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    gm[v] = 0.0;
    if (isa && nv[v] > 0) gm[v] = wg_sum((t < no && svr[t] == v) ? sz[t] : 0.0, part) / nv[v];
  }
  __syncthreads();
  if (!WEIGHTS && t != 0) return;

  // ---- multipliers, estimates and variances (one thread) -----------------------
  // mu_q = S^-1 (G_q - d_q), S = F'K^-1 F, G = F'K^-1 b, through the Cholesky
  // factor S = U'U as the reference's kriging_setup / kriging_solve_prepared
  // (solver.f90:224-300); a pivot that is not positive (a variable without
  // neighbours leaves a zero row in S under std_ck) sends the block to the
  // Bunch-Kaufman fallback like the reference's failed SPOTRF of S
  if (t == 0) {
    constexpr int PM = NV + ND;  // constraint rows at most
    double mu[NV][PM];  // mu[q][c]
    bool ok = true;
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
      for (int c = 0; c < PM; ++c) mu[q][c] = 0.0;
    if (p > 0) {
      double U[PM][PM];
#pragma unroll
      for (int i = 0; i < PM; ++i)
#pragma unroll
        for (int j = 0; j < PM; ++j) U[i][j] = (i < p && j < p) ? gram[ef + i][ef + j] : 0.0;
      bool spd = true;
#pragma unroll
      for (int j = 0; j < PM; ++j) {
        if (j >= p) break;
        double d = U[j][j];
#pragma unroll
        for (int k = 0; k < PM; ++k)
          if (k < j) d -= U[k][j] * U[k][j];
        spd = spd && d > 0.0;
        d = sqrt(d);
        U[j][j] = d;
#pragma unroll
        for (int i = 0; i < PM; ++i) {
          if (i <= j || i >= p) continue;
          double v = U[j][i];
#pragma unroll
          for (int k = 0; k < PM; ++k)
            if (k < j) v -= U[k][j] * U[k][i];
          U[j][i] = v / d;
        }
      }
      if (!spd) {
        a.flag[b] = 2;
        if (a.ctr) atomicAdd(a.ctr, 1);
        ok = false;
      }
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        if (!ok) break;
        double y[PM];
#pragma unroll
        for (int i = 0; i < PM; ++i) {
          if (i >= p) continue;
          double v = gram[ef + i][eb + q] - dval(i, q);
#pragma unroll
          for (int k = 0; k < PM; ++k)
            if (k < i) v -= U[k][i] * y[k];
          y[i] = v / U[i][i];
        }
#pragma unroll
        for (int i = PM - 1; i >= 0; --i) {
          if (i >= p) continue;
          double v = y[i];
#pragma unroll
          for (int k = 0; k < PM; ++k)
            if (k > i && k < p) v -= U[i][k] * mu[q][k];
          mu[q][i] = v / U[i][i];
        }
      }
    }
    if (ok && !a.weights_only) {
      // c' w_q = c'K^-1 b_q - sum_f mu_q[f] c'K^-1 F_f for an extra column c
      auto cw = [&](int e, int q) {
        double s = gram[e][eb + q];
#pragma unroll
        for (int f = 0; f < PM; ++f)
          if (f < p) s -= mu[q][f] * gram[e][ef + f];
        return s;
      };
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        double e = cw(ez, q);
        if (isa) {
#pragma unroll
          for (int v = 0; v < NV; ++v)
            if (v != q && nv[v] > 0) e += (gm[q] - gm[v]) * cw(ei + v, q);
        }
        if (!a.unbias && a.sk_mean[q] != 0.0) e += (1.0 - cw(ez + 1, q)) * a.sk_mean[q];
        e = e < a.bound_lo ? a.bound_lo : e;
        e = e > a.bound_hi ? a.bound_hi : e;
        a.est[b + (size_t)a.nb * q] = e;
      }
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          if (j < i) continue;
          const double base = ns > 1 ? sbase[i + NV * j] : a.models[i + nvar * j].cov0;
          double mud = 0.0;
#pragma unroll
          for (int c = 0; c < PM; ++c)
            if (c < p) mud += mu[i][c] * dval(c, j);
          double v = base - (cw(eb + j, i) + mud);
          v = v > 0.0 ? v : 0.0;  // diagonal and off-diagonal clamped at 0
          a.var[b + (size_t)a.nb * (i + nvar * j)] = v;
          a.var[b + (size_t)a.nb * (j + nvar * i)] = v;
        }
    }
    if (ok) a.flag[b] = 0;
    if (WEIGHTS) {
      sfail = ok ? 0 : 1;
#pragma unroll
      for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int c = 0; c < PM; ++c) smu[q][c] = mu[q][c];
    }
  }
  if (!WEIGHTS) return;
  __syncthreads();
  if (sfail) return;

  // ---- weights: L' w_q = D^-1 (y_bq - sum_f mu_q[f] y_Ff), backward ------------
  // Unit factor L(i, j) = A(i, j) / A(j, j), A(j, j) = d_j; an extra row
  // holds A(n + e, j) = d_j (L^-1 x_e)_j.  Thread j < n owns row j; after
  // w_i is final it is broadcast and every row j < i takes its share.
  double r[NV];
  double dj = 1.0;
#pragma unroll
  for (int q = 0; q < NV; ++q) r[q] = 0.0;
  if (t < n) {
    dj = sl[t * (t + 1) / 2 + t];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int rb_ = n + eb + q;
      double v = sl[rb_ * (rb_ + 1) / 2 + t];
#pragma unroll
      for (int f = 0; f < NV + ND; ++f) {
        if (f >= p) break;
        const int rf = n + ef + f;
        v -= smu[q][f] * sl[rf * (rf + 1) / 2 + t];
      }
      r[q] = v / dj;
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    if (t == i) {
#pragma unroll
      for (int q = 0; q < NV; ++q) sw[q][i] = r[q];
    }
    __syncthreads();
    if (t < i) {
      const double l = sl[i * (i + 1) / 2 + t] / dj;
#pragma unroll
      for (int q = 0; q < NV; ++q) r[q] -= l * sw[q][i];
    }
  }
  if (NV == 1 && a.wcorr) {
    // weight correction (nvar 1, kriging_base.F90:3388-3391,
    // kriging_clip_positive_normalize at :1995-2001): negative weights to 0,
    // the rest divided by their sum when it is positive; then the estimate
    // (estimate_block_base) and the variance base - (w'b + mu'd)
    // (calc_variance, kriging.F90:968-1028) with the corrected weights
    const double wp = (t < n && r[0] > 0.0) ? r[0] : 0.0;
    const double tot = wg_sum(wp, part);
    if (t < n) r[0] = tot > 0.0 ? wp / tot : wp;
    double bj = 0.0;
    if (t < n) {
      // the row's right-hand side, its partial sums added as in the assembly
      double w2[NPART];
#pragma unroll
      for (int h = 0; h < NPART; ++h) w2[h] = sb[h][0][t];
#pragma unroll
      for (int step = 1; step < NPART; step *= 2)
#pragma unroll
        for (int h = 0; h + step < NPART; h += 2 * step) w2[h] += w2[h + step];
      bj = w2[0];
    }
    const double est = wg_sum(t < n ? r[0] * sz[t] : 0.0, part);
    const double wb = wg_sum(t < n ? r[0] * bj : 0.0, part);
    const double tw = wg_sum(t < n ? r[0] : 0.0, part);
    if (t == 0 && !a.weights_only) {
      double mud = 0.0;
      for (int c = 0; c < p; ++c) mud += smu[0][c] * dval(c, 0);
      const double base = ns > 1 ? sbase[0] : a.models[0].cov0;
      double v = base - (wb + mud);
      v = v > 0.0 ? v : 0.0;
      double e = est;
      if (!a.unbias && a.sk_mean[0] != 0.0) e += (1.0 - tw) * a.sk_mean[0];
      e = e < a.bound_lo ? a.bound_lo : e;
      e = e > a.bound_hi ? a.bound_hi : e;
      a.est[b] = e;
      a.var[b] = v;
    }
  }
  if (t < no) {
    // weight(k, group v, target q) = x(q, istart(v) + k) (data rows only)
    const int v = svr[t];
    int s0 = 0;
#pragma unroll
    for (int u = 1; u < NV; ++u)
      if (u == v) s0 = st[u];
#pragma unroll
    for (int q = 0; q < NV; ++q)
      a.weights[(t - s0) + (size_t)a.mmax_w * (v + nvar * (q + (size_t)nvar * b))] = r[q];
  }
}

#if KK_MV_DIM == 0
// Bunch-Kaufman fallback for the blocks k_solve_mv flags (K or S not positive
// definite), one thread per block, matrices in HBM scratch: the reference's
// SSYTRF/SSYTRS on the full augmented system with one jittered retry, then
// NaN (kriging_base.F90:3299-3383).  The system is assembled in the
// reference's layout -- data rows by variable group, then the constraint
// rows -- with the right-hand sides of every target variable
// (kriging.F90:714-939); the variance and estimate follow calc_variance
// (kriging.F90:968-1028) and estimate_block_base (kriging_base.F90:3407-3503)
// in their summation order.
template <int DIM>
__global__ void k_fallback_mv(MVArgs a, const int* list, int nlist, double* scratch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int b = list[t];
  const int nvar = a.nvar;
  const int ld = a.rows + MAXP;
  int nsum = 0, ngt = 0;
#pragma unroll
  for (int v = 0; v < MAXVAR; ++v) {
    nsum += v < nvar ? a.nnear[v][b] : 0;
    ngt += v < nvar ? a.ngrad[v] : 0;
  }
  // nothing to solve (the weights-only pass lists every block)
  if (nsum == 0 || nsum + ngt < 2) return;
  double* K = scratch + (size_t)t * mv_fallback_words(a.rows);
  double* A = K + (size_t)ld * ld;
  double* B = A + (size_t)ld * ld;       // right-hand sides, then solutions
  double* B0 = B + (size_t)ld * MAXVAR;  // right-hand sides kept for the variance
  int* ipiv = reinterpret_cast<int*>(B0 + (size_t)ld * MAXVAR);
  int nv[MAXVAR], st[MAXVAR], n = 0;
  for (int v = 0; v < MAXVAR; ++v) {
    nv[v] = v < nvar ? a.nnear[v][b] : 0;
    st[v] = n;
    n += nv[v];
  }
  // increment rows after the data rows, variable by variable
  const int no = n;
  int gst[MAXVAR];
  for (int v = 0; v < MAXVAR; ++v) {
    gst[v] = n;
    n += v < nvar ? a.ngrad[v] : 0;
  }
  const int nd = a.ndrift;
  const int p = nd + (a.unbias ? (a.std_ck ? nvar : 1) : 0), m = n + p;
  const double rs = a.rangescale ? a.rangescale[b] : 1.0;
  const double ln = a.localnugget ? a.localnugget[b] : 0.0;
  const int ns = a.nsub;
  const double* x0 = a.bcoord + DIM * (size_t)b;
  // constraint row f of data row c (drift values, then unbiasedness), and its
  // right-hand side for target q
  auto frow = [&](int f, int vc, int ic) -> double {
    if (f < nd) return a.odrift[vc][f + nd * (size_t)ic];
    return a.std_ck ? (vc == f - nd ? 1.0 : 0.0) : 1.0;
  };
  auto dval = [&](int f, int q) -> double {
    if (f < nd) return a.bdrift[f + nd * ((size_t)q + nvar * (size_t)b)];
    return a.std_ck ? (f - nd == q ? 1.0 : 0.0) : 1.0;
  };
  // row r: variable (data and increment rows), observation index
  auto rowvar = [&](int r) {
    int v = 0;
    for (int u = 1; u < nvar; ++u)
      if (r >= (r < no ? st[u] : gst[u])) v = u;
    return v;
  };
  auto cv = [&](int va, int vb, const double* u, const double* w) {
    const double dx = (u[0] - w[0]) / rs;
    const double dy = DIM > 1 ? (u[1] - w[1]) / rs : 0.0;
    const double dz = DIM > 2 ? (u[2] - w[2]) / rs : 0.0;
    return mv_cov<DIM, false>(a.models, va, vb, nvar, dx, dy, dz);
  };
  auto rowx = [&](int r, int v) { return a.ocoord[v] + DIM * (size_t)a.inear[v][(size_t)b * a.mmax[v] + (r - st[v])]; };
  for (int c = no; c < n; ++c) {
    // increment row c (kriging.F90:796-840, 904-920)
    const int vc = rowvar(c), gc = c - gst[vc];
    const double* x1c = a.gx1[vc] + DIM * (size_t)gc;
    const double* x2c = a.gx2[vc] + DIM * (size_t)gc;
    K[c + ld * c] = (2.0 * a.models[vc + nvar * vc].cov0 - 2.0 * cv(vc, vc, x1c, x2c)) + a.gvar[vc][gc];
    for (int r = 0; r < c; ++r) {
      const int vr = rowvar(r);
      double v;
      if (r < no) {
        const double* xr = rowx(r, vr);
        v = cv(vc, vr, x1c, xr) - cv(vc, vr, x2c, xr);
      } else if (vr != vc) {
        v = 0.0;
      } else {
        // the pair (r, c), r first
        const double* x1r = a.gx1[vr] + DIM * (size_t)(r - gst[vr]);
        const double* x2r = a.gx2[vr] + DIM * (size_t)(r - gst[vr]);
        v = ((cv(vc, vc, x1r, x1c) - cv(vc, vc, x1r, x2c)) - cv(vc, vc, x2r, x1c)) + cv(vc, vc, x2r, x2c);
      }
      K[r + ld * c] = v;
      K[c + ld * r] = v;
    }
    for (int f = 0; f < p; ++f) {
      const double e = f < nd ? a.gdrift[vc][f + nd * (size_t)gc] : 0.0;
      K[(n + f) + ld * c] = e;
      K[c + ld * (n + f)] = e;
    }
    for (int q = 0; q < nvar; ++q) {
      double tmp = 0.0, tmp2 = 0.0;
      for (int k = 0; k < ns; ++k) {
        const double* g = ns > 1 ? a.scoord + DIM * ((size_t)b * ns + k) : x0;
        const double w = ns > 1 ? a.sweight[(size_t)b * ns + k] : 1.0;
        tmp = tmp + cv(vc, q, x1c, g) * w;
        tmp2 = tmp2 + cv(vc, q, x2c, g) * w;
      }
      B[c + ld * q] = tmp - tmp2;
    }
  }
  for (int c = 0; c < no; ++c) {
    const int vc = rowvar(c);
    const double* xc = rowx(c, vc);
    const int ic = a.inear[vc][(size_t)b * a.mmax[vc] + (c - st[vc])];
    K[c + ld * c] = (a.models[vc + nvar * vc].cov0 + a.ovar[vc][ic]) + ln;
    for (int r = 0; r < c; ++r) {
      const int vr = rowvar(r);
      const double* xr = rowx(r, vr);
      const double dx = (xr[0] - xc[0]) / rs;
      const double dy = DIM > 1 ? (xr[1] - xc[1]) / rs : 0.0;
      const double dz = DIM > 2 ? (xr[2] - xc[2]) / rs : 0.0;
      const double v = mv_cov<DIM, false>(a.models, vc, vr, nvar, dx, dy, dz);  // vgm(max, min)
      K[r + ld * c] = v;
      K[c + ld * r] = v;
    }
    for (int f = 0; f < p; ++f) {
      const double e = frow(f, vc, ic);
      K[(n + f) + ld * c] = e;
      K[c + ld * (n + f)] = e;
    }
    // right-hand sides: sum over the integration nodes of w_k C_{vgm(v_c, q)}
    for (int q = 0; q < nvar; ++q) {
      double tmp = 0.0;
      for (int k = 0; k < ns; ++k) {
        const double* g = ns > 1 ? a.scoord + DIM * ((size_t)b * ns + k) : x0;
        const double w = ns > 1 ? a.sweight[(size_t)b * ns + k] : 1.0;
        const double dx = (xc[0] - g[0]) / rs;
        const double dy = DIM > 1 ? (xc[1] - g[1]) / rs : 0.0;
        const double dz = DIM > 2 ? (xc[2] - g[2]) / rs : 0.0;
        tmp = tmp + mv_cov<DIM, false>(a.models, vc, q, nvar, dx, dy, dz) * w;
      }
      B[c + ld * q] = tmp;
    }
  }
  for (int f = 0; f < p; ++f) {
    for (int g = 0; g < p; ++g) K[(n + f) + ld * (n + g)] = 0.0;
    for (int q = 0; q < nvar; ++q) B[(n + f) + ld * q] = dval(f, q);
  }
  for (int q = 0; q < nvar; ++q)
    for (int r = 0; r < m; ++r) B0[r + ld * q] = B[r + ld * q];
  const int info = bk_fallback(K, A, ipiv, n, m, ld, B, nvar);
  if (!info && a.wcorr && nvar == 1) {
    // weight correction (kriging_base.F90:3388-3391, 1995-2001)
    double tot = 0.0;
    for (int k = 0; k < n; ++k) {
      B[k] = B[k] > 0.0 ? B[k] : 0.0;
      tot = tot + B[k];
    }
    if (tot > 0.0)
      for (int k = 0; k < n; ++k) B[k] = B[k] / tot;
  }
  if (a.weights) {
    // the weight store: weight(k, group v, target q) = x(q, istart(v) + k)
    for (int q = 0; q < nvar; ++q)
      for (int v = 0; v < nvar; ++v)
        for (int k = 0; k < nv[v]; ++k)
          a.weights[k + (size_t)a.mmax_w * (v + nvar * (q + (size_t)nvar * b))] =
              info ? __builtin_nan("") : B[(st[v] + k) + ld * q];
    if (a.weights_only) return;
  }
  if (info) {
    // x = NaN: NaN estimates, variances max(NaN, 0) = 0 (kriging.F90:1019-1025)
    for (int q = 0; q < nvar; ++q) a.est[b + (size_t)a.nb * q] = __builtin_nan("");
    for (int q = 0; q < nvar * nvar; ++q) a.var[b + (size_t)a.nb * q] = 0.0;
    a.flag[b] = 3;
    return;
  }
  // variances: base_ij - x_i . rhs_j over all m rows, i <= j, clamped at 0
  for (int i = 0; i < nvar; ++i)
    for (int j = i; j < nvar; ++j) {
      const Model& Mij = a.models[i + nvar * j];
      double base = Mij.cov0;
      if (ns > 1) {
        base = 0.0;
        for (int k = 0; k < ns; ++k) {
          const double* gk = a.scoord + DIM * ((size_t)b * ns + k);
          const double wk = a.sweight[(size_t)b * ns + k];
          base = base + Mij.cov0 * wk * wk;
          for (int l = k + 1; l < ns; ++l) {
            const double* gl = a.scoord + DIM * ((size_t)b * ns + l);
            const double dx = gk[0] - gl[0];
            const double dy = DIM > 1 ? gk[1] - gl[1] : 0.0;
            const double dz = DIM > 2 ? gk[2] - gl[2] : 0.0;
            base = base + mv_cov<DIM, false>(a.models, i, j, nvar, dx, dy, dz) * wk * a.sweight[(size_t)b * ns + l] * 2.0;
          }
        }
      }
      double dot = 0.0;
      for (int r = 0; r < m; ++r) dot = dot + B[r + ld * i] * B0[r + ld * j];
      double v = base - dot;
      v = v > 0.0 ? v : 0.0;
      a.var[b + (size_t)a.nb * (i + nvar * j)] = v;
      a.var[b + (size_t)a.nb * (j + nvar * i)] = v;
    }
  // estimates: per target q, the groups in order (Isaaks-Srivastava shift for
  // non-standard cokriging, simple-kriging mean term), clipped to the bounds
  const bool isa = !a.std_ck && a.unbias && nvar > 1;
  double tm[MAXVAR];
  for (int v = 0; v < nvar; ++v) {
    tm[v] = 0.0;
    if (!isa) continue;
    for (int k = 0; k < nv[v]; ++k) tm[v] = tm[v] + a.ovalue[v][a.inear[v][(size_t)b * a.mmax[v] + k]];
    if (nv[v] > 0) tm[v] = tm[v] / nv[v];
  }
  for (int q = 0; q < nvar; ++q) {
    double avg = 0.0, twsum = 0.0;
    for (int v = 0; v < nvar; ++v) {
      double tw = 0.0, dp = 0.0;
      for (int k = 0; k < nv[v]; ++k) {
        const double w = B[(st[v] + k) + ld * q];
        tw = tw + w;
        dp = dp + a.ovalue[v][a.inear[v][(size_t)b * a.mmax[v] + k]] * w;
      }
      avg = avg + dp;
      if (isa && v != q && nv[v] > 0) avg = avg + (tm[q] - tm[v]) * tw;
      twsum = twsum + tw;
    }
    // increments: their weights times their values (kriging_base.F90:3457-3466)
    for (int v = 0; v < nvar; ++v)
      for (int k = 0; k < a.ngrad[v]; ++k) avg = avg + B[(gst[v] + k) + ld * q] * a.gval[v][k];
    if (!a.unbias && a.sk_mean[q] != 0.0) avg = avg + (1.0 - twsum) * a.sk_mean[q];
    avg = avg < a.bound_lo ? a.bound_lo : avg;
    avg = avg > a.bound_hi ? a.bound_hi : avg;
    a.est[b + (size_t)a.nb * q] = avg;
  }
  a.flag[b] = 0;
}

hipError_t launch_fallback_mv(const MVArgs& a, const int* list, int nlist, double* scratch, hipStream_t st) {
  if (nlist <= 0) return hipSuccess;
  dim3 g((nlist + 63) / 64), blk(64);
  if (a.ndim == 1) hipLaunchKernelGGL(k_fallback_mv<1>, g, blk, 0, st, a, list, nlist, scratch);
  else if (a.ndim == 2) hipLaunchKernelGGL(k_fallback_mv<2>, g, blk, 0, st, a, list, nlist, scratch);
  else hipLaunchKernelGGL(k_fallback_mv<3>, g, blk, 0, st, a, list, nlist, scratch);
  return hipGetLastError();
}

int mv_extra_cols(int nvar, int std_ck, int unbias, int ndrift) {
  const int p = ndrift + (unbias ? (std_ck ? nvar : 1) : 0);
  const bool isa = !std_ck && unbias && nvar > 1;
  return nvar + p + 1 + (isa ? nvar : 0) + (unbias ? 0 : 1);
}

// FP32 + MFMA + refinement runs for systems up to 80 columns (weights,
// external drift and increment rows: FP64)
bool mv_fp32r_instance(const MVArgs& a) {
  const int cols = a.rows + mv_extra_cols(a.nvar, a.std_ck, a.unbias, a.ndrift);
  bool grad = false;
  for (int v = 0; v < a.nvar; ++v) grad = grad || a.ngrad[v] > 0;
  return a.prec == PREC_FP32R && !a.weights && !grad && a.ndrift == 0 && cols <= 80;
}

hipError_t launch_solve_mv(const MVArgs& a, hipStream_t st) {
  if (a.nb <= 0 || (a.list && a.nlist <= 0)) return hipSuccess;
  const int cols = a.rows + mv_extra_cols(a.nvar, a.std_ck, a.unbias, a.ndrift);
  if (a.ndrift < 0 || a.ndrift > MAXDRIFT || (a.wcorr && (a.nvar != 1 || !a.weights))) return hipErrorInvalidValue;
  if (a.rows > MVR || cols > MVR || a.nsub > MAXSUB || a.nvar < 1 || a.nvar > MAXVAR) return hipErrorInvalidValue;
  const int rb = cols <= 80 ? 5 : 8;
  // nvar 1: single-variable systems above the row-per-lane solver's 64 rows
  if (a.weights) {
    if (a.ndim == 1) return launch_solve_mv_d1_w1(a, rb, st);
    if (a.ndim == 2) return launch_solve_mv_d2_w1(a, rb, st);
    return launch_solve_mv_d3_w1(a, rb, st);
  }
  if (a.ndim == 1) return launch_solve_mv_d1_w0(a, rb, st);
  if (a.ndim == 2) return launch_solve_mv_d2_w0(a, rb, st);
  return launch_solve_mv_d3_w0(a, rb, st);
}

#endif  // KK_MV_DIM == 0

#if KK_MV_DIM > 0
template <int NV>
static void launch_nv(const MVArgs& a, int rb, hipStream_t st) {
  constexpr int D = KK_MV_DIM;
  dim3 g(a.list ? a.nlist : a.nb), blk(MVT);
  const bool drift = a.ndrift > 0;
#if KK_MV_W
  // the weight store: general covariance evaluation only
  if (drift) {
    if (rb == 5) hipLaunchKernelGGL((k_solve_mv<D, NV, 5, false, true, false, false, true>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_solve_mv<D, NV, 8, false, true, false, false, true>), g, blk, 0, st, a);
  } else {
    if (rb == 5) hipLaunchKernelGGL((k_solve_mv<D, NV, 5, false, true>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_solve_mv<D, NV, 8, false, true>), g, blk, 0, st, a);
  }
#else
  bool grad = false;
  for (int v = 0; v < a.nvar; ++v) grad = grad || a.ngrad[v] > 0;
  if (grad) {
    // increment rows: the FP64 general instance that holds them (either
    // precision; their systems are solved in FP64)
    if (rb == 5) hipLaunchKernelGGL((k_solve_mv<D, NV, 5, false, false, false, true>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_solve_mv<D, NV, 8, false, false, false, true>), g, blk, 0, st, a);
  } else if (drift) {
    // external drift rows: FP64 instances (either precision)
    if (rb == 5) {
      if (a.simple) hipLaunchKernelGGL((k_solve_mv<D, NV, 5, true, false, false, false, true>), g, blk, 0, st, a);
      else hipLaunchKernelGGL((k_solve_mv<D, NV, 5, false, false, false, false, true>), g, blk, 0, st, a);
    } else {
      if (a.simple) hipLaunchKernelGGL((k_solve_mv<D, NV, 8, true, false, false, false, true>), g, blk, 0, st, a);
      else hipLaunchKernelGGL((k_solve_mv<D, NV, 8, false, false, false, false, true>), g, blk, 0, st, a);
    }
  } else if (rb == 5 && a.prec == PREC_FP32R) {
    if (a.simple) hipLaunchKernelGGL((k_solve_mv<D, NV, 5, true, false, true>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_solve_mv<D, NV, 5, false, false, true>), g, blk, 0, st, a);
  } else if (rb == 5) {
    if (a.simple) hipLaunchKernelGGL((k_solve_mv<D, NV, 5, true, false>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_solve_mv<D, NV, 5, false, false>), g, blk, 0, st, a);
  } else {
    if (a.simple) hipLaunchKernelGGL((k_solve_mv<D, NV, 8, true, false>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_solve_mv<D, NV, 8, false, false>), g, blk, 0, st, a);
  }
#endif
}

#define KK_CAT2(x, y) x##y
#define KK_CAT(x, y) KK_CAT2(x, y)
hipError_t KK_CAT(KK_CAT(launch_solve_mv_d, KK_MV_DIM), KK_CAT(_w, KK_MV_W))(const MVArgs& a, int rb,
                                                                            hipStream_t st) {
  if (a.nvar == 1) launch_nv<1>(a, rb, st);
  else if (a.nvar == 2) launch_nv<2>(a, rb, st);
  else if (a.nvar == 3) launch_nv<3>(a, rb, st);
  else launch_nv<4>(a, rb, st);
  return hipGetLastError();
}
#endif  // KK_MV_DIM > 0

}  // namespace kk
