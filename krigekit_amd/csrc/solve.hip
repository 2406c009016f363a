// Per-node kriging solve on the GPU.
//
// One system per group of G lanes (G = 8/16/32/64) with R <= G matrix rows
// (by neighbour-count bucket: R = 8, 12, 16, 24, 32, 48 or 64), one matrix
// row per lane, held in registers:
//   1. assemble K (kriging.F90:714-849) and b (kriging.F90:864-939): each
//      lane evaluates R/2 pairs of a balanced "fold" (r, r+d mod R) and the
//      pair values are exchanged through LDS, so every covariance is
//      evaluated once;
//   2. factor K = L D L^T right-looking; column k is broadcast through LDS;
//   3. forward-substitute b, F (unbias rows) and the data values z;
//   4. the ordinary-kriging Schur step (solver.f90:224-300), the variance
//      (kriging.F90:968-1028) and the estimate (kriging_base.F90:3407-3503)
//      come out of five quadratic forms b'K^-1b, F'K^-1F, F'K^-1b, z'K^-1b,
//      z'K^-1F, so no backward substitution is needed for the outputs.
//
// PREC_FP32R factors a float copy of K and refines the two solves K^-1 b and
// K^-1 F once against the FP64 matrix (residual in FP64).  A system whose
// FP32 pivot ratio D_k / K_kk falls below pivot_tol is flagged and solved
// again by the FP64 kernel.  A system whose FP64 pivot is not positive is
// flagged for the Bunch-Kaufman fallback (the reference's SSYTRF fallback,
// kriging_base.F90:3299-3366).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kk_bk.hpp"
#include "kk_cov.hpp"
#include "kk_device.hpp"
#include "kk_types.hpp"
#include "kk_wave.hpp"

#ifndef KK_SOLVE_DIM
#define KK_SOLVE_DIM 0
#endif

namespace kk {

// The sum-metric joint term added to base = Cs + Ct (variogram_st.f90:178-194
// analytic, 241-261 tables): per spatial structure k, sill_st[k] times its
// shape at h_st = sqrt(h_k^2 + f(dt)^2), h_k the structure's anisotropic lag
// and f the temporal transform (f_time_vgm_st, :126-141).  With per-structure
// tables (Level A) and a positive sill the shape is the structure's table
// value divided by its sill, otherwise the analytic shape.
__device__ __noinline__ double st_sum_cov(const SolveArgs* ap, const Model* Mp, double base, double dx, double dy,
                                          double dz, double dt) {
  const SolveArgs& a = *ap;
  const Model& M = *Mp;
  const double adt = fabs(dt);
  double dw = 0.0;
  if (!(adt <= 2.220446049250313e-16)) {
    dw = a.t_nug + a.t_sill * (1.0 - corefunc(a.t_vtype, adt / a.t_at));
    dw = dw > 0.0 ? dw : 0.0;
  }
  double res = base;
  for (int k = 0; k < M.nstruct; ++k) {
    const double hs = aniso_h<3>(M.mat[k], dx, dy, dz);
    const double hst = sqrt(hs * hs + dw * dw);
    if (M.mode == COV_TABLE_A && M.sill[k] > 0.0)
      res = res + a.sill_st[k] * comp_tab_h(M, k, hst) / M.sill[k];
    else
      res = res + a.sill_st[k] * corefunc(M.vtype[k], hst);
  }
  return res;
}

// Covariance of one pair of rows.  DIM 4 is space-time (x, y, z, t): the
// product-sum k Cs(dx, dy, dz) Ct(|dt|) + Cs + Ct of the spatial model and
// the 1-D temporal model (variogram_st.f90:222-282), or the sum-metric.
// With TB both models are single Level-B tables and both lookups are inline
// (product-sum only: the host runs sum-metric models without TB).
template <int DIM, bool TB>
__device__ inline double pair_cov(const SolveArgs& a, const Model& M, double dx, double dy, double dz,
                                  double dt) {
  if (DIM <= 3) return cov_eval<DIM, TB>(M, dx, dy, dz);
  const double cs = cov_eval<3, TB>(M, dx, dy, dz);
  const double ct = cov_eval<1, TB>(*a.tmodel, fabs(dt), 0.0, 0.0);
  if (!TB && a.st_sum) return st_sum_cov(&a, &M, cs + ct, dx, dy, dz, dt);
  return a.k_ps * cs * ct + cs + ct;
}

// pair_cov of the TB kernels through the split Level-B lookup: the spatial
// and temporal entry loads are issued together.  `out` reports a lag beyond a
// table with an analytic tail; the caller then evaluates the pair with
// pair_cov, which includes the tail.
template <int DIM>
__device__ inline double pair_cov_tb(const SolveArgs& a, const TabB& Ts, const TabB& Tt, double dx, double dy,
                                     double dz, double dt, bool& out) {
  constexpr int SD = DIM == 4 ? 3 : DIM;
  const double hs = aniso_h<SD>(Ts.m, dx, dy, dz);
  double frs;
  const int is = tabB_index(hs, frs);
  const double s0 = Ts.tab[is], s1 = Ts.tab[is + 1];
  if (DIM <= 3) return tabB_pick(Ts, hs, frs, s0, s1, out);
  const double ht = aniso_h<1>(Tt.m, fabs(dt), 0.0, 0.0);
  double frt;
  const int it = tabB_index(ht, frt);
  const double t0 = Tt.tab[it], t1 = Tt.tab[it + 1];
  bool os, ot;
  const double cs = tabB_pick(Ts, hs, frs, s0, s1, os);
  const double ct = tabB_pick(Tt, ht, frt, t0, t1, ot);
  out = os || ot;
  return a.k_ps * cs * ct + cs + ct;
}

// G lanes per system, R <= G matrix rows (R < G: lanes R..G-1 hold no row;
// the 16-lane 12-row, 32-lane 24-row and 64-lane 48-row instances take the
// systems just above the next smaller width with rows a quarter shorter:
// fewer registers, fold offsets and trailing updates).
//
// Each group of G lanes takes one work item: a neighbour set (the rows of its
// representative block) and up to MAX_ITEM member blocks whose matrices are
// identical to it (or, with no item list, one block on its own).  K is
// assembled and factored once, with the unbias column F, the data values z
// and every member's right-hand side b eliminated alongside (the forward
// substitutions), so a block's arithmetic is the same whether it shares its
// factorisation or not.
template <int G, int DIM, int PREC, bool TB, int R = G>
__global__ __launch_bounds__(64) void k_solve(SolveArgs a) {
  constexpr int S = 64 / G;
  constexpr int MB = item_cap(R);  // member blocks per work item
  static_assert(R <= G && 2 * G <= 3 * R, "lanes past R store into the unread fold slots");
  __shared__ double sx[S][DIM][G];
  // fold exchange: sp[d][r] = C(row r, row (r + d) mod R) for d = 1..R/2 as
  // lane r computed it, sp[0][r] = diagonal; later the pivot row's extra
  // columns and (as floats) the FP32 factor
  __shared__ double sp[S][R / 2 + 1][R];
  __shared__ double col[S][G];
  // FP32+refine: the members' FP32 solutions, broadcast for the FP64 residuals
  __shared__ double xs[PREC == PREC_FP32R ? S : 1][MB][G];

  const int lane = threadIdx.x, s = lane / G, r = lane % G;
  // with a device-resident count (graph-captured passes) the grid is sized
  // for the largest possible list and workgroups past the count exit (a
  // loop over the list here would cost the instances a wave of occupancy in
  // registers)
  const int nlist = a.dn ? *a.dn : a.nlist;
  const int2* items0 = a.items ? a.items + (a.doff ? *a.doff : 0) : nullptr;
  const int* list0 = a.list + (a.doff && !a.items ? *a.doff : 0);
  const int base = blockIdx.x * S;
  if (base >= nlist) return;
  {
  // lane r holds row r; the slots (R/2, R/2..R-1) are never read, so lanes
  // past R store there instead of branching around their stores
  const bool row = R == G || r < R;
  const int rr = row ? r : 0;
  auto slot = [&](int d) -> double* { return row ? &sp[s][d][r] : &sp[s][R / 2][r - R / 2]; };
  auto wrap = [](int c) { return R == G ? (c & (G - 1)) : (c >= R ? c - R : c); };
  const int sys = base + s;
  const bool act = sys < nlist;
  // work item: representative block b, members m0 .. m0 + mc - 1 of the
  // member array (or b alone)
  int b = 0, m0 = 0, mc = 0;
  if (act) {
    if (items0) {
      const int2 it = items0[sys];
      b = it.x;
      m0 = it.y;
      mc = min(min(a.item_max, MB), a.setbase[b] + a.setcnt[b] - m0);
    } else {
      b = list0[sys];
      mc = 1;
    }
  }
  const int n = act ? a.nnear[b] : 0;
  const bool live = r < n;
  const int idx = live ? a.inear[(size_t)b * a.mmax + r] : 0;
  // largest system and largest member count in this wave; readfirstlane makes
  // them scalars, so the k/j loop guards below are scalar branches instead of
  // exec-mask regions
  const int nw = __builtin_amdgcn_readfirstlane(wmax(n));
  const int mw = __builtin_amdgcn_readfirstlane(wmax(mc));

  double x[4] = {0.0, 0.0, 0.0, 0.0};
  if (live) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) x[k] = a.ocoord[k + DIM * (size_t)idx];
  }
#pragma unroll
  for (int k = 0; k < DIM; ++k) sx[s][k][r] = x[k];
  const double z = live ? a.ovalue[idx] : 0.0;
  const double ov = live ? a.ovar[idx] : 0.0;
  // lags are divided by the block's range scale only when one is set (the
  // host passes none when every block's scale is 1); every member of a set
  // has the representative's scale and local nugget
  const bool scaled = a.rangescale != nullptr;
  const double rs = (act && scaled) ? a.rangescale[b] : 1.0;
  const double ln = (act && a.localnugget) ? a.localnugget[b] : 0.0;
  const Model& M = *a.model;
  const double cov0 = DIM == 4 ? a.st_cov0 : M.cov0;
  // TB: the tables' parameters, read once (kernel-uniform)
  TabB Ts{}, Tt{};
  if constexpr (TB) {
    Ts = tabB_view(M);
    if constexpr (DIM == 4) Tt = tabB_view(*a.tmodel);
  }
  __syncthreads();

  // lag from this lane's point to (u[0], u[DIM-1]), range-scaled when set
  auto lag_sc = [&](bool sc, const double* u, int stride, double& dx, double& dy, double& dz, double& dt) {
    dx = x[0] - u[0];
    dy = DIM >= 2 ? x[1] - u[stride] : 0.0;
    dz = DIM >= 3 ? x[2] - u[2 * stride] : 0.0;
    dt = DIM == 4 ? x[DIM - 1] - u[(DIM - 1) * stride] : 0.0;
    if (sc) {
      dx = dx / rs;
      dy = dy / rs;
      dz = dz / rs;
    }
  };
  auto lag = [&](const double* u, int stride, double& dx, double& dy, double& dz, double& dt) {
    lag_sc(scaled, u, stride, dx, dy, dz, dt);
  };

  // ---- assemble K by the fold (each pair once) --------------------------
  // TB: fold offsets whose lag lies beyond a table with an analytic tail are
  // recorded and evaluated after the loop, keeping the loop free of calls
  // offsets d >= nw pair no live rows: their slots are zero.  A slot
  // (R/2, r >= R/2) is never read, so every lane stores unconditionally; with
  // TB the loop body is branch-free and consecutive offsets' loads overlap
  // (the loop is instantiated with and without range scaling, so that no
  // branch splits its unrolled body)
  unsigned long long redo = 0;
  const int dmax = min(R / 2, nw - 1);
  // unrolled by 2 (by 4 in the 64-lane FP64 instances)
  constexpr int FOLD_UNROLL = (PREC == PREC_FP64 && G == 64) ? 4 : 2;
  auto fold = [&](auto sc) {
#pragma unroll FOLD_UNROLL
    for (int d = 1; d <= dmax; ++d) {
      const int c = wrap(r + d);
      const bool use = live && c < n && ((d < R / 2) || (r < R / 2));
      double v = 0.0;
      if (TB) {
        double dx, dy, dz, dt;
        lag_sc(decltype(sc)::value, &sx[s][0][c], G, dx, dy, dz, dt);
        bool out;
        v = pair_cov_tb<DIM>(a, Ts, Tt, dx, dy, dz, dt, out);
        if (use && out) redo |= 1ull << d;
        v = use ? v : 0.0;
      } else if (use) {
        double dx, dy, dz, dt;
        lag_sc(decltype(sc)::value, &sx[s][0][c], G, dx, dy, dz, dt);
        v = pair_cov<DIM, TB>(a, M, dx, dy, dz, dt);
      }
      *slot(d) = v;
    }
  };
  if (scaled)
    fold(std::true_type{});
  else
    fold(std::false_type{});
  for (int d = max(dmax, 0) + 1; d <= R / 2; ++d) *slot(d) = 0.0;
  if (TB) {
    while (redo) {
      const int d = __ffsll((long long)redo) - 1;
      redo &= redo - 1;
      const int c = wrap(r + d);
      double dx, dy, dz, dt;
      lag(&sx[s][0][c], G, dx, dy, dz, dt);
      sp[s][d][r] = pair_cov<DIM, TB>(a, M, dx, dy, dz, dt);
    }
  }
  if (row) sp[s][0][r] = live ? (cov0 + ov) + ln : 1.0;
  // right-hand side of this lane's row for block mb:
  // b_r = sum_k w_k C((x_r - g_k)/rs) over the block's integration points
  auto rhs_of = [&](int mb) -> double {
    double acc = 0.0;
    if (!live) return acc;
    for (int q = 0; q < a.nsub; ++q) {
      const double* g = a.nsub == 1 && !a.scoord ? a.bcoord + DIM * (size_t)mb
                                                 : a.scoord + DIM * ((size_t)mb * a.nsub + q);
      const double w = a.sweight ? a.sweight[(size_t)mb * a.nsub + q] : 1.0;
      double dx, dy, dz, dt;
      lag(g, 1, dx, dy, dz, dt);
      double v;
      if (TB) {
        bool out;
        v = pair_cov_tb<DIM>(a, Ts, Tt, dx, dy, dz, dt, out);
        if (out) v = pair_cov<DIM, TB>(a, M, dx, dy, dz, dt);
      } else {
        v = pair_cov<DIM, TB>(a, M, dx, dy, dz, dt);
      }
      acc = acc + v * w;
    }
    return acc;
  };
  // the item's member blocks (member j of the item; the block itself without
  // an item list) and their right-hand sides, evaluated before the matrix
  // rows are loaded; columns past the wave's largest member count are
  // skipped by scalar guards
  auto member = [&](int j) { return a.items ? a.members[m0 + j] : b; };
  int mb[MB];
  double bv[MB];
#pragma unroll
  for (int c = 0; c < MB; ++c) {
    const bool on = c < mc;
    mb[c] = on ? member(c) : b;
    bv[c] = (c < mw && on) ? rhs_of(mb[c]) : 0.0;
  }
  // ones on the live rows: the unbias row F for ordinary kriging, and the
  // weight sum 1^T K^-1 b that simple kriging needs for its mean term
  const double f = live ? 1.0 : 0.0;
  __syncthreads();
  // row r: entry j is the pair at offset d = (j - r) mod R, computed by lane
  // r when d <= R/2 (d == R/2 only by the lower lane), else by lane j; lanes
  // past R hold a zero row
  double A[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int d = wrap(j - rr + R);
    const bool own = d < R / 2 || (d == R / 2 && rr < R / 2);
    const double e = own ? sp[s][d][rr] : sp[s][R - d][j];
    A[j] = row ? e : 0.0;
  }
  const double kdiag = sp[s][0][rr];

  int fail = 0;
  // outputs of member block mb from the quadratic forms uu = F'K^-1F,
  // uy = F'K^-1b, yy = b'K^-1b, vy = z'K^-1b, vu = z'K^-1F
  auto write_out = [&](int mb, double uu, double uy, double yy, double vy, double vu) {
    if (n == 1) {
      // the reference skips the solve for a single neighbour (solve_base,
      // kriging_base.F90:2127): variance keeps its NaN fill; the estimate
      // would reuse another block's weights and is reported as NaN
      a.est[mb] = __builtin_nan("");
      a.var[mb] = __builtin_nan("");
      a.flag[mb] = 0;
      return;
    }
    if (fail & 1) {
      a.flag[mb] = 2;  // not positive definite: Bunch-Kaufman fallback
      return;
    }
    if (fail & 4) {
      a.flag[mb] = 1;  // FP32 pivot too small: FP64 re-solve
      return;
    }
    double var, est;
    const double base = a.base_cov ? a.base_cov[mb] : cov0;
    if (a.unbias) {
      const double lam = (uy - 1.0) / uu;
      var = (base - yy) + lam * (uy - 1.0);
      est = vy - lam * vu;
    } else {
      var = base - yy;
      est = vy;
      if (a.sk_mean != 0.0) est = est + (1.0 - uy) * a.sk_mean;
    }
    var = var > 0.0 ? var : 0.0;
    est = est < a.bound_lo ? a.bound_lo : est;
    est = est > a.bound_hi ? a.bound_hi : est;
    a.est[mb] = est;
    a.var[mb] = var;
    a.flag[mb] = 0;
  };
  // flagged work items and their blocks (rare: two atomics each)
  auto count_flagged = [&]() {
    if (act && r == 0 && n > 1 && (fail & 5) && a.ctr) {
      atomicAdd(&a.ctr[0], 1);
      atomicAdd(&a.ctr[1], mc);
    }
  };

  if (PREC == PREC_FP64) {
    // ---- L D L^T, row r in A[], with F, z and the members' b carried as
    // extra columns: eliminating them alongside K is the forward substitution
    // u = L^-1 F, v = L^-1 z, y_c = L^-1 b_c --------------------------------
    double D = 1.0;
    double u = f, v = z;
    double y[MB];
#pragma unroll
    for (int c = 0; c < MB; ++c) y[c] = bv[c];
    double* ext = &sp[s][0][0];  // pivot row's extra columns; sp is free after the row loads
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (k < nw) {
        col[s][r] = A[k];
        if (r == k) {
          ext[0] = u;
          ext[1] = v;
#pragma unroll
          for (int c = 0; c < MB; ++c) ext[2 + c] = y[c];
        }
        __syncthreads();
        const double akk = col[s][k];
        if (r == k) D = akk;
        fail |= (k < n) && !(akk > 0.0);
        const double t = (r > k) ? A[k] * recip(akk) : 0.0;
        // columns past nw hold zeros (padding rows): updating them is
        // harmless and keeps the unrolled loop branch-free
#pragma unroll
        for (int j = k + 1; j < R; ++j) A[j] = fma(-t, col[s][j], A[j]);
        u = fma(-t, ext[0], u);
        v = fma(-t, ext[1], v);
#pragma unroll
        // (columns past the wave's member count hold zeros: updating them
        // is harmless, and a guard here costs the wide instances a wave of
        // occupancy in registers)
        for (int c = 0; c < MB; ++c)
          y[c] = fma(-t, ext[2 + c], y[c]);
        __syncthreads();
      }
    }
    const double id = live ? 1.0 / D : 0.0;
    const double uu = gsum<G>(u * u * id);
    const double vu = gsum<G>(v * u * id);
    fail = gmax<G>(fail);
    count_flagged();
#pragma unroll
    for (int c = 0; c < MB; ++c) {
      if (c >= mw) continue;
      const double uy = gsum<G>(u * y[c] * id);
      const double yy = gsum<G>(y[c] * y[c] * id);
      const double vy = gsum<G>(v * y[c] * id);
      if (act && r == 0 && c < mc) write_out(mb[c], uu, uy, yy, vy, vu);
    }
  } else {
    // ---- FP32 L D L^T; F, z and the members' b carried as extra columns
    // (fused forward substitution), one FP32 back substitution for
    // x1 ~ K^-1 F, xz ~ K^-1 z, xb_c ~ K^-1 b_c.  Each quadratic form
    // c'K^-1 d is corrected with the FP64 residual of d:
    //   c'K^-1 d = c'x_d + x_c'(d - K x_d) - e_c' K e_d,
    // e = solve errors, so the dropped term is second order in the FP32
    // error (one refinement step's accuracy without a second solve) --------
    float* col32 = reinterpret_cast<float*>(&col[s][0]);
    float* ext32 = reinterpret_cast<float*>(&sp[s][0][0]);  // free after the row loads
    float L[R];
#pragma unroll
    for (int j = 0; j < R; ++j) L[j] = (float)A[j];
    float D = 1.0f;
    float pmin = 1.0f;
    float pu = (float)f, pz = (float)z;
    float py[MB];
#pragma unroll
    for (int c = 0; c < MB; ++c) py[c] = (float)bv[c];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (k < nw) {
        col32[r] = L[k];
        if (r == k) {
          ext32[0] = pu;
          ext32[1] = pz;
#pragma unroll
          for (int c = 0; c < MB; ++c) ext32[2 + c] = py[c];
        }
        __syncthreads();
        const float akk = col32[k];
        if (r == k) D = akk;
        fail |= (k < n) && !(akk > 0.0f);
        const float t = (r > k) ? L[k] * __builtin_amdgcn_rcpf(akk) : 0.0f;
#pragma unroll
        for (int j = k + 1; j < R; ++j) L[j] = fmaf(-t, col32[j], L[j]);
        pu = fmaf(-t, ext32[0], pu);
        pz = fmaf(-t, ext32[1], pz);
#pragma unroll
        for (int c = 0; c < MB; ++c) py[c] = fmaf(-t, ext32[2 + c], py[c]);
        L[k] = (r > k) ? t : L[k];
        __syncthreads();
      }
    }
    if (live) pmin = (float)((double)D / kdiag);
    // transpose L through LDS for the back substitution
    // (lanes past R read column r < G of the row-major R x (R+1) block, in
    // bounds, and never use it)
    static_assert(sizeof(float) * ((R - 1) * (R + 1) + G) <= sizeof(double) * (R / 2 + 1) * R, "transpose fits");
    float* sm32 = reinterpret_cast<float*>(&sp[s][0][0]);
    if (row) {
#pragma unroll
      for (int j = 0; j < R; ++j) sm32[r * (R + 1) + j] = L[j];
    }
    __syncthreads();
    float Lt[R];
#pragma unroll
    for (int j = 0; j < R; ++j) Lt[j] = sm32[j * (R + 1) + r];
    const float iD = 1.0f / D;
    pu *= iD;
    pz *= iD;
#pragma unroll
    for (int c = 0; c < MB; ++c) py[c] *= iD;
#pragma unroll
    for (int k = R - 1; k >= 0; --k) {
      if (k < nw) {
        const float l = (r < k) ? Lt[k] : 0.0f;
        const float uk = __shfl(pu, k, G), zk = __shfl(pz, k, G);
        pu = fmaf(-l, uk, pu);
        pz = fmaf(-l, zk, pz);
#pragma unroll
        for (int c = 0; c < MB; ++c) {
          if (c >= mw) continue;
          const float yk = __shfl(py[c], k, G);
          py[c] = fmaf(-l, yk, py[c]);
        }
      }
    }
    const double x1 = pu, xz = pz;
    // FP64 residuals r1 = F - K x1 and rb_c = b_c - K xb_c (row r of K is
    // A[], FP64)
    double r1 = f;
    __syncthreads();
    col[s][r] = x1;
#pragma unroll
    for (int c = 0; c < MB; ++c) xs[s][c][r] = py[c];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) r1 = fma(-A[j], col[s][j], r1);
    const double uu = gsum<G>(fma(x1, r1, f * x1));
    const double vu = gsum<G>(fma(xz, r1, z * x1));
    // pivot ratio check: re-solve in FP64 below the threshold
    float pm = live ? pmin : 1.0f;
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) pm = fminf(pm, __shfl_xor(pm, m, G));
    if (act && r == 0 && pm < (float)a.pivot_tol) fail |= 4;
    fail = gmax<G>(fail);
    count_flagged();
#pragma unroll
    for (int c = 0; c < MB; ++c) {
      if (c >= mw) continue;
      double rb = bv[c];
#pragma unroll
      for (int j = 0; j < R; ++j) rb = fma(-A[j], xs[s][c][j], rb);
      const double xb = py[c];
      const double uy = gsum<G>(fma(x1, rb, f * xb));
      const double yy = gsum<G>(fma(xb, rb, bv[c] * xb));
      const double vy = gsum<G>(fma(xz, rb, z * xb));
      if (act && r == 0 && c < mc) write_out(mb[c], uu, uy, yy, vy, vu);
    }
  }
  }
}

#if KK_SOLVE_DIM == 0
// blocks without neighbours: NaN outputs (kriging_base.F90:3421-3425)
__global__ void k_empty(SolveArgs a) {
  const int nl = a.dn ? *a.dn : a.nlist;
  const int* list = a.list + (a.doff ? *a.doff : 0);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x) {
    const int b = list[t];
    a.est[b] = __builtin_nan("");
    a.var[b] = __builtin_nan("");
    a.flag[b] = 0;
  }
}

// Bucket counts are aggregated in LDS per 1024-thread workgroup, so each
// workgroup issues one global atomic per bucket.
// Blocks 0..nblock-1 (or the blocks listed in `sel`) into per-size lists.
// A negative count marks a block whose neighbours are still being searched.
__global__ __launch_bounds__(1024) void k_bucket(const int* nnear, const int* sel, int nblock, int* counts,
                                                 int* lists, int cap, const int* dnsel, const int* skip) {
  // a device-resident count of the listed blocks (graph-captured passes;
  // the grid covers nblock)
  if (dnsel) nblock = *dnsel;
  // order-preserving within the workgroup: ballot ranks within a wave, an
  // exclusive scan over the 16 waves, one global atomic per bucket, so a
  // workgroup's blocks keep their order (and their run of the list is the
  // same on every pass)
  __shared__ int wcount[16][NBUCKET + 1];
  __shared__ int lbase[NBUCKET + 1];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int b = t < nblock ? (sel ? sel[t] : t) : -1;
  // blocks flagged for the exact search are left out whatever their count
  // reads: that search may be writing it on another stream
  const int n = (b >= 0 && !(skip && skip[b])) ? nnear[b] : -1;
  const int bk = n < 0 ? -1 : n == 0 ? 0 : n > 64 ? NBUCKET : bucket_of(n);
  int local = 0;
#pragma unroll
  for (int k = 0; k <= NBUCKET; ++k) {
    const unsigned long long m = __ballot(bk == k);
    if (bk == k) local = __popcll(m & ((1ull << ln) - 1ull));
    if (ln == 0) wcount[wv][k] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x <= NBUCKET) {
    const int k = threadIdx.x;
    int acc = 0;
    for (int w = 0; w < 16; ++w) {
      const int c = wcount[w][k];
      wcount[w][k] = acc;
      acc += c;
    }
    lbase[k] = acc ? atomicAdd(&counts[k], acc) : 0;
  }
  __syncthreads();
  if (bk >= 0) lists[(size_t)bk * cap + lbase[bk] + wcount[wv][bk] + local] = b;
}

// blocks with flag == code into list (atomic count); with perm, in the
// order perm lists the blocks
__global__ void k_collect(const int* flag, int nblock, int code, int* count, int* list, const int* perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int t = i < nblock ? (perm ? perm[i] : i) : 0;
  const bool hit = i < nblock && flag[t] == code;
  const unsigned long long m = __ballot(hit);
  if (m == 0) return;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (hit) list[base + __popcll(m & below)] = t;
}

// The reference's fallback when the Cholesky path fails, one thread per
// system with the matrices in HBM scratch: Bunch-Kaufman factorisation of the
// full augmented system (SSYTRF/SSYTRS), one retry with the data diagonal
// raised by 1e-6*mean(diag) when it is exactly singular, then NaN
// (kriging_base.F90:3299-3383).
template <int DIM>
__device__ inline double fb_cov(const SolveArgs& a, const Model& M, const double* xi, const double* xj, double rs) {
  const double dx = (xi[0] - xj[0]) / rs, dy = DIM >= 2 ? (xi[1] - xj[1]) / rs : 0.0;
  const double dz = DIM >= 3 ? (xi[2] - xj[2]) / rs : 0.0;
  if (DIM <= 3) return cov<DIM>(M, dx, dy, dz);
  const double cs = cov<3>(M, dx, dy, dz);
  const double dt = xi[DIM - 1] - xj[DIM - 1];
  const double ct = cov<1>(*a.tmodel, fabs(dt), 0.0, 0.0);
  if (a.st_sum) return st_sum_cov(&a, &M, cs + ct, dx, dy, dz, dt);
  return a.k_ps * cs * ct + cs + ct;
}

// block b, with fallback_words(mmax) doubles of scratch at `scratch`
template <int DIM>
__device__ void fallback_one(const SolveArgs& a, int b, double* scratch) {
  const int n = a.nnear[b];
  const int p = a.unbias ? 1 : 0, m = n + p;
  const int ld = a.mmax + 1;
  double* K = scratch;
  double* A = K + ld * ld;
  double* rhs = A + ld * ld;
  double* xs = rhs + ld;
  int* ipiv = reinterpret_cast<int*>(xs + ld);
  const int* in = a.inear + (size_t)b * a.mmax;
  const double rs = a.rangescale ? a.rangescale[b] : 1.0;
  const double ln = a.localnugget ? a.localnugget[b] : 0.0;
  const Model& M = *a.model;
  for (int i = 0; i < n; ++i) {
    const double* xi = a.ocoord + DIM * (size_t)in[i];
    K[i + ld * i] = ((DIM == 4 ? a.st_cov0 : M.cov0) + a.ovar[in[i]]) + ln;
    for (int j = i + 1; j < n; ++j) {
      const double* xj = a.ocoord + DIM * (size_t)in[j];
      const double c = fb_cov<DIM>(a, M, xi, xj, rs);
      K[j + ld * i] = c;
      K[i + ld * j] = c;
    }
    if (p) {
      K[n + ld * i] = 1.0;
      K[i + ld * n] = 1.0;
    }
    double tmp = 0.0;
    for (int q = 0; q < a.nsub; ++q) {
      const double* g = a.nsub == 1 && !a.scoord ? a.bcoord + DIM * (size_t)b
                                                 : a.scoord + DIM * ((size_t)b * a.nsub + q);
      const double w = a.sweight ? a.sweight[(size_t)b * a.nsub + q] : 1.0;
      tmp = tmp + fb_cov<DIM>(a, M, xi, g, rs) * w;
    }
    rhs[i] = tmp;
  }
  if (p) {
    K[n + ld * n] = 0.0;
    rhs[n] = 1.0;
  }
  for (int i = 0; i < m; ++i) xs[i] = rhs[i];
  const int info = bk_fallback(K, A, ipiv, n, m, ld, xs, 1);
  if (a.weights) {
    // the weight store: x(1:n), NaN when the system stays singular
    for (int i = 0; i < n; ++i) a.weights[(size_t)b * a.mmax + i] = info ? __builtin_nan("") : xs[i];
    if (a.weights_only) return;
  }
  a.flag[b] = info ? 3 : 0;
  if (info) {
    // x = NaN (kriging_base.F90:3369-3383): NaN estimate; the variance is
    // max(NaN, 0) = 0 after calc_variance's clamp (kriging.F90:1019-1025)
    a.est[b] = __builtin_nan("");
    a.var[b] = 0.0;
    return;
  }
  double dot = 0.0, est = 0.0, tw = 0.0;
  for (int i = 0; i < m; ++i) dot = dot + xs[i] * rhs[i];
  for (int i = 0; i < n; ++i) {
    est = est + a.ovalue[in[i]] * xs[i];
    tw = tw + xs[i];
  }
  const double base = a.base_cov ? a.base_cov[b] : (DIM == 4 ? a.st_cov0 : M.cov0);
  double var = base - dot;
  if (!a.unbias && a.sk_mean != 0.0) est = est + (1.0 - tw) * a.sk_mean;
  var = var > 0.0 ? var : 0.0;
  est = est < a.bound_lo ? a.bound_lo : est;
  est = est > a.bound_hi ? a.bound_hi : est;
  a.est[b] = est;
  a.var[b] = var;
}

// one thread per listed block; with a device-resident count the grid loops,
// each thread reusing its own scratch slot
template <int DIM>
__global__ void k_fallback(SolveArgs a, double* scratch) {
  const int nl = a.dn ? *a.dn : a.nlist;
  const int* list = a.list + (a.doff ? *a.doff : 0);
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (int t = t0; t < nl; t += gridDim.x * blockDim.x)
    fallback_one<DIM>(a, list[t], scratch + (size_t)t0 * fallback_words(a.mmax));
}

// blocks with more than limit neighbours into list (atomic count)
__global__ void k_collect_above(const int* nnear, int nblock, int limit, int* count, int* list) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool hit = i < nblock && nnear[i] > limit;
  const unsigned long long m = __ballot(hit);
  if (m == 0) return;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (hit) list[base + __popcll(m & below)] = i;
}

// plain copy of n doubles (reads 8n bytes, writes 8n bytes): calibrates HBM
// traffic counters against a kernel of known size
__global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy(const double* src, double* dst, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

int bucket_width(int bk) { return bk >= 1 && bk < NBUCKET ? BUCKET_ROWS[bk - 1] : 0; }

int solve_width(int n) { return BUCKET_ROWS[bucket_of(n) - 1]; }

hipError_t launch_bucket(const int* nnear, const int* sel, int nblock, int* counts, int* lists, int cap,
                         hipStream_t st, const int* dnsel, const int* skip) {
  if (nblock <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bucket, dim3((nblock + 1023) / 1024), dim3(1024), 0, st, nnear, sel, nblock, counts,
                     lists, cap, dnsel, skip);
  return hipGetLastError();
}

// graph-captured passes: any counter above the bound its launches were
// sized for sets *overflow
__global__ void k_check_bounds(const int* counts, BoundCheck c, int* overflow) {
  const int i = threadIdx.x;
  if (i < c.n && counts[c.idx[i]] > c.bound[i]) atomicOr(overflow, 1);
}

hipError_t launch_check_bounds(const int* counts, const BoundCheck& c, int* overflow, hipStream_t st) {
  hipLaunchKernelGGL(k_check_bounds, dim3(1), dim3(64), 0, st, counts, c, overflow);
  return hipGetLastError();
}

__global__ void k_fill(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill(int* p, int n, int v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_collect_above(const int* nnear, int nblock, int limit, int* count, int* list, hipStream_t st) {
  if (nblock <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_collect_above, dim3((nblock + 255) / 256), dim3(256), 0, st, nnear, nblock, limit, count, list);
  return hipGetLastError();
}

hipError_t launch_collect(const int* flag, int nblock, int code, int* count, int* list,
                          hipStream_t st, const int* perm) {
  hipLaunchKernelGGL(k_collect, dim3((nblock + 255) / 256), dim3(256), 0, st, flag, nblock, code,
                     count, list, perm);
  return hipGetLastError();
}

#endif  // KK_SOLVE_DIM == 0

// k_solve launchers for one coordinate count, one per translation unit
// (solve.hip compiled with -DKK_SOLVE_DIM=1..4) so the 16 instances of each
// compile in parallel; the unit without KK_SOLVE_DIM holds the dispatch.
hipError_t launch_solve_d1(const SolveArgs& a, int width, int prec, bool tableB, hipStream_t st);
hipError_t launch_solve_d2(const SolveArgs& a, int width, int prec, bool tableB, hipStream_t st);
hipError_t launch_solve_d3(const SolveArgs& a, int width, int prec, bool tableB, hipStream_t st);
hipError_t launch_solve_d4(const SolveArgs& a, int width, int prec, bool tableB, hipStream_t st);

#if KK_SOLVE_DIM > 0
template <int G, int PREC, int R = G>
static void launch_g(const SolveArgs& a, bool tb, hipStream_t st) {
  constexpr int S = 64 / G;
  dim3 grd((a.nlist + S - 1) / S), blk(64);
  if (tb) hipLaunchKernelGGL((k_solve<G, KK_SOLVE_DIM, PREC, true, R>), grd, blk, 0, st, a);
  else hipLaunchKernelGGL((k_solve<G, KK_SOLVE_DIM, PREC, false, R>), grd, blk, 0, st, a);
}

#define KK_CAT2(x, y) x##y
#define KK_CAT(x, y) KK_CAT2(x, y)
hipError_t KK_CAT(launch_solve_d, KK_SOLVE_DIM)(const SolveArgs& a, int width, int prec, bool tableB,
                                                hipStream_t st) {
  if (prec == PREC_FP64) {
    switch (width) {
      case 8: launch_g<8, PREC_FP64>(a, tableB, st); break;
      case 12: launch_g<16, PREC_FP64, 12>(a, tableB, st); break;
      case 16: launch_g<16, PREC_FP64>(a, tableB, st); break;
      case 24: launch_g<32, PREC_FP64, 24>(a, tableB, st); break;
      case 32: launch_g<32, PREC_FP64>(a, tableB, st); break;
      case 48: launch_g<64, PREC_FP64, 48>(a, tableB, st); break;
      case 64: launch_g<64, PREC_FP64>(a, tableB, st); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (width) {
      case 8: launch_g<8, PREC_FP32R>(a, tableB, st); break;
      case 12: launch_g<16, PREC_FP32R, 12>(a, tableB, st); break;
      case 16: launch_g<16, PREC_FP32R>(a, tableB, st); break;
      case 24: launch_g<32, PREC_FP32R, 24>(a, tableB, st); break;
      case 32: launch_g<32, PREC_FP32R>(a, tableB, st); break;
      case 48: launch_g<64, PREC_FP32R, 48>(a, tableB, st); break;
      case 64: launch_g<64, PREC_FP32R>(a, tableB, st); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}
#else
hipError_t launch_solve(const SolveArgs& a, int width, int prec, bool tableB, hipStream_t st) {
  if (a.nlist <= 0) return hipSuccess;
  switch (a.ndim) {
    case 1: return launch_solve_d1(a, width, prec, tableB, st);
    case 2: return launch_solve_d2(a, width, prec, tableB, st);
    case 3: return launch_solve_d3(a, width, prec, tableB, st);
    default: return launch_solve_d4(a, width, prec, tableB, st);  // space-time (x, y, z, t)
  }
}

hipError_t launch_empty(const SolveArgs& a, hipStream_t st) {
  if (a.nlist <= 0) return hipSuccess;
  int ng = (a.nlist + 255) / 256;
  if (a.grid_cap > 0 && ng > a.grid_cap) ng = a.grid_cap;
  hipLaunchKernelGGL(k_empty, dim3(ng), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_fallback(const SolveArgs& a, double* scratch, hipStream_t st) {
  if (a.nlist <= 0) return hipSuccess;
  if (a.ndim == 1) hipLaunchKernelGGL(k_fallback<1>, dim3((a.nlist + 63) / 64), dim3(64), 0, st, a, scratch);
  else if (a.ndim == 2) hipLaunchKernelGGL(k_fallback<2>, dim3((a.nlist + 63) / 64), dim3(64), 0, st, a, scratch);
  else if (a.ndim == 3) hipLaunchKernelGGL(k_fallback<3>, dim3((a.nlist + 63) / 64), dim3(64), 0, st, a, scratch);
  else hipLaunchKernelGGL(k_fallback<4>, dim3((a.nlist + 63) / 64), dim3(64), 0, st, a, scratch);
  return hipGetLastError();
}

#endif  // KK_SOLVE_DIM

}  // namespace kk
