// Device-side covariance evaluation.
//
// C(lag) for one variable pair follows the reference's COV macro
// (src/libkriging/cov_dispatch.fh:14-19):
//   table mode  -> struct_cov_tab   (src/libkriging/variogram.f90:750-771)
//   analytic    -> struct_cov_lag   (src/libkriging/variogram.f90:731-747)
// The lookup reproduces cov_struct_tab_h (variogram.f90:682-710) exactly:
// zone search over h_bounds, index clamped to n_zone-1, fraction taken from
// the unclamped integer part, entry 0 = C(0) including the nugget.
#pragma once
#include <hip/hip_runtime.h>

#include "kk_types.hpp"

namespace kk {

// correlation shapes, src/libkriging/vgmfunc.f90:48-87
__device__ inline double corefunc(int t, double r) {
  const double pi = 3.141592653589793238462643383279502884;
  switch (t) {
    case SPH: return (r < 1.0) ? (1.0 - 1.5 * r) + 0.5 * (r * r * r) : 0.0;
    case EXP: return exp(-3.0 * r);
    case GAU: return exp(-3.0625 * (r * r));
    case HOL: return cos(pi * r);
    case POW: return (r < 1.0) ? 1.0 - pow(r, 1.5) : 0.0;
    case BSQ: { double a = 1.0 - r * r; return (r < 1.0) ? a * a : 0.0; }
    case CIR: return (r < 1.0) ? 1.0 - (2.0 * r * sqrt(1.0 - r * r) + 2.0 * asin(r)) / pi : 0.0;
    case LIN: return (r < 1.0) ? 1.0 - r : 0.0;
    case CYC: { double s = sin(pi * r); return exp(-2.0 * (s * s)); }
    case DCO: return exp(-3.0 * r) * cos(pi * r);
    default: return 0.0;
  }
}

__device__ inline bool has_tail(int t) {
  return t == EXP || t == GAU || t == HOL || t == CYC || t == DCO;
}

// h = |mat * lag|, src/libkriging/variogram.f90:194-210
template <int DIM>
__device__ inline double aniso_h(const double* m, double dx, double dy, double dz) {
  if (DIM == 1) {
    return fabs(dx * m[0]);
  } else if (DIM == 2) {
    double rx = m[0] * dx + m[3] * dy;
    double ry = m[1] * dx + m[4] * dy;
    return sqrt(rx * rx + ry * ry);
  } else {
    double rx = m[0] * dx + m[3] * dy + m[6] * dz;
    double ry = m[1] * dx + m[4] * dy + m[7] * dz;
    double rz = m[2] * dx + m[5] * dy + m[8] * dz;
    return sqrt(rx * rx + ry * ry + rz * rz);
  }
}

// comp_cov_h, variogram.f90:327-337
__device__ inline double comp_cov_h(const Model& M, int c, double h) {
  const double eps = 2.2250738585072014e-308 * 1.0e3;
  return (h > eps) ? M.sill[c] * corefunc(M.vtype[c], h) : M.sill[c] + M.nugget[c];
}

// struct_cov_h, variogram.f90:712-728 (one h for every structure)
__device__ inline double struct_cov_h(const Model& M, double h) {
  double res = 0.0;
  int iv = 0;
  while (iv < M.nstruct) {
    double g = comp_cov_h(M, iv, h);
    while (iv + 1 < M.nstruct && M.is_product[iv + 1]) {
      ++iv;
      g = g * comp_cov_h(M, iv, h);
    }
    res = res + g;
    ++iv;
  }
  return res;
}

// Table entry index and fraction of h, as cov_struct_tab_h computes them
// (variogram.f90:682-710): zone k = (h >= .1) + (h >= .5), fi = (h - zb_k) /
// step_k, index zoff_k + min(int(fi), n_k - 1), fraction fi - int(fi).  The
// zones are those of TAB_HB, which every table is built on, so they are
// constants chosen by selects.  The index stays in [0, TAB_N - 1] for every
// h, NaN and h beyond the table included.
__device__ inline int tabB_index(double h, double& fr) {
  constexpr double zb1 = TAB_HB[1], zb2 = TAB_HB[2];
  constexpr double zi0 = tab_zinv(0), zi1 = tab_zinv(1), zi2 = tab_zinv(2);
  constexpr int zo0 = tab_zoff(0), zo1 = tab_zoff(1), zo2 = tab_zoff(2);
  constexpr int zn0 = tab_zn(0), zn1 = tab_zn(1), zn2 = tab_zn(2);
  const bool k1 = h >= zb1, k2 = h >= zb2;
  const double zb = k2 ? zb2 : k1 ? zb1 : TAB_HB[0];
  const double zi = k2 ? zi2 : k1 ? zi1 : zi0;
  const int zo = k2 ? zo2 : k1 ? zo1 : zo0;
  const int zn = k2 ? zn2 : k1 ? zn1 : zn0;
  double fi = (h - zb) * zi;
  const int ifi = (int)fmin(fi, 1073741824.0);
  fr = fi - (double)ifi;
  return max(zo + min(ifi, zn - 1), 0);
}

// in-table part of cov_struct_tab_h / comp_cov_tab_h; h in (0, 3.5)
__device__ inline double table_interp_c(const double* tab, double h) {
  double fr;
  const int i = tabB_index(h, fr);
  const double t0 = tab[i], t1 = tab[i + 1];
  return t0 + fr * (t1 - t0);
}

// Inline evaluation of any mode for models whose shapes are all "simple"
// (nugget, spherical, exponential, Gaussian, bounded-square, linear): the
// same arithmetic as cov<DIM>, without the trigonometric and power shapes
// whose FP64 library code would inflate every caller's registers.
__device__ inline bool simple_shape(int t) { return t == NUG || t == SPH || t == EXP || t == GAU || t == BSQ || t == LIN; }

__device__ inline double corefunc_simple(int t, double r) {
  switch (t) {
    case SPH: return (r < 1.0) ? (1.0 - 1.5 * r) + 0.5 * (r * r * r) : 0.0;
    case EXP: return exp(-3.0 * r);
    case GAU: return exp(-3.0625 * (r * r));
    case BSQ: { double a = 1.0 - r * r; return (r < 1.0) ? a * a : 0.0; }
    case LIN: return (r < 1.0) ? 1.0 - r : 0.0;
    default: return 0.0;
  }
}

__device__ inline double comp_cov_h_simple(const Model& M, int c, double h) {
  const double eps = 2.2250738585072014e-308 * 1.0e3;
  return (h > eps) ? M.sill[c] * corefunc_simple(M.vtype[c], h) : M.sill[c] + M.nugget[c];
}

// struct_cov_h for models of simple shapes only, inline
__device__ inline double struct_cov_h_simple(const Model& M, double h) {
  double res = 0.0;
  int iv = 0;
  while (iv < M.nstruct) {
    double g = comp_cov_h_simple(M, iv, h);
    while (iv + 1 < M.nstruct && M.is_product[iv + 1]) {
      ++iv;
      g = g * comp_cov_h_simple(M, iv, h);
    }
    res = res + g;
    ++iv;
  }
  return res;
}

// analytic tail beyond the table (h >= 3.5): inline for models of simple
// shapes (the common exponential and Gaussian tails: a few structures, one
// exp each); otherwise out of line, as the general shape switch would
// inflate every caller's registers.  M.simple is kernel-uniform, so the
// choice is a scalar branch.
__device__ __noinline__ double struct_cov_h_tail(const Model* M, double h) { return struct_cov_h(*M, h); }
__device__ inline double tabB_tail(const Model& M, double h) {
  if (M.simple) return struct_cov_h_simple(M, h);
  return struct_cov_h_tail(&M, h);
}

// Level-B table evaluation: one anisotropy transform and one lookup
template <int DIM>
__device__ inline double cov_tableB(const Model& M, double dx, double dy, double dz) {
  const double h = aniso_h<DIM>(M.bmat, dx, dy, dz);
  if (h >= TAB_HB[3]) return M.b_tail ? tabB_tail(M, h) : 0.0;
  if (h <= 0.0) return M.tab[0];
  return table_interp_c(M.tab, h);
}

// Level-B lookup split into index, load and pick, for loops that evaluate
// many lags: the zone parameters are compile-time constants chosen by selects
// (no per-lane gathers), the index is kept inside the table for every h, so
// the two entry loads are unconditional and several lookups can be in flight.
// Lags beyond the table are reported, not evaluated: the caller re-evaluates
// them with cov_tableB.  In-table values are those of cov_tableB.
struct TabB {
  double m[9];
  int tail;
  const __attribute__((address_space(1))) double* tab;
};

__device__ inline TabB tabB_view(const Model& M) {
  TabB T;
#pragma unroll
  for (int k = 0; k < 9; ++k) T.m[k] = M.bmat[k];
  T.tail = M.b_tail;
  T.tab = (const __attribute__((address_space(1))) double*)M.tab;
  return T;
}

// C(h) from the two loaded entries.  At h = 0 the index is entry 0 with
// fraction 0, so the interpolation returns C(0) itself.  Beyond the table it
// returns 0, which is final when the model has no analytic tail; otherwise
// `out` is set and the caller evaluates the tail.  The select is marked
// unpredictable so that it stays a select: as a branch it would sink the
// entry loads into it and serialise consecutive lookups.
__device__ inline double tabB_pick(const TabB& T, double h, double fr, double e0, double e1, bool& out) {
  const bool beyond = h >= TAB_HB[3];
  out = beyond && T.tail;
  const double v = e0 + fr * (e1 - e0);
  return __builtin_unpredictable(beyond) ? 0.0 : v;
}

template <int DIM>
__device__ __noinline__ double cov_general(const Model* Mp, double dx, double dy, double dz);

template <int DIM>
__device__ inline double cov(const Model& M, double dx, double dy, double dz) {
  if (M.mode == COV_TABLE_B) return cov_tableB<DIM>(M, dx, dy, dz);
  double res = 0.0;
  int iv = 0;
  while (iv < M.nstruct) {
    double g = 1.0;
    bool first = true;
    do {
      if (!first) ++iv;
      double h = aniso_h<DIM>(M.mat[iv], dx, dy, dz);
      double c;
      if (M.mode == COV_TABLE_A) {
        const double* tab = M.tab + (size_t)iv * (TAB_N + 1);
        if (h >= TAB_HB[3]) c = has_tail(M.vtype[iv]) ? comp_cov_h(M, iv, h) : 0.0;
        else if (h <= 0.0) c = tab[0];
        else c = table_interp_c(tab, h);
      } else {
        c = comp_cov_h(M, iv, h);
      }
      g = first ? c : g * c;
      first = false;
    } while (iv + 1 < M.nstruct && M.is_product[iv + 1]);
    res = res + g;
    ++iv;
  }
  return res;
}

// one structure's own table at a dimensionless lag h (Level A tables), with
// its analytic tail beyond the table: comp_cov_tab_h, variogram.f90:353-382
__device__ inline double comp_tab_h(const Model& M, int c, double h) {
  const double* tab = M.tab + (size_t)c * (TAB_N + 1);
  if (h >= TAB_HB[3]) return has_tail(M.vtype[c]) ? comp_cov_h(M, c, h) : 0.0;
  if (h <= 0.0) return tab[0];
  return table_interp_c(tab, h);
}

// every mode (Level A tables, analytic), out of line
template <int DIM>
__device__ __noinline__ double cov_general(const Model* Mp, double dx, double dy, double dz) {
  return cov<DIM>(*Mp, dx, dy, dz);
}

template <int DIM>
__device__ inline double cov_simple(const Model& M, double dx, double dy, double dz) {
  if (M.mode == COV_TABLE_B) {
    const double h = aniso_h<DIM>(M.bmat, dx, dy, dz);
    if (h >= TAB_HB[3]) return M.b_tail ? struct_cov_h_simple(M, h) : 0.0;
    if (h <= 0.0) return M.tab[0];
    return table_interp_c(M.tab, h);
  }
  double res = 0.0;
  int iv = 0;
  while (iv < M.nstruct) {
    double g = 1.0;
    bool first = true;
    do {
      if (!first) ++iv;
      double h = aniso_h<DIM>(M.mat[iv], dx, dy, dz);
      double c;
      if (M.mode == COV_TABLE_A) {
        const double* tab = M.tab + (size_t)iv * (TAB_N + 1);
        if (h >= TAB_HB[3]) c = has_tail(M.vtype[iv]) ? comp_cov_h_simple(M, iv, h) : 0.0;
        else if (h <= 0.0) c = tab[0];
        else c = table_interp_c(tab, h);
      } else {
        c = comp_cov_h_simple(M, iv, h);
      }
      g = first ? c : g * c;
      first = false;
    } while (iv + 1 < M.nstruct && M.is_product[iv + 1]);
    res = res + g;
    ++iv;
  }
  return res;
}

// kernels specialise on the model mode: TB = Level-B table inline,
// otherwise the general path out of line
template <int DIM, bool TB>
__device__ inline double cov_eval(const Model& M, double dx, double dy, double dz) {
  if (TB) return cov_tableB<DIM>(M, dx, dy, dz);
  return cov_general<DIM>(&M, dx, dy, dz);
}

}  // namespace kk
