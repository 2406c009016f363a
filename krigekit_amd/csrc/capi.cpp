// C-ABI of the MI355X kriging backend: the krige_* entry points that
// krigekit's ctypes wrapper binds (src/krigekit/kriging.py:142-286), with the
// reference's handle registry and first-error conventions
// (src/libkriging/kriging_capi_common.F90:61-65,627-695;
//  src/libkriging/kriging_err.f90:15-118).
//
// Setup calls copy caller arrays (kriging_base.F90:600).  solve() builds the
// kd-tree (at set_search, as the reference does) and covariance tables on the
// host, keeps inputs resident in HBM across solves, runs search -> bucket ->
// per-node solve on the GPU and copies estimates and variances back.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/krigekit_amd.h"
#include "kk_device.hpp"
#include "kk_types.hpp"

using kk::Model;

namespace {

// ---------------------------------------------------------------------------
// first-error slot (kriging_err.f90): every entry clears it; the first error
// raised in a call chain is kept.
// ---------------------------------------------------------------------------
int g_ierr = 0;
std::string g_msg;

void clear_error() {
  g_ierr = 0;
  g_msg.clear();
}
void set_error(const char* where, const std::string& msg, long iblock = -1) {
  if (g_ierr) return;
  g_ierr = 1;
  g_msg = "\nKRIGING ERROR: Location: ";
  g_msg += where;
  if (iblock >= 0) g_msg += "; Block: " + std::to_string(iblock);
  g_msg += "\nDescription: " + msg;
}
bool failed() { return g_ierr != 0; }

#define HIPCHK(expr, where)                                                        \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(where, std::string("HIP error: ") + hipGetErrorString(e_));        \
      return false;                                                                \
    }                                                                              \
  } while (0)

// as HIPCHK, in functions that return the error status
#define HIPCHK_I(expr, where)                                                      \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(where, std::string("HIP error: ") + hipGetErrorString(e_));        \
      return ierr();                                                               \
    }                                                                              \
  } while (0)

// ---------------------------------------------------------------------------
// model: variogram.f90 structures, anisotropy (rotation.f90:51-95), cov0
// (variogram.f90:853-869), piecewise tables (variogram.f90:223-325,497-680)
// ---------------------------------------------------------------------------
struct VStruct {
  int vtype, is_product;
  double nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge;
};

const double kPI = 4.0 * std::atan(1.0);

void matmul3(const double* a, const double* b, double* c) {
  double t[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + a[i + 3 * k] * b[k + 3 * j];
      t[i + 3 * j] = s;
    }
  std::memcpy(c, t, sizeof t);
}

// calc_rotmat: diag(1/anis1, 1, 1/anis2) * Ry(plunge) * Rx(dip) * Rz(azimuth)
void calc_rotmat(double ang1, double ang2, double ang3, double anis1, double anis2, double* R) {
  const double d2r = kPI / 180.0, eps = 1.e-10;
  std::fill(R, R + 9, 0.0);
  R[0] = 1.0 / anis1;
  R[4] = 1.0;
  R[8] = 1.0 / anis2;
  if (std::fabs(ang3) > eps) {
    double a = -ang3 * d2r, s = std::sin(a), c = std::cos(a);
    double T[9] = {c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c};
    matmul3(R, T, R);
  }
  if (std::fabs(ang2) > eps) {
    double a = -ang2 * d2r, s = std::sin(a), c = std::cos(a);
    double T[9] = {1.0, 0.0, 0.0, 0.0, c, s, 0.0, -s, c};
    matmul3(R, T, R);
  }
  if (std::fabs(ang1) > eps) {
    double a = ang1 * d2r, s = std::sin(a), c = std::cos(a);
    double T[9] = {c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0};
    matmul3(R, T, R);
  }
}

// Derivatives of the anisotropy matrix M = calc_rotmat(...) / a_major with
// respect to (a_major, a_minor1, a_minor2, azimuth, dip, plunge).  Written as
// M = diag(1/a_minor1, 1/a_major, 1/a_minor2) * Ty(plunge) Tx(dip) Tz(azimuth)
// (the reference's product, rotation.f90:51-95, with the scaling folded in).
void aniso_dmat(const VStruct& v, double dm[6][9]) {
  const double d2r = kPI / 180.0;
  auto rot = [&](int axis, double ang, bool deriv, double* T) {
    // column-major rotation matrices of calc_rotmat and their derivatives
    // with respect to the angle in degrees
    const double a = (axis == 2 ? 1.0 : -1.0) * ang * d2r, da = (axis == 2 ? 1.0 : -1.0) * d2r;
    const double s = std::sin(a), c = std::cos(a);
    if (axis == 0) {  // Y (plunge)
      const double t[9] = {c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c};
      const double d[9] = {-s * da, 0.0, c * da, 0.0, 0.0, 0.0, -c * da, 0.0, -s * da};
      std::memcpy(T, deriv ? d : t, sizeof t);
    } else if (axis == 1) {  // X (dip)
      const double t[9] = {1.0, 0.0, 0.0, 0.0, c, s, 0.0, -s, c};
      const double d[9] = {0.0, 0.0, 0.0, 0.0, -s * da, c * da, 0.0, -c * da, -s * da};
      std::memcpy(T, deriv ? d : t, sizeof t);
    } else {  // Z (azimuth)
      const double t[9] = {c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0};
      const double d[9] = {-s * da, c * da, 0.0, -c * da, -s * da, 0.0, 0.0, 0.0, 0.0};
      std::memcpy(T, deriv ? d : t, sizeof t);
    }
  };
  const double ang[3] = {v.plunge, v.dip, v.azimuth};
  auto chain = [&](int dax, double* out) {  // Ty Tx Tz with factor dax differentiated
    double A[9], B[9], C[9], t[9];
    rot(0, ang[0], dax == 0, A);
    rot(1, ang[1], dax == 1, B);
    rot(2, ang[2], dax == 2, C);
    matmul3(A, B, t);
    matmul3(t, C, out);
  };
  const double sc[3] = {1.0 / v.a_minor1, 1.0 / v.a_major, 1.0 / v.a_minor2};
  auto scale_rows = [](const double* f, const double* T, double* out) {
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) out[r + 3 * c] = f[r] * T[r + 3 * c];
  };
  double T[9];
  chain(-1, T);
  const double fa[3] = {0.0, -sc[1] * sc[1], 0.0};
  const double f1[3] = {-sc[0] * sc[0], 0.0, 0.0};
  const double f2[3] = {0.0, 0.0, -sc[2] * sc[2]};
  scale_rows(fa, T, dm[0]);
  scale_rows(f1, T, dm[1]);
  scale_rows(f2, T, dm[2]);
  double dT[9];
  chain(2, dT);
  scale_rows(sc, dT, dm[3]);
  chain(1, dT);
  scale_rows(sc, dT, dm[4]);
  chain(0, dT);
  scale_rows(sc, dT, dm[5]);
}

double corefunc_host(int t, double r) {
  switch (t) {
    case kk::SPH: return (r < 1.0) ? (1.0 - 1.5 * r) + 0.5 * (r * r * r) : 0.0;
    case kk::EXP: return std::exp(-3.0 * r);
    case kk::GAU: return std::exp(-3.0625 * (r * r));
    case kk::HOL: return std::cos(kPI * r);
    case kk::POW: return (r < 1.0) ? 1.0 - std::pow(r, 1.5) : 0.0;
    case kk::BSQ: { double a = 1.0 - r * r; return (r < 1.0) ? a * a : 0.0; }
    case kk::CIR: return (r < 1.0) ? 1.0 - (2.0 * r * std::sqrt(1.0 - r * r) + 2.0 * std::asin(r)) / kPI : 0.0;
    case kk::LIN: return (r < 1.0) ? 1.0 - r : 0.0;
    case kk::CYC: { double s = std::sin(kPI * r); return std::exp(-2.0 * (s * s)); }
    case kk::DCO: return std::exp(-3.0 * r) * std::cos(kPI * r);
    default: return 0.0;
  }
}
bool has_tail_host(int t) { return t == kk::EXP || t == kk::GAU || t == kk::HOL || t == kk::CYC || t == kk::DCO; }

int vtype_from_str(const char* s) {
  static const char* names[] = {"nug", "sph", "exp", "gau", "hol", "pow", "bsq", "cir", "lin", "cyc", "dco"};
  for (int i = 0; i < 11; ++i)
    if (strncasecmp(s, names[i], 3) == 0) return i;
  return -1;
}

// Builds the kernel Model plus its table storage.  table_mode follows the
// reference's default build (USE_COV_TABLE); analytic is the alternative.
bool build_model(const std::vector<VStruct>& vs, int ndim, bool table_mode, Model& M,
                 std::vector<double>& tab) {
  std::memset(&M, 0, sizeof M);
  M.ndim = ndim;
  M.nstruct = (int)vs.size();
  if (M.nstruct < 1 || M.nstruct > kk::MAXS) return false;
  M.simple = 1;
  for (int i = 0; i < M.nstruct; ++i) {
    const int t = vs[i].vtype;
    M.simple &= (t == kk::NUG || t == kk::SPH || t == kk::EXP || t == kk::GAU || t == kk::BSQ || t == kk::LIN);
    M.vtype[i] = vs[i].vtype;
    M.is_product[i] = vs[i].is_product;
    M.sill[i] = vs[i].sill;
    M.nugget[i] = vs[i].nugget;
    double R[9];
    calc_rotmat(vs[i].azimuth, vs[i].dip, vs[i].plunge, vs[i].a_minor1 / vs[i].a_major,
                vs[i].a_minor2 / vs[i].a_major, R);
    for (int k = 0; k < 9; ++k) M.mat[i][k] = R[k] / vs[i].a_major;
  }
  double tot = 0.0;
  int iv = 0;
  while (iv < M.nstruct) {
    double g = vs[iv].sill + vs[iv].nugget;
    while (iv + 1 < M.nstruct && vs[iv + 1].is_product) {
      ++iv;
      g = g * (vs[iv].sill + vs[iv].nugget);
    }
    tot = tot + g;
    ++iv;
  }
  M.cov0 = tot;
  M.mode = kk::COV_ANALYTIC;
  tab.clear();
  if (!table_mode) return true;
  const double* hb = kk::TAB_HB;
  for (int k = 0; k < 3; ++k) {
    M.zn[k] = kk::tab_zn(k);
    M.zoff[k] = kk::tab_zoff(k);
    M.zinv[k] = kk::tab_zinv(k);
  }
  for (int k = 0; k < 4; ++k) M.zb[k] = hb[k];
  M.ntab = kk::TAB_N;
  // Level B if all non-nugget structures share one matrix (MAT_TOL 1e-6;
  // product members only need equal angles, ANG_TOL 1e-4)
  int ref = -1;
  bool levelB = true;
  for (int i = 0; i < M.nstruct; ++i) {
    if (vs[i].vtype == kk::NUG) continue;
    if (ref < 0) {
      ref = i;
      continue;
    }
    if (vs[i].is_product) {
      if (std::fabs(vs[i].azimuth - vs[ref].azimuth) > 1e-4 || std::fabs(vs[i].dip - vs[ref].dip) > 1e-4 ||
          std::fabs(vs[i].plunge - vs[ref].plunge) > 1e-4) {
        levelB = false;
        break;
      }
    } else {
      double mx = 0.0;
      for (int k = 0; k < 9; ++k) mx = std::max(mx, std::fabs(M.mat[i][k] - M.mat[ref][k]));
      if (mx > 1.0e-6) {
        levelB = false;
        break;
      }
    }
  }
  if (levelB && ref < 0) return true;  // all nuggets: no table is built, analytic path
  if (levelB) {
    M.mode = kk::COV_TABLE_B;
    std::memcpy(M.bmat, M.mat[ref], sizeof M.bmat);
    const double aref = vs[ref].a_major;
    tab.assign(M.ntab + 1, 0.0);
    tab[0] = M.cov0;
    for (int k = 0; k < 3; ++k) {
      const double hlo = hb[k], step = 1.0 / M.zinv[k];
      for (int i = 1; i <= M.zn[k]; ++i) {
        const double rdist = hlo + (double)i * step;
        double val = 0.0;
        int j = 0;
        while (j < M.nstruct) {
          if (vs[j].vtype == kk::NUG) {
            ++j;
            continue;
          }
          double rk = rdist * aref / vs[j].a_major;
          double g = vs[j].sill * corefunc_host(vs[j].vtype, rk);
          while (j + 1 < M.nstruct && vs[j + 1].is_product) {
            ++j;
            rk = rdist * aref / vs[j].a_major;
            g = g * vs[j].sill * corefunc_host(vs[j].vtype, rk);
          }
          val = val + g;
          ++j;
        }
        tab[M.zoff[k] + i] = val;
      }
    }
    int tail = 0, j = 0;
    while (j < M.nstruct) {
      if (has_tail_host(vs[j].vtype) && vs[j].vtype != kk::NUG) {
        int gt = has_tail_host(vs[j].vtype);
        while (j + 1 < M.nstruct && vs[j + 1].is_product) {
          ++j;
          gt = gt && has_tail_host(vs[j].vtype);
        }
        tail = tail || gt;
      }
      ++j;
    }
    M.b_tail = tail;
  } else {
    M.mode = kk::COV_TABLE_A;
    tab.assign((size_t)M.nstruct * (M.ntab + 1), 0.0);
    for (int c = 0; c < M.nstruct; ++c) {
      double* t = tab.data() + (size_t)c * (M.ntab + 1);
      t[0] = vs[c].sill + vs[c].nugget;
      for (int k = 0; k < 3; ++k) {
        const double hlo = hb[k], step = 1.0 / M.zinv[k];
        for (int i = 1; i <= M.zn[k]; ++i)
          t[M.zoff[k] + i] = vs[c].sill * corefunc_host(vs[c].vtype, hlo + (double)i * step);
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// kd-tree build on the host (kdtree2_maxidx.f90:629-998): mean split on the
// widest dimension, buckets of at most 13 points, boxes of the non-cut
// dimensions inherited from the parent, rearranged coordinate copy.
// Flattened into node records with parent links for the GPU traversal.
// ---------------------------------------------------------------------------
struct HostTree {
  int dim = 0, n = 0, root = -1;
  std::vector<int> ind;
  std::vector<double> rdata;
  std::vector<int> ni;     // NODE_I per node
  std::vector<double> nd;  // NODE_D per node
  int nnodes() const { return (int)(ni.size() / kk::NODE_I); }
};

struct TreeBuilder {
  HostTree& t;
  const double* data;
  int dim;
  double lo(int node, int k) const { return t.nd[(size_t)node * kk::NODE_D + kk::NODE_LO + k]; }
  double hi(int node, int k) const { return t.nd[(size_t)node * kk::NODE_D + kk::NODE_HI + k]; }
  void set_box(int node, int k, double l, double h) {
    t.nd[(size_t)node * kk::NODE_D + kk::NODE_LO + k] = l;
    t.nd[(size_t)node * kk::NODE_D + kk::NODE_HI + k] = h;
  }
  void spread(int c, int l, int u, double& mn, double& mx) const {
    mn = mx = data[c + dim * (size_t)t.ind[l]];
    for (int i = l + 1; i <= u; ++i) {
      double v = data[c + dim * (size_t)t.ind[i]];
      mn = std::min(mn, v);
      mx = std::max(mx, v);
    }
  }
  int add_node() {
    int id = t.nnodes();
    t.ni.resize(t.ni.size() + kk::NODE_I, -1);
    t.nd.resize(t.nd.size() + kk::NODE_D, 0.0);
    return id;
  }
  int build(int l, int u, int parent, int depth) {
    if (u < l) return -1;
    if (depth > 256) return -2;
    const int id = add_node();
    int* ni = &t.ni[(size_t)id * kk::NODE_I];
    ni[1] = l;
    ni[2] = u;
    ni[5] = parent;
    if (u - l <= 12) {
      for (int k = 0; k < dim; ++k) {
        double a, b;
        spread(k, l, u, a, b);
        set_box(id, k, a, b);
      }
      ni[0] = -1;
      return id;
    }
    const int pcut = parent >= 0 ? t.ni[(size_t)parent * kk::NODE_I] : -1;
    for (int k = 0; k < dim; ++k) {
      if (parent < 0 || k == pcut) {
        double a, b;
        spread(k, l, u, a, b);
        set_box(id, k, a, b);
      } else {
        set_box(id, k, lo(parent, k), hi(parent, k));
      }
    }
    int c = 0;
    double best = hi(id, 0) - lo(id, 0);
    for (int k = 1; k < dim; ++k)
      if (hi(id, k) - lo(id, k) > best) {
        best = hi(id, k) - lo(id, k);
        c = k;
      }
    double sum = 0.0;
    for (int j = l; j <= u; ++j) sum = sum + data[c + dim * (size_t)t.ind[j]];
    const double average = sum / (double)(u - l + 1);
    // partition: points with coordinate <= average to the left
    int lb = l, rb = u;
    while (lb < rb) {
      if (data[c + dim * (size_t)t.ind[lb]] <= average) {
        ++lb;
      } else {
        std::swap(t.ind[lb], t.ind[rb]);
        --rb;
      }
    }
    const int m = (data[c + dim * (size_t)t.ind[lb]] <= average) ? lb : lb - 1;
    t.ni[(size_t)id * kk::NODE_I + 0] = c;
    t.nd[(size_t)id * kk::NODE_D + 0] = average;
    const int left = build(l, m, id, depth + 1);
    if (left == -2) return -2;
    const int right = build(m + 1, u, id, depth + 1);
    if (right == -2) return -2;
    ni = &t.ni[(size_t)id * kk::NODE_I];
    double* nd = &t.nd[(size_t)id * kk::NODE_D];
    ni[3] = left;
    ni[4] = right;
    if (right < 0) {
      for (int k = 0; k < dim; ++k) set_box(id, k, lo(left, k), hi(left, k));
      nd[1] = hi(left, c);
      nd[0] = nd[1];
    } else if (left < 0) {
      for (int k = 0; k < dim; ++k) set_box(id, k, lo(right, k), hi(right, k));
      nd[2] = lo(right, c);
      nd[0] = nd[2];
    } else {
      nd[2] = lo(right, c);
      nd[1] = hi(left, c);
      nd[0] = (nd[1] + nd[2]) / 2;
      for (int k = 0; k < dim; ++k)
        set_box(id, k, std::min(lo(left, k), lo(right, k)), std::max(hi(left, k), hi(right, k)));
    }
    return id;
  }
};

bool build_tree(HostTree& t, int dim, int n, const double* data) {
  t = HostTree();
  t.dim = dim;
  t.n = n;
  t.ind.resize(n);
  for (int j = 0; j < n; ++j) t.ind[j] = j;
  t.ni.reserve((size_t)kk::NODE_I * (2 * n / 6 + 8));
  t.nd.reserve((size_t)kk::NODE_D * (2 * n / 6 + 8));
  TreeBuilder b{t, data, dim};
  t.root = b.build(0, n - 1, -1, 0);
  if (t.root < 0) return false;
  t.rdata.resize((size_t)dim * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < dim; ++k) t.rdata[k + dim * (size_t)i] = data[k + dim * (size_t)t.ind[i]];
  return true;
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  bool ensure(size_t nb) {
    if (nb <= bytes && p) return true;
    release();
    if (hipMalloc(&p, std::max<size_t>(nb, 16)) != hipSuccess) {
      p = nullptr;
      return false;
    }
    bytes = std::max<size_t>(nb, 16);
    return true;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// counters of one pipeline pass (device block and pinned host copy)
constexpr int PASS_COUNTERS = 128;

// ---------------------------------------------------------------------------
// kriging object
// ---------------------------------------------------------------------------
struct Obs {
  int n = 0;
  std::vector<double> coord, value, var;
  int nmax = INT_MAX;
  int nmax_in = INT_MAX;  // as given to set_obs (nmax is normalised at set_search)
  double maxdist2 = HUGE_VAL;
  double sk_mean = 0.0;
  bool sector = false, search_set = false, need_search = false, aniso = false;
  double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::vector<double> rcoord;  // search coordinates: rotated (aniso search), space-time t * time_at
  double time_at = 1.0;        // space-time search scaling of the time axis
  std::vector<double> drift;   // (ndrift, n): external drift values (zeros until set_obs_drift)
  HostTree tree;
  // increment observations z(x1) - z(x2) of this variable (set_grad):
  // coordinates (ndim, ngrad), values, variances, drift values (ndrift, ngrad)
  int ngrad = 0;
  std::vector<double> gx1, gx2, gval, gvar, gdrift;
};

// device copies of one variable's observations, tree and neighbour lists
// (multi-variable solves)
struct DevVar {
  DBuf ocoord, socoord, ovalue, ovar, rdata, ind, ni, nd, qcoord, nnear, inear, sflag, xlist, hdis, hidx, odrift;
  DBuf gx1, gx2, gval, gvar, gdrift;  // increment observations (set_grad)
  int mmax = 0;
};

// The weight store in the reference's layouts (kriging_base.F90:284-295,
// 2703-2746): order[nb], nnear[g + ng b], inear[k + nm (g + ng b)] (1-based,
// zero past nnear), weight[k + nm (g + ng (q + nv b))] (zero past nnear),
// var[i + nv (j + nv b)]; group g holds the neighbours of variable g.
struct WStore {
  bool alloc = false;
  int nm = 0, ng = 0, nv = 0, nb = 0;
  std::vector<int> order, nnear, inear;
  std::vector<double> weight, var;
  void init(int nm_, int ng_, int nv_, int nb_) {
    alloc = true;
    nm = nm_;
    ng = ng_;
    nv = nv_;
    nb = nb_;
    order.resize(nb);
    for (int b = 0; b < nb; ++b) order[b] = b + 1;
    nnear.assign((size_t)ng * nb, 0);
    inear.assign((size_t)nm * ng * nb, 0);
    weight.assign((size_t)nm * ng * nv * nb, 0.0);
    var.assign((size_t)nv * nv * nb, 0.0);
  }
  // any neighbour stored (t_weight_store%stored)
  bool stored() const {
    for (int n : nnear)
      if (n != 0) return true;
    return false;
  }
};

struct Krige {
  int ndim = 2, nvar = 1, ndrift = 0, unbias = 1, nsim = 0;
  // space-time (krige_st_*): coordinates are (x, y, z, t), ndim = 4 rows,
  // spatial structures in vgm, temporal ones in vgm_t; the product-sum
  // coefficient k_ps, or for the sum-metric model (the reference's default,
  // kriging_st.F90:52-56) the joint sills (one per spatial structure) and the
  // temporal transform f(dt)
  bool st = false;
  int st_model = 0;  // 0: sum_metric, 1: product_sum
  int t_vtype = 8;   // transform shape (linear)
  double t_nug = 0.0, t_sill = 1.0, t_at = 1.0;
  std::vector<double> sill_st;
  int mv_simple = 0;  // cokriging models use only inline-evaluated shapes
  double k_ps = 0.0, st_cov0 = 0.0;
  std::vector<VStruct> vgm_t;
  Model tmodel{};
  std::vector<double> ttab;
  bool aniso_search = false, weight_correction = false, use_old_weight = false, store_weight = false;
  bool cross_validation = false, write_mat = false, neglect_error = true, varying_vgm = false;
  bool std_ck = false, verbose = false, pf_cache = false;
  double bounds[2] = {-DBL_MAX, DBL_MAX};
  int seed = 0;
  bool initialized = false;
  std::vector<Obs> obs;
  std::vector<std::vector<VStruct>> vgm;  // nvar*nvar, (ivar-1)+nvar*(jvar-1)
  // blocks
  int nblock = 0;
  std::vector<double> bcoord;             // (ndim, nblock) centres
  std::vector<int> soff;                  // nblock+1 offsets into sub-points (empty: point kriging)
  std::vector<double> scoord, sweight;
  std::vector<double> rangescale, localnugget;
  // options
  int precision = kk::PREC_FP64;
  double pivot_tol = 1.0e-3;
  bool table_mode = true;
  // results
  bool solved = false;
  std::vector<double> est, var;
  int mmax = 0;
  // weight store (store_weight / set_weights / use_old_weight): host copy,
  // its device copy for the replay (ws_dev: up to date), the weights pass
  // output and flags, the observation values of a replay
  WStore ws;
  bool ws_dev = false;
  std::string weight_file;
  DBuf d_wnn, d_win, d_ww, d_wvar, d_wout, d_wflag;
  DBuf d_rz[kk::MAXVAR];
  int mmax_w = 0;  // neighbour stride of d_wout
  int stats[5] = {0, 0, 0, 0, 0};
  double timing[16] = {0};
  std::string info;
  // device residency
  bool dev_inputs_ok = false;
  bool model_ok = false;
  // d_nnear / d_inear hold the single-variable lists of the current
  // observations, blocks and search (cleared with dev_inputs_ok): the
  // leave-one-out gradient re-evaluates a new variogram on them
  bool lists_ok = false;
  int list_rows = 0;  // largest neighbour count in those lists
  // blocks without neighbours in the last single-variable pass (counted by
  // its bucketing; -1 unknown)
  long pass_empty = -1;
  DBuf d_ocoord, d_socoord, d_ovalue, d_ovar, d_rdata, d_ind, d_ni, d_nd, d_qcoord, d_bcoord;
  DBuf d_rs, d_ln, d_soff, d_scoord, d_sweight, d_basecov, d_tab;
  DBuf d_model, d_sflag, d_perm, d_xlist, d_gz, d_gdz, d_gred, d_gcount, d_tmodel, d_ttab, d_owork;
  // factor reuse over identical neighbour sets: -1 auto (on unless
  // leave-one-out, whose sets are all distinct), 0 off, 1 on; members per
  // work item
  int reuse_mode = -1;
  int item_max = kk::MAX_ITEM;
  int search_mode = 2;          // 0 exact, 1 per-lane fast, 2 wave-cooperative fast
  // sequential simulation (nsim > 0): the path (0-based block of each path
  // position), the standard normal draws (draw s of path position p at
  // s + nsim p) and the realisations (block b, realisation s at b + nblock s);
  // simlim >= 0 marks the inner pass of a simulation (its search sees the
  // points before each block on the path, simlim = the observation count)
  std::vector<int> sgs_path;
  std::vector<double> sgs_sample, est_sim;
  bool sim_set = false;
  int simlim = -1;
  DBuf d_sgs_val, d_sgs_sample, d_sgs_order;
  bool reorder = true;          // Morton order of queries for the wave-cooperative search
  bool chunked = true;          // with Morton order: two search chunks, the first one's solve overlapped
  bool perm_ok = false;
  DBuf d_nnear, d_inear, d_est, d_var, d_flag, d_hdis, d_hidx, d_counts, d_lists, d_fb;
  DBuf d_sets, d_items, d_table;  // factor reuse: per-block set data, work items, hash table
  // auto mode: a pass in which fewer than a quarter of the solved blocks
  // shared another's factorisation turns the grouping off for the next passes
  // over the same inputs (results are the same either way)
  bool reuse_low = false;
  bool reuse_on() const { return reuse_mode > 0 || (reuse_mode < 0 && !cross_validation && !reuse_low); }
  Model model{};
  std::vector<double> tab;
  // multi-variable path
  DevVar dv[kk::MAXVAR];
  DBuf d_models, d_mvtab, d_mvest, d_mvvar;
  std::vector<double> est_mv, var_mv;  // [nb * nvar], [nb * nvar * nvar]
  // external drift at the blocks, (ndrift, nvar, nb): target variable q's
  // right-hand side (zeros until set_grid_drift)
  std::vector<double> bdrift;
  DBuf d_bdrift;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // exact search of flagged queries, overlapped with the solve
  hipStream_t stream3 = nullptr;  // solve of the first search chunk, overlapped with the second
  int* h_counts = nullptr;        // pinned: bucket counts read while the GPU works on
  hipEvent_t ev[14] = {nullptr};
  // graph-captured pass (kk_graph_capture / kk_graph_launch): valid while the
  // inputs of the pass it was captured from are unchanged (solved stays true)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t gev[4] = {nullptr};
  const void* graph_key[3] = {nullptr, nullptr, nullptr};  // buffers the graph reads
  int* graph_overflow = nullptr;  // device flag: a count exceeded its captured bound
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  ~Krige() {
    drop_graph();
    for (auto& e : gev)
      if (e) (void)hipEventDestroy(e);
    if (h_counts) (void)hipHostFree(h_counts);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream3) (void)hipStreamDestroy(stream3);
  }
};

// The results of the last pass live in the multi-variable buffers: cokriging,
// and single-variable kriging with external drift or weight correction
// (both solved by the bordered solver); stored-weight replays keep the
// single-variable ones.
int grad_total(const Krige& K) {
  int n = 0;
  for (const Obs& O : K.obs) n += O.ngrad;
  return n;
}

bool mv_results(const Krige& K) {
  return K.nvar > 1 || (!K.use_old_weight && (K.ndrift > 0 || K.weight_correction || grad_total(K) > 0));
}

std::vector<Krige*> g_reg;

Krige* get_obj(int64_t h) {
  if (h == 0) {
    set_error("kriging_capi", "Null kriging object handle");
    return nullptr;
  }
  if (h < 1 || h > (int64_t)g_reg.size() || !g_reg[h - 1]) {
    set_error("kriging_capi", "Invalid kriging object handle");
    return nullptr;
  }
  return g_reg[h - 1];
}

int ierr() { return g_ierr; }

// normalise nmax like kriging_normalize_nmax (kriging_base.F90:1966-1971)
int normalize_nmax(int nmax, int ntot) {
  int v = std::min(nmax, ntot);
  return v <= 0 ? ntot : v;
}

// ---------------------------------------------------------------------------
// device pipeline
// ---------------------------------------------------------------------------
template <class T>
bool upload(DBuf& b, const std::vector<T>& v, const char* where) {
  if (!b.ensure(v.size() * sizeof(T))) {
    set_error(where, "device allocation failed");
    return false;
  }
  if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), where);
  return true;
}

bool ensure_stream(Krige& K) {
  if (!K.stream) {
    HIPCHK(hipStreamCreateWithFlags(&K.stream, hipStreamNonBlocking), "solve");
    HIPCHK(hipStreamCreateWithFlags(&K.stream2, hipStreamNonBlocking), "solve");
    HIPCHK(hipStreamCreateWithFlags(&K.stream3, hipStreamNonBlocking), "solve");
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&K.h_counts), sizeof(int) * PASS_COUNTERS, hipHostMallocDefault),
           "solve");
    for (auto& e : K.ev) HIPCHK(hipEventCreate(&e), "solve");
  }
  return true;
}

// Uploads everything solve() reads.  Inputs stay resident until a setter
// changes them.
bool stage_inputs(Krige& K) {
  const char* W = "t_kriging_base%solve_base";
  Obs& O = K.obs[0];
  const int nd = K.ndim;
  if (!K.model_ok) {
    if (!build_model(K.vgm[0], K.st ? 3 : nd, K.table_mode, K.model, K.tab)) {
      set_error(W, "invalid variogram model");
      return false;
    }
    if (!upload(K.d_tab, K.tab, W)) return false;
    K.model.tab = K.tab.empty() ? nullptr : K.d_tab.as<double>();
    if (!K.d_model.ensure(sizeof(Model))) {
      set_error(W, "device allocation failed");
      return false;
    }
    HIPCHK(hipMemcpy(K.d_model.p, &K.model, sizeof(Model), hipMemcpyHostToDevice), W);
    if (K.st) {
      // temporal structures: 1-D, isotropic range at_k; C(0, 0) of the
      // product-sum model (variogram_st.f90:344-375)
      if (!build_model(K.vgm_t, 1, K.table_mode, K.tmodel, K.ttab)) {
        set_error(W, "invalid temporal variogram model");
        return false;
      }
      if (!upload(K.d_ttab, K.ttab, W)) return false;
      K.tmodel.tab = K.ttab.empty() ? nullptr : K.d_ttab.as<double>();
      if (!K.d_tmodel.ensure(sizeof(Model))) {
        set_error(W, "device allocation failed");
        return false;
      }
      HIPCHK(hipMemcpy(K.d_tmodel.p, &K.tmodel, sizeof(Model), hipMemcpyHostToDevice), W);
      if (K.st_model == 1) {
        K.st_cov0 = K.k_ps * K.model.cov0 * K.tmodel.cov0 + K.model.cov0 + K.tmodel.cov0;
      } else {
        // sum-metric: Cs(0) + Ct(0) + the joint sills of the non-nugget
        // structures (a nugget shape is 0 at every lag)
        K.st_cov0 = K.model.cov0 + K.tmodel.cov0;
        for (size_t k = 0; k < K.sill_st.size(); ++k)
          if (K.vgm[0][k].vtype != kk::NUG) K.st_cov0 = K.st_cov0 + K.sill_st[k];
      }
    }
    K.model_ok = true;
  }
  if (K.dev_inputs_ok) return true;
  if (!upload(K.d_ocoord, O.coord, W) || !upload(K.d_ovalue, O.value, W) || !upload(K.d_ovar, O.var, W))
    return false;
  if (!upload(K.d_socoord, (O.aniso || K.st) ? O.rcoord : O.coord, W)) return false;
  if (O.need_search) {
    if (!upload(K.d_rdata, O.tree.rdata, W) || !upload(K.d_ind, O.tree.ind, W) ||
        !upload(K.d_ni, O.tree.ni, W) || !upload(K.d_nd, O.tree.nd, W))
      return false;
  }
  // block centres, and their rotated copy for the search
  const std::vector<double>& bc = K.cross_validation ? O.coord : K.bcoord;
  if (!upload(K.d_bcoord, bc, W)) return false;
  if (O.aniso || K.st) {
    // queries in search coordinates (rotated; space-time t * time_at)
    std::vector<double> q(bc.size());
    kk::Rot3 R;
    std::memcpy(R.m, O.rot, sizeof R.m);
    for (int b = 0; b < K.nblock; ++b) {
      if (nd == 1) R.apply<1>(&bc[(size_t)nd * b], &q[(size_t)nd * b]);
      else if (nd == 2) R.apply<2>(&bc[(size_t)nd * b], &q[(size_t)nd * b]);
      else R.apply<3>(&bc[(size_t)nd * b], &q[(size_t)nd * b]);
      if (K.st) q[(size_t)nd * b + 3] = bc[(size_t)nd * b + 3] * O.time_at;
    }
    if (!upload(K.d_qcoord, q, W)) return false;
  }
  if (!upload(K.d_rs, K.rangescale, W) || !upload(K.d_ln, K.localnugget, W)) return false;
  if (!K.soff.empty()) {
    if (!upload(K.d_soff, K.soff, W) || !upload(K.d_scoord, K.scoord, W) || !upload(K.d_sweight, K.sweight, W))
      return false;
  }
  K.dev_inputs_ok = true;
  K.reuse_low = false;
  return true;
}

// Morton (Z-order) permutation of the query points, so that the 64 queries
// of a wave are spatially compact; built once per grid on the host.
bool morton_perm(Krige& K) {
  if (K.perm_ok) return true;
  const char* W = "t_kriging_base%solve_base";
  const int nd = K.ndim, nb = K.nblock;
  // the queries (cross-validation: the observations) are resident in
  // d_bcoord; space-time queries are ordered in the search metric (t * time_at)
  const double tsc = K.st ? K.obs[0].time_at : 1.0;
  const size_t wb = kk::order_work_bytes(nb, nd);
  if (K.d_bcoord.bytes < sizeof(double) * (size_t)nd * nb || !K.d_perm.ensure(sizeof(int) * (size_t)nb) ||
      !K.d_owork.ensure(wb)) {
    set_error(W, "device allocation failed (query order)");
    return false;
  }
  HIPCHK(kk::launch_order(K.d_bcoord.as<double>(), nb, nd, tsc, K.d_perm.as<int>(), K.d_owork.p, wb, K.stream), W);
  K.perm_ok = true;
  return true;
}

// Counter block of a pass (device, and its pinned host copy): [0] and [8]
// block counts of the flag collections, [16] set members placed, [17]/[18]
// flagged work items/blocks of the grouped solves, [20]/[21] those of the
// FP64 re-solve, [24] blocks above 64 rows, [25] those sent to the fallback;
// per round 32 counters at ROUND_CTR[round] (bucket or
// set-grouping counts of the first search chunk, the rest, the exact-search
// blocks).
constexpr int ROUND_CTR[3] = {32, 64, 96};
// hash table slots for nsel blocks: a power of two of at least 2 nsel
size_t set_table_slots(int nsel) {
  size_t ts = 16;
  while (ts < 2 * (size_t)nsel) ts <<= 1;
  return ts;
}

// Device memory one pass of run_pipeline works in.  By default the handle's
// buffers, grown as needed; the allocation-free entry point passes views of
// its fixed scratch instead (every capacity checked before the pass).
struct PassBufs {
  int *nnear = nullptr, *inear = nullptr, *sflag = nullptr, *xlist = nullptr, *flag = nullptr;
  int *counts = nullptr;   // PASS_COUNTERS ints (layout at run_pipeline)
  int *lists = nullptr;
  // factor reuse: per-block set data, the hash table (tcap slots), work items
  // and blocks without neighbours (nb each)
  int *rep = nullptr, *rank = nullptr, *setcnt = nullptr, *setbase = nullptr, *setitem = nullptr,
      *members = nullptr;
  unsigned long long* table = nullptr;
  size_t tcap = 0;
  unsigned long long* hash = nullptr;  // per block
  int2* items = nullptr;
  int* empties = nullptr;
  double *est = nullptr, *var = nullptr;
  double* hdis = nullptr;  // exact-search heaps: hcap queries of k entries
  int* hidx = nullptr;
  int hcap = 0;
  double* fb = nullptr;    // fallback matrices: fbcap blocks of fallback_words(mmax)
  int fbcap = 0;
};

// neighbour-list stride and candidate count of a single-variable pass: with
// sector search 2^ndim nmax rows (space-time: octants of the spatial axes)
// from 4x as many candidates (kriging_base.F90:3005-3011, kriging.F90:623)
void pass_dims(const Krige& K, int& mmax, int& k) {
  const Obs& O = K.obs[0];
  const bool sector = O.sector && O.need_search;
  const int nsec = 1 << (K.st ? 3 : K.ndim);
  mmax = sector ? nsec * O.nmax : O.nmax;
  k = sector ? std::min(O.n, nsec * O.nmax * 4) : O.nmax;
}

// ---------------------------------------------------------------------------
// Per-bucket solve schedules.  A bucket's systems (at most BUCKET_ROWS[bk]
// rows) run on a k_solve instance of some width W >= that: G lanes (a
// quarter, half or whole wave) per system, 64 / G systems per wave, rows a
// quarter shorter for the 12/24/48-row instances.  The table is indexed by
// coordinate count, precision and covariance mode; an entry with width 0
// keeps the bucket's own width.  (The file format keeps a grid column for
// looping grids; looping k_solve over its list cost the instances a wave of
// occupancy in registers, so every solve runs one workgroup per 64 / G
// systems and the column is written as 0.)
// kk_autotune times candidate schedules on a solved handle's work lists and
// writes the winners to a text file; the library reads such a file at load
// (KRIGEKIT_AMD_TUNING, else tuning.txt beside the library).
// ---------------------------------------------------------------------------
struct TuneEntry {
  int width = 0;
  int grid = 0;
  float ms = 0.0f;
};
TuneEntry g_tune[5][2][2][kk::NBUCKET];

TuneEntry& tune_entry(int dim, int prec, bool tb, int bk) {
  return g_tune[dim < 1 || dim > 4 ? 0 : dim][prec ? 1 : 0][tb ? 1 : 0][bk];
}
// instance width for bucket bk: the tuned one when it holds the bucket's rows
int tuned_width(int dim, int prec, bool tb, int bk) {
  const int w = tune_entry(dim, prec, tb, bk).width;
  return w >= kk::bucket_width(bk) && w <= 64 ? w : kk::bucket_width(bk);
}

// lines "dim prec tb bucket width grid ms"; '#' starts a comment
int read_tuning(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return 0;
  int n = 0;
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int d, p, t, bk, w, g;
    float ms = 0.0f;
    if (std::sscanf(line, "%d %d %d %d %d %d %f", &d, &p, &t, &bk, &w, &g, &ms) >= 6 && d >= 1 && d <= 4 &&
        (p == 0 || p == 1) && (t == 0 || t == 1) && bk >= 1 && bk < kk::NBUCKET && w >= kk::bucket_width(bk) &&
        w <= 64 && kk::solve_width(w) == w && g >= 0) {
      TuneEntry& e = tune_entry(d, p, t, bk);
      e.width = w;
      e.grid = g;
      e.ms = ms;
      ++n;
    }
  }
  std::fclose(f);
  return n;
}

int write_tuning(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return 1;
  std::fprintf(f, "# krigekit_amd per-bucket solve schedules (kk_autotune): dim prec tb bucket width grid ms\n"
                  "# bucket b holds systems of at most 8/12/16/24/32/48/64 rows (b = 1..7); width = k_solve\n"
                  "# rows per system; grid = workgroups looping over the list (0 = one per 64/G systems)\n");
  for (int d = 1; d <= 4; ++d)
    for (int p = 0; p < 2; ++p)
      for (int t = 0; t < 2; ++t)
        for (int bk = 1; bk < kk::NBUCKET; ++bk) {
          const TuneEntry& e = tune_entry(d, p, t, bk);
          if (e.width > 0) std::fprintf(f, "%d %d %d %d %d %d %.4f\n", d, p, t, bk, e.width, e.grid, e.ms);
        }
  std::fclose(f);
  return 0;
}

// the tuning file read when the library is loaded
struct TuneLoader {
  TuneLoader() {
    std::string path;
    if (const char* env = std::getenv("KRIGEKIT_AMD_TUNING")) {
      path = env;
    } else {
      Dl_info info;
      if (dladdr(reinterpret_cast<void*>(&read_tuning), &info) && info.dli_fname) {
        path = info.dli_fname;
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning.txt";
      }
    }
    if (!path.empty()) read_tuning(path);
  }
} g_tune_loader;

// The device buffers and per-node solve arguments of one single-variable
// pass (the handle's buffers, grown as needed, or the views in ext).
bool setup_pass(Krige& K, const PassBufs* ext, int mmax, PassBufs& B, kk::SolveArgs& va, bool& tb) {
  const char* W = "t_kriging_base%solve_base";
  Obs& O = K.obs[0];
  const int nd = K.ndim, nb = K.nblock;
  if (ext) {
    B = *ext;
  } else {
    if (!K.d_nnear.ensure(sizeof(int) * nb) || !K.d_inear.ensure(sizeof(int) * (size_t)nb * mmax) ||
        !K.d_est.ensure(sizeof(double) * nb) || !K.d_var.ensure(sizeof(double) * nb) ||
        !K.d_flag.ensure(sizeof(int) * nb) || !K.d_counts.ensure(sizeof(int) * PASS_COUNTERS) ||
        !K.d_lists.ensure(sizeof(int) * (size_t)(kk::NBUCKET + 1) * nb * 2 + 64) ||
        !K.d_sflag.ensure(sizeof(int) * nb) || !K.d_xlist.ensure(sizeof(int) * nb)) {
      set_error(W, "device allocation failed");
      return false;
    }
    if (K.reuse_on()) {
      const size_t ts = set_table_slots(nb);
      if (!K.d_sets.ensure(sizeof(int) * (size_t)nb * 7) || !K.d_items.ensure(sizeof(int2) * (size_t)nb) ||
          !K.d_table.ensure(sizeof(unsigned long long) * (ts + (size_t)nb))) {
        set_error(W, "device allocation failed (neighbour sets)");
        return false;
      }
      int* sb = K.d_sets.as<int>();
      B.rep = sb;
      B.rank = sb + nb;
      B.setcnt = sb + 2 * (size_t)nb;
      B.setbase = sb + 3 * (size_t)nb;
      B.members = sb + 4 * (size_t)nb;
      B.empties = sb + 5 * (size_t)nb;
      B.setitem = sb + 6 * (size_t)nb;
      B.items = K.d_items.as<int2>();
      B.table = K.d_table.as<unsigned long long>();
      B.tcap = ts;
      B.hash = B.table + ts;
    }
    B.nnear = K.d_nnear.as<int>();
    B.inear = K.d_inear.as<int>();
    B.est = K.d_est.as<double>();
    B.var = K.d_var.as<double>();
    B.flag = K.d_flag.as<int>();
    B.counts = K.d_counts.as<int>();
    B.lists = K.d_lists.as<int>();
    B.sflag = K.d_sflag.as<int>();
    B.xlist = K.d_xlist.as<int>();
  }
  // ---- per-node solve arguments ------------------------------------------------
  va = kk::SolveArgs{};
  va.model = K.d_model.as<Model>();
  va.ndim = nd;
  va.unbias = K.unbias;
  va.sk_mean = O.sk_mean;
  va.bound_lo = K.bounds[0];
  va.bound_hi = K.bounds[1];
  va.nsub = 1;
  va.bcoord = K.d_bcoord.as<double>();
  // the space-time assembly does not range-scale lags (kriging_st.F90:805)
  va.rangescale = (K.rangescale.empty() || K.st) ? nullptr : K.d_rs.as<double>();
  va.localnugget = K.localnugget.empty() ? nullptr : K.d_ln.as<double>();
  if (K.st) {
    va.tmodel = K.d_tmodel.as<Model>();
    va.k_ps = K.k_ps;
    va.st_cov0 = K.st_cov0;
    va.st_sum = K.st_model == 0;
    va.t_vtype = K.t_vtype;
    va.t_nug = K.t_nug;
    va.t_sill = K.t_sill;
    va.t_at = K.t_at;
    for (size_t k = 0; k < K.sill_st.size() && k < (size_t)kk::MAXS; ++k) va.sill_st[k] = K.sill_st[k];
  }
  va.ocoord = K.d_ocoord.as<double>();
  va.ovalue = K.d_ovalue.as<double>();
  va.ovar = K.d_ovar.as<double>();
  va.nnear = B.nnear;
  va.inear = B.inear;
  va.mmax = mmax;
  va.pivot_tol = K.pivot_tol;
  va.est = B.est;
  va.var = B.var;
  va.flag = B.flag;
  if (!K.soff.empty()) {
    // uniform sub-point count per block (Gauss-Legendre blocks)
    va.nsub = K.soff[1] - K.soff[0];
    va.scoord = K.d_scoord.as<double>();
    va.sweight = K.d_sweight.as<double>();
    va.base_cov = K.d_basecov.as<double>();
  }
  // the inline Level-B kernels need every model on one composite table: for
  // space-time both the spatial and the temporal model, and the product-sum
  // (the sum-metric joint term is evaluated out of line)
  tb = K.model.mode == kk::COV_TABLE_B && (!K.st || (K.tmodel.mode == kk::COV_TABLE_B && K.st_model == 1));
  return true;
}

bool run_pipeline(Krige& K, const PassBufs* ext = nullptr) {
  const char* W = "t_kriging_base%solve_base";
  if (!ext && !ensure_stream(K)) return false;
  hipStream_t st = K.stream;
  Obs& O = K.obs[0];
  const int nd = K.ndim, nb = K.nblock;
  int nmax = O.nmax;  // already normalised at set_search
  const bool sector = O.sector && O.need_search;
  // octants are over the spatial axes only: 2^3 for space-time
  // (kriging_st.F90:587, obs_nmax in kriging_base.F90:3005-3011)
  int mmax, kcand;
  pass_dims(K, mmax, kcand);
  K.mmax = mmax;
  // systems above 64 rows go to the one-workgroup bordered solver, which
  // holds up to 128 rows and extra columns (not for space-time)
  const int big_rows_max = 128 - kk::mv_extra_cols(1, 0, K.unbias);
  const int nsub_pass = K.soff.empty() ? 1 : K.soff[1] - K.soff[0];
  if (mmax > 64 && (K.st || mmax > big_rows_max || nsub_pass > 64)) {
    set_error(W, "neighbourhoods larger than " + std::to_string(K.st ? 64 : big_rows_max) +
                     " data rows are not supported by the GPU solver (mmax=" + std::to_string(mmax) + ")");
    return false;
  }
  PassBufs B;
  kk::SolveArgs va;
  bool tb = false;
  if (!setup_pass(K, ext, mmax, B, va, tb)) return false;
  // A round groups (factor reuse) or buckets the blocks sel[0..nsel) (all
  // when sel is null) by neighbour count on stream s, copies its counters
  // (round area ro of the counter block) to pinned host memory between
  // events e0 and e1, then -- once the host has them -- launches one solve
  // per bucket.  Work lists: plain block lists of capacity cap per bucket at
  // lists, or work items and empty blocks at items/empties.
  const bool reuse = B.items != nullptr;
  float bucket_ms = 0.0f, bucket_ms_first = 0.0f;  // all bucketing; the first chunk's (inside "search")
  int nempty = 0;
  long n_items = 0, n1_items = 0, n1_blocks = 0;  // work items, those and blocks of single-neighbour sets
  struct Round {
    const int* sel;
    int nsel;
    int ro;
    const int* skip;
    int* lists;
    int cap;
    int2* items;
    int* empties;
  };
  auto prepare = [&](hipStream_t s, const Round& rd, hipEvent_t e0, hipEvent_t e1) -> bool {
    HIPCHK(hipEventRecord(e0, s), W);
    if (reuse) {
      kk::SetGroupArgs g{};
      g.nnear = B.nnear;
      g.inear = B.inear;
      g.mmax = mmax;
      g.rs = va.rangescale;
      g.ln = va.localnugget;
      g.sel = rd.sel;
      g.nsel = rd.nsel;
      g.skip = rd.skip;
      g.table = B.table;
      g.hash = B.hash;
      g.tsize = std::min(B.tcap, set_table_slots(rd.nsel));
      g.item_max = K.item_max;
      g.rep = B.rep;
      g.rank = B.rank;
      g.cnt = B.setcnt;
      g.setbase = B.setbase;
      g.setitem = B.setitem;
      g.members = B.members;
      g.ctr = B.counts + rd.ro;
      g.total = B.counts + 16;
      g.items = rd.items;
      g.empties = rd.empties;
      HIPCHK(hipMemsetAsync(g.ctr, 0, sizeof(int) * kk::SET_NCTR, s), W);
      HIPCHK(kk::launch_set_group(g, s), W);
      HIPCHK(hipMemcpyAsync(K.h_counts + rd.ro, g.ctr, sizeof(int) * kk::SET_NCTR, hipMemcpyDeviceToHost, s), W);
    } else {
      HIPCHK(hipMemsetAsync(B.counts + rd.ro, 0, sizeof(int) * 16, s), W);
      HIPCHK(kk::launch_bucket(B.nnear, rd.sel, rd.nsel, B.counts + rd.ro, rd.lists, rd.cap, s, nullptr, rd.skip),
             W);
      HIPCHK(hipMemcpyAsync(K.h_counts + rd.ro, B.counts + rd.ro, sizeof(int) * 16, hipMemcpyDeviceToHost, s), W);
    }
    HIPCHK(hipEventRecord(e1, s), W);
    return true;
  };
  auto solve_round = [&](hipStream_t s, const Round& rd) -> bool {
    // (blocks above 64 rows, in the overflow bucket or left out of the sets,
    // are solved after the rounds)
    const int* c = K.h_counts + rd.ro;
    nempty += c[0];
    kk::SolveArgs v = va;
    v.ctr = B.counts + 17;
    v.list = reuse ? rd.empties : rd.lists;
    v.nlist = c[0];
    HIPCHK(kk::launch_empty(v, s), W);
    if (reuse) {
      v.setcnt = B.setcnt;
      v.setbase = B.setbase;
      v.members = B.members;
      v.item_max = K.item_max;
      for (int bk = 1; bk < kk::NBUCKET; ++bk) n_items += c[bk];
      n1_items += c[kk::SET_N1_ITEMS];
      n1_blocks += c[kk::SET_N1_BLOCKS];
    }
    for (int bk = 1; bk < kk::NBUCKET; ++bk) {
      if (reuse) v.items = rd.items + c[kk::SET_OFF + bk];
      else v.list = rd.lists + (size_t)bk * rd.cap;
      v.nlist = c[bk];
      HIPCHK(kk::launch_solve(v, tuned_width(nd, K.precision, tb, bk), K.precision, tb, s), W);
    }
    return true;
  };
  auto prepare_solve = [&](hipStream_t s, const Round& rd) -> bool {
    if (!prepare(s, rd, K.ev[5], K.ev[6])) return false;
    HIPCHK(hipEventSynchronize(K.ev[6]), W);
    float bms = 0.0f;
    HIPCHK(hipEventElapsedTime(&bms, K.ev[5], K.ev[6]), W);
    bucket_ms += bms;
    return solve_round(s, rd);
  };
  HIPCHK(hipMemsetAsync(B.counts, 0, sizeof(int) * PASS_COUNTERS, st), W);
  HIPCHK(hipEventRecord(K.ev[0], st), W);
  bool overlap = false;  // exact search of flagged queries runs on stream2
  bool exact_used = false;  // the fast search ran and flagged queries for the exact kernel
  int nx = 0;            // blocks solved in the second round
  // with a Morton order and enough blocks the search runs in two chunks, and
  // the first chunk's blocks are bucketed and solved on stream3 while the
  // second chunk is searched
  int split = 0;
  Round first{};
  // ---- search -----------------------------------------------------------
  kk::SearchArgs sa{};
  sa.ndim = nd;
  sa.qcoord = (O.aniso || K.st) ? K.d_qcoord.as<double>() : K.d_bcoord.as<double>();
  sa.ocoord = K.d_socoord.as<double>();
  sa.nmax = nmax;
  sa.cv = K.cross_validation;
  sa.maxdist2 = O.maxdist2;
  sa.mmax = mmax;
  sa.nnear = B.nnear;
  sa.inear = B.inear;
  sa.simlim = K.simlim;
  if (!O.need_search) {
    // rotated_dists rotates (obs - query) of the unrotated coordinates;
    // space-time adds the scaled time difference (kriging_st.F90:609-612)
    kk::Rot3 R;
    std::memcpy(R.m, O.rot, sizeof R.m);
    sa.nq = nb;
    sa.q0 = 0;
    if (K.st && O.aniso) {
      set_error(W, "space-time search anisotropy with nmax >= nobs is not supported by this backend");
      return false;
    }
    sa.qcoord = K.st ? K.d_qcoord.as<double>() : K.d_bcoord.as<double>();
    HIPCHK(kk::launch_search_all(sa, O.n, K.st ? K.d_socoord.as<double>() : K.d_ocoord.as<double>(), R, st), W);
  } else {
    sa.tree.dim = nd;
    sa.tree.n = O.n;
    sa.tree.nnodes = O.tree.nnodes();
    sa.tree.root = O.tree.root;
    sa.tree.rdata = K.d_rdata.as<double>();
    sa.tree.ind = K.d_ind.as<int>();
    sa.tree.ni = K.d_ni.as<int>();
    sa.tree.nd = K.d_nd.as<double>();
    sa.sort = sector;
    sa.k = kcand;
    const size_t lds = kk::fast_search_lds(nd, sector, sa.k, nmax);
    int nexact = nb;
    const int* exact_list = nullptr;
    exact_used = false;
    // sectored space-time searches run the exact kernel for every query
    if (K.search_mode > 0 && lds <= 64 * 1024 && !(K.st && sector)) {
      sa.qflag = B.sflag;
      sa.q0 = 0;
      sa.nq = nb;
      HIPCHK(hipEventRecord(K.ev[8], st), W);
      if (K.search_mode == 2) {
        if (K.reorder && !morton_perm(K)) return false;
        const int* perm = K.reorder ? K.d_perm.as<int>() : nullptr;
        // first chunk 60% of the blocks (measured: 11.15 ms at 60/40, about
        // 11.35 ms at 50/50 and 70/30 on the 3-D workload)
        // With factor reuse the search runs in one launch: the set grouping
        // of a first chunk, beside the second search, delayed that chunk's
        // solves more than the overlap saved (measured at 4.2M nodes: 11.4
        // ms per pass chunked against 11.1 ms in one launch)
        split = (perm && K.chunked && !reuse && nb >= (1 << 17)) ? (int)((size_t)nb * 3 / 5 + 63) / 64 * 64 : 0;
        sa.qlist = perm;
        if (split) {
          sa.nq = split;
          HIPCHK(kk::launch_search_wave(sa, st), W);
          // the first chunk's decided blocks (flagged ones carry a negative
          // count) are bucketed right behind its search, into the second half
          // of the list area: queued behind the next search they would wait
          // for its waves to retire
          first = Round{perm, split, ROUND_CTR[0], B.sflag, B.lists + (size_t)(kk::NBUCKET + 1) * nb, split,
                        B.items, B.empties};
          if (reuse) {
            // the set grouping takes longer than plain bucketing: it runs on
            // stream3 beside the second search instead of delaying it
            HIPCHK(hipEventRecord(K.ev[7], st), W);
            HIPCHK(hipStreamWaitEvent(K.stream3, K.ev[7], 0), W);
          }
          if (!prepare(reuse ? K.stream3 : st, first, K.ev[12], K.ev[13])) return false;
          sa.qlist = perm + split;
          sa.nq = nb - split;
        }
        HIPCHK(kk::launch_search_wave(sa, st), W);
        sa.qlist = nullptr;
        sa.nq = nb;
      } else {
        HIPCHK(kk::launch_search_fast(sa, st), W);
      }
      HIPCHK(hipEventRecord(K.ev[9], st), W);
      if (split) {
        // the first chunk's solves run on stream3 beside the second search
        HIPCHK(hipEventSynchronize(K.ev[13]), W);
        HIPCHK(hipEventElapsedTime(&bucket_ms_first, K.ev[12], K.ev[13]), W);
        bucket_ms += bucket_ms_first;
        HIPCHK(hipStreamWaitEvent(K.stream3, K.ev[13], 0), W);
        if (!solve_round(K.stream3, first)) return false;
        HIPCHK(hipEventRecord(K.ev[12], K.stream3), W);
      }
      HIPCHK(hipMemsetAsync(B.counts, 0, sizeof(int) * 16, st), W);
      // in Morton order when the search used it: the exact kernel's lanes
      // then take neighbouring queries and follow similar tree paths
      HIPCHK(kk::launch_collect(B.sflag, nb, 1, B.counts + 8, B.xlist, st,
                                K.search_mode == 2 && K.reorder ? K.d_perm.as<int>() : nullptr), W);
      HIPCHK(hipMemcpyAsync(&nexact, B.counts + 8, sizeof(int), hipMemcpyDeviceToHost, st), W);
      HIPCHK(hipStreamSynchronize(st), W);
      exact_list = B.xlist;
      exact_used = true;
    }
    K.timing[7] = exact_list ? nexact : -1;
    K.timing[8] = exact_list ? (split ? 2 : 1) : 0;  // fast-search launches between ev[8] and ev[9]
    // the queries the fast kernel left undecided run on the second stream,
    // overlapped with the solve of every other block
    overlap = exact_list != nullptr && nexact > 0;
    hipStream_t xs = overlap ? K.stream2 : st;
    if (overlap) HIPCHK(hipStreamWaitEvent(xs, K.ev[9], 0), W);
    HIPCHK(hipEventRecord(K.ev[10], xs), W);
    if (nexact > 0) {
      const int chunk = std::min(nexact, ext ? B.hcap : 1 << 20);
      if (ext) {
        sa.hdis = B.hdis;
        sa.hidx = B.hidx;
      } else {
        if (!K.d_hdis.ensure(sizeof(double) * (size_t)sa.k * chunk) ||
            !K.d_hidx.ensure(sizeof(int) * (size_t)sa.k * chunk)) {
          set_error(W, "device allocation failed (search scratch)");
          return false;
        }
        sa.hdis = K.d_hdis.as<double>();
        sa.hidx = K.d_hidx.as<int>();
      }
      sa.hstride = chunk;
      for (int q0 = 0; q0 < nexact; q0 += chunk) {
        sa.q0 = exact_list ? 0 : q0;
        sa.qlist = exact_list ? exact_list + q0 : nullptr;
        sa.nq = std::min(chunk, nexact - q0);
        HIPCHK(kk::launch_search(sa, xs), W);
      }
    }
    HIPCHK(hipEventRecord(K.ev[11], xs), W);
    nx = overlap ? nexact : 0;
    // simulation pass: drop simulated neighbours that duplicate an
    // observation of the same row (the exact search ran on st)
    if (K.simlim >= 0)
      HIPCHK(kk::launch_sgsim_lists(B.nnear, B.inear, mmax, nb, K.simlim, nd, K.d_ocoord.as<double>(), st), W);
  }
  HIPCHK(hipEventRecord(K.ev[1], st), W);
  HIPCHK(hipEventRecord(K.ev[2], st), W);
  // ---- bucket by neighbour count, solve per bucket ---------------------------
  // round 0: every block the search has decided (the second chunk's, when
  // the first was solved on stream3); round 1: the blocks the exact search
  // decided on the second stream
  // (in the first two, blocks left to the exact search are skipped by their
  // search flag: their lists may still be in flight)
  const int* skip = exact_used ? B.sflag : nullptr;
  for (int round = 0; round < (nx > 0 ? 2 : 1); ++round) {
    if (round == 1) HIPCHK(hipStreamWaitEvent(st, K.ev[11], 0), W);
    Round rd{};
    rd.sel = round ? B.xlist : (split ? K.d_perm.as<int>() + split : nullptr);
    rd.nsel = round ? nx : nb - split;
    rd.ro = ROUND_CTR[1 + round];
    rd.skip = round ? nullptr : skip;
    rd.lists = B.lists;
    rd.cap = nb;
    // the exact-search round starts its work lists at 0: the first chunk's
    // solves on stream3, which read that area, have finished by then
    const size_t lo = round ? 0 : (size_t)split;
    if (round == 1 && split) HIPCHK(hipStreamWaitEvent(st, K.ev[12], 0), W);
    rd.items = reuse ? B.items + lo : nullptr;
    rd.empties = reuse ? B.empties + lo : nullptr;
    if (!prepare_solve(st, rd)) return false;
  }
  if (split) HIPCHK(hipStreamWaitEvent(st, K.ev[12], 0), W);
  // ---- systems above 64 rows: one workgroup each (FP64) --------------------
  int nbig = 0;
  // the one-workgroup bordered solver on a list of blocks (nvar 1)
  auto single_mv = [&](const int* list, int nlist) {
    kk::MVArgs m{};
    m.models = K.d_model.as<Model>();
    m.ndim = nd;
    m.nvar = 1;
    m.nb = nb;
    m.std_ck = 0;
    m.unbias = K.unbias;
    m.nsub = va.nsub;
    m.rows = mmax;
    m.simple = 1;
    for (int i = 0; i < K.model.nstruct; ++i) {
      const int t = K.model.vtype[i];
      m.simple &= (t == kk::NUG || t == kk::SPH || t == kk::EXP || t == kk::GAU || t == kk::BSQ || t == kk::LIN);
    }
    m.sk_mean[0] = O.sk_mean;
    m.bound_lo = K.bounds[0];
    m.bound_hi = K.bounds[1];
    m.bcoord = va.bcoord;
    m.scoord = va.nsub > 1 ? va.scoord : nullptr;
    m.sweight = va.nsub > 1 ? va.sweight : nullptr;
    m.rangescale = va.rangescale;
    m.localnugget = va.localnugget;
    m.ocoord[0] = va.ocoord;
    m.ovalue[0] = va.ovalue;
    m.ovar[0] = va.ovar;
    m.nnear[0] = B.nnear;
    m.inear[0] = B.inear;
    m.mmax[0] = mmax;
    m.list = list;
    m.nlist = nlist;
    m.est = B.est;
    m.var = B.var;
    m.flag = B.flag;
    return m;
  };
  if (mmax > 64) {
    HIPCHK(hipMemsetAsync(B.counts + 24, 0, sizeof(int) * 2, st), W);
    HIPCHK(kk::launch_collect_above(B.nnear, nb, 64, B.counts + 24, B.xlist, st), W);
    HIPCHK(hipMemcpyAsync(&nbig, B.counts + 24, sizeof(int), hipMemcpyDeviceToHost, st), W);
    HIPCHK(hipStreamSynchronize(st), W);
    if (nbig > 0) {
      kk::MVArgs m = single_mv(B.xlist, nbig);
      m.ctr = B.counts + 25;
      HIPCHK(kk::launch_solve_mv(m, st), W);
    }
  }
  HIPCHK(hipEventRecord(K.ev[3], st), W);
  // ---- FP64 re-solve of flagged FP32 systems, Bunch-Kaufman fallback -------
  int nref = 0, nfb = 0;
  if (K.precision == kk::PREC_FP32R) {
    HIPCHK(hipMemsetAsync(B.counts, 0, sizeof(int) * 16, st), W);
    HIPCHK(kk::launch_collect(B.flag, nb, 1, B.counts, B.lists, st), W);
    HIPCHK(hipMemcpyAsync(&nref, B.counts, sizeof(int), hipMemcpyDeviceToHost, st), W);
    HIPCHK(hipStreamSynchronize(st), W);
    if (nref > 0) {
      va.list = B.lists;
      va.nlist = nref;
      va.ctr = B.counts + 20;
      HIPCHK(kk::launch_solve(va, kk::solve_width(mmax), kk::PREC_FP64, tb, st), W);
    }
  }
  HIPCHK(hipMemsetAsync(B.counts, 0, sizeof(int) * 16, st), W);
  HIPCHK(kk::launch_collect(B.flag, nb, 2, B.counts, B.lists, st), W);
  HIPCHK(hipMemcpyAsync(&nfb, B.counts, sizeof(int), hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  if (nfb > 0) {
    // at most 256 MiB of fallback matrices per launch
    const int fchunk = ext ? B.fbcap
                           : (int)std::max<size_t>(1, std::min<size_t>(4096, (size_t(256) << 20) /
                                                                             (8 * kk::fallback_words(mmax))));
    double* fb = B.fb;
    if (!ext) {
      if (!K.d_fb.ensure(sizeof(double) * fchunk * kk::fallback_words(mmax))) {
        set_error(W, "device allocation failed (fallback scratch)");
        return false;
      }
      fb = K.d_fb.as<double>();
    }
    for (int f0 = 0; f0 < nfb; f0 += fchunk) {
      va.list = B.lists + f0;
      va.nlist = std::min(fchunk, nfb - f0);
      HIPCHK(kk::launch_fallback(va, fb, st), W);
    }
  }
  HIPCHK(hipEventRecord(K.ev[4], st), W);
  HIPCHK(hipMemcpyAsync(K.h_counts + 16, B.counts + 16, sizeof(int) * 16, hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  float ms;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[0], K.ev[1]), W);
  K.timing[0] = ms;
  K.timing[1] = bucket_ms;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[2], K.ev[3]), W);
  K.timing[2] = ms - (bucket_ms - bucket_ms_first);
  HIPCHK(hipEventElapsedTime(&ms, K.ev[3], K.ev[4]), W);
  K.timing[3] = ms;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[0], K.ev[4]), W);
  K.timing[4] = ms;
  K.timing[9] = K.timing[10] = 0.0;
  if (K.timing[8] > 0.0) {
    HIPCHK(hipEventElapsedTime(&ms, K.ev[8], K.ev[9]), W);
    // plain bucketing of the first chunk runs between the two launches (the
    // set grouping runs on stream3)
    K.timing[9] = reuse ? ms : ms - bucket_ms_first;
  }
  if (O.need_search) {
    HIPCHK(hipEventElapsedTime(&ms, K.ev[10], K.ev[11]), W);
    K.timing[10] = ms;
  }
  // counters with the reference's meaning (kriging_base.F90:3260-3330):
  // Cholesky factorisations, blocks solved from another block's factorisation,
  // blocks of the symmetric-indefinite fallback (failures are split off when
  // the results are fetched)
  const int* hc = K.h_counts;
  K.stats[0] = 0;
  if (reuse) {
    // blocks above 64 rows are factored each on their own
    const long items_ok = (n_items - n1_items) - hc[17];
    const long blocks_ok = ((long)hc[16] - n1_blocks) - hc[18];
    K.stats[1] = (int)(items_ok + (nref - hc[20]) + (nbig - hc[25]));
    K.stats[2] = (int)(blocks_ok - items_ok);
    if (K.reuse_mode < 0 && 4L * K.stats[2] < (long)K.stats[1] + K.stats[2]) K.reuse_low = true;
  } else {
    K.stats[1] = nb - nempty - nfb;
    K.stats[2] = 0;
  }
  K.pass_empty = nempty;
  K.stats[3] = nfb;
  K.stats[4] = 0;
  K.timing[5] = nref;
  K.timing[6] = nfb;
  K.timing[12] = 0.0;
  if (K.store_weight && !ext) {
    // ---- the weight store: explicit weights of every block with two or
    // more neighbours, FP64, in the bordered solver (a block whose K is not
    // positive definite: the Bunch-Kaufman fallback, as in the solve); for
    // space-time and blocks of more than 64 integration nodes the
    // Bunch-Kaufman path of each block
    K.mmax_w = mmax;
    if (!K.d_wout.ensure(sizeof(double) * (size_t)nb * mmax) || !K.d_wflag.ensure(sizeof(int) * nb)) {
      set_error(W, "device allocation failed (weight store)");
      return false;
    }
    HIPCHK(hipEventRecord(K.ev[5], st), W);
    int nw = 0, nwf = 0;
    HIPCHK(hipMemsetAsync(B.counts, 0, sizeof(int) * 16, st), W);
    HIPCHK(kk::launch_collect_above(B.nnear, nb, 1, B.counts, B.xlist, st), W);
    HIPCHK(hipMemcpyAsync(&nw, B.counts, sizeof(int), hipMemcpyDeviceToHost, st), W);
    HIPCHK(hipStreamSynchronize(st), W);
    const bool mv_ok = !K.st && va.nsub <= 64;
    if (nw > 0 && mv_ok) {
      HIPCHK(hipMemsetAsync(K.d_wflag.p, 0, sizeof(int) * nb, st), W);
      kk::MVArgs m = single_mv(B.xlist, nw);
      m.weights = K.d_wout.as<double>();
      m.mmax_w = mmax;
      m.weights_only = 1;
      m.flag = K.d_wflag.as<int>();
      HIPCHK(kk::launch_solve_mv(m, st), W);
      HIPCHK(hipMemsetAsync(B.counts, 0, sizeof(int) * 16, st), W);
      HIPCHK(kk::launch_collect(K.d_wflag.as<int>(), nb, 2, B.counts, B.lists, st), W);
      HIPCHK(hipMemcpyAsync(&nwf, B.counts, sizeof(int), hipMemcpyDeviceToHost, st), W);
      HIPCHK(hipStreamSynchronize(st), W);
      if (nwf > 0) {
        const int fchunk = (int)std::max<size_t>(1, std::min<size_t>(4096, (size_t(256) << 20) /
                                                                          (8 * kk::mv_fallback_words(mmax))));
        if (!K.d_fb.ensure(sizeof(double) * fchunk * kk::mv_fallback_words(mmax))) {
          set_error(W, "device allocation failed (fallback scratch)");
          return false;
        }
        for (int f0 = 0; f0 < nwf; f0 += fchunk)
          HIPCHK(kk::launch_fallback_mv(m, B.lists + f0, std::min(fchunk, nwf - f0), K.d_fb.as<double>(), st), W);
      }
    } else if (nw > 0) {
      const int fchunk = (int)std::max<size_t>(1, std::min<size_t>(4096, (size_t(256) << 20) /
                                                                        (8 * kk::fallback_words(mmax))));
      if (!K.d_fb.ensure(sizeof(double) * fchunk * kk::fallback_words(mmax))) {
        set_error(W, "device allocation failed (fallback scratch)");
        return false;
      }
      kk::SolveArgs vw = va;
      vw.weights = K.d_wout.as<double>();
      vw.weights_only = 1;
      for (int f0 = 0; f0 < nw; f0 += fchunk) {
        vw.list = B.xlist + f0;
        vw.nlist = std::min(fchunk, nw - f0);
        HIPCHK(kk::launch_fallback(vw, K.d_fb.as<double>(), st), W);
      }
    }
    HIPCHK(hipEventRecord(K.ev[6], st), W);
    HIPCHK(hipStreamSynchronize(st), W);
    HIPCHK(hipEventElapsedTime(&ms, K.ev[5], K.ev[6]), W);
    K.timing[12] = ms;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Graph-captured pass.  The same search -> (group | bucket) -> solve per
// bucket -> exact search of flagged queries -> their solves -> FP64 re-solve
// -> fallback sequence as run_pipeline, with every count consumed on the
// device: the solve, empty-block, exact-search and fallback kernels read
// their list lengths (and work-item offsets) from the counter block and loop
// over them with a fixed grid, and counters are cleared by a fill kernel.
// So the whole pass has no host synchronisation and is captured once into a
// HIP graph, then replayed.  The exact search runs on the second stream,
// forked after the collection of flagged queries and joined before their
// solves.  Supported: single-variable passes with the wave-cooperative
// search, at most 64 rows per system, and the exact search's heap in LDS.
// Counter slots: [0] FP64 re-solves, [1] fallbacks, [8] flagged queries,
// [16..] set members, rounds at ROUND_CTR[1] (decided blocks) and
// ROUND_CTR[2] (flagged blocks).
// ---------------------------------------------------------------------------
bool capture_pass(Krige& K) {
  const char* W = "kk_graph_capture";
  Obs& O = K.obs[0];
  const int nd = K.ndim, nb = K.nblock;
  int mmax, kcand;
  pass_dims(K, mmax, kcand);
  const bool sector = O.sector && O.need_search;
  if (mv_results(K) || !O.need_search || mmax > 64 || K.search_mode != 2 || (K.st && sector) ||
      kk::fast_search_lds(nd, sector, kcand, O.nmax) > 64 * 1024 || !kk::exact_search_fits_lds(kcand) ||
      K.store_weight || K.use_old_weight) {
    set_error(W, "graph capture needs a single-variable pass with the wave-cooperative search, at most 64 "
                 "rows per system and an exact-search heap in LDS");
    return false;
  }
  if (!ensure_stream(K) || !morton_perm(K)) return false;
  for (auto& e : K.gev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming), W);
  PassBufs B;
  kk::SolveArgs va;
  bool tb = false;
  if (!setup_pass(K, nullptr, mmax, B, va, tb)) return false;
  const bool reuse = B.items != nullptr;
  // fallback scratch for a fixed grid of threads
  constexpr int FB_THREADS = 1024;
  if (!K.d_fb.ensure(sizeof(double) * FB_THREADS * kk::fallback_words(mmax))) {
    set_error(W, "device allocation failed (fallback scratch)");
    return false;
  }
  hipStream_t st = K.stream, s2 = K.stream2;
  const int* perm = K.d_perm.as<int>();
  int* C = B.counts;
  int* overflow = C + PASS_COUNTERS - 1;
  K.drop_graph();
  // The pass, enqueued on st (and s2).  bnd: launch bounds per count (the
  // list lengths of an earlier run of the same pass), or null for bounds of
  // nb everywhere.  Kernels read the actual counts on the device and set the
  // overflow slot if one exceeds its bound.
  struct Bounds {
    int r0[kk::NBUCKET + 1], r1[kk::NBUCKET + 1], nexact, nres;
  };
  auto enqueue = [&](const Bounds* bnd) -> bool {
    if (kk::launch_fill(C, PASS_COUNTERS, 0, st) != hipSuccess) return false;
    // search, one launch over the Morton order
    kk::SearchArgs sa{};
    sa.ndim = nd;
    sa.qcoord = (O.aniso || K.st) ? K.d_qcoord.as<double>() : K.d_bcoord.as<double>();
    sa.ocoord = K.d_socoord.as<double>();
    sa.nmax = O.nmax;
    sa.cv = K.cross_validation;
    sa.maxdist2 = O.maxdist2;
    sa.mmax = mmax;
    sa.nnear = B.nnear;
    sa.inear = B.inear;
    sa.tree.dim = nd;
    sa.tree.n = O.n;
    sa.tree.nnodes = O.tree.nnodes();
    sa.tree.root = O.tree.root;
    sa.tree.rdata = K.d_rdata.as<double>();
    sa.tree.ind = K.d_ind.as<int>();
    sa.tree.ni = K.d_ni.as<int>();
    sa.tree.nd = K.d_nd.as<double>();
    sa.sort = sector;
    sa.k = kcand;
    sa.qflag = B.sflag;
    sa.q0 = 0;
    sa.nq = nb;
    sa.qlist = perm;
    if (kk::launch_search_wave(sa, st) != hipSuccess) return false;
    // flagged queries in Morton order; the exact search on the second stream
    if (kk::launch_collect(B.sflag, nb, 1, C + 8, B.xlist, st, perm) != hipSuccess) return false;
    if (hipEventRecord(K.gev[0], st) != hipSuccess || hipStreamWaitEvent(s2, K.gev[0], 0) != hipSuccess)
      return false;
    kk::SearchArgs sx = sa;
    sx.qlist = B.xlist;
    sx.dnq = C + 8;
    sx.qflag = nullptr;
    sx.nq = bnd ? std::max(bnd->nexact, 1) : nb;
    if (kk::launch_search_dn(sx, s2) != hipSuccess) return false;
    if (hipEventRecord(K.gev[1], s2) != hipSuccess) return false;
    // one solve per bucket, lengths and offsets read on the device
    auto solve_buckets = [&](const int* ctr, const int* bound, const int* lists, const int2* items,
                             const int* empties) -> bool {
      kk::SolveArgs v = va;
      v.nlist = 256 * 64;
      v.grid_cap = 64;
      v.list = items ? empties : lists;
      v.dn = ctr;
      v.doff = nullptr;
      if (kk::launch_empty(v, st) != hipSuccess) return false;
      v.grid_cap = 0;
      if (items) {
        v.setcnt = B.setcnt;
        v.setbase = B.setbase;
        v.members = B.members;
        v.item_max = K.item_max;
      }
      for (int bk = 1; bk < kk::NBUCKET; ++bk) {
        // bounded by the captured length (at least 1, so that a count that
        // grew is still seen); workgroups past the device count exit
        v.nlist = bound ? std::max(bound[bk], 1) : nb;
        v.dn = ctr + bk;
        if (items) {
          v.items = items;
          v.doff = ctr + kk::SET_OFF + bk;
          v.list = nullptr;
        } else {
          v.items = nullptr;
          v.doff = nullptr;
          v.list = lists + (size_t)bk * nb;
        }
        if (kk::launch_solve(v, tuned_width(nd, K.precision, tb, bk), K.precision, tb, st) != hipSuccess)
          return false;
      }
      return true;
    };
    // round 0: the blocks the fast search decided (flagged ones are skipped)
    int* r0 = C + ROUND_CTR[1];
    if (reuse) {
      kk::SetGroupArgs g{};
      g.nnear = B.nnear;
      g.inear = B.inear;
      g.mmax = mmax;
      g.rs = va.rangescale;
      g.ln = va.localnugget;
      g.sel = nullptr;
      g.nsel = nb;
      g.skip = B.sflag;
      g.table = B.table;
      g.hash = B.hash;
      g.tsize = std::min(B.tcap, set_table_slots(nb));
      g.item_max = K.item_max;
      g.rep = B.rep;
      g.rank = B.rank;
      g.cnt = B.setcnt;
      g.setbase = B.setbase;
      g.setitem = B.setitem;
      g.members = B.members;
      g.ctr = r0;
      g.total = C + 16;
      g.items = B.items;
      g.empties = B.empties;
      if (kk::launch_set_group(g, st) != hipSuccess) return false;
      if (!solve_buckets(r0, bnd ? bnd->r0 : nullptr, nullptr, B.items, B.empties)) return false;
    } else {
      if (kk::launch_bucket(B.nnear, nullptr, nb, r0, B.lists, nb, st, nullptr, B.sflag) != hipSuccess) return false;
      if (!solve_buckets(r0, bnd ? bnd->r0 : nullptr, B.lists, nullptr, nullptr)) return false;
    }
    // round 1: the exact search's blocks, bucketed on the device count
    if (hipStreamWaitEvent(st, K.gev[1], 0) != hipSuccess) return false;
    int* r1 = C + ROUND_CTR[2];
    int* l1 = B.lists + (size_t)(kk::NBUCKET + 1) * nb;
    const int nx = bnd ? std::max(bnd->nexact, 1) : nb;
    if (kk::launch_bucket(B.nnear, B.xlist, nx, r1, l1, nb, st, C + 8) != hipSuccess) return false;
    if (!solve_buckets(r1, bnd ? bnd->r1 : nullptr, l1, nullptr, nullptr)) return false;
    // FP64 re-solve of the FP32 systems whose pivot ratio fell below pivot_tol
    kk::SolveArgs v = va;
    if (K.precision == kk::PREC_FP32R) {
      if (kk::launch_collect(B.flag, nb, 1, C + 0, B.lists, st) != hipSuccess) return false;
      v.list = B.lists;
      v.dn = C + 0;
      v.nlist = bnd ? std::max(bnd->nres, 1) : nb;
      if (kk::launch_solve(v, kk::solve_width(mmax), kk::PREC_FP64, tb, st) != hipSuccess) return false;
    }
    // Bunch-Kaufman fallback (a fixed grid looping over the count)
    if (kk::launch_collect(B.flag, nb, 2, C + 1, B.lists + nb, st) != hipSuccess) return false;
    v.list = B.lists + nb;
    v.dn = C + 1;
    v.nlist = FB_THREADS;
    v.grid_cap = 0;
    if (kk::launch_fallback(v, K.d_fb.as<double>(), st) != hipSuccess) return false;
    if (bnd) {
      // every count whose launches were sized by a bound
      kk::BoundCheck c{};
      auto add = [&](int idx, int bound) {
        c.idx[c.n] = idx;
        c.bound[c.n] = bound;
        ++c.n;
      };
      for (int bk = 1; bk < kk::NBUCKET; ++bk) {
        add(ROUND_CTR[1] + bk, std::max(bnd->r0[bk], 1));
        add(ROUND_CTR[2] + bk, std::max(bnd->r1[bk], 1));
      }
      add(8, std::max(bnd->nexact, 1));
      if (K.precision == kk::PREC_FP32R) add(0, std::max(bnd->nres, 1));
      if (kk::launch_check_bounds(C, c, overflow, st) != hipSuccess) return false;
    }
    return true;
  };
  // one run with bounds of nb, for the pass's list lengths
  if (!enqueue(nullptr)) {
    set_error(W, "HIP error in the pass to be captured");
    return false;
  }
  std::vector<int> hc(PASS_COUNTERS);
  HIPCHK(hipMemcpyAsync(hc.data(), C, sizeof(int) * PASS_COUNTERS, hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  Bounds bd{};
  for (int bk = 0; bk <= kk::NBUCKET; ++bk) {
    bd.r0[bk] = hc[ROUND_CTR[1] + bk];
    bd.r1[bk] = hc[ROUND_CTR[2] + bk];
  }
  bd.nexact = hc[8];
  bd.nres = hc[0];
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), W);
  const bool ok = enqueue(&bd);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (!ok || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    set_error(W, "stream capture of the pass failed");
    return false;
  }
  K.graph = g;
  HIPCHK(hipGraphInstantiate(&K.gexec, K.graph, nullptr, nullptr, 0), W);
  K.graph_overflow = overflow;
  K.graph_key[0] = B.nnear;
  K.graph_key[1] = K.d_fb.p;
  K.graph_key[2] = reuse ? (const void*)B.items : (const void*)B.lists;
  return true;
}

bool fetch_results(Krige& K) {
  const char* W = "t_kriging_base%solve_base";
  const int nb = K.nblock;
  K.est.resize(nb);
  K.var.resize(nb);
  HIPCHK(hipMemcpy(K.est.data(), K.d_est.p, sizeof(double) * nb, hipMemcpyDeviceToHost), W);
  HIPCHK(hipMemcpy(K.var.data(), K.d_var.p, sizeof(double) * nb, hipMemcpyDeviceToHost), W);
  // fallback failures: NaN estimates, counted in n_fail (kriging_base.F90:3369-3383)
  std::vector<int> flag;
  if (K.stats[3] > 0) {
    flag.resize(nb);
    HIPCHK(hipMemcpy(flag.data(), K.d_flag.p, sizeof(int) * nb, hipMemcpyDeviceToHost), W);
    int nf = 0;
    long first = -1;
    for (int b = 0; b < nb; ++b)
      if (flag[b] == 3) {
        ++nf;
        if (first < 0) first = b + 1;
      }
    K.stats[0] = nf;
    K.stats[3] -= nf;
    if (nf > 0 && !K.neglect_error) {
      set_error("t_kriging_base%solve_linear_system", "Singular matrix", first);
      return false;
    }
  }
  // blocks without neighbours are an error unless neglect_error
  // (kriging_base.F90:3198-3207); the pass's bucketing counted them, so the
  // counts are read back only to name the first
  if (!K.neglect_error && K.pass_empty != 0) {
    std::vector<int> nn(nb);
    HIPCHK(hipMemcpy(nn.data(), K.d_nnear.p, sizeof(int) * nb, hipMemcpyDeviceToHost), W);
    for (int b = 0; b < nb; ++b)
      if (nn[b] == 0) {
        set_error("t_kriging_base%assemble_linear_system", "not enough neighbors for kriging at block", b + 1);
        return false;
      }
  }
  return true;
}

bool validate_for_solve(Krige& K) {
  const char* W = "t_kriging%prepare";
  if (!K.initialized) {
    set_error(W, "Call initialize() before solve.");
    return false;
  }
  if (K.nvar < 1 || K.nvar > kk::MAXVAR || K.nsim != 0 || K.ndrift < 0 || K.ndrift > kk::MAXDRIFT) {
    set_error(W, "this backend solves kriging and cokriging with nvar <= 4, nsim=0, ndrift <= 4");
    return false;
  }
  if (grad_total(K) > 0 && (K.st || K.store_weight || K.weight_correction || K.cross_validation ||
                            K.varying_vgm)) {
    set_error(W, "increment (gradient) constraints are provided without a weight store, weight correction, "
                 "cross-validation or space-time");
    return false;
  }
  if (K.st) {
    if (K.nvar != 1) {
      set_error(W, "space-time kriging with nvar > 1 is not supported by this backend");
      return false;
    }
    if (K.st_model == 0) {
      // is_valid_vgm_st, variogram_st.f90:388-402
      if (K.sill_st.empty()) {
        set_error("is_valid_vgm_st", "WARNING vgm_struct_st(1,1): sill_st not set for sum-metric model");
        return false;
      }
      if (K.sill_st.size() != K.vgm[0].size()) {
        set_error("is_valid_vgm_st", "WARNING vgm_struct_st(1,1): one joint sill per spatial structure is needed");
        return false;
      }
      for (size_t k = 0; k < K.sill_st.size(); ++k)
        if (K.sill_st[k] < 0.0) {
          set_error("is_valid_vgm_st",
                    "WARNING vgm_struct_st(1,1): negative joint sill at structure " + std::to_string(k + 1));
          return false;
        }
    }
    if (K.vgm_t.empty()) {
      set_error("t_kriging_st%prepare", "Invalid ST variogram: no temporal structure (set_vgm_temporal)");
      return false;
    }
    if (!K.soff.empty()) {
      set_error(W, "space-time block kriging is not supported by this backend");
      return false;
    }
  }
  if (mv_results(K) && K.cross_validation) {
    set_error(W, "cross validation with nvar > 1, external drift or weight correction is not supported by this "
                 "backend");
    return false;
  }
  if (K.st && (K.ndrift > 0 || K.weight_correction)) {
    set_error(W, "space-time kriging with external drift or weight correction is not supported by this backend");
    return false;
  }
  for (int v = 1; v < K.nvar; ++v) {
    if (K.obs[v].n == 0) {
      set_error(W, "set_obs() needs to be called for every variable before solve().");
      return false;
    }
    if (!K.obs[v].search_set) {
      set_error(W, "set_search() needs to be called for every variable before solve().");
      return false;
    }
  }
  for (int i = 0; i < K.nvar; ++i)
    for (int j = i; j < K.nvar; ++j) {
      const auto& vs = K.vgm[(size_t)i + (size_t)K.nvar * j];
      if (vs.empty()) {
        set_error("t_kriging%validate_vgm", "variogram not set for ivar=" + std::to_string(i + 1) +
                                                ", jvar=" + std::to_string(j + 1));
        return false;
      }
      for (const auto& st : vs)
        if (st.nugget < 0.0 || st.sill < 0.0 || st.a_major <= 0.0 || st.a_minor1 <= 0.0 || st.a_minor2 <= 0.0) {
          set_error("t_kriging%validate_vgm", "variogram is not valid for ivar=" + std::to_string(i + 1) +
                                                  ", jvar=" + std::to_string(j + 1));
          return false;
        }
    }
  if (K.varying_vgm) {
    set_error(W, "varying_vgm is not supported by this backend");
    return false;
  }
  if (K.obs[0].n == 0) {
    set_error(W, "set_obs() needs to be called before solve().");
    return false;
  }
  if (K.vgm[0].empty()) {
    set_error("t_kriging_sva%validate_vgm",
              "t_kriging_sva: variogram not set for block 1, ivar=1, jvar=1. Call set_vgm_block() or "
              "set_vgm_block_all().");
    return false;
  }
  for (const auto& s : K.vgm[0]) {
    if (s.nugget < 0.0 || s.sill < 0.0 || s.a_major <= 0.0 || s.a_minor1 <= 0.0 || s.a_minor2 <= 0.0 ||
        (s.vtype == kk::HOL && K.ndim == 3)) {
      set_error("t_kriging_sva%validate_vgm",
                "t_kriging_sva: variogram is not valid for block 1, ivar=1, jvar=1. Check your variogram "
                "parameters.");
      return false;
    }
  }
  if (K.nblock == 0 && !K.cross_validation) {
    set_error(W, "set_grid() needs to be called before solve().");
    return false;
  }
  if (!K.obs[0].search_set) {
    set_error(W, "set_search() needs to be called before solve().");
    return false;
  }
  return true;
}

bool compute_base_cov(Krige& K) {
  // prior covariance of each block over its integration points
  // (kriging.F90:992-1008), evaluated once per block on the host
  const char* W = "t_kriging%calc_variance";
  if (K.soff.empty()) return true;
  std::vector<double> base(K.nblock);
  const int nd = K.ndim;
  // identical relative layouts share one value: hoisted when the block
  // geometry repeats (uniform Gauss-Legendre blocks)
  std::vector<double> tabcopy = K.tab;
  Model M = K.model;
  M.tab = tabcopy.empty() ? nullptr : tabcopy.data();
  auto cov_host = [&](const double* lag) -> double {
    // host evaluation through the same model description
    auto aniso = [&](const double* m) {
      double rx, ry, rz = 0.0;
      if (nd == 1) return std::fabs(lag[0] * m[0]);
      if (nd == 2) {
        rx = m[0] * lag[0] + m[3] * lag[1];
        ry = m[1] * lag[0] + m[4] * lag[1];
        return std::sqrt(rx * rx + ry * ry);
      }
      rx = m[0] * lag[0] + m[3] * lag[1] + m[6] * lag[2];
      ry = m[1] * lag[0] + m[4] * lag[1] + m[7] * lag[2];
      rz = m[2] * lag[0] + m[5] * lag[1] + m[8] * lag[2];
      return std::sqrt(rx * rx + ry * ry + rz * rz);
    };
    auto comp_h = [&](int c, double h) {
      return (h > DBL_MIN * 1.0e3) ? M.sill[c] * corefunc_host(M.vtype[c], h) : M.sill[c] + M.nugget[c];
    };
    auto interp = [&](const double* tab, double h) {
      int k = (h >= M.zb[1]) + (h >= M.zb[2]);
      double fi = (h - M.zb[k]) * M.zinv[k];
      int ifi = (int)fi;
      int i = M.zoff[k] + std::min(ifi, M.zn[k] - 1);
      fi = fi - (double)ifi;
      return tab[i] + fi * (tab[i + 1] - tab[i]);
    };
    if (M.mode == kk::COV_TABLE_B) {
      double h = aniso(M.bmat);
      if (h >= M.zb[3]) {
        if (!M.b_tail) return 0.0;
        double res = 0.0;
        int iv = 0;
        while (iv < M.nstruct) {
          double g = comp_h(iv, h);
          while (iv + 1 < M.nstruct && M.is_product[iv + 1]) {
            ++iv;
            g = g * comp_h(iv, h);
          }
          res = res + g;
          ++iv;
        }
        return res;
      }
      if (h <= 0.0) return M.tab[0];
      return interp(M.tab, h);
    }
    double res = 0.0;
    int iv = 0;
    while (iv < M.nstruct) {
      double g = 0.0;
      bool first = true;
      do {
        if (!first) ++iv;
        double h = aniso(M.mat[iv]), c;
        if (M.mode == kk::COV_TABLE_A) {
          const double* tab = M.tab + (size_t)iv * (M.ntab + 1);
          if (h >= M.zb[3]) c = has_tail_host(M.vtype[iv]) ? comp_h(iv, h) : 0.0;
          else if (h <= 0.0) c = tab[0];
          else c = interp(tab, h);
        } else {
          c = comp_h(iv, h);
        }
        g = first ? c : g * c;
        first = false;
      } while (iv + 1 < M.nstruct && M.is_product[iv + 1]);
      res = res + g;
      ++iv;
    }
    return res;
  };
  std::vector<double> prev_rel;
  double prev_val = 0.0;
  for (int b = 0; b < K.nblock; ++b) {
    const int s0 = K.soff[b], s1 = K.soff[b + 1], ns = s1 - s0;
    std::vector<double> rel((size_t)ns * (nd + 1));
    for (int i = 0; i < ns; ++i) {
      for (int k = 0; k < nd; ++k)
        rel[(size_t)i * (nd + 1) + k] = K.scoord[(size_t)nd * (s0 + i) + k] - K.scoord[(size_t)nd * s0 + k];
      rel[(size_t)i * (nd + 1) + nd] = K.sweight[s0 + i];
    }
    if (b > 0 && rel == prev_rel) {
      base[b] = prev_val;
      continue;
    }
    double v = 0.0;
    for (int i = 0; i < ns; ++i) {
      const double wi = K.sweight[s0 + i];
      v = v + K.model.cov0 * wi * wi;
      for (int j = i + 1; j < ns; ++j) {
        double lag[3] = {0, 0, 0};
        for (int k = 0; k < nd; ++k) lag[k] = K.scoord[(size_t)nd * (s0 + i) + k] - K.scoord[(size_t)nd * (s0 + j) + k];
        v = v + cov_host(lag) * wi * K.sweight[s0 + j] * 2.0;
      }
    }
    base[b] = v;
    prev_rel.swap(rel);
    prev_val = v;
  }
  return upload(K.d_basecov, base, W);
}

// ---------------------------------------------------------------------------
// multi-variable (co)kriging: one neighbour search per variable, then one
// workgroup per block solving the bordered system (cokrige.hip)
// ---------------------------------------------------------------------------
bool search_var(Krige& K, int v, hipStream_t st) {
  const char* W = "t_kriging_base%solve_base";
  Obs& O = K.obs[v];
  DevVar& D = K.dv[v];
  const int nd = K.ndim, nb = K.nblock;
  const bool sector = O.sector && O.need_search;
  const int nmax = O.nmax;
  D.mmax = sector ? (1 << nd) * nmax : nmax;
  if (!D.nnear.ensure(sizeof(int) * nb) || !D.inear.ensure(sizeof(int) * (size_t)nb * D.mmax)) {
    set_error(W, "device allocation failed");
    return false;
  }
  kk::SearchArgs sa{};
  sa.ndim = nd;
  sa.qcoord = O.aniso ? D.qcoord.as<double>() : K.d_bcoord.as<double>();
  sa.ocoord = D.socoord.as<double>();
  sa.nmax = nmax;
  sa.cv = 0;
  sa.maxdist2 = O.maxdist2;
  sa.mmax = D.mmax;
  sa.nnear = D.nnear.as<int>();
  sa.inear = D.inear.as<int>();
  if (!O.need_search) {
    kk::Rot3 R;
    std::memcpy(R.m, O.rot, sizeof R.m);
    sa.nq = nb;
    HIPCHK(kk::launch_search_all(sa, O.n, D.ocoord.as<double>(), R, st), W);
    return true;
  }
  sa.tree.dim = nd;
  sa.tree.n = O.n;
  sa.tree.nnodes = O.tree.nnodes();
  sa.tree.root = O.tree.root;
  sa.tree.rdata = D.rdata.as<double>();
  sa.tree.ind = D.ind.as<int>();
  sa.tree.ni = D.ni.as<int>();
  sa.tree.nd = D.nd.as<double>();
  sa.sort = sector;
  sa.k = sector ? std::min(O.n, (1 << nd) * nmax * 4) : nmax;
  int nexact = nb;
  const int* list = nullptr;
  if (K.search_mode > 0 && kk::fast_search_lds(nd, sector, sa.k, nmax) <= 64 * 1024) {
    if (!D.sflag.ensure(sizeof(int) * nb) || !D.xlist.ensure(sizeof(int) * nb)) {
      set_error(W, "device allocation failed (search flags)");
      return false;
    }
    sa.qflag = D.sflag.as<int>();
    sa.nq = nb;
    if (K.search_mode == 2) {
      // queries in Morton order: a wave's queries are close together
      if (K.reorder && !morton_perm(K)) return false;
      sa.qlist = K.reorder ? K.d_perm.as<int>() : nullptr;
      HIPCHK(kk::launch_search_wave(sa, st), W);
      sa.qlist = nullptr;
    } else {
      HIPCHK(kk::launch_search_fast(sa, st), W);
    }
    HIPCHK(hipMemsetAsync(K.d_counts.p, 0, sizeof(int) * 16, st), W);
    HIPCHK(kk::launch_collect(D.sflag.as<int>(), nb, 1, K.d_counts.as<int>() + 8, D.xlist.as<int>(), st,
                              K.search_mode == 2 && K.reorder ? K.d_perm.as<int>() : nullptr), W);
    HIPCHK(hipMemcpyAsync(&nexact, K.d_counts.as<int>() + 8, sizeof(int), hipMemcpyDeviceToHost, st), W);
    HIPCHK(hipStreamSynchronize(st), W);
    list = D.xlist.as<int>();
  }
  if (nexact > 0) {
    const int chunk = std::min(nexact, 1 << 20);
    if (!D.hdis.ensure(sizeof(double) * (size_t)sa.k * chunk) || !D.hidx.ensure(sizeof(int) * (size_t)sa.k * chunk)) {
      set_error(W, "device allocation failed (search scratch)");
      return false;
    }
    sa.hdis = D.hdis.as<double>();
    sa.hidx = D.hidx.as<int>();
    sa.hstride = chunk;
    sa.qflag = nullptr;
    for (int q0 = 0; q0 < nexact; q0 += chunk) {
      sa.q0 = list ? 0 : q0;
      sa.qlist = list ? list + q0 : nullptr;
      sa.nq = std::min(chunk, nexact - q0);
      HIPCHK(kk::launch_search(sa, st), W);
    }
  }
  return true;
}

// estimates, variances and failure counts of a multi-variable pass, from
// the device (the reference's n_fail / Singular matrix rules)
bool fetch_results_mv(Krige& K) {
  const char* W = "t_kriging_base%solve_base";
  const int nb = K.nblock, nvar = K.nvar;
  // blocks sent to the fallback by the last pass
  const int nfb = (int)K.timing[6];
  K.est_mv.resize((size_t)nb * nvar);
  K.var_mv.resize((size_t)nb * nvar * nvar);
  HIPCHK(hipMemcpy(K.est_mv.data(), K.d_mvest.p, sizeof(double) * K.est_mv.size(), hipMemcpyDeviceToHost), W);
  HIPCHK(hipMemcpy(K.var_mv.data(), K.d_mvvar.p, sizeof(double) * K.var_mv.size(), hipMemcpyDeviceToHost), W);
  std::vector<int> flag(nb), tot(nb, 0), nn(nb);
  HIPCHK(hipMemcpy(flag.data(), K.d_flag.p, sizeof(int) * nb, hipMemcpyDeviceToHost), W);
  for (int v = 0; v < nvar; ++v) {
    HIPCHK(hipMemcpy(nn.data(), K.dv[v].nnear.p, sizeof(int) * nb, hipMemcpyDeviceToHost), W);
    for (int b = 0; b < nb; ++b) tot[b] += nn[b];
  }
  int nfail = 0, nsolved = 0;
  long first = -1;
  for (int b = 0; b < nb; ++b) {
    if (flag[b] == 3) {
      ++nfail;
      if (first < 0) first = b + 1;
    } else if (tot[b] > 1) {
      ++nsolved;  // a single neighbour is not solved (kriging_base.F90:2127)
    }
    if (tot[b] == 0 && !K.neglect_error) {
      set_error("t_kriging_base%assemble_linear_system", "not enough neighbors for kriging at block", b + 1);
      return false;
    }
  }
  K.stats[0] = nfail;
  K.stats[1] = nsolved - (nfb - nfail);
  K.stats[2] = K.stats[4] = 0;
  K.stats[3] = nfb - nfail;
  if (nfail > 0 && !K.neglect_error) {
    set_error("t_kriging_base%solve_linear_system", "Singular matrix", first);
    return false;
  }
  return true;
}


bool do_solve_mv(Krige& K, bool fetch) {
  const char* W = "t_kriging_base%solve_base";
  if (!ensure_stream(K)) return false;
  hipStream_t st = K.stream;
  const int nd = K.ndim, nb = K.nblock, nvar = K.nvar;
  // models of every ordered pair; in table mode only vgm(a, b) with a <= b
  // carries tables (the pairs the reference builds them for,
  // kriging.F90:410-420), the others are evaluated analytically
  if (!K.model_ok) {
    std::vector<Model> ms((size_t)nvar * nvar);
    std::vector<std::vector<double>> tabs(ms.size());
    size_t total = 0;
    for (int a = 0; a < nvar; ++a)
      for (int b = 0; b < nvar; ++b) {
        const size_t i = (size_t)a + (size_t)nvar * b;
        if (!build_model(K.vgm[i], nd, K.table_mode && a <= b, ms[i], tabs[i])) {
          set_error(W, "invalid variogram model");
          return false;
        }
        total += tabs[i].size();
      }
    std::vector<double> all;
    all.reserve(total);
    std::vector<size_t> off(ms.size());
    for (size_t i = 0; i < ms.size(); ++i) {
      off[i] = all.size();
      all.insert(all.end(), tabs[i].begin(), tabs[i].end());
    }
    if (!upload(K.d_mvtab, all, W)) return false;
    for (size_t i = 0; i < ms.size(); ++i) ms[i].tab = tabs[i].empty() ? nullptr : K.d_mvtab.as<double>() + off[i];
    if (!upload(K.d_models, ms, W)) return false;
    // shapes with no trigonometric or power term evaluate inline
    K.mv_simple = 1;
    for (const Model& m : ms)
      for (int i = 0; i < m.nstruct; ++i) {
        const int t = m.vtype[i];
        K.mv_simple &= (t == kk::NUG || t == kk::SPH || t == kk::EXP || t == kk::GAU || t == kk::BSQ || t == kk::LIN);
      }
    K.model_ok = true;
  }
  if (!K.dev_inputs_ok) {
    if (!upload(K.d_bcoord, K.bcoord, W) || !upload(K.d_rs, K.rangescale, W) || !upload(K.d_ln, K.localnugget, W))
      return false;
    if (K.ndrift > 0 && !upload(K.d_bdrift, K.bdrift, W)) return false;
    if (!K.soff.empty() && (!upload(K.d_scoord, K.scoord, W) || !upload(K.d_sweight, K.sweight, W))) return false;
    for (int v = 0; v < nvar; ++v) {
      Obs& O = K.obs[v];
      DevVar& D = K.dv[v];
      if (!upload(D.ocoord, O.coord, W) || !upload(D.ovalue, O.value, W) || !upload(D.ovar, O.var, W) ||
          !upload(D.socoord, O.aniso ? O.rcoord : O.coord, W))
        return false;
      if (K.ndrift > 0 && !upload(D.odrift, O.drift, W)) return false;
      if (O.ngrad > 0 && (!upload(D.gx1, O.gx1, W) || !upload(D.gx2, O.gx2, W) || !upload(D.gval, O.gval, W) ||
                          !upload(D.gvar, O.gvar, W) || (K.ndrift > 0 && !upload(D.gdrift, O.gdrift, W))))
        return false;
      if (O.need_search && (!upload(D.rdata, O.tree.rdata, W) || !upload(D.ind, O.tree.ind, W) ||
                            !upload(D.ni, O.tree.ni, W) || !upload(D.nd, O.tree.nd, W)))
        return false;
      if (O.aniso) {
        std::vector<double> q(K.bcoord.size());
        kk::Rot3 R;
        std::memcpy(R.m, O.rot, sizeof R.m);
        for (int b = 0; b < nb; ++b) {
          if (nd == 1) R.apply<1>(&K.bcoord[(size_t)nd * b], &q[(size_t)nd * b]);
          else if (nd == 2) R.apply<2>(&K.bcoord[(size_t)nd * b], &q[(size_t)nd * b]);
          else R.apply<3>(&K.bcoord[(size_t)nd * b], &q[(size_t)nd * b]);
        }
        if (!upload(D.qcoord, q, W)) return false;
      }
    }
    K.dev_inputs_ok = true;
  }
  if (!K.d_counts.ensure(sizeof(int) * 16) || !K.d_mvest.ensure(sizeof(double) * (size_t)nb * nvar) ||
      !K.d_mvvar.ensure(sizeof(double) * (size_t)nb * nvar * nvar) || !K.d_flag.ensure(sizeof(int) * nb)) {
    set_error(W, "device allocation failed");
    return false;
  }
  HIPCHK(hipEventRecord(K.ev[0], st), W);
  int rows = 0;
  for (int v = 0; v < nvar; ++v) {
    if (!search_var(K, v, st)) return false;
    rows += K.dv[v].mmax;
  }
  HIPCHK(hipEventRecord(K.ev[1], st), W);
  // increment rows join every block's system
  rows += grad_total(K);
  const int cols = rows + kk::mv_extra_cols(nvar, K.std_ck, K.unbias, K.ndrift);
  if (cols > 128) {
    set_error(W, "cokriging systems larger than 128 rows + right-hand sides are not supported by the GPU solver (" +
                     std::to_string(cols) + ")");
    return false;
  }
  const int nsub = K.soff.empty() ? 1 : K.soff[1] - K.soff[0];
  if (nsub > 64) {
    set_error(W, "blocks with more than 64 integration nodes are not supported with nvar > 1");
    return false;
  }
  kk::MVArgs a{};
  a.models = K.d_models.as<Model>();
  a.ndim = nd;
  a.nvar = nvar;
  a.nb = nb;
  a.std_ck = K.std_ck;
  a.unbias = K.unbias;
  a.nsub = nsub;
  a.rows = rows;
  // increment rows are assembled by the general-shape instances only
  a.simple = K.mv_simple && grad_total(K) == 0;
  for (int v = 0; v < nvar; ++v) a.sk_mean[v] = K.obs[v].sk_mean;
  a.bound_lo = K.bounds[0];
  a.bound_hi = K.bounds[1];
  a.bcoord = K.d_bcoord.as<double>();
  a.scoord = nsub > 1 ? K.d_scoord.as<double>() : nullptr;
  a.sweight = nsub > 1 ? K.d_sweight.as<double>() : nullptr;
  a.rangescale = K.rangescale.empty() ? nullptr : K.d_rs.as<double>();
  a.localnugget = K.localnugget.empty() ? nullptr : K.d_ln.as<double>();
  for (int v = 0; v < nvar; ++v) {
    a.ocoord[v] = K.dv[v].ocoord.as<double>();
    a.ovalue[v] = K.dv[v].ovalue.as<double>();
    a.ovar[v] = K.dv[v].ovar.as<double>();
    a.nnear[v] = K.dv[v].nnear.as<int>();
    a.inear[v] = K.dv[v].inear.as<int>();
    a.mmax[v] = K.dv[v].mmax;
  }
  a.est = K.d_mvest.as<double>();
  a.var = K.d_mvvar.as<double>();
  a.flag = K.d_flag.as<int>();
  a.ndrift = K.ndrift;
  if (K.ndrift > 0) {
    for (int v = 0; v < nvar; ++v) a.odrift[v] = K.dv[v].odrift.as<double>();
    a.bdrift = K.d_bdrift.as<double>();
  }
  for (int v = 0; v < nvar; ++v) {
    a.ngrad[v] = K.obs[v].ngrad;
    if (a.ngrad[v] > 0) {
      a.gx1[v] = K.dv[v].gx1.as<double>();
      a.gx2[v] = K.dv[v].gx2.as<double>();
      a.gval[v] = K.dv[v].gval.as<double>();
      a.gvar[v] = K.dv[v].gvar.as<double>();
      a.gdrift[v] = K.ndrift > 0 ? K.dv[v].gdrift.as<double>() : nullptr;
    }
  }
  // weight correction needs the weights: the weights instance
  a.wcorr = K.weight_correction && nvar == 1;
  if (K.store_weight || a.wcorr) {
    // the weight store: the solver's weights instance, the fallback writes
    // the weights of the blocks it solves
    K.mmax_w = 0;
    for (int v = 0; v < nvar; ++v) K.mmax_w = std::max(K.mmax_w, K.dv[v].mmax);
    if (!K.d_wout.ensure(sizeof(double) * (size_t)K.mmax_w * nvar * nvar * nb)) {
      set_error(W, "device allocation failed (weight store)");
      return false;
    }
    a.weights = K.d_wout.as<double>();
    a.mmax_w = K.mmax_w;
  }
  a.prec = K.precision;
  a.pivot_tol = K.pivot_tol;
  HIPCHK(kk::launch_solve_mv(a, st), W);
  int nfb = 0, nres = 0;
  if (!K.d_lists.ensure(sizeof(int) * (size_t)nb)) {
    set_error(W, "device allocation failed");
    return false;
  }
  // FP32 + refinement: blocks whose FP32 pivot ratio fell below pivot_tol
  // (or whose FP32 pivot was not positive) are solved again in FP64
  if (kk::mv_fp32r_instance(a)) {
    HIPCHK(hipMemsetAsync(K.d_counts.p, 0, sizeof(int) * 16, st), W);
    HIPCHK(kk::launch_collect(K.d_flag.as<int>(), nb, 1, K.d_counts.as<int>(), K.d_lists.as<int>(), st), W);
    HIPCHK(hipMemcpyAsync(&nres, K.d_counts.p, sizeof(int), hipMemcpyDeviceToHost, st), W);
    HIPCHK(hipStreamSynchronize(st), W);
    if (nres > 0) {
      kk::MVArgs r = a;
      r.prec = kk::PREC_FP64;
      r.list = K.d_lists.as<int>();
      r.nlist = nres;
      HIPCHK(kk::launch_solve_mv(r, st), W);
    }
  }
  // blocks whose K or Schur complement is not positive definite: the
  // Bunch-Kaufman fallback on the full augmented system
  HIPCHK(hipMemsetAsync(K.d_counts.p, 0, sizeof(int) * 16, st), W);
  HIPCHK(kk::launch_collect(K.d_flag.as<int>(), nb, 2, K.d_counts.as<int>(), K.d_lists.as<int>(), st), W);
  HIPCHK(hipMemcpyAsync(&nfb, K.d_counts.p, sizeof(int), hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  if (nfb > 0) {
    const int fchunk = 256;
    if (!K.d_fb.ensure(sizeof(double) * fchunk * kk::mv_fallback_words(rows))) {
      set_error(W, "device allocation failed (fallback scratch)");
      return false;
    }
    for (int f0 = 0; f0 < nfb; f0 += fchunk)
      HIPCHK(kk::launch_fallback_mv(a, K.d_lists.as<int>() + f0, std::min(fchunk, nfb - f0), K.d_fb.as<double>(), st),
             W);
  }
  HIPCHK(hipEventRecord(K.ev[4], st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[0], K.ev[1]), W);
  for (double& x : K.timing) x = 0.0;
  K.timing[0] = ms;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[1], K.ev[4]), W);
  K.timing[2] = ms;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[0], K.ev[4]), W);
  K.timing[4] = ms;
  K.timing[5] = nres;
  K.timing[6] = nfb;
  K.timing[7] = -1;
  if (!fetch) return true;
  return fetch_results_mv(K);
}

// ---------------------------------------------------------------------------
// weight store (kriging_base.F90:2703-2977, kriging_capi_common.F90:491-619)
// ---------------------------------------------------------------------------
// obs_nmax of variable v: 2^ndim nmax under sector search (octants over the
// spatial axes: 2^3 for space-time), else nmax (kriging_base.F90:3003-3011)
int obs_nmax(const Krige& K, int v) {
  const Obs& O = K.obs[v];
  const long long nsec = 1LL << (K.st ? 3 : K.ndim);
  return (int)std::min<long long>(INT_MAX, O.sector ? nsec * O.nmax : O.nmax);
}

int store_nmax(const Krige& K) {
  int nm = 0;
  for (int v = 0; v < K.nvar; ++v) nm = std::max(nm, obs_nmax(K, v));
  return nm;
}

// The store after a solve with store_weight (save_block_weights,
// kriging_base.F90:2760-2782, called for every block after its solve): the
// neighbour lists and weights of every group, the block's variance.  Blocks
// the fallback could not solve keep one neighbour per group with a NaN
// weight (ctx%nnear = 1 and x = NaN, kriging_base.F90:3369-3383).  A block
// with a single neighbour is not solved by the reference, which leaves the
// previous block's solution in its weight; it is NaN here, like its
// estimate.  Reads the device results directly (fetched or not).
bool store_from_device(Krige& K) {
  const char* W = "t_kriging_base%save_block_weights";
  const int nb = K.nblock, nvar = K.nvar;
  const bool mv = mv_results(K);
  const int nm = store_nmax(K);
  const int mmw = K.mmax_w;
  std::vector<int> flag(nb);
  std::vector<double> w((size_t)mmw * nvar * nvar * nb), vout((size_t)nb * nvar * nvar);
  HIPCHK(hipMemcpy(flag.data(), K.d_flag.p, sizeof(int) * nb, hipMemcpyDeviceToHost), W);
  HIPCHK(hipMemcpy(w.data(), K.d_wout.p, sizeof(double) * w.size(), hipMemcpyDeviceToHost), W);
  HIPCHK(hipMemcpy(vout.data(), !mv ? K.d_var.p : K.d_mvvar.p, sizeof(double) * vout.size(),
                   hipMemcpyDeviceToHost), W);
  std::vector<std::vector<int>> nn(nvar, std::vector<int>(nb));
  std::vector<int> mm(nvar);
  for (int v = 0; v < nvar; ++v) {
    mm[v] = !mv ? K.mmax : K.dv[v].mmax;
    HIPCHK(hipMemcpy(nn[v].data(), !mv ? K.d_nnear.p : K.dv[v].nnear.p, sizeof(int) * nb,
                     hipMemcpyDeviceToHost), W);
  }
  std::vector<int> tot(nb, 0);
  for (int v = 0; v < nvar; ++v)
    for (int b = 0; b < nb; ++b) tot[b] += nn[v][b];
  WStore& S = K.ws;
  S.init(nm, nvar, nvar, nb);
  const double qnan = std::nan("");
  std::vector<int> in;
  for (int g = 0; g < nvar; ++g) {
    in.resize((size_t)nb * mm[g]);
    HIPCHK(hipMemcpy(in.data(), !mv ? K.d_inear.p : K.dv[g].inear.p, sizeof(int) * in.size(),
                     hipMemcpyDeviceToHost), W);
    for (int b = 0; b < nb; ++b) {
      const bool failed = flag[b] == 3;
      const int n = nn[g][b];
      const int ns = failed ? 1 : n;
      S.nnear[g + (size_t)nvar * b] = ns;
      int* si = &S.inear[(size_t)nm * (g + (size_t)nvar * b)];
      for (int k = 0; k < std::min(ns, n); ++k) si[k] = in[(size_t)mm[g] * b + k] + 1;
      for (int q = 0; q < nvar; ++q) {
        double* sw = &S.weight[(size_t)nm * (g + (size_t)nvar * (q + (size_t)nvar * b))];
        for (int k = 0; k < ns; ++k)
          sw[k] = (failed || tot[b] < 2) ? qnan : w[k + (size_t)mmw * (g + (size_t)nvar * (q + (size_t)nvar * b))];
      }
    }
  }
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < nvar; ++j)
      for (int i = 0; i < nvar; ++i)
        S.var[i + (size_t)nvar * (j + (size_t)nvar * b)] = vout[b + (size_t)nb * (i + (size_t)nvar * j)];
  K.ws_dev = false;
  return true;
}

// The text factor file (write_weight_store / read_weight_to_store,
// kriging_base.F90:2838-2866, 2896-2917).  Header: nblock, nvar and
// obs_nmax of every variable; then per block the records
//   order nnear(1:ngroups)
//   inear of every group, concatenated
//   variance (nvar x nvar, column-major)
//   weights of every group, concatenated: one record per target variable.
// Weights carry ten decimals (F0.10) and variances ten significant digits,
// as the reference writes them.
bool write_weight_file(const Krige& K) {
  const char* W = "t_kriging_base%write_weight_store";
  FILE* f = std::fopen(K.weight_file.c_str(), "w");
  if (!f) {
    set_error(W, "cannot open weight_file for writing: " + K.weight_file);
    return false;
  }
  const WStore& S = K.ws;
  const int ng = S.ng, nv = S.nv, nm = S.nm;
  auto num = [&](double x, bool fixed) {
    if (std::isnan(x)) std::fputs("  NaN", f);
    else if (std::isinf(x)) std::fputs(x > 0 ? "  Infinity" : "  -Infinity", f);
    else std::fprintf(f, fixed ? "  %.10f" : "  %.9E", x);
  };
  std::fprintf(f, "%12d%12d", S.nb, K.nvar);
  for (int v = 0; v < K.nvar; ++v) std::fprintf(f, "%12d", obs_nmax(K, v));
  std::fputc('\n', f);
  for (int b = 0; b < S.nb; ++b) {
    const int* nn = &S.nnear[(size_t)ng * b];
    std::fprintf(f, "%d", S.order[b]);
    for (int g = 0; g < ng; ++g) std::fprintf(f, "  %d", nn[g]);
    std::fputc('\n', f);
    for (int g = 0; g < ng; ++g)
      for (int k = 0; k < nn[g]; ++k) std::fprintf(f, "  %d", S.inear[k + (size_t)nm * (g + (size_t)ng * b)]);
    std::fputc('\n', f);
    for (int j = 0; j < nv; ++j)
      for (int i = 0; i < nv; ++i) num(S.var[i + (size_t)nv * (j + (size_t)nv * b)], false);
    std::fputc('\n', f);
    for (int q = 0; q < nv; ++q) {
      for (int g = 0; g < ng; ++g)
        for (int k = 0; k < nn[g]; ++k) num(S.weight[k + (size_t)nm * (g + (size_t)ng * (q + (size_t)nv * b))], true);
      std::fputc('\n', f);
    }
  }
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad) {
    set_error(W, "write error on weight_file " + K.weight_file);
    return false;
  }
  return true;
}

// list-directed reading: every read starts on a new record and takes items
// from as many records as it needs; the rest of its last record is skipped,
// and a read of no items consumes one record
struct FactorRecords {
  std::vector<std::string> lines;
  size_t at = 0;
  bool read(size_t n, std::vector<std::string>& tok) {
    tok.clear();
    if (n == 0) return at < lines.size() ? (++at, true) : false;
    while (tok.size() < n) {
      if (at >= lines.size()) return false;
      const std::string& l = lines[at++];
      size_t i = 0;
      while (i < l.size()) {
        while (i < l.size() && (l[i] == ' ' || l[i] == '\t' || l[i] == ',' || l[i] == '\r')) ++i;
        size_t j = i;
        while (j < l.size() && !(l[j] == ' ' || l[j] == '\t' || l[j] == ',' || l[j] == '\r')) ++j;
        if (j > i) tok.push_back(l.substr(i, j - i));
        i = j;
      }
    }
    tok.resize(n);
    return true;
  }
};

bool parse_int(const std::string& t, int& v) {
  char* e = nullptr;
  const long x = std::strtol(t.c_str(), &e, 10);
  if (!e || *e != 0 || x < INT_MIN || x > INT_MAX) return false;
  v = (int)x;
  return true;
}

bool parse_real(std::string t, double& v) {
  for (char& c : t)
    if (c == 'D' || c == 'd') c = 'E';  // Fortran double-precision exponents
  char* e = nullptr;
  v = std::strtod(t.c_str(), &e);
  return e && *e == 0 && !t.empty();
}

// use_old_weight with a weight_file: the whole file is read into the store
// before the block loop (prepare_common, kriging_base.F90:3074-3095)
bool read_weight_file(Krige& K) {
  const char* W = "t_kriging%prepare";
  const char* WR = "t_kriging_base%read_weight_to_store";
  FILE* f = std::fopen(K.weight_file.c_str(), "r");
  if (!f) {
    set_error(W, "cannot open weight_file: " + K.weight_file);
    return false;
  }
  FactorRecords R;
  {
    std::string cur;
    int c;
    while ((c = std::fgetc(f)) != EOF) {
      if (c == '\n') {
        R.lines.push_back(cur);
        cur.clear();
      } else {
        cur.push_back((char)c);
      }
    }
    if (!cur.empty()) R.lines.push_back(cur);
    std::fclose(f);
  }
  const int nvar = K.nvar, nb = K.nblock;
  std::vector<std::string> tok;
  int hb = 0, hv = 0;
  if (!R.read(2 + nvar, tok) || !parse_int(tok[0], hb) || !parse_int(tok[1], hv)) {
    set_error(W, "Failed to read weight_file header.");
    return false;
  }
  bool same = hb == nb && hv == nvar;
  for (int v = 0; v < nvar && same; ++v) {
    int x = 0;
    same = parse_int(tok[2 + v], x) && x == obs_nmax(K, v);
  }
  if (!same) {
    set_error(W, "weight_file dimensions do not match this kriging object.");
    return false;
  }
  WStore& S = K.ws;
  const int nm = store_nmax(K), ng = nvar;
  S.init(nm, ng, nvar, nb);
  auto fail = [&]() {
    set_error(WR, "Unexpected end-of-file or read error in factor file.");
    S = WStore();
    return false;
  };
  for (int b = 0; b < nb; ++b) {
    if (!R.read(1 + ng, tok) || !parse_int(tok[0], S.order[b])) return fail();
    int tot = 0;
    for (int g = 0; g < ng; ++g) {
      int n = 0;
      if (!parse_int(tok[1 + g], n) || n < 0 || n > nm) return fail();
      S.nnear[g + (size_t)ng * b] = n;
      tot += n;
    }
    if (!R.read(tot, tok)) return fail();
    size_t t = 0;
    for (int g = 0; g < ng; ++g)
      for (int k = 0; k < S.nnear[g + (size_t)ng * b]; ++k)
        if (!parse_int(tok[t++], S.inear[k + (size_t)nm * (g + (size_t)ng * b)])) return fail();
    if (!R.read((size_t)nvar * nvar, tok)) return fail();
    for (int j = 0; j < nvar; ++j)
      for (int i = 0; i < nvar; ++i)
        if (!parse_real(tok[i + (size_t)nvar * j], S.var[i + (size_t)nvar * (j + (size_t)nvar * b)])) return fail();
    for (int q = 0; q < nvar; ++q) {
      if (!R.read(tot, tok)) return fail();
      t = 0;
      for (int g = 0; g < ng; ++g)
        for (int k = 0; k < S.nnear[g + (size_t)ng * b]; ++k)
          if (!parse_real(tok[t++], S.weight[k + (size_t)nm * (g + (size_t)ng * (q + (size_t)nvar * b))]))
            return fail();
    }
  }
  K.ws_dev = false;
  return true;
}

// the block loop with use_old_weight: estimates from the stored weights and
// the current observation values, variances from the store
// (kriging_base.F90:2125-2126, 2792-2812, 3407-3503), on the GPU
bool do_replay(Krige& K, bool fetch) {
  const char* W = "t_kriging%prepare";
  if (!K.initialized) {
    set_error(W, "Call initialize() before solve.");
    return false;
  }
  if (K.nvar < 1 || K.nvar > kk::MAXVAR || K.nsim != 0 || K.ndrift < 0 || K.ndrift > kk::MAXDRIFT) {
    set_error(W, "this backend solves kriging and cokriging with nvar <= 4, nsim=0, ndrift <= 4");
    return false;
  }
  for (int v = 0; v < K.nvar; ++v)
    if (K.obs[v].n == 0) {
      set_error(W, "set_obs() needs to be called before solve().");
      return false;
    }
  if (K.nblock == 0) {
    set_error(W, "set_grid() needs to be called before solve().");
    return false;
  }
  if (!K.weight_file.empty()) {
    if (!read_weight_file(K)) return false;
  } else {
    if (!K.ws.alloc) {
      set_error(W, "use_old_weight with no weight_file requires set_weights() to be called first.");
      return false;
    }
    if (!K.ws.stored()) {
      set_error(W, "in-memory weight store is empty. Call set_weights() first.");
      return false;
    }
  }
  WStore& S = K.ws;
  const int nb = K.nblock, nvar = K.nvar;
  if (S.nb != nb || S.ng != nvar || S.nv != nvar) {
    set_error(W, "the weight store does not match this kriging object (nblock, nvar)");
    return false;
  }
  if (!ensure_stream(K)) return false;
  hipStream_t st = K.stream;
  if (!K.ws_dev) {
    if (!upload(K.d_wnn, S.nnear, W) || !upload(K.d_win, S.inear, W) || !upload(K.d_ww, S.weight, W) ||
        !upload(K.d_wvar, S.var, W))
      return false;
    K.ws_dev = true;
  }
  for (int v = 0; v < nvar; ++v)
    if (!upload(K.d_rz[v], K.obs[v].value, W)) return false;
  DBuf& de = nvar == 1 ? K.d_est : K.d_mvest;
  DBuf& dv = nvar == 1 ? K.d_var : K.d_mvvar;
  if (!de.ensure(sizeof(double) * (size_t)nb * nvar) || !dv.ensure(sizeof(double) * (size_t)nb * nvar * nvar)) {
    set_error(W, "device allocation failed");
    return false;
  }
  kk::ReplayArgs r{};
  r.nb = nb;
  r.nm = S.nm;
  r.ng = nvar;
  r.nvar = nvar;
  r.nnear = K.d_wnn.as<int>();
  r.inear = K.d_win.as<int>();
  r.weight = K.d_ww.as<double>();
  r.wvar = K.d_wvar.as<double>();
  for (int v = 0; v < nvar; ++v) {
    r.ovalue[v] = K.d_rz[v].as<double>();
    r.nobs[v] = K.obs[v].n;
    r.sk_mean[v] = K.obs[v].sk_mean;
  }
  r.std_ck = K.std_ck;
  r.unbias = K.unbias;
  r.bound_lo = K.bounds[0];
  r.bound_hi = K.bounds[1];
  r.est = de.as<double>();
  r.var = dv.as<double>();
  HIPCHK(hipEventRecord(K.ev[0], st), W);
  HIPCHK(kk::launch_replay(r, st), W);
  HIPCHK(hipEventRecord(K.ev[4], st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, K.ev[0], K.ev[4]), W);
  for (double& x : K.timing) x = 0.0;
  K.timing[4] = K.timing[12] = ms;
  K.timing[7] = -1;
  for (int& x : K.stats) x = 0;
  if (!fetch) return true;
  if (nvar == 1) {
    K.est.resize(nb);
    K.var.resize(nb);
    HIPCHK(hipMemcpy(K.est.data(), K.d_est.p, sizeof(double) * nb, hipMemcpyDeviceToHost), W);
    HIPCHK(hipMemcpy(K.var.data(), K.d_var.p, sizeof(double) * nb, hipMemcpyDeviceToHost), W);
  } else {
    K.est_mv.resize((size_t)nb * nvar);
    K.var_mv.resize((size_t)nb * nvar * nvar);
    HIPCHK(hipMemcpy(K.est_mv.data(), K.d_mvest.p, sizeof(double) * K.est_mv.size(), hipMemcpyDeviceToHost), W);
    HIPCHK(hipMemcpy(K.var_mv.data(), K.d_mvvar.p, sizeof(double) * K.var_mv.size(), hipMemcpyDeviceToHost), W);
  }
  return true;
}


// Sequential Gaussian simulation, one variable (kriging_base.F90:1226-1316,
// 2055-2181, 3405-3505; kriging.F90:535-612).  The reference's search pool
// is the observations followed by the blocks in path order, each block
// searching the points before it (kdtree2_n_nearest_maxidx); its kriging
// weights depend only on that geometry.  So one inner pass of the ordinary
// pipeline -- a handle of its own over that pool, the blocks in path order
// as targets, the exact search limited per block, co-located simulated
// neighbours dropped, FP64 weights stored -- gives every block's weights and
// variance, and the realisations follow by the recurrence over dependency
// levels (k_sgsim_level).
bool build_search(Krige* K, Obs& O, bool force_tree);

bool do_solve_sgsim(Krige& K, bool fetch) {
  const char* W = "t_kriging_base%solve_base";
  if (K.nvar != 1 || K.st || !K.soff.empty() || K.ndrift != 0 || K.weight_correction || K.store_weight ||
      K.varying_vgm) {
    set_error(W, "sequential simulation is provided for one variable and point targets, without drift, weight "
                 "correction or a weight store");
    return false;
  }
  if (!K.sim_set) {
    set_error("t_kriging%set_search", "set_sim() needs to be called before set_search().");
    return false;
  }
  Obs& O = K.obs[0];
  if (O.n == 0 || !O.search_set || K.nblock == 0 || K.vgm[0].empty()) {
    set_error(W, "set_obs, set_grid, set_vgm and set_search are needed before solve()");
    return false;
  }
  const int nd = K.ndim, nb = K.nblock, nobs = O.n, ns = K.nsim;
  // the pool: observations, then the blocks in path order (no observation
  // error, value 0: their part of each estimate is added per realisation)
  const int np = nobs + nb;
  std::vector<double> pc((size_t)nd * np), pv(np, 0.0), pe(np, 0.0), bc((size_t)nd * nb), rs(nb), ln(nb);
  std::copy(O.coord.begin(), O.coord.end(), pc.begin());
  std::copy(O.value.begin(), O.value.end(), pv.begin());
  std::copy(O.var.begin(), O.var.end(), pe.begin());
  for (int p = 0; p < nb; ++p) {
    const int b = K.sgs_path[p];
    for (int k = 0; k < nd; ++k) {
      bc[k + (size_t)nd * p] = K.bcoord[k + (size_t)nd * b];
      pc[k + (size_t)nd * (nobs + p)] = K.bcoord[k + (size_t)nd * b];
    }
    rs[p] = K.rangescale.empty() ? 1.0 : K.rangescale[b];
    ln[p] = K.localnugget.empty() ? 0.0 : K.localnugget[b];
  }
  int64_t hi = 0;
  if (krige_create(&hi)) return false;
  // the inner handle goes when this returns, keeping any error set here
  struct Drop {
    int64_t h;
    ~Drop() {
      const int e = g_ierr;
      const std::string m = g_msg;
      krige_destroy(&h);
      g_ierr = e;
      g_msg = m;
    }
  } drop{hi};
  const double huge_b[2] = {-DBL_MAX, DBL_MAX};
  if (krige_initialize(hi, nd, 1, 0, K.unbias, 0, K.aniso_search, 0, 0, 1, 0, 0, K.neglect_error, 0, 1, 0, 0, "",
                       huge_b, K.seed))
    return false;
  Krige& I = *get_obj(hi);
  I.precision = kk::PREC_FP64;
  I.pivot_tol = K.pivot_tol;
  I.table_mode = K.table_mode;
  I.search_mode = 0;
  I.reorder = false;
  I.reuse_mode = 0;
  // the pool is set directly: a block may lie on an observation (set_sim
  // extends the coordinates without set_obs's duplicate check)
  Obs& IO = I.obs[0];
  IO = Obs();
  IO.n = np;
  IO.coord = std::move(pc);
  IO.value = std::move(pv);
  IO.var = std::move(pe);
  IO.nmax = IO.nmax_in = O.nmax_in;
  IO.maxdist2 = O.maxdist2;
  IO.sk_mean = O.sk_mean;
  if (krige_set_grid(hi, nb, nd, bc.data(), rs.data(), ln.data())) return false;
  I.vgm = K.vgm;
  IO.sector = O.sector;
  IO.aniso = O.aniso;
  std::memcpy(IO.rot, O.rot, sizeof IO.rot);
  if (!build_search(&I, IO, true)) return false;
  I.simlim = nobs;
  if (!validate_for_solve(I) || !stage_inputs(I) || !run_pipeline(I)) return false;
  // dependency levels: a block's level is one above its simulated neighbours'
  const int mmax = I.mmax;
  std::vector<int> nn(nb), in((size_t)nb * mmax), lev(nb), order(nb);
  hipStream_t st = I.stream;
  HIPCHK(hipMemcpyAsync(nn.data(), I.d_nnear.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipMemcpyAsync(in.data(), I.d_inear.p, sizeof(int) * in.size(), hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  int nlev = 0;
  for (int p = 0; p < nb; ++p) {
    int l = 0;
    for (int k = 0; k < nn[p]; ++k) {
      const int c = in[(size_t)p * mmax + k];
      if (c >= nobs) l = std::max(l, lev[c - nobs] + 1);
    }
    lev[p] = l;
    nlev = std::max(nlev, l + 1);
  }
  std::vector<int> loff(nlev + 1, 0);
  for (int p = 0; p < nb; ++p) ++loff[lev[p] + 1];
  for (int l = 0; l < nlev; ++l) loff[l + 1] += loff[l];
  {
    std::vector<int> fill(loff.begin(), loff.end() - 1);
    for (int p = 0; p < nb; ++p) order[fill[lev[p]]++] = p;
  }
  if (!upload(K.d_sgs_order, order, W) || !upload(K.d_sgs_sample, K.sgs_sample, W) ||
      !K.d_sgs_val.ensure(sizeof(double) * (size_t)ns * nb)) {
    if (!failed()) set_error(W, "device allocation failed (simulation)");
    return false;
  }
  kk::SgsimArgs a{};
  a.nb = nb;
  a.nobs = nobs;
  a.nsim = ns;
  a.mmax = mmax;
  a.unbias = K.unbias;
  a.sk_mean = O.sk_mean;
  a.bound_lo = K.bounds[0];
  a.bound_hi = K.bounds[1];
  a.order = K.d_sgs_order.as<int>();
  a.nnear = I.d_nnear.as<int>();
  a.inear = I.d_inear.as<int>();
  a.weights = I.d_wout.as<double>();
  a.var = I.d_var.as<double>();
  a.ovalue = I.d_ovalue.as<double>();
  a.sample = K.d_sgs_sample.as<double>();
  a.val = K.d_sgs_val.as<double>();
  HIPCHK(hipEventRecord(I.ev[3], st), W);
  for (int l = 0; l < nlev; ++l) HIPCHK(kk::launch_sgsim_level(a, loff[l], loff[l + 1], st), W);
  HIPCHK(hipEventRecord(I.ev[4], st), W);
  std::vector<double> val((size_t)ns * nb), var(nb);
  HIPCHK(hipMemcpyAsync(val.data(), a.val, sizeof(double) * val.size(), hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipMemcpyAsync(var.data(), a.var, sizeof(double) * nb, hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  float sim_ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&sim_ms, I.ev[3], I.ev[4]), W);
  // back to block order (reorder_sgsim_base, kriging_base.F90:2268-2282)
  K.est_sim.assign((size_t)ns * nb, 0.0);
  K.var.assign(nb, 0.0);
  K.est.assign(nb, 0.0);
  for (int p = 0; p < nb; ++p) {
    const int b = K.sgs_path[p];
    for (int s2 = 0; s2 < ns; ++s2) K.est_sim[b + (size_t)nb * s2] = val[s2 + (size_t)ns * p];
    K.var[b] = var[p];
    K.est[b] = val[(size_t)ns * p];
  }
  for (int i = 0; i < (int)(sizeof K.timing / sizeof K.timing[0]); ++i) K.timing[i] = I.timing[i];
  K.timing[13] = sim_ms;
  K.timing[14] = nlev;
  for (int i = 0; i < 5; ++i) K.stats[i] = I.stats[i];
  (void)fetch;
  return true;
}

bool do_solve(Krige& K, bool fetch) {
  K.solved = false;
  if (K.nsim > 0) {
    if (!do_solve_sgsim(K, fetch)) return false;
    K.solved = true;
    return true;
  }
  if (K.use_old_weight) {
    if (!do_replay(K, fetch)) return false;
    K.solved = true;
    return true;
  }
  if (!validate_for_solve(K)) return false;
  if (mv_results(K)) {
    if (!do_solve_mv(K, fetch)) return false;
  } else {
    if (!stage_inputs(K)) return false;
    if (!compute_base_cov(K)) return false;
    if (!run_pipeline(K)) return false;
    if (fetch && !fetch_results(K)) return false;
  }
  // store_weight: the store, then the factor file when one is named
  // (kriging_base.F90:2134, 2161-2162)
  if (K.store_weight && (!store_from_device(K) || (!K.weight_file.empty() && !write_weight_file(K))))
    return false;
  K.solved = true;
  return true;
}

Krige* checked(int64_t h) { return get_obj(h); }

bool check_solved(Krige* K, const char* caller) {
  if (!K->solved) {
    set_error(caller, "solve() has not been called; call solve() before retrieving results");
    return false;
  }
  return true;
}

int unsupported(const char* name, const char* what) {
  set_error(name, what);
  return ierr();
}

// The search structures of one variable (its sector, anisotropy and rotation
// already set): nmax against the point count, search coordinates and the
// kd-tree; force_tree builds the tree even when every point fits in nmax.
bool build_search(Krige* K, Obs& O, bool force_tree) {
  O.nmax = normalize_nmax(O.nmax, O.n);
  O.need_search = O.n > O.nmax || force_tree;
  O.tree = HostTree();
  O.rcoord.clear();
  if (O.aniso || K->st) {
    // search coordinates: spatial part rotated/scaled, space-time time axis
    // scaled by time_at (kriging_st.F90:387-395)
    O.rcoord.resize(O.coord.size());
    kk::Rot3 R;
    std::memcpy(R.m, O.rot, sizeof R.m);
    const int nd = K->ndim;
    for (int i = 0; i < O.n; ++i) {
      const double* c = &O.coord[(size_t)nd * i];
      double* r = &O.rcoord[(size_t)nd * i];
      if (nd == 1) R.apply<1>(c, r);
      else if (nd == 2) R.apply<2>(c, r);
      else R.apply<3>(c, r);
      if (K->st) r[3] = c[3] * O.time_at;
    }
  }
  if (O.need_search) {
    if (!build_tree(O.tree, K->ndim, O.n, (O.aniso || K->st) ? O.rcoord.data() : O.coord.data())) {
      set_error("kdtree2_create", "kd-tree construction did not terminate (degenerate coordinates)");
      return false;
    }
  }
  O.search_set = true;
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// allocation-free pass: the device memory of kk_solve_resident_noalloc is
// this fixed array, one copy per GPU, held until the process ends.  Host code
// addresses it only through the pointer hipGetSymbolAddress returns for the
// current device (the array's own name is a host-side stand-in).
// ---------------------------------------------------------------------------
constexpr size_t KK_NOALLOC_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[KK_NOALLOC_BYTES];

namespace {

constexpr int KK_MAX_DEVICES = 64;
void* g_scratch_dev[KK_MAX_DEVICES] = {nullptr};

// the current device's copy of the scratch array (looked up once per device)
bool scratch_ptr(void** out, const char* W) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev), W);
  if (dev < 0 || dev >= KK_MAX_DEVICES) {
    set_error(W, "device ordinal out of range");
    return false;
  }
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), W);
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return true;
}

// carves the pass buffers of nb blocks (list stride mmax, k search
// candidates) out of base; returns the bytes used (base may be null to size)
size_t carve_pass(char* base, int nb, int mmax, int k, bool reuse, PassBufs& B) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  B.nnear = reinterpret_cast<int*>(take(sizeof(int) * nb));
  B.inear = reinterpret_cast<int*>(take(sizeof(int) * (size_t)nb * mmax));
  B.sflag = reinterpret_cast<int*>(take(sizeof(int) * nb));
  B.xlist = reinterpret_cast<int*>(take(sizeof(int) * nb));
  B.flag = reinterpret_cast<int*>(take(sizeof(int) * nb));
  B.counts = reinterpret_cast<int*>(take(sizeof(int) * PASS_COUNTERS));
  B.lists = reinterpret_cast<int*>(take(sizeof(int) * ((size_t)(kk::NBUCKET + 1) * nb * 2 + 64)));
  B.hcap = std::max(1, std::min(nb, 4096));
  B.hdis = reinterpret_cast<double*>(take(sizeof(double) * (size_t)k * B.hcap));
  B.hidx = reinterpret_cast<int*>(take(sizeof(int) * (size_t)k * B.hcap));
  B.fbcap = (int)std::max<size_t>(1, std::min<size_t>(std::min(nb, 256), (size_t(4) << 20) /
                                                                             (8 * kk::fallback_words(mmax))));
  B.fb = reinterpret_cast<double*>(take(sizeof(double) * B.fbcap * kk::fallback_words(mmax)));
  if (reuse) {
    int* sb = reinterpret_cast<int*>(take(sizeof(int) * (size_t)nb * 7));
    B.rep = sb;
    B.rank = sb ? sb + nb : nullptr;
    B.setcnt = sb ? sb + 2 * (size_t)nb : nullptr;
    B.setbase = sb ? sb + 3 * (size_t)nb : nullptr;
    B.members = sb ? sb + 4 * (size_t)nb : nullptr;
    B.empties = sb ? sb + 5 * (size_t)nb : nullptr;
    B.setitem = sb ? sb + 6 * (size_t)nb : nullptr;
    B.items = reinterpret_cast<int2*>(take(sizeof(int2) * (size_t)nb));
    B.tcap = set_table_slots(nb);
    B.table = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * B.tcap));
    B.hash = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * (size_t)nb));
  }
  return off;
}

}  // namespace

// ===========================================================================
// exported C ABI (everything else in the library has hidden visibility)
// ===========================================================================
#pragma GCC visibility push(default)
extern "C" {

int krige_create(int64_t* handle) {
  clear_error();
  Krige* K = new Krige();
  for (size_t i = 0; i < g_reg.size(); ++i)
    if (!g_reg[i]) {
      g_reg[i] = K;
      *handle = (int64_t)i + 1;
      return 0;
    }
  g_reg.push_back(K);
  *handle = (int64_t)g_reg.size();
  return 0;
}

int krige_ind_create(int64_t* handle) {
  clear_error();
  *handle = 0;
  return unsupported("krige_ind_create", "indicator kriging is not provided by this backend");
}

int krige_ind_set_ncat(int64_t, int) {
  clear_error();
  return unsupported("krige_ind_set_ncat", "indicator kriging is not provided by this backend");
}

int krige_destroy(int64_t* handle) {
  clear_error();
  if (!handle) return 0;
  const int64_t h = *handle;
  if (h >= 1 && h <= (int64_t)g_reg.size() && g_reg[h - 1]) {
    delete g_reg[h - 1];
    g_reg[h - 1] = nullptr;
  }
  *handle = 0;
  return 0;
}

int krige_initialize(int64_t h, int ndim, int nvar, int ndrift, int unbias, int nsim, int anisotropic_search,
                     int weight_correction, int use_old_weight, int store_weight, int cross_validation,
                     int write_mat, int neglect_error, int varying_vgm, int std_ck, int verbose, int pf_cache,
                     const char* weight_file, const double* bounds, int seed) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%initialize";
  if (store_weight && use_old_weight) {
    set_error(W, "store_weight and use_old_weight are mutually exclusive");
    return ierr();
  }
  if (pf_cache && use_old_weight) {
    set_error(W, "pf_cache and use_old_weight are mutually exclusive");
    return ierr();
  }
  if (cross_validation && nsim > 0) {
    set_error(W, "cross_validation and nsim>0 are mutually exclusive");
    return ierr();
  }
  if (ndim < 1 || ndim > 3 || nvar < 1) {
    set_error(W, "ndim must be 1, 2 or 3 and nvar >= 1");
    return ierr();
  }
  const int prec = K->precision;
  const double ptol = K->pivot_tol;
  const bool tmode = K->table_mode;
  // a fresh object; the streams, events and pinned counters move over
  Krige* fresh = new Krige();
  std::swap(fresh->stream, K->stream);
  std::swap(fresh->stream2, K->stream2);
  std::swap(fresh->stream3, K->stream3);
  std::swap(fresh->h_counts, K->h_counts);
  for (size_t i = 0; i < sizeof K->ev / sizeof K->ev[0]; ++i) std::swap(fresh->ev[i], K->ev[i]);
  delete K;
  K = fresh;
  g_reg[h - 1] = K;
  K->precision = prec;
  K->pivot_tol = ptol;
  K->table_mode = tmode;
  K->ndim = ndim;
  K->nvar = nvar;
  K->ndrift = ndrift;
  K->unbias = unbias;
  K->nsim = nsim;
  K->aniso_search = anisotropic_search;
  K->weight_correction = weight_correction;
  K->use_old_weight = use_old_weight;
  K->store_weight = store_weight;
  K->cross_validation = cross_validation;
  K->write_mat = write_mat;
  K->neglect_error = neglect_error;
  K->varying_vgm = varying_vgm;
  K->std_ck = std_ck;
  K->verbose = verbose;
  K->pf_cache = pf_cache;
  // Fortran trim(): trailing blanks are not part of the name
  K->weight_file = weight_file ? weight_file : "";
  while (!K->weight_file.empty() && K->weight_file.back() == ' ') K->weight_file.pop_back();
  if (bounds) {
    K->bounds[0] = bounds[0];
    K->bounds[1] = bounds[1];
  }
  K->seed = seed;
  K->obs.assign(nvar, Obs());
  K->vgm.assign((size_t)nvar * nvar, {});
  K->initialized = true;
  return 0;
}

int krige_set_obs(int64_t h, int ivar, int nobs, int ndim_c, const double* coord, const double* value,
                  const double* variance, int nmax, double maxdist, double sk_mean) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%set_obs";
  if (!K->initialized) {
    set_error(W, "Call initialize() before set_obs.");
    return ierr();
  }
  if (ivar < 1 || ivar > K->nvar) {
    set_error(W, "ivar out of range");
    return ierr();
  }
  if (ndim_c != K->ndim) {
    set_error(W, "nlag /= size(coord, 1) for obs");
    return ierr();
  }
  if (nobs < 1) {
    set_error(W, "coord column count != nobs");
    return ierr();
  }
  // duplicate coordinate tuples are rejected (utils.F90:43-133)
  std::vector<int> perm(nobs);
  for (int i = 0; i < nobs; ++i) perm[i] = i;
  std::sort(perm.begin(), perm.end(), [&](int a, int b) {
    for (int k = 0; k < ndim_c; ++k) {
      double x = coord[k + (size_t)ndim_c * a], y = coord[k + (size_t)ndim_c * b];
      if (x != y) return x < y;
    }
    return a < b;
  });
  for (int i = 0; i + 1 < nobs; ++i) {
    bool same = true;
    for (int k = 0; k < ndim_c && same; ++k)
      same = coord[k + (size_t)ndim_c * perm[i]] == coord[k + (size_t)ndim_c * perm[i + 1]];
    if (same) {
      set_error(W, "ERROR: Duplicate coordinate found! Station " + std::to_string(perm[i] + 1) + " and Station " +
                       std::to_string(perm[i + 1] + 1) + " share identical coordinates.");
      return ierr();
    }
  }
  Obs& O = K->obs[ivar - 1];
  {
    // the variable's increment observations are set apart (set_grad)
    Obs fresh;
    fresh.ngrad = O.ngrad;
    fresh.gx1.swap(O.gx1);
    fresh.gx2.swap(O.gx2);
    fresh.gval.swap(O.gval);
    fresh.gvar.swap(O.gvar);
    fresh.gdrift.swap(O.gdrift);
    O = std::move(fresh);
  }
  O.n = nobs;
  K->perm_ok = false;
  O.coord.assign(coord, coord + (size_t)ndim_c * nobs);
  O.value.assign(value, value + nobs);
  O.var.assign(nobs, 0.0);
  if (variance) O.var.assign(variance, variance + nobs);
  O.nmax = nmax;
  O.nmax_in = nmax;
  O.maxdist2 = maxdist * maxdist;
  O.sk_mean = sk_mean;
  if (K->ndrift > 0) O.drift.assign((size_t)K->ndrift * nobs, 0.0);
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  K->solved = false;
  return 0;
}

// external drift values at variable ivar's observations, (ndrift, nobs)
// column-major (set_obs_drift_base, kriging_base.F90:756-791)
int krige_set_obs_drift(int64_t h, int ivar, int ndrift_c, int nobs, const double* drift) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%set_obs_drift";
  if (!K->initialized) {
    set_error(W, "Call initialize() before set_obs_drift.");
    return ierr();
  }
  if (ivar < 1 || ivar > K->nvar) {
    set_error(W, "ivar out of range: " + std::to_string(ivar) + " not in 1.." + std::to_string(K->nvar));
    return ierr();
  }
  Obs& O = K->obs[ivar - 1];
  if (O.n == 0) {
    set_error(W, "Observation needs to be set before adding drift.");
    return ierr();
  }
  if (K->ndrift == 0) {
    set_error(W, "Observation drift is specified but ndrift==0");
    return ierr();
  }
  if (ndrift_c != K->ndrift) {
    set_error(W, "size(drift, 1) /= ndrift");
    return ierr();
  }
  if (nobs != O.n) {
    set_error(W, "size(drift, 2) /= nobs");
    return ierr();
  }
  O.drift.assign(drift, drift + (size_t)ndrift_c * nobs);
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  K->solved = false;
  return 0;
}

int krige_update_obs_value(int64_t h, int ivar, int nobs, const double* value) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (ivar < 1 || ivar > K->nvar || K->obs[ivar - 1].n != nobs) {
    set_error("krige_update_obs_value", "nobs does not match the previous set_obs call");
    return ierr();
  }
  K->obs[ivar - 1].value.assign(value, value + nobs);
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  // a filled weight store and a finished solve switch the next solve to
  // the stored weights (kriging_base.F90:669-678)
  if (K->ws.alloc && K->ws.stored() && K->solved) {
    const std::vector<double>& e = !mv_results(*K) ? K->est : K->est_mv;
    bool all_nan = !e.empty();
    for (double x : e) all_nan = all_nan && std::isnan(x);
    if (!all_nan) {
      K->use_old_weight = true;
      K->store_weight = false;
    }
  }
  return 0;
}

int krige_set_nscore(int64_t h, int, double, double, int, int, double, double, int, const double*) {
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_set_nscore", "normal-score transform (SGSIM) is not provided by this backend");
}

int krige_reset_vgm(int64_t h, int ivar, int jvar) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (ivar < 1 || jvar < 1 || ivar > K->nvar || jvar > K->nvar) return 0;
  K->vgm[(ivar - 1) + (size_t)K->nvar * (jvar - 1)].clear();
  K->vgm[(jvar - 1) + (size_t)K->nvar * (ivar - 1)].clear();
  K->model_ok = false;
  return 0;
}

int krige_set_vgm(int64_t h, int ivar, int jvar, const char* vtype, double nugget, double sill, double a_major,
                  double a_minor1, double a_minor2, double azimuth, double dip, double plunge, int is_product) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging%set_vgm";
  if (!K->initialized) {
    set_error(W, "Call initialize() before set_vgm.");
    return ierr();
  }
  if (ivar < 1 || jvar < 1 || ivar > K->nvar || jvar > K->nvar) {
    set_error(W, "variable index out of range");
    return ierr();
  }
  if (jvar < ivar) {
    set_error(W, "jvar must be >= ivar to set the upper triangle of the variogram matrix");
    return ierr();
  }
  const int id = vtype_from_str(vtype ? vtype : "sph");
  if (id < 0) {
    set_error("vtype_from_str", std::string("unknown vtype: ") + (vtype ? vtype : ""));
    return ierr();
  }
  auto& v = K->vgm[(ivar - 1) + (size_t)K->nvar * (jvar - 1)];
  if (is_product && v.empty()) {
    set_error("vgm_struct%add",
              "first structure cannot be a product member (product=.true. requires a preceding structure)");
    return ierr();
  }
  if ((int)v.size() >= kk::MAXS) {
    set_error("vgm_struct%add", "exceeded maxvgm nested structures");
    return ierr();
  }
  VStruct s{id, is_product ? 1 : 0, nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge};
  v.push_back(s);
  if (ivar != jvar) K->vgm[(jvar - 1) + (size_t)K->nvar * (ivar - 1)].push_back(s);
  K->solved = false;
  K->model_ok = false;
  return 0;
}

int krige_set_vgm_block(int64_t h, int, int, int, const char*, double, double, double, double, double, double,
                        double, double, int) {
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_set_vgm_block", "spatially varying variograms are not provided by this backend");
}

static int set_grid_common(Krige* K, int nblock, const double* rangescale, const double* localnugget) {
  K->nblock = nblock;
  K->rangescale.clear();
  K->localnugget.clear();
  bool rs_unit = true, ln_zero = true;
  for (int b = 0; rangescale && b < nblock; ++b) rs_unit &= rangescale[b] == 1.0;
  for (int b = 0; localnugget && b < nblock; ++b) ln_zero &= localnugget[b] == 0.0;
  if (rangescale && !rs_unit) K->rangescale.assign(rangescale, rangescale + nblock);
  if (localnugget && !ln_zero) K->localnugget.assign(localnugget, localnugget + nblock);
  K->cross_validation = false;
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  K->perm_ok = false;
  K->solved = false;
  K->bdrift.assign((size_t)K->ndrift * K->nvar * nblock, 0.0);
  return 0;
}

int krige_set_grid(int64_t h, int ngrid, int ndim_c, const double* coord, const double* rangescale,
                   const double* localnugget) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%set_grid_point";
  if (!K->initialized || K->obs.empty() || K->obs[0].n == 0) {
    set_error(W, "set_obs() needs to be called before set_grid().");
    return ierr();
  }
  if (ndim_c != K->ndim) {
    set_error(W, "nlag /= size(coord, 1) for self%grid");
    return ierr();
  }
  K->bcoord.assign(coord, coord + (size_t)ndim_c * ngrid);
  K->soff.clear();
  K->scoord.clear();
  K->sweight.clear();
  return set_grid_common(K, ngrid, rangescale, localnugget);
}

int krige_set_grid_block(int64_t h, int block_type, int ngrid, int ndim_c, const double* coord, int nblock,
                         const int* nblockpnt, const double* pointweight, const double* blocksize,
                         const double* rangescale, const double* localnugget) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%set_grid_block";
  if (!K->initialized || K->obs.empty() || K->obs[0].n == 0) {
    set_error(W, "set_obs() needs to be called before set_grid().");
    return ierr();
  }
  if (ndim_c != K->ndim) {
    set_error(W, "nlag /= size(coord, 1) for self%grid");
    return ierr();
  }
  const int nd = ndim_c;
  if (block_type == -4) {
    // Gauss-Legendre 4^ndim sub-points (gaussian_quadrature.f90:3-41; the
    // decimal-truncated constants are the reference's)
    const double frac[4] = {0.4305681558, 0.1699905218, -0.1699905218, -0.4305681558};
    const double wgq4[4] = {0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226};
    const int ngq = nd == 1 ? 4 : nd == 2 ? 16 : 64;
    K->bcoord.assign(coord, coord + (size_t)nd * ngrid);
    K->soff.resize(ngrid + 1);
    K->scoord.resize((size_t)nd * ngrid * ngq);
    K->sweight.resize((size_t)ngrid * ngq);
    for (int b = 0; b < ngrid; ++b) {
      K->soff[b] = b * ngq;
      for (int g = 0; g < ngq; ++g) {
        const int ii[3] = {g % 4, (g / 4) % 4, g / 16};
        double w = 1.0;
        for (int k = 0; k < nd; ++k) {
          K->scoord[(size_t)nd * ((size_t)b * ngq + g) + k] =
              coord[k + (size_t)nd * b] + frac[ii[k]] * blocksize[k + (size_t)nd * b];
          w = w * wgq4[ii[k]];
        }
        K->sweight[(size_t)b * ngq + g] = w;
      }
    }
    K->soff[ngrid] = ngrid * ngq;
    return set_grid_common(K, ngrid, rangescale, localnugget);
  }
  // user-supplied sub-points: centroid = weighted mean (kriging_base.F90:1500-1563)
  int tot = 0;
  for (int b = 0; b < nblock; ++b) {
    if (nblockpnt[b] <= 0) {
      set_error("set_grid_centroid_common", "nblockpnt entries must be positive");
      return ierr();
    }
    tot += nblockpnt[b];
  }
  if (tot != ngrid) {
    set_error("set_grid_centroid_common", "gridcoord columns != sum(nblockpnt)");
    return ierr();
  }
  for (int b = 1; b < nblock; ++b)
    if (nblockpnt[b] != nblockpnt[0]) {
      set_error(W, "blocks with different sub-point counts are not supported by this backend");
      return ierr();
    }
  K->soff.resize(nblock + 1);
  K->scoord.assign(coord, coord + (size_t)nd * ngrid);
  K->sweight.resize(ngrid);
  K->bcoord.resize((size_t)nd * nblock);
  int off = 0;
  for (int b = 0; b < nblock; ++b) {
    const int nn = nblockpnt[b];
    K->soff[b] = off;
    double denom = 0.0;
    for (int i = 0; i < nn; ++i) {
      K->sweight[off + i] = pointweight ? pointweight[off + i] : 1.0 / nn;
      denom = denom + K->sweight[off + i];
    }
    if (!(denom > 0.0)) {
      set_error("set_grid_centroid_common", "sum(pointweight) must be positive for every block");
      return ierr();
    }
    for (int k = 0; k < nd; ++k) {
      double s = 0.0;
      for (int i = 0; i < nn; ++i) s = s + coord[k + (size_t)nd * (off + i)] * K->sweight[off + i];
      K->bcoord[k + (size_t)nd * b] = s / denom;
    }
    off += nn;
  }
  K->soff[nblock] = off;
  if (nblockpnt[0] == 1) {
    K->soff.clear();
    K->scoord.clear();
    K->sweight.clear();
  }
  return set_grid_common(K, nblock, rangescale, localnugget);
}

int krige_set_grid_cv(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%set_grid_cv";
  if (!K->initialized || K->obs.empty() || K->obs[0].n == 0) {
    set_error(W, "set_obs() needs to be called before set_grid_cv().");
    return ierr();
  }
  set_grid_common(K, K->obs[0].n, nullptr, nullptr);
  K->cross_validation = true;
  K->perm_ok = false;
  K->bcoord.clear();
  K->soff.clear();
  // kriging_base.F90:1167: one extra neighbour so self can be excluded
  Obs& O = K->obs[0];
  if (O.nmax > 0) O.nmax = (int)((unsigned)O.nmax + 1u);
  // the blocks are the observations: their drift is variable 1's, for every
  // target (finish_grid_setup_common with copy_cv_drift, kriging_base.F90:1358-1366)
  for (int b = 0; b < O.n; ++b)
    for (int q = 0; q < K->nvar; ++q)
      for (int k = 0; k < K->ndrift; ++k)
        K->bdrift[k + (size_t)K->ndrift * (q + (size_t)K->nvar * b)] = O.drift[k + (size_t)K->ndrift * b];
  return 0;
}

// external drift values at the blocks, (ndrift, nblocks) column-major, for
// target variable ivar, or for every target when ivar < 0
// (set_grid_drift_base, kriging_base.F90:1173-1220)
int krige_set_grid_drift(int64_t h, int ivar, int ndrift_c, int nblocks, const double* drift) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_base%set_grid_drift";
  if (!K->initialized) {
    set_error(W, "Call initialize() before set_grid_drift.");
    return ierr();
  }
  if (K->nblock == 0) {
    set_error(W, "Grid needs to be set before adding drift.");
    return ierr();
  }
  if (K->ndrift == 0) {
    set_error(W, "grid/block drift is specified but ndrift==0");
    return ierr();
  }
  if (ndrift_c != K->ndrift) {
    set_error(W, "size(drift, 1) /= ndrift");
    return ierr();
  }
  if (nblocks != K->nblock) {
    set_error(W, "size(drift, 2) /= block%n; one drift value per block, not per grid node");
    return ierr();
  }
  if (ivar == 0 || ivar > K->nvar) {
    set_error(W, "ivar must be 1..nvar or < 0 (broadcast)");
    return ierr();
  }
  const int nd = K->ndrift, nv = K->nvar;
  for (int b = 0; b < nblocks; ++b)
    for (int q = 0; q < nv; ++q)
      if (ivar < 0 || q == ivar - 1)
        for (int k = 0; k < nd; ++k) K->bdrift[k + (size_t)nd * (q + (size_t)nv * b)] = drift[k + (size_t)nd * b];
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  K->solved = false;
  return 0;
}

int krige_set_sim(int64_t h, int nblocks, const void* randpath, int nsim_c, int nvar_c, const void* sample) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  // set_sim_base, kriging_base.F90:1226-1316
  const char* W = "t_kriging_base%set_sim";
  if (K->nsim == 0) return 0;
  if (!K->initialized) {
    set_error(W, "call initialize() before set_sim().");
    return ierr();
  }
  if (K->nblock == 0) {
    set_error(W, "Grid needs to be set first.");
    return ierr();
  }
  for (const Obs& O : K->obs) {
    if (O.n == 0) {
      set_error(W, "Observations need to be set first.");
      return ierr();
    }
    if (O.nmax_in > (int)(INT_MAX * 0.001)) {
      set_error(W, "`nmax` must be set for all observations for SGSIM.");
      return ierr();
    }
  }
  const int nb = K->nblock, ns = K->nsim;
  // the library's own generator when the caller gives no path or draws
  // (xoshiro256** seeded by splitmix64 from seed; normals by Box-Muller): the
  // reference's Fortran generator is not reproduced
  uint64_t st[4], z = (uint64_t)(int64_t)K->seed;
  for (auto& w : st) {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    w = x ^ (x >> 31);
  }
  auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
  auto next = [&]() {
    const uint64_t r = rotl(st[1] * 5, 7) * 9, t = st[1] << 17;
    st[2] ^= st[0];
    st[3] ^= st[1];
    st[1] ^= st[2];
    st[0] ^= st[3];
    st[2] ^= t;
    st[3] = rotl(st[3], 45);
    return r;
  };
  auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  std::vector<int> path(nb);
  if (randpath) {
    if (nblocks != nb) {
      set_error(W, "size(randpath) /= nblock");
      return ierr();
    }
    const int* rp = static_cast<const int*>(randpath);
    std::vector<char> seen(nb, 0);
    for (int p = 0; p < nb; ++p) {
      if (rp[p] < 1 || rp[p] > nb || seen[rp[p] - 1]) {
        set_error(W, "randpath must be a 1-based permutation of 1..nblocks");
        return ierr();
      }
      seen[rp[p] - 1] = 1;
      path[p] = rp[p] - 1;
    }
  } else {
    for (int p = 0; p < nb; ++p) path[p] = p;
    for (int p = nb - 1; p > 0; --p) std::swap(path[p], path[(int)(unif() * (p + 1)) % (p + 1)]);
  }
  std::vector<double> draws((size_t)ns * nb);
  if (sample) {
    if (nsim_c != ns) {
      set_error(W, "size(sample, 1) /= nsim");
      return ierr();
    }
    if (nvar_c != K->nvar) {
      set_error(W, "size(sample, 2) /= nvar");
      return ierr();
    }
    if (nblocks != nb) {
      set_error(W, "size(sample, 3) /= nblock");
      return ierr();
    }
    // (nsim, nvar, nblock) column-major: the first variable's draws
    const double* sp = static_cast<const double*>(sample);
    for (int p = 0; p < nb; ++p)
      for (int s2 = 0; s2 < ns; ++s2) draws[s2 + (size_t)ns * p] = sp[s2 + (size_t)ns * K->nvar * p];
  } else {
    for (size_t i = 0; i < draws.size(); ++i) {
      const double u1 = unif(), u2 = unif();
      draws[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPI * u2);
    }
  }
  K->sgs_path = std::move(path);
  K->sgs_sample = std::move(draws);
  K->sim_set = true;
  K->solved = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  return 0;
}

int krige_set_search(int64_t h, int ivar, double anis1, double anis2, double azimuth, double dip, double plunge,
                     int sector_search) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging%set_search";
  if (!K->initialized) {
    set_error(W, "Call initialize() before set_search.");
    return ierr();
  }
  if (ivar < 1 || ivar > K->nvar) {
    set_error(W, "ivar out of range");
    return ierr();
  }
  Obs& O = K->obs[ivar - 1];
  if (O.n == 0) {
    set_error(W, "set_obs() needs to be called before set_search().");
    return ierr();
  }
  if (sector_search && K->ndim < 2) {
    set_error(W, "sector search needs ndim >= 2");
    return ierr();
  }
  O.sector = sector_search != 0;
  O.aniso = (std::fabs(anis1 - 1.0) > 1e-10 || std::fabs(anis2 - 1.0) > 1e-10) && K->aniso_search;
  std::fill(O.rot, O.rot + 9, 0.0);
  O.rot[0] = O.rot[4] = O.rot[8] = 1.0;
  if (O.aniso) calc_rotmat(azimuth, dip, plunge, anis1, anis2, O.rot);
  if (!build_search(K, O, false)) return ierr();
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  K->solved = false;
  return 0;
}


int krige_set_grad(int64_t h, int ivar, int ngrad, int ndim_c, const double* coord1, const double* coord2,
                   const double* grad_val, const double* variance, int ndrift_c, const double* drift_ext) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  // kriging_capi.F90:472-506 -> set_grad / reset_grad, kriging_base.F90:2591-2690
  const char* W = "t_kriging_base%set_grad";
  if (!K->initialized) {
    set_error(W, "Call initialize() before set_grad.");
    return ierr();
  }
  if (ivar < 1 || ivar > K->nvar) {
    set_error(W, "ivar out of range: " + std::to_string(ivar) + " not in 1.." + std::to_string(K->nvar));
    return ierr();
  }
  Obs& O = K->obs[ivar - 1];
  K->dev_inputs_ok = false;
  K->lists_ok = false;
  K->pass_empty = -1;
  K->solved = false;
  if (ngrad <= 0) {
    O.ngrad = 0;
    O.gx1.clear();
    O.gx2.clear();
    O.gval.clear();
    O.gvar.clear();
    O.gdrift.clear();
    return 0;
  }
  if (ndim_c != K->ndim) {
    set_error(W, "size(coord,1) /= nlag or size(coord2,1) /= nlag");
    return ierr();
  }
  if (ndrift_c > 0 && ndrift_c != K->ndrift) {
    set_error(W, "size(drift_ext,1) /= ndrift or size(drift_ext,2) /= ngrad");
    return ierr();
  }
  if (K->ndrift > 0 && ndrift_c <= 0) {
    set_error(W, "drift_ext must be provided when ndrift > 0");
    return ierr();
  }
  if (grad_total(*K) - O.ngrad + ngrad > kk::MAXGRAD) {
    set_error(W, "more than " + std::to_string(kk::MAXGRAD) +
                     " increment observations over all variables are not supported by the GPU solver");
    return ierr();
  }
  O.ngrad = ngrad;
  O.gx1.assign(coord1, coord1 + (size_t)ndim_c * ngrad);
  O.gx2.assign(coord2, coord2 + (size_t)ndim_c * ngrad);
  O.gval.assign(grad_val, grad_val + ngrad);
  O.gvar.assign(ngrad, 0.0);
  if (variance) O.gvar.assign(variance, variance + ngrad);
  O.gdrift.assign((size_t)std::max(K->ndrift, 1) * ngrad, 0.0);
  if (K->ndrift > 0) O.gdrift.assign(drift_ext, drift_ext + (size_t)K->ndrift * ngrad);
  return 0;
}

int krige_prepare(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  validate_for_solve(*K);
  return ierr();
}

int krige_solve(int64_t h, int nthread, int ncache) {
  clear_error();
  (void)nthread;
  (void)ncache;
  Krige* K = checked(h);
  if (!K) return ierr();
  do_solve(*K, true);
  return ierr();
}

int krige_get_nblocks(int64_t h, int* n) {
  clear_error();
  *n = 0;
  Krige* K = checked(h);
  if (!K) return ierr();
  *n = K->nblock;
  return 0;
}

int krige_get_nsim(int64_t h, int* n) {
  clear_error();
  *n = 0;
  Krige* K = checked(h);
  if (!K) return ierr();
  *n = K->nsim;
  return 0;
}

int krige_get_block_coord(int64_t h, int ndim_c, int nblocks, double* out) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const std::vector<double>& bc = K->cross_validation ? K->obs[0].coord : K->bcoord;
  for (int b = 0; b < nblocks && b < K->nblock; ++b)
    for (int k = 0; k < ndim_c; ++k) out[k + (size_t)ndim_c * b] = bc[k + (size_t)K->ndim * b];
  return 0;
}

int krige_get_estimate(int64_t h, int nsim_c, int nblocks, double* out) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!check_solved(K, "krige_get_estimate")) return ierr();
  for (int s = 0; s < nsim_c; ++s)
    for (int b = 0; b < nblocks; ++b)
      out[b + (size_t)nblocks * s] = K->nsim > 0 ? K->est_sim[b + (size_t)K->nblock * std::min(s, K->nsim - 1)]
                                    : mv_results(*K) ? K->est_mv[b] : K->est[b];
  return 0;
}

int krige_get_estimate_all(int64_t h, int nblocks, int nvar_c, int nsim_c, double* out) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!check_solved(K, "krige_get_estimate_all")) return ierr();
  for (int s = 0; s < nsim_c; ++s)
    for (int v = 0; v < nvar_c; ++v)
      for (int b = 0; b < nblocks; ++b)
        out[b + (size_t)nblocks * (v + (size_t)nvar_c * s)] =
            K->nsim > 0 ? K->est_sim[b + (size_t)K->nblock * std::min(s, K->nsim - 1)]
            : mv_results(*K) ? K->est_mv[b + (size_t)nblocks * v] : K->est[b];
  return 0;
}

int krige_get_variance(int64_t h, int nblocks, double* out) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!check_solved(K, "krige_get_variance")) return ierr();
  for (int b = 0; b < nblocks; ++b) out[b] = mv_results(*K) ? K->var_mv[b] : K->var[b];
  return 0;
}

int krige_get_variance_all(int64_t h, int nblocks, int nvar_c, double* out) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!check_solved(K, "krige_get_variance_all")) return ierr();
  for (int i = 0; i < nvar_c; ++i)
    for (int j = 0; j < nvar_c; ++j)
      for (int b = 0; b < nblocks; ++b)
        out[b + (size_t)nblocks * (i + (size_t)nvar_c * j)] =
            mv_results(*K) ? K->var_mv[b + (size_t)nblocks * (i + (size_t)K->nvar * j)] : K->var[b];
  return 0;
}

int krige_get_factor_info(int64_t h, int*, int*, int*) {
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_get_factor_info", "persistent factor cache is not provided by this backend");
}

int krige_get_factor_matrices(int64_t h, int, int, double*, double*, double*) {
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_get_factor_matrices", "persistent factor cache is not provided by this backend");
}

int krige_get_factor_system(int64_t h, int, int, int, double*, double*) {
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_get_factor_system", "persistent factor cache is not provided by this backend");
}

int krige_get_last_error(char* buffer, int nbuf) {
  if (nbuf <= 0) return g_ierr;
  const int n = std::min((int)g_msg.size(), nbuf - 1);
  std::memcpy(buffer, g_msg.data(), n);
  std::memset(buffer + n, 0, nbuf - n);
  return g_ierr;
}

void krige_solver_stats(int64_t h, int* out) {
  if (h < 1 || h > (int64_t)g_reg.size() || !g_reg[h - 1]) {
    for (int i = 0; i < 5; ++i) out[i] = 0;
    return;
  }
  for (int i = 0; i < 5; ++i) out[i] = g_reg[h - 1]->stats[i];
}

intptr_t krige_to_str(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return 0;
  char buf[512];
  std::snprintf(buf, sizeof buf,
                " MI355X kriging backend\n ndim=%d nvar=%d unbias=%d nblock=%d nobs=%d precision=%s table=%d\n",
                K->ndim, K->nvar, K->unbias, K->nblock, K->obs.empty() ? 0 : K->obs[0].n,
                K->precision == kk::PREC_FP64 ? "fp64" : "fp32+refine", (int)K->table_mode);
  K->info = buf;
  return (intptr_t)K->info.c_str();
}

void krige_get_max_threads(int* n) { *n = 1; }
void krige_get_num_threads(int* n) { *n = 1; }

// ---- weight-store views (neighbour lists of the last solve) ------------
int krige_free_weight_store(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  K->ws = WStore();
  K->ws_dev = false;
  K->d_wnn.release();
  K->d_win.release();
  K->d_ww.release();
  K->d_wvar.release();
  return 0;
}

// the getters copy out(1:n1, 1:n2, ...) of the store, as the reference does;
// dimensions beyond the store's are an error here
static bool store_ready(Krige* K, const char* W) {
  if (!K->ws.alloc) {
    set_error(W, "Weight store not allocated");
    return false;
  }
  return true;
}

static bool within(const char* W, std::initializer_list<std::pair<int, int>> dims) {
  for (const auto& d : dims)
    if (d.first < 0 || d.first > d.second) {
      set_error(W, "requested dimensions exceed the weight store");
      return false;
    }
  return true;
}

int krige_get_weight_dims(int64_t h, int* nm, int* ng, int* nb) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!store_ready(K, "krige_get_weight_dims")) return ierr();
  *nm = K->ws.nm;
  *ng = K->ws.ng;
  *nb = K->ws.nb;
  return 0;
}

int krige_get_weight_nnear(int64_t h, int ngroups_c, int nblock_c, int* out) {
  clear_error();
  const char* W = "krige_get_weight_nnear";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!store_ready(K, W)) return ierr();
  const WStore& S = K->ws;
  if (!within(W, {{ngroups_c, S.ng}, {nblock_c, S.nb}})) return ierr();
  for (int b = 0; b < nblock_c; ++b)
    for (int g = 0; g < ngroups_c; ++g) out[g + (size_t)ngroups_c * b] = S.nnear[g + (size_t)S.ng * b];
  return 0;
}

int krige_get_weight_inear(int64_t h, int nmax_c, int ngroups_c, int nblock_c, int* out) {
  clear_error();
  const char* W = "krige_get_weight_inear";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!store_ready(K, W)) return ierr();
  const WStore& S = K->ws;
  if (!within(W, {{nmax_c, S.nm}, {ngroups_c, S.ng}, {nblock_c, S.nb}})) return ierr();
  for (int b = 0; b < nblock_c; ++b)
    for (int g = 0; g < ngroups_c; ++g)
      for (int k = 0; k < nmax_c; ++k)
        out[k + (size_t)nmax_c * (g + (size_t)ngroups_c * b)] = S.inear[k + (size_t)S.nm * (g + (size_t)S.ng * b)];
  return 0;
}

int krige_get_weight_data(int64_t h, int nmax_c, int ngroups_c, int nvar_c, int nblock_c, double* out) {
  clear_error();
  const char* W = "krige_get_weight_data";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!store_ready(K, W)) return ierr();
  const WStore& S = K->ws;
  if (!within(W, {{nmax_c, S.nm}, {ngroups_c, S.ng}, {nvar_c, S.nv}, {nblock_c, S.nb}})) return ierr();
  for (int b = 0; b < nblock_c; ++b)
    for (int q = 0; q < nvar_c; ++q)
      for (int g = 0; g < ngroups_c; ++g)
        for (int k = 0; k < nmax_c; ++k)
          out[k + (size_t)nmax_c * (g + (size_t)ngroups_c * (q + (size_t)nvar_c * b))] =
              S.weight[k + (size_t)S.nm * (g + (size_t)S.ng * (q + (size_t)S.nv * b))];
  return 0;
}

int krige_get_weight_var(int64_t h, int nvar_c, int nblock_c, double* out) {
  clear_error();
  const char* W = "krige_get_weight_var";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!store_ready(K, W)) return ierr();
  const WStore& S = K->ws;
  if (!within(W, {{nvar_c, S.nv}, {nblock_c, S.nb}})) return ierr();
  for (int b = 0; b < nblock_c; ++b)
    for (int j = 0; j < nvar_c; ++j)
      for (int i = 0; i < nvar_c; ++i)
        out[i + (size_t)nvar_c * (j + (size_t)nvar_c * b)] = S.var[i + (size_t)S.nv * (j + (size_t)S.nv * b)];
  return 0;
}

// alloc_weight_store's preconditions and set_weights' shape checks
// (kriging_base.F90:2703-2733, 2930-2977), then the store takes the arrays
// as given and the next solve replays them (use_old_weight).  Neighbour
// counts outside 0..nmax are rejected here (the replay runs on the GPU);
// the neighbour ids are checked against the observations at the solve.
int krige_set_weights(int64_t h, int nmax_c, int ngroups_c, int nvar_c, int nblock_c, const int* nnear,
                      const int* inear, const double* weight, const int* order, const double* var) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* WA = "t_kriging_base%alloc_weight_store";
  const char* W = "t_kriging_base%set_weights";
  if (K->nblock == 0) {
    set_error(WA, "call set_grid() before alloc_weight_store()");
    return ierr();
  }
  if (!K->initialized) {
    set_error(WA, "call initialize() before alloc_weight_store()");
    return ierr();
  }
  const int nm = K->ws.alloc ? K->ws.nm : store_nmax(*K);
  const int ng = K->nvar, nb = K->nblock, nv = K->nvar;
  if (nm <= 0) {
    set_error(WA, "call set_obs() before alloc_weight_store() so nmax is set");
    return ierr();
  }
  if (ngroups_c != ng || nblock_c != nb) {
    set_error(W, "nnear shape mismatch");
    return ierr();
  }
  if (nmax_c != nm) {
    set_error(W, "inear shape mismatch");
    return ierr();
  }
  if (nvar_c != nv) {
    set_error(W, "weight shape mismatch");
    return ierr();
  }
  if (!nnear || !inear || !weight) {
    set_error(W, "nnear, inear and weight are required");
    return ierr();
  }
  for (size_t i = 0; i < (size_t)ng * nb; ++i)
    if (nnear[i] < 0 || nnear[i] > nm) {
      set_error(W, "nnear out of range 0.." + std::to_string(nm) + " at block " + std::to_string(i / ng + 1));
      return ierr();
    }
  WStore& S = K->ws;
  S.init(nm, ng, nv, nb);
  std::memcpy(S.nnear.data(), nnear, sizeof(int) * S.nnear.size());
  std::memcpy(S.inear.data(), inear, sizeof(int) * S.inear.size());
  std::memcpy(S.weight.data(), weight, sizeof(double) * S.weight.size());
  if (order) std::memcpy(S.order.data(), order, sizeof(int) * S.order.size());
  if (var) std::memcpy(S.var.data(), var, sizeof(double) * S.var.size());
  K->ws_dev = false;
  K->use_old_weight = true;
  return 0;
}

// ---- extensions of this backend (no counterpart in the reference) -------

// precision: 0 = FP64 solve, 1 = FP32 factorisation + FP64 refinement.
// pivot_tol: FP32 pivot ratio below which a node is re-solved in FP64.
// table_mode: 1 = lookup-table covariance (the reference default), 0 = analytic.
// ---- space-time (kriging_st_capi.f90) --------------------------------------
// The space-time object shares the handle registry; its coordinates are
// (x, y, z, t): spatial ndim 3 (smaller ndim is zero-padded), time last.

int krige_st_create(int64_t* handle) {
  if (krige_create(handle)) return ierr();
  Krige* K = checked(*handle);
  K->st = true;
  K->ndim = 4;
  return 0;
}

int krige_st_destroy(int64_t* handle) { return krige_destroy(handle); }

int krige_st_initialize(int64_t h, int nvar, int ndrift, int unbias, int nsim, int anisotropic_search,
                        int weight_correction, int use_old_weight, int store_weight, int cross_validation,
                        int write_mat, int neglect_error, int verbose, const char* weight_file,
                        const double* bounds, int seed) {
  clear_error();
  if (!checked(h)) return ierr();
  if (krige_initialize(h, 3, nvar, ndrift, unbias, nsim, anisotropic_search, weight_correction, use_old_weight,
                       store_weight, cross_validation, write_mat, neglect_error, 0, 1, verbose, 0, weight_file,
                       bounds, seed))
    return ierr();
  Krige* K = checked(h);
  K->st = true;
  K->ndim = 4;
  return 0;
}

int krige_st_set_st_model(int64_t h, const char* model, const char* transform, double at, double time_nugget,
                          double time_sill, double k_ps) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const std::string m = model ? model : "";
  if (m != "product_sum" && m != "sum_metric") {
    set_error("krige_st_set_st_model", "model must be \"sum_metric\" or \"product_sum\", got: " + m);
    return ierr();
  }
  // transform id and checks: kriging_st_capi.f90:139-150, kriging_st.F90:143-165
  const int tid = vtype_from_str(transform ? transform : "");
  if (tid < 0) {
    set_error("vtype_from_str", std::string("unknown vtype: ") + (transform ? transform : ""));
    return ierr();
  }
  if (tid < kk::NUG || tid > kk::LIN || tid == kk::HOL) {
    set_error("set_st_model", "transform must be a monotone vgmfunc id: nug, sph, exp, gau, pow, bsq, cir, or lin.");
    return ierr();
  }
  if (!(at > 1.e-10)) {
    set_error("set_st_model", "at must be positive.");
    return ierr();
  }
  if (time_nugget < 0.0 || time_sill < 0.0) {
    set_error("set_st_model", "time_nugget and time_sill must be non-negative.");
    return ierr();
  }
  K->st_model = m == "product_sum" ? 1 : 0;
  K->t_vtype = tid;
  K->t_nug = time_nugget;
  K->t_sill = time_sill;
  K->t_at = at;
  K->k_ps = k_ps;
  K->model_ok = false;
  K->solved = false;
  return 0;
}

int krige_st_set_obs(int64_t h, int ivar, int nobs, int ndim, const double* coord, const double* value,
                     const double* variance, int nmax, double maxdist, double sk_mean) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->st) {
    set_error("t_kriging_base%set_obs", "space-time observations need a krige_st handle");
    return ierr();
  }
  // coord is (ndim+1, nobs) with time last; the space-time object always has
  // 3 spatial rows + time, so any other ndim fails the row-count check
  // (kriging_st_capi.f90:167-189 -> set_obs_base)
  return krige_set_obs(h, ivar, nobs, ndim + 1, coord, value, variance, nmax, maxdist, sk_mean);
}

int krige_st_set_grad(int64_t h, int ivar, int ngrad, int ndim, const double* coord1, const double* coord2,
                      const double* grad_val, const double* variance, int ndrift_c, const double* drift_ext) {
  (void)ivar, (void)ngrad, (void)ndim, (void)coord1, (void)coord2, (void)grad_val, (void)variance;
  (void)ndrift_c, (void)drift_ext;
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_st_set_grad", "gradient constraints are not provided by this backend");
}

int krige_st_reset_vgm(int64_t h, int ivar, int jvar) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (krige_reset_vgm(h, ivar, jvar)) return ierr();
  K->vgm_t.clear();
  K->sill_st.clear();
  K->model_ok = false;
  return 0;
}

int krige_st_set_vgm(int64_t h, int ivar, int jvar, const char* vtype, double nugget, double sill, double a_major,
                     double a_minor1, double a_minor2, double azimuth, double dip, double plunge) {
  return krige_set_vgm(h, ivar, jvar, vtype, nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge, 0);
}

int krige_st_set_vgm_temporal(int64_t h, int ivar, int jvar, const char* vtype, double nugget, double sill,
                              double at_k) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  const char* W = "t_kriging_st%set_vgm_temporal";
  if (!K->st || ivar != 1 || jvar != 1) {
    set_error(W, "temporal structures need a krige_st handle with nvar = 1");
    return ierr();
  }
  const int id = vtype_from_str(vtype ? vtype : "sph");
  if (id < 0) {
    set_error("vtype_from_str", std::string("unknown vtype: ") + (vtype ? vtype : ""));
    return ierr();
  }
  if (!(at_k > 0.0) || nugget < 0.0 || sill < 0.0 || (int)K->vgm_t.size() >= kk::MAXS) {
    set_error(W, "invalid temporal structure (at_k > 0, nugget >= 0, sill >= 0)");
    return ierr();
  }
  // isotropic 1-D geometry: a_major = a_minor1 = a_minor2 = at_k (variogram_st.f90:302-317)
  K->vgm_t.push_back(VStruct{id, 0, nugget, sill, at_k, at_k, at_k, 0.0, 0.0, 0.0});
  K->model_ok = false;
  K->solved = false;
  return 0;
}

int krige_st_set_vgm_joint_sills(int64_t h, int ivar, int jvar, int n, const double* sills) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  // kriging_st.F90:267-288 -> variogram_st.f90:325-337: one joint sill per
  // spatial structure added so far
  if (!K->st || ivar != 1 || jvar != 1) {
    set_error("set_vgm_joint_sills", "joint sills need a krige_st handle with nvar = 1");
    return ierr();
  }
  if (n != (int)K->vgm[0].size() || (n > 0 && !sills)) {
    set_error("set_joint_sills_vgm_st", "vgm_struct_st%set_joint_sills: length of sills must equal cs%nstruct");
    return ierr();
  }
  K->sill_st.assign(sills, sills + n);
  K->model_ok = false;
  K->solved = false;
  return 0;
}

int krige_st_set_grid(int64_t h, int ngrid, const double* coord, const double* time, const double* rangescale,
                      const double* localnugget) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->st) {
    set_error("t_kriging_st%set_grid_point", "space-time grids need a krige_st handle");
    return ierr();
  }
  std::vector<double> c((size_t)4 * ngrid);
  for (int i = 0; i < ngrid; ++i) {
    for (int k = 0; k < 3; ++k) c[k + (size_t)4 * i] = coord[k + (size_t)3 * i];
    c[3 + (size_t)4 * i] = time[i];
  }
  return krige_set_grid(h, ngrid, 4, c.data(), rangescale, localnugget);
}

int krige_st_set_grid_block(int64_t h, int nblocks, const double* coord, const double* time, const int* nblockpnt,
                            int npnts_total, const double* blockcoord, const double* blocktime,
                            const double* pointweight, const double* rangescale, const double* localnugget) {
  (void)nblocks, (void)coord, (void)time, (void)nblockpnt, (void)npnts_total, (void)blockcoord;
  (void)blocktime, (void)pointweight, (void)rangescale, (void)localnugget;
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_st_set_grid_block", "space-time block kriging is not provided by this backend");
}

int krige_st_set_grid_drift(int64_t h, int ndrift_c, int nblocks, const double* drift) {
  (void)ndrift_c, (void)nblocks, (void)drift;
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_st_set_grid_drift", "external drift is not provided by this backend");
}

int krige_st_set_sim(int64_t h, int nblocks, const void* randpath, int nsim_c, const void* sample) {
  (void)nblocks, (void)randpath, (void)nsim_c, (void)sample;
  clear_error();
  if (!checked(h)) return ierr();
  return unsupported("krige_st_set_sim", "sequential simulation is not provided by this backend");
}

int krige_st_set_search(int64_t h, int ivar, double time_at, double anis1, double anis2, double azimuth, double dip,
                        double plunge, int sector_search) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->st || ivar < 1 || ivar > K->nvar) {
    set_error("set_search", "space-time search needs a krige_st handle and a valid ivar");
    return ierr();
  }
  if (!(time_at > 1.e-10)) {
    set_error("set_search", "time_at must be positive.");
    return ierr();
  }
  K->obs[ivar - 1].time_at = time_at;
  K->perm_ok = false;
  return krige_set_search(h, ivar, anis1, anis2, azimuth, dip, plunge, sector_search);
}

void krige_st_get_max_threads(int* n) { *n = 1; }
void krige_st_get_num_threads(int* n) { *n = 1; }

int kk_set_options(int64_t h, int precision, double pivot_tol, int table_mode) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (precision != kk::PREC_FP64 && precision != kk::PREC_FP32R) {
    set_error("kk_set_options", "precision must be 0 (fp64) or 1 (fp32 + refinement)");
    return ierr();
  }
  K->precision = precision;
  K->pivot_tol = pivot_tol;
  K->table_mode = table_mode != 0;
  K->solved = false;
  K->model_ok = false;
  return 0;
}

// mode 0: exact-emulation search kernel for every query; 1: per-lane two-pass
// search; 2: wave-cooperative two-pass search (default).  reorder: 0 query
// order as given, 1 Morton order in two chunks with the first chunk's solve
// overlapping the second chunk's search (default), 2 Morton order in one
// launch (every phase timed on its own).
int kk_set_factor_reuse(int64_t h, int mode, int item_max) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (mode < -1 || mode > 1 || item_max < 1 || item_max > kk::MAX_ITEM) {
    set_error("kk_set_factor_reuse", "mode must be -1, 0 or 1 and item_max in 1.." + std::to_string(kk::MAX_ITEM));
    return ierr();
  }
  K->reuse_mode = mode;
  K->item_max = item_max;
  return 0;
}

int kk_set_search_mode(int64_t h, int mode, int reorder) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (mode < 0 || mode > 2) {
    set_error("kk_set_search_mode", "mode must be 0, 1 or 2");
    return ierr();
  }
  K->search_mode = mode;
  K->reorder = reorder != 0;
  K->chunked = reorder == 1;
  K->solved = false;
  return 0;
}

// Runs the device pipeline only: inputs resident in HBM, results left there.
int kk_solve_resident(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  do_solve(*K, false);
  return ierr();
}

// Captures the device pass of a solved single-variable handle into a HIP
// graph (no host synchronisation inside: every count is consumed on the
// device); kk_graph_launch replays it over the same resident inputs.
int kk_graph_capture(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->solved) {
    set_error("kk_graph_capture", "solve the handle first: the graph replays a solved pass's inputs");
    return ierr();
  }
  capture_pass(*K);
  return ierr();
}

// Replays the captured pass: results stay in HBM (kk_fetch copies them back).
// The solver statistics are those of the solve the graph was captured after
// (same inputs, same outcome).
int kk_graph_launch(int64_t h) {
  clear_error();
  const char* W = "kk_graph_launch";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->gexec || !K->solved || K->graph_key[0] != K->d_nnear.p || K->graph_key[1] != K->d_fb.p ||
      (K->graph_key[2] != K->d_items.p && K->graph_key[2] != K->d_lists.p)) {
    K->drop_graph();
    set_error(W, "no valid captured pass: inputs or options changed since kk_graph_capture");
    return ierr();
  }
  HIPCHK(hipEventRecord(K->ev[0], K->stream), W);
  HIPCHK(hipGraphLaunch(K->gexec, K->stream), W);
  HIPCHK(hipEventRecord(K->ev[4], K->stream), W);
  int over = 0;
  HIPCHK(hipMemcpyAsync(&over, K->graph_overflow, sizeof(int), hipMemcpyDeviceToHost, K->stream), W);
  HIPCHK(hipStreamSynchronize(K->stream), W);
  if (over) {
    K->drop_graph();
    K->solved = false;
    set_error(W, "a list of the replayed pass outgrew the captured launch bounds (inputs changed); capture again");
    return ierr();
  }
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, K->ev[0], K->ev[4]), W);
  for (double& x : K->timing) x = 0.0;
  K->timing[4] = ms;
  K->timing[7] = -1;
  return ierr();
}

// Times candidate solve schedules (instance width and looping grid) for every
// non-empty bucket on the solved handle's own work lists, keeps the fastest
// in the library's table (used by later passes and captures) and, with a
// path, writes the table there.  reps timed launches per candidate.
int kk_autotune(int64_t h, const char* path, int reps) {
  clear_error();
  const char* W = "kk_autotune";
  Krige* K = checked(h);
  if (!K) return ierr();
  int mmax, kcand;
  pass_dims(*K, mmax, kcand);
  if (!K->solved || mv_results(*K) || mmax > 64) {
    set_error(W, "autotuning needs a solved single-variable handle with at most 64 rows per system");
    return ierr();
  }
  reps = std::max(1, std::min(reps, 50));
  PassBufs B;
  kk::SolveArgs va;
  bool tb = false;
  if (!setup_pass(*K, nullptr, mmax, B, va, tb)) return ierr();
  hipStream_t st = K->stream;
  const int nb = K->nblock, nd = K->ndim, prec = K->precision;
  const bool reuse = B.items != nullptr;
  int* ctr = B.counts + ROUND_CTR[0];
  // the whole grid's work lists, as one round
  if (reuse) {
    kk::SetGroupArgs g{};
    g.nnear = B.nnear;
    g.inear = B.inear;
    g.mmax = mmax;
    g.rs = va.rangescale;
    g.ln = va.localnugget;
    g.nsel = nb;
    g.table = B.table;
    g.hash = B.hash;
    g.tsize = std::min(B.tcap, set_table_slots(nb));
    g.item_max = K->item_max;
    g.rep = B.rep;
    g.rank = B.rank;
    g.cnt = B.setcnt;
    g.setbase = B.setbase;
    g.setitem = B.setitem;
    g.members = B.members;
    g.ctr = ctr;
    g.total = B.counts + 16;
    g.items = B.items;
    g.empties = B.empties;
    HIPCHK(kk::launch_fill(B.counts + 16, 1, 0, st), W);
    HIPCHK(kk::launch_set_group(g, st), W);
  } else {
    HIPCHK(kk::launch_fill(ctr, 32, 0, st), W);
    HIPCHK(kk::launch_bucket(B.nnear, nullptr, nb, ctr, B.lists, nb, st), W);
  }
  int hc[32];
  HIPCHK(hipMemcpyAsync(hc, ctr, sizeof hc, hipMemcpyDeviceToHost, st), W);
  HIPCHK(hipStreamSynchronize(st), W);
  static const int grids[] = {0};
  for (int bk = 1; bk < kk::NBUCKET; ++bk) {
    if (hc[bk] <= 0) continue;
    kk::SolveArgs v = va;
    v.nlist = hc[bk];
    if (reuse) {
      v.items = B.items + hc[kk::SET_OFF + bk];
      v.setcnt = B.setcnt;
      v.setbase = B.setbase;
      v.members = B.members;
      v.item_max = K->item_max;
    } else {
      v.list = B.lists + (size_t)bk * nb;
    }
    TuneEntry best;
    best.ms = FLT_MAX;
    int nw = 0;
    for (int w : kk::BUCKET_ROWS) {
      if (w < kk::bucket_width(bk) || nw == 3) continue;
      ++nw;
      for (int g : grids) {
        (void)g;
        HIPCHK(kk::launch_solve(v, w, prec, tb, st), W);  // warm-up
        std::vector<float> t(reps);
        for (int r = 0; r < reps; ++r) {
          HIPCHK(hipEventRecord(K->ev[0], st), W);
          HIPCHK(kk::launch_solve(v, w, prec, tb, st), W);
          HIPCHK(hipEventRecord(K->ev[1], st), W);
          HIPCHK(hipEventSynchronize(K->ev[1]), W);
          HIPCHK(hipEventElapsedTime(&t[r], K->ev[0], K->ev[1]), W);
        }
        std::sort(t.begin(), t.end());
        const float med = t[reps / 2];
        if (std::getenv("KRIGEKIT_AMD_TUNE_VERBOSE"))
          std::fprintf(stderr, "kk_autotune dim %d prec %d bucket %d items %d width %d grid %d: %.4f ms\n", nd, prec,
                       bk, hc[bk], w, g, med);
        if (med < best.ms) {
          best.ms = med;
          best.width = w;
          best.grid = g;
        }
      }
    }
    tune_entry(nd, prec, tb, bk) = best;
  }
  K->drop_graph();
  if (path && path[0] && write_tuning(path)) set_error(W, std::string("cannot write ") + path);
  return ierr();
}

// Replaces the library's schedule table by the file at path (NULL or empty:
// every bucket back to its own width and a full grid).
int kk_set_tuning(const char* path) {
  clear_error();
  for (auto& a : g_tune)
    for (auto& b : a)
      for (auto& c : b)
        for (auto& e : c) e = TuneEntry{};
  if (path && path[0] && read_tuning(path) == 0) set_error("kk_set_tuning", std::string("no schedules read from ") + path);
  for (Krige* K : g_reg)
    if (K) K->drop_graph();
  return ierr();
}

// Copies the results of the last kk_solve_resident back to the host.
int kk_fetch(int64_t h) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!(mv_results(*K) ? fetch_results_mv(*K) : fetch_results(*K))) K->solved = false;
  return ierr();
}

// Leave-one-out MSE and its gradient with respect to the variogram parameters
// (nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge) of the
// single structure of an ordinary-kriging cross-validation setup.  Runs the
// search and forward solve, then the adjoint kernel.  zhat (nblock) and dz
// (nblock x 8, block-major) are optional per-block outputs.
int kk_loo_grad(int64_t h, double* mse, double* grad, double* zhat, double* dz) {
  clear_error();
  const char* W = "kk_loo_grad";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->cross_validation || K->nvar != 1 || K->unbias != 1) {
    set_error(W, "the gradient needs single-variable ordinary kriging with cross_validation=True");
    return ierr();
  }
  if (K->vgm.empty() || K->vgm[0].size() != 1 || K->vgm[0][0].vtype == kk::NUG) {
    set_error(W, "the gradient supports one variogram structure (with its nugget)");
    return ierr();
  }
  if (!K->rangescale.empty() || !K->localnugget.empty()) {
    set_error(W, "the gradient does not support per-block range scale or local nugget");
    return ierr();
  }
  if (K->nsim > 0 || K->use_old_weight) {
    set_error(W, "the gradient needs its own kriging pass (no simulation, no stored-weight replay)");
    return ierr();
  }
  // The neighbour lists do not depend on the variogram: after the first
  // evaluation (one full pass) only the model is rebuilt, and the adjoint
  // kernel solves every block's system once for both the estimate and the
  // gradient.
  const int nb = K->nblock, nd = K->ndim;
  if (K->lists_ok) {
    if (!validate_for_solve(*K) || !stage_inputs(*K)) return ierr();
  } else {
    if (!do_solve(*K, false)) return ierr();
    // the largest system decides how many blocks share a wave
    int rows = 0;
    if (!K->d_gcount.ensure(sizeof(int))) {
      set_error(W, "device allocation failed");
      return ierr();
    }
    HIPCHK_I(kk::launch_max_count(K->d_nnear.as<int>(), nb, K->d_gcount.as<int>(), K->stream), W);
    HIPCHK_I(hipMemcpyAsync(&rows, K->d_gcount.p, sizeof(int), hipMemcpyDeviceToHost, K->stream), W);
    HIPCHK_I(hipStreamSynchronize(K->stream), W);
    K->list_rows = rows;
    K->lists_ok = true;
  }
  if (!K->d_gz.ensure(sizeof(double) * nb) || !K->d_gdz.ensure(sizeof(double) * (size_t)nb * kk::NGRAD) ||
      !K->d_gred.ensure(sizeof(double) * kk::loo_reduce_words())) {
    set_error(W, "device allocation failed");
    return ierr();
  }
  kk::GradArgs ga{};
  ga.model = K->d_model.as<Model>();
  ga.ndim = nd;
  ga.nb = nb;
  ga.mmax = K->mmax;
  ga.rows = K->list_rows;
  ga.bcoord = K->d_bcoord.as<double>();
  ga.ocoord = K->d_ocoord.as<double>();
  ga.ovalue = K->d_ovalue.as<double>();
  ga.ovar = K->d_ovar.as<double>();
  ga.nnear = K->d_nnear.as<int>();
  ga.inear = K->d_inear.as<int>();
  aniso_dmat(K->vgm[0][0], ga.dmat);
  ga.zhat = K->d_gz.as<double>();
  ga.dz = K->d_gdz.as<double>();
  hipStream_t st = K->stream;
  HIPCHK_I(hipEventRecord(K->ev[5], st), W);
  HIPCHK_I(kk::launch_loo_grad(ga, st), W);
  HIPCHK_I(hipEventRecord(K->ev[6], st), W);
  // squared error and gradient sums on the device; the per-block values
  // travel only when the caller asks for them
  constexpr int NS = kk::NGRAD + 2;
  double sums[NS];
  HIPCHK_I(kk::launch_loo_reduce(K->d_gz.as<double>(), K->d_gdz.as<double>(), K->d_ovalue.as<double>(), nb,
                                 K->d_gred.as<double>(), st), W);
  HIPCHK_I(hipMemcpyAsync(sums, K->d_gred.as<double>() + (kk::loo_reduce_words() - NS), sizeof sums,
                          hipMemcpyDeviceToHost, st), W);
  if (zhat) HIPCHK_I(hipMemcpyAsync(zhat, K->d_gz.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st), W);
  if (dz)
    HIPCHK_I(hipMemcpyAsync(dz, K->d_gdz.p, sizeof(double) * (size_t)nb * kk::NGRAD, hipMemcpyDeviceToHost, st), W);
  HIPCHK_I(hipStreamSynchronize(st), W);
  float ms = 0.0f;
  HIPCHK_I(hipEventElapsedTime(&ms, K->ev[5], K->ev[6]), W);
  K->timing[11] = ms;
  const double cnt = sums[NS - 1];
  *mse = cnt > 0.0 ? sums[0] / cnt : std::nan("");
  for (int p = 0; p < kk::NGRAD; ++p) grad[p] = cnt > 0.0 ? 2.0 * sums[1 + p] / cnt : std::nan("");
  return 0;
}

// out[0..12]: search ms, bucket ms, solve ms, re-solve/fallback ms, total ms,
// FP64 re-solved nodes, symmetric-indefinite fallback nodes, queries decided by the exact
// search kernel (-1 when the fast search was not used), fast search used
// (1/0), fast search kernel ms, exact search kernel ms, LOO gradient kernel ms,
// weight store pass or stored-weight replay ms
int kk_get_timing(int64_t h, double* out) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  for (int i = 0; i < 15; ++i) out[i] = K->timing[i];
  return 0;
}

// Copies nbytes between two device buffers with a plain copy kernel, reps
// times; *ms receives the mean kernel time (HIP events).  Used to calibrate
// HBM traffic counters against a kernel of known size.
int kk_copy_probe(double nbytes, int reps, double* ms) {
  clear_error();
  const size_t n = (size_t)(nbytes / 8.0);
  double *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) {
    set_error("kk_copy_probe", "device allocation failed");
    if (a) (void)hipFree(a);
    return ierr();
  }
  (void)hipMemset(a, 0, n * 8);
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)kk::launch_copy(a, b, n, nullptr);
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps; ++r) (void)kk::launch_copy(a, b, n, nullptr);
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float t = 0.0f;
  (void)hipEventElapsedTime(&t, e0, e1);
  *ms = t / std::max(reps, 1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

// kk_solve_resident without any device allocation: the pass's neighbour
// lists, flags, bucket lists, search heaps and fallback matrices live in the
// fixed scratch array; est and var are the caller's device buffers of nblock
// doubles.  The handle must hold a finished single-variable solve of its
// current inputs (krige_solve or kk_solve_resident), so that its inputs,
// streams and query order are already resident.  The handle's own results
// are left as they were.
int kk_solve_resident_noalloc(int64_t h, double* est, double* var) {
  clear_error();
  const char* W = "kk_solve_resident_noalloc";
  Krige* K = checked(h);
  if (!K) return ierr();
  if (mv_results(*K) || !K->solved || !K->dev_inputs_ok || !K->model_ok || !K->stream ||
      (K->search_mode == 2 && K->reorder && !K->perm_ok)) {
    set_error(W, "needs a finished single-variable solve of the handle's current inputs "
                 "(krige_solve or kk_solve_resident) before the allocation-free pass");
    return ierr();
  }
  if (K->store_weight || K->use_old_weight) {
    set_error(W, "the allocation-free pass neither stores nor replays kriging weights");
    return ierr();
  }
  if (!est || !var) {
    set_error(W, "est and var must be device buffers of nblock doubles");
    return ierr();
  }
  int mmax = 0, k = 0;
  pass_dims(*K, mmax, k);
  PassBufs B;
  const size_t need = carve_pass(nullptr, K->nblock, mmax, k, K->reuse_on(), B);
  if (need > KK_NOALLOC_BYTES) {
    set_error(W, "nblock=" + std::to_string(K->nblock) + " with " + std::to_string(mmax) +
                     " rows needs " + std::to_string(need >> 20) + " MiB of scratch, above the " +
                     std::to_string(KK_NOALLOC_BYTES >> 20) + " MiB of the allocation-free entry point");
    return ierr();
  }
  void* base = nullptr;
  if (!scratch_ptr(&base, W)) return ierr();
  carve_pass(static_cast<char*>(base), K->nblock, mmax, k, K->reuse_on(), B);
  B.est = est;
  B.var = var;
  run_pipeline(*K, &B);
  return ierr();
}

int kk_device_synchronize(void) {
  clear_error();
  if (hipDeviceSynchronize() != hipSuccess) {
    set_error("kk_device_synchronize", "hipDeviceSynchronize failed");
    return ierr();
  }
  return 0;
}

// neighbour lists of the last solve straight from HBM (0-based ids, ascending)
int kk_get_neighbours(int64_t h, int nblock, int mmax, int* nnear, int* inear) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  if (!K->d_nnear.p || nblock != K->nblock || mmax != K->mmax) {
    set_error("kk_get_neighbours", "no neighbour lists of matching shape");
    return ierr();
  }
  if (hipMemcpy(nnear, K->d_nnear.p, sizeof(int) * nblock, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(inear, K->d_inear.p, sizeof(int) * (size_t)nblock * mmax, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("kk_get_neighbours", "device copy failed");
    return ierr();
  }
  return 0;
}

int kk_get_mmax(int64_t h, int* mmax) {
  clear_error();
  Krige* K = checked(h);
  if (!K) return ierr();
  *mmax = K->mmax;
  return 0;
}

}  // extern "C"
#pragma GCC visibility pop
