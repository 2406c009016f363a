// Neighbour search on the GPU: exact kdtree2 semantics, one lane per query.
//
// Reproduces, per query, what the reference's search_neighbors computes
// (src/libkriging/kriging.F90:490-666):
//   * kdtree2_n_nearest traversal order and pruning
//     (src/libkriging/kdtree2_maxidx.f90:1383-1457), bounded max-heap with
//     replace-on-equal (:307-446, :1510-1613)
//   * sector mode: unstable heapsort of the candidates (:1860-1914) and the
//     quadrant/octant caps (kriging.F90:1093-1154)
//   * leave-one-out self exclusion (kriging.F90:647-657), maxdist filter
//     (kriging_base.F90:2491-2505) and the ascending index sort
//     (kriging_base.F90:3510-3524)
// so neighbour lists, including which of equal-distance points is kept, are
// bit-identical to the reference.
//
// Traversal is stackless: the tree carries parent links, and whether a lane
// returns from the closer or the farther child is recomputed from the query.
// Squared distances accumulate dimension by dimension with separately
// rounded multiply and add (this file is built with -ffp-contract=off), which
// is how the reference's x86-64 build rounds them.
//
// The heap of each lane lives in HBM scratch, lane-interleaved
// (entry p of lane t at p*hstride + t), so the top of the heap, touched on
// every insert, is read by a wave as one contiguous segment.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kk_device.hpp"
#include "kk_types.hpp"

namespace kk {

#pragma clang fp contract(off)

struct Heap {
  double* d;
  int* i;
  size_t s;
  __device__ double& D(int p) const { return d[(size_t)(p - 1) * s]; }
  __device__ int& I(int p) const { return i[(size_t)(p - 1) * s]; }
};

// pq_insert, kdtree2_maxidx.f90:307-346
__device__ inline double pq_insert(const Heap& h, int& size, double dis, int idx) {
  size += 1;
  int i = size;
  while (i > 1) {
    int ip = i >> 1;
    double pd = h.D(ip);
    if (dis > pd) {
      h.D(i) = pd;
      h.I(i) = h.I(ip);
      i = ip;
    } else {
      break;
    }
  }
  h.D(i) = dis;
  h.I(i) = idx;
  return h.D(1);
}

// pq_replace_max, kdtree2_maxidx.f90:392-446
__device__ inline double pq_replace_max(const Heap& h, int N, double dis, int idx) {
  int parent = 1, child = 2;
  while (child <= N) {
    double pc = h.D(child);
    if (child < N) {
      double pc1 = h.D(child + 1);
      if (pc < pc1) {
        child = child + 1;
        pc = pc1;
      }
    }
    if (dis >= pc) break;
    h.D(parent) = pc;
    h.I(parent) = h.I(child);
    parent = child;
    child = 2 * parent;
  }
  h.D(parent) = dis;
  h.I(parent) = idx;
  return h.D(1);
}

// heapsort_struct, kdtree2_maxidx.f90:1860-1914 (unstable, ascending)
__device__ inline void heapsort(const Heap& a, int n) {
  if (n <= 1) return;
  int ileft = n / 2 + 1, iright = n;
  double vd;
  int vi;
  for (;;) {
    if (ileft > 1) {
      ileft = ileft - 1;
      vd = a.D(ileft);
      vi = a.I(ileft);
    } else {
      vd = a.D(iright);
      vi = a.I(iright);
      a.D(iright) = a.D(1);
      a.I(iright) = a.I(1);
      iright = iright - 1;
      if (iright == 1) {
        a.D(1) = vd;
        a.I(1) = vi;
        return;
      }
    }
    int i = ileft, j = 2 * ileft;
    while (j <= iright) {
      if (j < iright && a.D(j) < a.D(j + 1)) j = j + 1;
      double aj = a.D(j);
      if (vd < aj) {
        a.D(i) = aj;
        a.I(i) = a.I(j);
        i = j;
        j = j + j;
      } else {
        j = iright + 1;
      }
    }
    a.D(i) = vd;
    a.I(i) = vi;
  }
}

__device__ inline double sq(double x) { return __dmul_rn(x, x); }

__device__ inline double dis2_bnd(double x, double lo, double hi) {
  if (x > hi) return sq(__dsub_rn(x, hi));
  if (x < lo) return sq(__dsub_rn(lo, x));
  return 0.0;
}

// ascending insertion sort of out[0..n) in global memory
__device__ inline void isort(int* a, int n) {
  for (int i = 1; i < n; ++i) {
    int t = a[i], j = i - 1;
    while (j >= 0 && a[j] > t) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = t;
  }
}

// Stackless kdtree2 traversal (kdtree2_maxidx.f90:1383-1457): closer child
// first; the farther child is visited when its box lower bound is <= the
// current ball.  leaf(l, u) processes a bucket and may shrink `ball`.
template <int DIM, class Leaf>
__device__ inline void traverse(const Tree& T, const double* q, double& ball, Leaf&& leaf) {
  int node = T.root, prev = -1;
  while (node >= 0) {
    const int* ni = T.ni + (size_t)node * NODE_I;
    const int cut = ni[0], left = ni[3], right = ni[4], parent = ni[5];
    if (prev == parent) {
      if (left < 0 || right < 0) {
        leaf(ni[1], ni[2]);
        prev = node;
        node = parent;
        continue;
      }
      const double cv = T.nd[(size_t)node * NODE_D + 0];
      prev = node;
      node = (q[cut] < cv) ? left : right;
      continue;
    }
    const double* nd = T.nd + (size_t)node * NODE_D;
    const bool goes_left = q[cut] < nd[0];
    const int closer = goes_left ? left : right;
    if (prev == closer) {
      const int farther = goes_left ? right : left;
      double dis = goes_left ? sq(__dsub_rn(nd[2], q[cut])) : sq(__dsub_rn(nd[1], q[cut]));
      bool visit = dis <= ball;
      if (visit) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          if (i != cut && visit) {
            dis = __dadd_rn(dis, dis2_bnd(q[i], nd[3 + i], nd[6 + i]));
            if (dis > ball) visit = false;
          }
        }
      }
      if (visit) {
        prev = node;
        node = farther;
        continue;
      }
    }
    prev = node;
    node = parent;
  }
}

// squared distance accumulated like process_terminal_node, with early exit
template <int DIM>
__device__ inline bool point_sd(const Tree& T, int i, const double* q, double ball, double& sd) {
  sd = 0.0;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    sd = __dadd_rn(sd, sq(__dsub_rn(T.rdata[k + DIM * (size_t)i], q[k])));
    if (sd > ball) return false;
  }
  return true;
}

template <int DIM>
__device__ inline int octant(const double* o, const double* q) {
  int sec = (__dsub_rn(o[0], q[0]) < 0.0);
  if (DIM >= 2) sec += 2 * (__dsub_rn(o[1], q[1]) < 0.0);
  if (DIM == 3) sec += 4 * (__dsub_rn(o[2], q[2]) < 0.0);
  return sec;
}

// Exact emulation of the reference search for one query per lane (heap in
// HBM scratch).  Runs over every query, or over the queries the fast kernel
// could not decide (a.qlist).
// LDSHEAP: the heap lives in LDS (lane-interleaved, 32-lane workgroups),
// used when only the queries the fast kernel flagged are left.
template <int DIM, bool LDSHEAP>
__global__ __launch_bounds__(256) void k_search(SearchArgs a) {
  extern __shared__ double lds_heap[];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const int qid = a.qlist ? a.qlist[t] : a.q0 + t;
  double q[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) q[k] = a.qcoord[k + DIM * (size_t)qid];
  int* out = a.inear + (size_t)qid * a.mmax;
  const Tree& T = a.tree;
  Heap h = LDSHEAP ? Heap{lds_heap + threadIdx.x,
                          reinterpret_cast<int*>(lds_heap + (size_t)a.k * blockDim.x) + threadIdx.x,
                          (size_t)blockDim.x}
                   : Heap{a.hdis + t, a.hidx + t, (size_t)a.hstride};

  double ballsize = 1.7976931348623157e308;  // huge(1.0)
  int nfound = 0;
  const int nn = a.k;
  traverse<DIM>(T, q, ballsize, [&](int l, int u) {
    // process_terminal_node, kdtree2_maxidx.f90:1510-1613
    for (int i = l; i <= u; ++i) {
      double sd;
      if (!point_sd<DIM>(T, i, q, ballsize, sd)) continue;
      const int idx = T.ind[i];
      if (nfound < nn) {
        double newpri = pq_insert(h, nfound, sd, idx);
        if (nfound == nn) ballsize = newpri;
      } else {
        ballsize = pq_replace_max(h, nn, sd, idx);
      }
    }
  });

  int cnt = 0;
  if (a.sort) {
    heapsort(h, nn);
    // apply_sector_filter, kriging.F90:1093-1154
    int counts[1 << DIM];
#pragma unroll
    for (int s = 0; s < (1 << DIM); ++s) counts[s] = 0;
    const int skip_idx = a.cv ? qid : -1;
    for (int c = 1; c <= nn; ++c) {
      const int idx = h.I(c);
      if (idx == skip_idx) continue;
      if (h.D(c) > a.maxdist2) continue;
      const int sec = octant<DIM>(a.ocoord + DIM * (size_t)idx, q);
      int cs = 0;
#pragma unroll
      for (int s = 0; s < (1 << DIM); ++s) cs = (s == sec) ? counts[s] : cs;
      if (cs < a.nmax) {
#pragma unroll
        for (int s = 0; s < (1 << DIM); ++s) counts[s] += (s == sec);
        out[cnt++] = idx;
      }
    }
  } else {
    bool removed = !a.cv;
    for (int c = 1; c <= nn; ++c) {
      const int idx = h.I(c);
      if (!removed && idx == qid) {
        removed = true;
        continue;
      }
      if (h.D(c) <= a.maxdist2) out[cnt++] = idx;
    }
  }
  isort(out, cnt);
  a.nnear[qid] = cnt;
}

__device__ inline float f_next(float f) { return __int_as_float(__float_as_int(f) + 1); }
__device__ inline float f_prev(float f) { return f > 0.0f ? __int_as_float(__float_as_int(f) - 1) : 0.0f; }

// Fast search, one query per lane, all per-lane state in LDS (lane-
// interleaved words, no cross-lane traffic).
//
// Pass 1: f = the k-th smallest of RN_fp32(d^2) over all observations, with a
//   float max-heap of k keys.  Rounding is monotonic, so f = RN_fp32(r_k)
//   exactly, where r_k is the k-th smallest FP64 squared distance.
// Pass 2: fixed-radius traversal.  Points with key < prev(f) are certainly
//   among the k nearest; points with key in [prev(f), next(f)] form the
//   boundary set.  If (#certain + #boundary) == k the k-nearest set is known
//   exactly; otherwise equal or adjacent distances at rank k make the answer
//   depend on the reference's visit order, and the query is flagged for the
//   exact kernel.  In sector mode each octant keeps its nmax+1 nearest (exact
//   FP64 distances); an exact tie between rank nmax and nmax+1 of an octant
//   (where the reference's unstable heapsort decides) also flags the query.
template <int DIM, bool SECTOR>
__global__ __launch_bounds__(64) void k_search_fast(SearchArgs a) {
  extern __shared__ double lds_d[];
  const int lane = threadIdx.x;
  const int t = blockIdx.x * 64 + lane;
  if (t >= a.nq) return;
  const int qid = a.q0 + t;
  double q[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) q[k] = a.qcoord[k + DIM * (size_t)qid];
  const Tree& T = a.tree;
  const int K = a.k;
  float* hk = reinterpret_cast<float*>(lds_d);
  auto HK = [&](int p) -> float& { return hk[(p - 1) * 64 + lane]; };

  // ---- pass 1: k-th smallest FP32 key ---------------------------------
  double ball = 1.7976931348623157e308;
  int nf = 0;
  traverse<DIM>(T, q, ball, [&](int l, int u) {
    for (int i = l; i <= u; ++i) {
      double sd;
      if (!point_sd<DIM>(T, i, q, ball, sd)) continue;
      const float key = (float)sd;
      if (nf < K) {
        int c = ++nf;
        while (c > 1) {
          const int pp = c >> 1;
          const float pk = HK(pp);
          if (key > pk) {
            HK(c) = pk;
            c = pp;
          } else {
            break;
          }
        }
        HK(c) = key;
        if (nf == K) ball = (double)f_next(HK(1));
      } else if (key < HK(1)) {
        int par = 1, ch = 2;
        while (ch <= K) {
          float ck = HK(ch);
          if (ch < K) {
            const float c1 = HK(ch + 1);
            if (ck < c1) {
              ++ch;
              ck = c1;
            }
          }
          if (key >= ck) break;
          HK(par) = ck;
          par = ch;
          ch = 2 * par;
        }
        HK(par) = key;
        ball = (double)f_next(HK(1));
      }
    }
  });
  const float f = HK(1);
  const float flo = f_prev(f), fhi = f_next(f);
  double bound = (double)f_next(fhi);

  // ---- pass 2: exact collection ------------------------------------------
  constexpr int NOCT = 1 << DIM;
  const int C = a.nmax + 1;  // per-octant capacity (sector mode)
  int* ocnt = reinterpret_cast<int*>(lds_d);
  double* od = lds_d + (size_t)(NOCT / 2 + (NOCT & 1)) * 64;
  int* oi = reinterpret_cast<int*>(od + (size_t)NOCT * C * 64);
  if (SECTOR) {
#pragma unroll
    for (int s = 0; s < NOCT; ++s) ocnt[s * 64 + lane] = 0;
  }
  int* out = a.inear + (size_t)qid * a.mmax;
  int nin = 0, nb = 0, cnt = 0;
  bool flag = false;
  traverse<DIM>(T, q, bound, [&](int l, int u) {
    for (int i = l; i <= u; ++i) {
      double sd;
      if (!point_sd<DIM>(T, i, q, bound, sd)) continue;
      const float key = (float)sd;
      if (key > fhi) continue;
      if (key < flo) ++nin;
      else ++nb;
      const int idx = T.ind[i];
      if (a.cv && idx == qid) continue;
      if (sd > a.maxdist2) continue;
      if (SECTOR) {
        const int s = octant<DIM>(a.ocoord + DIM * (size_t)idx, q);
        const int c = ocnt[s * 64 + lane];
        double* dd = od + (size_t)s * C * 64 + lane;
        int* ii = oi + (size_t)s * C * 64 + lane;
        if (c == C && sd >= dd[(size_t)(C - 1) * 64]) continue;
        int j = c < C ? c : C - 1;
        while (j > 0 && dd[(size_t)(j - 1) * 64] > sd) {
          dd[(size_t)j * 64] = dd[(size_t)(j - 1) * 64];
          ii[(size_t)j * 64] = ii[(size_t)(j - 1) * 64];
          --j;
        }
        dd[(size_t)j * 64] = sd;
        ii[(size_t)j * 64] = idx;
        ocnt[s * 64 + lane] = c < C ? c + 1 : C;
      } else {
        if (cnt < a.mmax) out[cnt++] = idx;
        else flag = true;
      }
    }
  });
  if (nin + nb != K) flag = true;
  if (SECTOR) {
    for (int s = 0; s < NOCT; ++s) {
      const int c = ocnt[s * 64 + lane];
      const double* dd = od + (size_t)s * C * 64 + lane;
      const int* ii = oi + (size_t)s * C * 64 + lane;
      if (c == C && dd[(size_t)(C - 2) * 64] == dd[(size_t)(C - 1) * 64]) flag = true;
      const int take = c < a.nmax ? c : a.nmax;
      for (int j = 0; j < take; ++j) out[cnt++] = ii[(size_t)j * 64];
    }
  }
  if (flag) {
    a.qflag[qid] = 1;
    return;
  }
  isort(out, cnt);
  a.nnear[qid] = cnt;
  a.qflag[qid] = 0;
}

// Wave-cooperative variant of the fast search: the 64 lanes of a wave hold 64
// spatially coherent queries and walk the tree together.  Control flow is
// uniform: a node is entered when ANY lane's exact box lower bound is within
// that lane's current ball (a conservative test for every lane), children
// are ordered by the wave's first query, and leaf points are read with
// uniform (scalar) loads and offered to every lane.  Each lane keeps its own
// LDS heap / octant lists, so its result is the same as the per-lane
// kernel's; the exactness flags are the same.
template <int DIM>
__device__ inline double box_lb(const double* nd, const double* q) {
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < DIM; ++k) d = __dadd_rn(d, dis2_bnd(q[k], nd[3 + k], nd[6 + k]));
  return d;
}

template <int DIM, class Leaf, class Ball>
__device__ inline void traverse_wave(const Tree& T, const double* q, const double* q0, Ball&& ball,
                                     Leaf&& leaf) {
  int node = T.root, prev = -1;
  while (node >= 0) {
    node = __builtin_amdgcn_readfirstlane(node);
    const int* ni = T.ni + (size_t)node * NODE_I;
    const int cut = ni[0], left = ni[3], right = ni[4], parent = ni[5];
    const double* nd = T.nd + (size_t)node * NODE_D;
    if (prev == parent) {
      // inflate by 1e-12 relative: rounding in the bound must never prune a
      // point whose squared distance sits at the ball
      const bool need = box_lb<DIM>(nd, q) <= ball() * (1.0 + 1e-12);
      if (__ballot(need) == 0) {
        prev = node;
        node = parent;
        continue;
      }
      if (left < 0 || right < 0) {
        leaf(ni[1], ni[2], node);
        prev = node;
        node = parent;
        continue;
      }
      prev = node;
      node = (q0[cut] < nd[0]) ? left : right;
      continue;
    }
    const int closer = (q0[cut] < nd[0]) ? left : right;
    const int farther = closer == left ? right : left;
    if (prev == closer) {
      prev = node;
      node = farther;
      continue;
    }
    prev = node;
    node = parent;
  }
}

// Boundary candidates of pass 2 (FP32 key within one ulp of the k-th key),
// kept sorted by exact FP64 distance in registers.
constexpr int BCAP = 4;
struct Boundary {
  double d[BCAP];
  int i[BCAP];
  int n = 0;
  __device__ void add(double sd, int idx, bool& overflow) {
    if (n == BCAP) {
      overflow = true;
      return;
    }
    // insertion into the sorted prefix by compare-select (static indices)
    double cd = sd;
    int ci = idx;
#pragma unroll
    for (int j = 0; j < BCAP; ++j) {
      if (j < n) {
        if (d[j] > cd) {
          const double td = d[j];
          const int ti = i[j];
          d[j] = cd;
          i[j] = ci;
          cd = td;
          ci = ti;
        }
      } else if (j == n) {
        d[j] = cd;
        i[j] = ci;
      }
    }
    ++n;
  }
};

constexpr int LEAFCAP = 1024;  // leaves recorded per wave in pass 1

template <int DIM, bool SECTOR>
__global__ __launch_bounds__(64) void k_search_wave(SearchArgs a) {
  extern __shared__ double lds_d[];
  __shared__ int leaves[LEAFCAP];
  const int lane = threadIdx.x;
  const int t = blockIdx.x * 64 + lane;
  const bool valid = t < a.nq;
  const int qid = a.qlist ? a.qlist[valid ? t : a.nq - 1] : a.q0 + (valid ? t : a.nq - 1);
  double q[DIM], q0[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    q[k] = a.qcoord[k + DIM * (size_t)qid];
    q0[k] = __builtin_amdgcn_readfirstlane(q[k]);
  }
  const Tree& T = a.tree;
  const int K = a.k;
  float* hk = reinterpret_cast<float*>(lds_d);
  auto HK = [&](int p) -> float& { return hk[(p - 1) * 64 + lane]; };

  // ---- pass 1: k-th smallest FP32 key; records the leaves it visits --------
  // The ball is next(next(max key)) so the recorded leaves cover every point
  // pass 2 needs (its bound is next(next(f))).
  double ball = 1.7976931348623157e308;
  int nf = 0, nleaf = 0;
  traverse_wave<DIM>(T, q, q0, [&]() { return ball; }, [&](int l, int u, int node) {
    if (nleaf < LEAFCAP && lane == 0) leaves[nleaf] = node;
    ++nleaf;
    for (int i = l; i <= u; ++i) {
      double sd = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(T.rdata[k + DIM * (size_t)i], q[k])));
      const float key = (float)sd;
      if (nf < K) {
        int c = ++nf;
        while (c > 1) {
          const int pp = c >> 1;
          const float pk = HK(pp);
          if (key > pk) {
            HK(c) = pk;
            c = pp;
          } else {
            break;
          }
        }
        HK(c) = key;
        if (nf == K) ball = (double)f_next(f_next(HK(1)));
      } else if (key < HK(1)) {
        int par = 1, ch = 2;
        while (ch <= K) {
          float ck = HK(ch);
          if (ch < K) {
            const float c1 = HK(ch + 1);
            if (ck < c1) {
              ++ch;
              ck = c1;
            }
          }
          if (key >= ck) break;
          HK(par) = ck;
          par = ch;
          ch = 2 * par;
        }
        HK(par) = key;
        ball = (double)f_next(f_next(HK(1)));
      }
    }
  });
  const float f = HK(1);
  const float flo = f_prev(f), fhi = f_next(f);
  const double bound = (double)f_next(fhi);

  // ---- pass 2: exact collection over the recorded leaves -------------------
  constexpr int NOCT = 1 << DIM;
  const int C = a.nmax + 1;
  int* ocnt = reinterpret_cast<int*>(lds_d);
  double* od = lds_d + (size_t)(NOCT / 2 + (NOCT & 1)) * 64;
  int* oi = reinterpret_cast<int*>(od + (size_t)NOCT * C * 64);
  if (SECTOR) {
#pragma unroll
    for (int s = 0; s < NOCT; ++s) ocnt[s * 64 + lane] = 0;
  }
  int* out = a.inear + (size_t)qid * a.mmax;
  int nin = 0, cnt = 0;
  bool flag = false;
  Boundary B;
  auto keep = [&](double sd, int idx) {
    // a point of the k-nearest set: sector caps or the plain list
    if (a.cv && idx == qid) return;
    if (sd > a.maxdist2) return;
    if (SECTOR) {
      const int s = octant<DIM>(a.ocoord + DIM * (size_t)idx, q);
      const int c = ocnt[s * 64 + lane];
      double* dd = od + (size_t)s * C * 64 + lane;
      int* ii = oi + (size_t)s * C * 64 + lane;
      if (c == C && sd >= dd[(size_t)(C - 1) * 64]) return;
      int j = c < C ? c : C - 1;
      while (j > 0 && dd[(size_t)(j - 1) * 64] > sd) {
        dd[(size_t)j * 64] = dd[(size_t)(j - 1) * 64];
        ii[(size_t)j * 64] = ii[(size_t)(j - 1) * 64];
        --j;
      }
      dd[(size_t)j * 64] = sd;
      ii[(size_t)j * 64] = idx;
      ocnt[s * 64 + lane] = c < C ? c + 1 : C;
    } else {
      if (cnt < a.mmax) out[cnt++] = idx;
      else flag = true;
    }
  };
  auto visit = [&](int l, int u, int) {
    for (int i = l; i <= u; ++i) {
      double sd = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(T.rdata[k + DIM * (size_t)i], q[k])));
      if (!(sd <= bound)) continue;
      const float key = (float)sd;
      if (key > fhi) continue;
      const int idx = T.ind[i];
      if (key < flo) {
        ++nin;
        keep(sd, idx);
      } else {
        B.add(sd, idx, flag);
      }
    }
  };
  __syncthreads();
  if (nleaf <= LEAFCAP) {
    for (int j = 0; j < nleaf; ++j) {
      const int node = __builtin_amdgcn_readfirstlane(leaves[j]);
      const int* ni = T.ni + (size_t)node * NODE_I;
      const bool need_leaf = box_lb<DIM>(T.nd + (size_t)node * NODE_D, q) <= bound * (1.0 + 1e-12);
      if (__ballot(need_leaf) == 0) continue;
      visit(ni[1], ni[2], node);
    }
  } else {
    traverse_wave<DIM>(T, q, q0, [&]() { return bound; }, visit);
  }
  // the k-nearest set: all certain points plus the need closest boundary points
  const int need = K - nin;
  if (need < 0 || need > B.n) {
    flag = true;
  } else if (need < B.n) {
    // cut inside the boundary set: exact FP64 order decides; a (near-)tie at
    // the cut is left to the exact kernel
    double dcut = 0.0, dnext = 0.0;
#pragma unroll
    for (int j = 0; j < BCAP; ++j) {
      if (j == need - 1) dcut = B.d[j];
      if (j == need) dnext = B.d[j];
    }
    if (need > 0 && !(dnext > dcut * (1.0 + 1e-12))) flag = true;
    if (need == 0 && !(B.d[0] > 0.0)) flag = true;
  }
#pragma unroll
  for (int j = 0; j < BCAP; ++j)
    if (j < B.n && j < need) keep(B.d[j], B.i[j]);
  if (!valid) return;
  if (SECTOR) {
    for (int s = 0; s < NOCT; ++s) {
      const int c = ocnt[s * 64 + lane];
      const double* dd = od + (size_t)s * C * 64 + lane;
      const int* ii = oi + (size_t)s * C * 64 + lane;
      if (c == C && dd[(size_t)(C - 2) * 64] == dd[(size_t)(C - 1) * 64]) flag = true;
      const int take = c < a.nmax ? c : a.nmax;
      for (int j = 0; j < take; ++j) out[cnt++] = ii[(size_t)j * 64];
    }
  }
  if (flag) {
    a.qflag[qid] = 1;
    return;
  }
  isort(out, cnt);
  a.nnear[qid] = cnt;
  a.qflag[qid] = 0;
}

// every observation is a neighbour (nmax >= nobs): kriging.F90:640-657
template <int DIM>
__global__ __launch_bounds__(256) void k_search_all(SearchArgs a, int nobs, const double* ocoord_orig,
                                                    Rot3 srot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const int qid = a.q0 + t;
  double q[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) q[k] = a.qcoord[k + DIM * (size_t)qid];
  int* out = a.inear + (size_t)qid * a.mmax;
  int cnt = 0;
  bool removed = !a.cv;
  for (int i = 0; i < nobs; ++i) {
    if (!removed && i == qid) {
      removed = true;
      continue;
    }
    // rotated_dists (src/libkriging/rotation.f90:193-218): rotate (obs - query)
    double df[3] = {0.0, 0.0, 0.0}, r[3];
#pragma unroll
    for (int k = 0; k < DIM; ++k) df[k] = __dsub_rn(ocoord_orig[k + DIM * (size_t)i], q[k]);
    srot.apply<DIM>(df, r);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) s = __dadd_rn(s, sq(r[k]));
    if (s <= a.maxdist2) out[cnt++] = i;
  }
  a.nnear[qid] = cnt;
}

size_t fast_search_lds(int ndim, bool sector, int k, int nmax) {
  const int noct = 1 << ndim;
  const int words = sector ? std::max(k, 2 * ((noct + 1) / 2) + 3 * noct * (nmax + 1)) : k;
  return (size_t)64 * 4 * (size_t)words;
}

hipError_t launch_search_wave(const SearchArgs& a, hipStream_t st) {
  const size_t lds = fast_search_lds(a.ndim, a.sort != 0, a.k, a.nmax);
  dim3 blk(64), grd((a.nq + 63) / 64);
  if (a.sort) {
    if (a.ndim == 2) hipLaunchKernelGGL((k_search_wave<2, true>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_wave<3, true>), grd, blk, lds, st, a);
  } else {
    if (a.ndim == 1) hipLaunchKernelGGL((k_search_wave<1, false>), grd, blk, lds, st, a);
    else if (a.ndim == 2) hipLaunchKernelGGL((k_search_wave<2, false>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_wave<3, false>), grd, blk, lds, st, a);
  }
  return hipGetLastError();
}

hipError_t launch_search_fast(const SearchArgs& a, hipStream_t st) {
  const size_t lds = fast_search_lds(a.ndim, a.sort != 0, a.k, a.nmax);
  dim3 blk(64), grd((a.nq + 63) / 64);
  if (a.sort) {
    if (a.ndim == 2) hipLaunchKernelGGL((k_search_fast<2, true>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_fast<3, true>), grd, blk, lds, st, a);
  } else {
    if (a.ndim == 1) hipLaunchKernelGGL((k_search_fast<1, false>), grd, blk, lds, st, a);
    else if (a.ndim == 2) hipLaunchKernelGGL((k_search_fast<2, false>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_fast<3, false>), grd, blk, lds, st, a);
  }
  return hipGetLastError();
}

bool exact_search_fits_lds(int k) { return (size_t)k * 12 * 32 <= 64 * 1024; }

hipError_t launch_search(const SearchArgs& a, hipStream_t st) {
  if (a.qlist && exact_search_fits_lds(a.k)) {
    const size_t lds = (size_t)a.k * 12 * 32;
    dim3 blk(32), grd((a.nq + 31) / 32);
    if (a.ndim == 1) hipLaunchKernelGGL((k_search<1, true>), grd, blk, lds, st, a);
    else if (a.ndim == 2) hipLaunchKernelGGL((k_search<2, true>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search<3, true>), grd, blk, lds, st, a);
    return hipGetLastError();
  }
  dim3 blk(256), grd((a.nq + 255) / 256);
  if (a.ndim == 1) hipLaunchKernelGGL((k_search<1, false>), grd, blk, 0, st, a);
  else if (a.ndim == 2) hipLaunchKernelGGL((k_search<2, false>), grd, blk, 0, st, a);
  else hipLaunchKernelGGL((k_search<3, false>), grd, blk, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_search_all(const SearchArgs& a, int nobs, const double* ocoord_orig, const Rot3& r,
                             hipStream_t st) {
  dim3 blk(256), grd((a.nq + 255) / 256);
  if (a.ndim == 1) hipLaunchKernelGGL(k_search_all<1>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  else if (a.ndim == 2) hipLaunchKernelGGL(k_search_all<2>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  else hipLaunchKernelGGL(k_search_all<3>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  return hipGetLastError();
}

}  // namespace kk
