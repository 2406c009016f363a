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
            dis = __dadd_rn(dis, dis2_bnd(q[i], nd[NODE_LO + i], nd[NODE_HI + i]));
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

// sector of a point relative to the query over the spatial dimensions (the
// time coordinate of a space-time search is not sectored, kriging_st.F90:573-604)
template <int DIM>
__device__ inline int octant(const double* o, const double* q) {
  int sec = (__dsub_rn(o[0], q[0]) < 0.0);
  if (DIM >= 2) sec += 2 * (__dsub_rn(o[1], q[1]) < 0.0);
  if (DIM >= 3) sec += 4 * (__dsub_rn(o[2], q[2]) < 0.0);
  return sec;
}

// Exact emulation of the reference search for one query per lane (heap in
// HBM scratch).  Runs over every query, or over the queries the fast kernel
// could not decide (a.qlist).
// LDSHEAP: the heap lives in LDS (lane-interleaved, 32-lane workgroups),
// used when only the queries the fast kernel flagged are left.
constexpr int EXACT_ONE = 4096;
template <int DIM, bool LDSHEAP>
__global__ __launch_bounds__(256) void k_search(SearchArgs a) {
  extern __shared__ double lds_heap[];
  // with a device-resident count (graph-captured passes, a grid sized for
  // the largest count): up to EXACT_ONE queries one per workgroup (lane 0),
  // more one per lane; workgroups and lanes past the count exit
  const int nq = a.dnq ? *a.dnq : a.nq;
  const bool one = a.dnq && nq <= EXACT_ONE;
  const int t = one ? (int)blockIdx.x : (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if ((one && threadIdx.x != 0) || t >= nq) return;
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
  // sequential simulation: kdtree2_n_nearest_maxidx (kdtree2_maxidx.f90:
  // 1928-1964, 2062-2108) over the points before this block on the path,
  // asking for no more candidates than there are (kriging.F90:535-557)
  const int lim = a.simlim >= 0 ? a.simlim + qid : 0x7fffffff;
  const int nn = a.k < lim ? a.k : lim;
  traverse<DIM>(T, q, ballsize, [&](int l, int u) {
    // process_terminal_node, kdtree2_maxidx.f90:1510-1613.  The bucket's
    // points (at most 13, contiguous in tree order) are loaded up front so
    // their latencies overlap; a partial sum above the ball rejects a point
    // exactly when the full sum does, so the full squared distance is formed.
    // (a node with one empty child -- 14 or more coincident points -- is
    // processed as a bucket too, in chunks of 13)
    constexpr int BK = 13;
    for (int c0 = l; c0 <= u; c0 += BK) {
      const int cu = c0 + BK - 1 < u ? c0 + BK - 1 : u;
      double sds[BK];
      int ids[BK];
#pragma unroll
      for (int p = 0; p < BK; ++p) {
        const int i = c0 + p <= cu ? c0 + p : cu;
        double sd = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(T.rdata[k + DIM * (size_t)i], q[k])));
        sds[p] = sd;
        ids[p] = T.ind[i];
      }
      for (int p = 0; p <= cu - c0; ++p) {
        double sd = sds[0];
        int idx = ids[0];
#pragma unroll
        for (int r = 1; r < BK; ++r) {
          sd = p == r ? sds[r] : sd;
          idx = p == r ? ids[r] : idx;
        }
        if (idx >= lim || sd > ballsize) continue;
        if (nfound < nn) {
          double newpri = pq_insert(h, nfound, sd, idx);
          if (nfound == nn) ballsize = newpri;
        } else {
          ballsize = pq_replace_max(h, nn, sd, idx);
        }
      }
    }
  });

  // with the heap in LDS the list is compacted into the heap's index slots
  // (position <= the slot being read), sorted there and written once with
  // streaming stores; otherwise it is built and sorted in the output row
  int cnt = 0;
  auto emit = [&](int idx) {
    if (LDSHEAP) h.I(++cnt) = idx;
    else out[cnt++] = idx;
  };
  if (a.sort) {
    heapsort(h, nn);
    // apply_sector_filter, kriging.F90:1093-1154
    constexpr int NSEC = 1 << (DIM < 3 ? DIM : 3);
    int counts[NSEC];
#pragma unroll
    for (int s = 0; s < NSEC; ++s) counts[s] = 0;
    const int skip_idx = a.cv ? qid : -1;
    for (int c = 1; c <= nn; ++c) {
      const int idx = h.I(c);
      if (idx == skip_idx) continue;
      if (h.D(c) > a.maxdist2) continue;
      const int sec = octant<DIM>(a.ocoord + DIM * (size_t)idx, q);
      int cs = 0;
#pragma unroll
      for (int s = 0; s < NSEC; ++s) cs = (s == sec) ? counts[s] : cs;
      if (cs < a.nmax) {
#pragma unroll
        for (int s = 0; s < NSEC; ++s) counts[s] += (s == sec);
        emit(idx);
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
      if (h.D(c) <= a.maxdist2) emit(idx);
    }
  }
  if (LDSHEAP) {
    for (int i = 2; i <= cnt; ++i) {
      const int t = h.I(i);
      int j = i - 1;
      while (j >= 1 && h.I(j) > t) {
        h.I(j + 1) = h.I(j);
        --j;
      }
      h.I(j + 1) = t;
    }
    for (int i = 1; i <= cnt; ++i) __builtin_nontemporal_store(h.I(i), &out[i - 1]);
  } else {
    isort(out, cnt);
  }
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
    a.nnear[qid] = -1;  // decided by the exact kernel
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
  for (int k = 0; k < DIM; ++k) d = __dadd_rn(d, dis2_bnd(q[k], nd[NODE_LO + k], nd[NODE_HI + k]));
  return d;
}

// squared distance from q to the farthest corner of a node's box
template <int DIM>
__device__ inline double box_ub(const double* nd, const double* q) {
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const double a = fabs(q[k] - nd[NODE_LO + k]), b = fabs(nd[NODE_HI + k] - q[k]);
    const double m = a > b ? a : b;
    d = d + m * m;
  }
  return d;
}

// The deepest node on the query's path that still holds >= k points has them
// all inside its box, so the k-th squared distance is at most the squared
// distance to that box's farthest corner (inflated like the box test);
// huge when the tree holds fewer than k points.
template <int DIM>
__device__ inline double kpoint_bound(const Tree& T, const double* q, int K) {
  int node = T.root;
  for (;;) {
    const int* ni = T.ni + (size_t)node * NODE_I;
    const int left = ni[3], right = ni[4];
    if (left < 0 || right < 0) break;
    const int child = (q[ni[0]] < T.nd[(size_t)node * NODE_D]) ? left : right;
    const int* nc = T.ni + (size_t)child * NODE_I;
    if (nc[2] - nc[1] + 1 < K) break;
    node = child;
  }
  const int* ni = T.ni + (size_t)node * NODE_I;
  if (ni[2] - ni[1] + 1 < K) return 1.7976931348623157e308;
  return box_ub<DIM>(T.nd + (size_t)node * NODE_D, q) * (1.0 + 1e-12);
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
  // pass 2 needs (its bound is next(next(f))).  From the start it is capped
  // by the lane's k-point node bound, inflated by 1e-6 so that the cap also
  // covers next(next(f)) (f = RN_fp32 of the k-th squared distance).
  const double ucap = kpoint_bound<DIM>(a.tree, q, a.k) * (1.0 + 1e-6);
  double ball = ucap;
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
        if (nf == K) ball = fmin((double)f_next(f_next(HK(1))), ucap);
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
        ball = fmin((double)f_next(f_next(HK(1))), ucap);
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
  // plain lists are collected in this lane's column of the pass-1 heap (k >=
  // mmax words, free now), sorted there and written once with streaming stores
  int* col = reinterpret_cast<int*>(lds_d) + lane;
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
      if (cnt < a.mmax) col[64 * cnt++] = idx;
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
    a.nnear[qid] = -1;  // decided by the exact kernel
    return;
  }
  if (SECTOR) {
    isort(out, cnt);
  } else {
    for (int i = 1; i < cnt; ++i) {
      const int t = col[64 * i];
      int j = i - 1;
      while (j >= 0 && col[64 * j] > t) {
        col[64 * (j + 1)] = col[64 * j];
        --j;
      }
      col[64 * (j + 1)] = t;
    }
    for (int i = 0; i < cnt; ++i) __builtin_nontemporal_store(col[64 * i], &out[i]);
  }
  a.nnear[qid] = cnt;
  a.qflag[qid] = 0;
}


// ---------------------------------------------------------------------------
// Sector (octant/quadrant) search in one traversal, without a k-heap.
//
// The reference keeps the k = min(n, 2^ndim * nmax * 4) nearest points, sorts
// them and walks them in ascending order keeping at most nmax per sector
// (kriging.F90:1093-1154).  For a sector s that equals: the nmax nearest
// points of s overall, restricted to those among the k nearest.  So each lane
// keeps, in registers, the C = nmax+1 nearest points of every sector (float
// key RN_fp32(d^2) and tree position), after the maxdist and leave-one-out
// filters, and decides membership in the k-nearest set from a per-lane
// histogram of the keys of every point it has seen:
//   * 64 bins of the key's float bits relative to a per-lane base: eighth-
//     octave bins ending at the key of the lane's initial bound when that
//     bound is tight, otherwise quarter-octave bins (bin 0 catches everything
//     below, and bin 63 everything above unless it holds the bound's key),
//     16-bit counters packed two to a word (a wave that visits more than
//     65000 points leaves its queries to the exact kernel);
//   * jK = the lowest bin whose cumulative count reaches k, tracked as points
//     arrive; its upper edge E bounds the k-th key, so boxes beyond E are
//     pruned, and so are boxes beyond every sector's (nmax+1)-th key;
//   * all counts are of distinct visited points within the lane's bound at
//     the time (the bound only shrinks), hence lower bounds of the true
//     counts below any key, and exact below the final bound, where every
//     kept entry lies; that is all the decisions below use:
//       entry bin < jK, or bin == jK with cum(jK) <= k : #(d2 <= d2_e) <= k, in;
//       entry bin > jK                                   : #(d2 <  d2_e) >= k, out;
//       otherwise the rank of the entry straddles k: a walk over the
//       recorded leaves splits bin jK into 32 sub-bins and counts
//       #(d2 < d2_e) and #(d2 == d2_e) in FP64 for up to NAMB entries, and a
//       last walk counts the entries still undecided.
// Anything the reference's visit order could decide -- equal FP64 distances
// straddling rank k, equal keys at a sector's nmax / nmax+1 cut -- flags the
// query for the exact kernel.
// Every lane takes the same point of a leaf (contiguous in tree order), so
// points are read with uniform (scalar) loads.
// ---------------------------------------------------------------------------
constexpr int NBIN = 64;      // bins per lane, 16-bit counters
// per-lane bin width: key bits >> bsh = exponent and 3 mantissa bits (bsh 20,
// eighth-octave bins, 8 octaves) or 2 (bsh 21, quarter-octave, 16 octaves);
// the 32 sub-bins of a bin are the next 5 key bits
constexpr int SLEAFCAP = 512;
constexpr int MAXVISIT = 65000;
constexpr int NAMB = 4;  // straddling entries resolved per lane

__device__ inline double readlane_d(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int C>
__device__ inline void sector_insert(float (&k)[C], int (&p)[C], float key, int pos) {
#pragma unroll
  for (int j = C - 1; j > 0; --j) {
    if (key < k[j - 1]) {
      k[j] = k[j - 1];
      p[j] = p[j - 1];
    } else if (key < k[j]) {
      k[j] = key;
      p[j] = pos;
    }
  }
  if (key < k[0]) {
    k[0] = key;
    p[0] = pos;
  }
}

__device__ inline int key_bin(float key, unsigned base, unsigned bsh) {
  const int d = (int)(__float_as_uint(key) >> bsh) - (int)(base >> bsh);
  return d < 0 ? 0 : (d >= NBIN ? NBIN - 1 : d);
}

template <int DIM, int C>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_search_sector(SearchArgs a) {
  __shared__ unsigned hist[NBIN / 2 * 64];
  __shared__ int leaves[SLEAFCAP];
  constexpr int NOCT = 1 << DIM;
  constexpr float FINF = __builtin_huge_valf();
  constexpr double BIG = 1.7976931348623157e308;
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
  float ok[NOCT][C];
  int op[NOCT][C];
#pragma unroll
  for (int s = 0; s < NOCT; ++s)
#pragma unroll
    for (int j = 0; j < C; ++j) {
      ok[s][j] = FINF;
      op[s][j] = -1;
    }
#pragma unroll
  for (int b = 0; b < NBIN / 2; ++b) hist[b * 64 + lane] = 0u;
  auto hcount = [&](int b) { return (int)((hist[(b >> 1) * 64 + lane] >> (16 * (b & 1))) & 0xffffu); };

  unsigned base = 0;
  bool have_base = false;
  int jK = NBIN - 1, tot = 0, nleaf = 0, nvisit = 0;
  // before any point is counted: the lane's k-point node bound.  Points
  // beyond it rank below k for this lane, so it caps the bound from the start.
  const double ubound = kpoint_bound<DIM>(T, q, K);
  double ball = ubound;
  // Bin range, set at the first leaf before any point is counted, from mk =
  // the largest key of that leaf:
  //   * the bound within 5 octaves above mk: eighth-octave bins ending at the
  //     bin of the bound's key (every counted point lies at or below it, so
  //     the top bin is exact);
  //   * within 10 octaves: quarter-octave bins ending there (16 octaves);
  //   * otherwise (a loose bound, or none): quarter-octave bins from 8
  //     octaves below mk, the top bin also counting every larger key.
  const unsigned ukey = ubound < BIG ? __float_as_uint((float)ubound) : 0xffffffffu;
  bool anchored = false;
  unsigned bsh = 21;

  // pos = tree position; the observation index is read only for the
  // leave-one-out exclusion
  auto point = [&](const double* x, int pos) {
    double sd = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(x[k], q[k])));
    // beyond this lane's current bound (which only shrinks): not among its k
    // nearest and not among the C nearest of its sector at the end, so it
    // changes no decision; every point within the final bound is still
    // counted, which is what the in/out rules below rely on
    if (!(sd <= ball)) return;
    const float key = (float)sd;
    const int bin = key_bin(key, base, bsh);
    __hip_atomic_fetch_add(&hist[(bin >> 1) * 64 + lane], 1u << (16 * (bin & 1)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    tot += (bin <= jK);
    if (!(sd <= a.maxdist2)) return;
    if (a.cv && T.ind[pos] == qid) return;
    // the octant's (nmax+1)-th key by a select tree on the three sign tests
    // (octant 1 + ix + 2 iy + 4 iz with ix = (dx < 0)); the octant index
    // itself is only formed for an insert
    const bool b0 = __dsub_rn(x[0], q[0]) < 0.0;
    const bool b1 = DIM >= 2 && __dsub_rn(x[1], q[1]) < 0.0;
    const bool b2 = DIM == 3 && __dsub_rn(x[2], q[2]) < 0.0;
    float thr;
    if (DIM == 1) {
      thr = b0 ? ok[1][C - 1] : ok[0][C - 1];
    } else {
      float t[NOCT / 2];
#pragma unroll
      for (int s = 0; s < NOCT / 2; ++s) t[s] = b0 ? ok[2 * s + 1][C - 1] : ok[2 * s][C - 1];
      if (DIM == 2) {
        thr = b1 ? t[1] : t[0];
      } else {
        const float lo = b1 ? t[1] : t[0], hi = b1 ? t[3 % (NOCT / 2)] : t[2 % (NOCT / 2)];
        thr = b2 ? hi : lo;
      }
    }
    if (!(key < thr)) return;
    const int oct = (int)b0 + 2 * (int)b1 + 4 * (int)b2;
#pragma unroll
    for (int s = 0; s < NOCT; ++s)
      if (oct == s) sector_insert<C>(ok[s], op[s], key, pos);
  };

  // a bucket holds at most 13 points; a node with one empty child (14 or
  // more coincident points) is processed as a bucket, in chunks of 64/DIM
  constexpr int CHK = 64 / DIM;
  traverse_wave<DIM>(T, q, q0, [&]() { return ball; }, [&](int l0, int u, int node) {
    if (lane == 0 && nleaf < SLEAFCAP) leaves[nleaf] = node;
    ++nleaf;
    nvisit += u - l0 + 1;
    for (int l = l0; l <= u; l += CHK) {
      const int m = u - l + 1 < CHK ? u - l + 1 : CHK;
      const double v = lane < DIM * m ? T.rdata[DIM * (size_t)l + lane] : 0.0;
      const int iv = lane < m ? T.ind[l + lane] : 0;
      if (!have_base) {
        float mk = 0.0f;
        for (int p = 0; p < m; ++p) {
          double sd = 0.0;
#pragma unroll
          for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(readlane_d(v, DIM * p + k), q[k])));
          mk = fmaxf(mk, (float)sd);
        }
        const unsigned mkey = __float_as_uint(mk);
        if (ukey != 0xffffffffu && (ukey >> 20) <= (mkey >> 20) + 40u) {
          bsh = 20;
          anchored = true;
        } else if (ukey != 0xffffffffu && (ukey >> 21) <= (mkey >> 21) + 40u) {
          anchored = true;
        }
        if (anchored) {
          const unsigned top = ukey >> bsh;
          base = top >= (unsigned)(NBIN - 1) ? (top - (NBIN - 1)) << bsh : 0u;
        } else {
          const unsigned mb = mkey > (1u << 23) ? mkey : (1u << 23);
          base = mb > (32u << 21) ? ((mb - (32u << 21)) >> 21) << 21 : 0u;
        }
        have_base = true;
      }
      // every lane takes the same point: uniform (scalar) loads
      for (int p = 0; p < m; ++p) {
        const int pp = __builtin_amdgcn_readfirstlane(l + p);
        double x[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) x[k] = T.rdata[DIM * (size_t)pp + k];
        point(x, pp);
      }
    }
    // lowest bin reaching k, and the ball: min(upper edge of that bin,
    // next(largest (nmax+1)-th sector key))
    while (jK > 0) {
      const int h = hcount(jK);
      if (tot - h < K) break;
      tot -= h;
      --jK;
    }
    double bE = BIG;
    if (tot >= K && jK < NBIN - 1) {
      const unsigned eb = base + ((unsigned)(jK + 1) << bsh);
      if (eb < 0x7f800000u) bE = (double)__uint_as_float(eb);
    }
    float mo = 0.0f;
#pragma unroll
    for (int s = 0; s < NOCT; ++s) mo = fmaxf(mo, ok[s][C - 1]);
    const double bO = mo < FINF ? (double)f_next(mo) : BIG;
    ball = bE < bO ? bE : bO;
    ball = ball < ubound ? ball : ubound;
  });

  // ---- k-nearest membership of every kept entry ------------------------------
  // state per entry (bit s*C+j): in, or undecided; neither = out
  static_assert(NOCT * C <= 64, "sector lists exceed the state masks");
  unsigned long long in_m = 0, und_m = 0;
  auto bit = [](int s, int j) { return 1ull << (s * C + j); };
  bool flag = nvisit > MAXVISIT;  // 16-bit bin counters could have wrapped
  auto entry_sd = [&](int pos) {
    double sd = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(T.rdata[k + DIM * (size_t)pos], q[k])));
    return sd;
  };
#pragma unroll
  for (int s = 0; s < NOCT; ++s)
#pragma unroll
    for (int j = 0; j < C; ++j) {
      if (op[s][j] < 0) continue;
      const int b = key_bin(ok[s][j], base, bsh);
      if (b < jK || (b == jK && tot <= K)) in_m |= bit(s, j);
      else if (b == jK) und_m |= bit(s, j);
    }
  // only the first nmax entries of an octant are output; the (nmax+1)-th
  // matters only for the key-tie check against the nmax-th, so when their
  // keys differ it is left out unresolved (out: the prefix and tie checks
  // then come out as they would for either membership)
#pragma unroll
  for (int s = 0; s < NOCT; ++s)
    if ((und_m & bit(s, C - 1)) && ok[s][C - 2] != ok[s][C - 1]) und_m &= ~bit(s, C - 1);
  if (__ballot(und_m != 0)) {
    // walks over the points of the recorded leaves (every point with a key
    // below the final ball) whose box is within some lane's bound
    // A recorded leaf is walked for a lane when its box reaches within the
    // lane's bound and is not wholly inside the lane's inner radius (inner
    // > 0: every point of such a leaf has a key below the bin being
    // resolved, so it is already in the count below that bin; the 1e-6
    // margin keeps a point whose key would round up to the bin's edge).
    auto walk = [&](double bound, double inner, auto&& fn) {
      auto leaf = [&](int l0, int mm) {
        for (int p = 0; p < mm; ++p) {
          const int pp = __builtin_amdgcn_readfirstlane(l0 + p);
          double sd = 0.0;
#pragma unroll
          for (int k = 0; k < DIM; ++k) sd = __dadd_rn(sd, sq(__dsub_rn(T.rdata[DIM * (size_t)pp + k], q[k])));
          fn(sd);
        }
      };
      if (nleaf <= SLEAFCAP) {
        for (int j = 0; j < nleaf; ++j) {
          const int node = __builtin_amdgcn_readfirstlane(leaves[j]);
          const double* nd = T.nd + (size_t)node * NODE_D;
          const bool need = box_lb<DIM>(nd, q) <= bound * (1.0 + 1e-12) && !(box_ub<DIM>(nd, q) * (1.0 + 1e-6) < inner);
          if (__ballot(need) == 0) continue;
          const int* ni = T.ni + (size_t)node * NODE_I;
          leaf(ni[1], ni[2] - ni[1] + 1);
        }
      } else {
        traverse_wave<DIM>(T, q, q0, [&]() { return bound; }, [&](int l, int u, int) { leaf(l, u - l + 1); });
      }
    };
    // exact FP64 ranks for up to NAMB distinct undecided distances per walk
    double amb[NAMB];
    int clt[NAMB], ceq[NAMB], namb = 0;
    auto pick = [&]() {
      namb = 0;
#pragma unroll
      for (int z = 0; z < NAMB; ++z) {
        amb[z] = -1.0;
        clt[z] = ceq[z] = 0;
      }
#pragma unroll
      for (int s = 0; s < NOCT; ++s)
#pragma unroll
        for (int j = 0; j < C; ++j) {
          if (!(und_m & bit(s, j))) continue;
          const double sd = entry_sd(op[s][j]);
          bool known = false;
#pragma unroll
          for (int z = 0; z < NAMB; ++z) known |= (z < namb && amb[z] == sd);
          if (!known && namb < NAMB) {
#pragma unroll
            for (int z = 0; z < NAMB; ++z)
              if (z == namb) amb[z] = sd;
            ++namb;
          }
        }
    };
    // in if every point at or below d2_e fits in k, out if k points are
    // strictly closer, otherwise equal distances straddle rank k
    auto settle = [&]() {
#pragma unroll
      for (int s = 0; s < NOCT; ++s)
#pragma unroll
        for (int j = 0; j < C; ++j) {
          if (!(und_m & bit(s, j))) continue;
          const double sd = entry_sd(op[s][j]);
#pragma unroll
          for (int z = 0; z < NAMB; ++z)
            if (z < namb && amb[z] == sd) {
              if (clt[z] + ceq[z] <= K) {
                in_m |= bit(s, j);
                und_m &= ~bit(s, j);
              } else if (clt[z] >= K) {
                und_m &= ~bit(s, j);
              } else {
                flag = true;
              }
            }
        }
    };
    // walk 1: sub-histogram of bin jK (the 5 key bits below the bin's: 1/256
    // or 1/128 octave, 32-bit counters) plus exact ranks for the first NAMB
    // undecided distances.  Undecided entries lie in bin jK, so with an exact
    // bin (subok) their rank is cb plus a count over bin jK alone, and leaves
    // wholly below the bin's lower edge are skipped.
    const int hK = hcount(jK);
    const int cb = tot - hK;  // counted points in bins below jK
    // (bin 0 also holds every smaller key; the top bin is an exact bin when
    // the bins end at the initial bound, otherwise it holds every larger key)
    const bool subok = jK > 0 && (jK < NBIN - 1 || anchored);
    const unsigned ekb = base + ((unsigned)(jK + 1) << bsh);
    const double eK = !subok ? 0.0 : (ekb < 0x7f800000u ? (double)__uint_as_float(ekb) : BIG);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 32; ++b) hist[b * 64 + lane] = 0u;
    pick();
    // a lane with nothing undecided takes no leaf into the walk (bound -1)
    double b1 = und_m ? eK : -1.0;
#pragma unroll
    for (int z = 0; z < NAMB; ++z) b1 = fmax(b1, amb[z]);
    const unsigned lkb = base + ((unsigned)jK << bsh);
    const double inner = (subok && und_m) ? (double)__uint_as_float(lkb) : 0.0;
    walk(b1, inner, [&](double sd) {
      const float key = (float)sd;
      // unclamped bin: the walk also sees points beyond the bound
      const bool inbin = (int)(__float_as_uint(key) >> bsh) - (int)(base >> bsh) == jK;
      if (subok && inbin)
        __hip_atomic_fetch_add(&hist[((__float_as_uint(key) >> (bsh - 5)) & 31) * 64 + lane], 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
      if (!subok || inbin) {
#pragma unroll
        for (int z = 0; z < NAMB; ++z) {
          clt[z] += (sd < amb[z]);
          ceq[z] += (sd == amb[z]);
        }
      }
    });
    if (subok) {
#pragma unroll
      for (int z = 0; z < NAMB; ++z) clt[z] += cb;
    }
    settle();
    if (subok) {
      int c = cb, jK2 = 31, tot2 = cb;
      for (int b = 0; b < 32; ++b) {
        c += (int)hist[b * 64 + lane];
        if (c >= K) {
          jK2 = b;
          tot2 = c;
          break;
        }
        tot2 = c;
      }
#pragma unroll
      for (int s = 0; s < NOCT; ++s)
#pragma unroll
        for (int j = 0; j < C; ++j) {
          if (!(und_m & bit(s, j))) continue;
          const int b = (int)((__float_as_uint(ok[s][j]) >> (bsh - 5)) & 31);
          if (b < jK2 || (b == jK2 && tot2 <= K)) {
            in_m |= bit(s, j);
            und_m &= ~bit(s, j);
          } else if (b > jK2) {
            und_m &= ~bit(s, j);
          }
        }
    }
    // walk 2: exact ranks for what is still undecided
    if (__ballot(und_m != 0)) {
      pick();
      double b2 = -1.0;
#pragma unroll
      for (int z = 0; z < NAMB; ++z) b2 = fmax(b2, amb[z]);
      walk(b2, 0.0, [&](double sd) {
#pragma unroll
        for (int z = 0; z < NAMB; ++z) {
          clt[z] += (sd < amb[z]);
          ceq[z] += (sd == amb[z]);
        }
      });
      settle();
      flag |= und_m != 0;
    }
  }
  // per sector: the in-entries must be a prefix of the key order (only equal
  // keys could break that), take the first nmax, and a key tie at the cut
  // between two in-entries is left to the exact kernel.
  // The list is sorted in this lane's column of the (now free) histogram;
  // the wave then writes its rows with streaming stores, two rows per store
  // (32 lanes per row), so the output rows neither round-trip through L2 for
  // the sort nor evict the tree from it, and each row leaves as one
  // contiguous write.
  constexpr bool LDS_OUT = NOCT * (C - 1) <= NBIN / 2;
  static_assert(!LDS_OUT || NOCT * (C - 1) <= 32, "a row fits half a wave");
  if (!LDS_OUT && !valid) return;
  int* out = a.inear + (size_t)qid * a.mmax;
  int* col = reinterpret_cast<int*>(hist) + lane;
  int cnt = 0;
#pragma unroll
  for (int s = 0; s < NOCT; ++s) {
    int nin = 0;
    bool gap = false;
#pragma unroll
    for (int j = 0; j < C; ++j) {
      if (in_m & bit(s, j)) {
        if (gap) flag = true;
        ++nin;
      } else {
        gap = true;
      }
    }
    if (nin == C && ok[s][C - 2] == ok[s][C - 1]) flag = true;
#pragma unroll
    for (int j = 0; j < C - 1; ++j)
      if (j < nin) {
        const int v = T.ind[op[s][j]];
        if (LDS_OUT)
          col[64 * cnt] = v;
        else
          out[cnt] = v;
        ++cnt;
      }
  }
  if (valid && flag) {
    a.qflag[qid] = 1;
    a.nnear[qid] = -1;  // decided by the exact kernel
  }
  if (LDS_OUT) {
    for (int i = 1; i < cnt; ++i) {
      const int t = col[64 * i];
      int j = i - 1;
      while (j >= 0 && col[64 * j] > t) {
        col[64 * (j + 1)] = col[64 * j];
        --j;
      }
      col[64 * (j + 1)] = t;
    }
    const int wc = (valid && !flag) ? cnt : 0;
    const int* cols = reinterpret_cast<const int*>(hist);
    __syncthreads();  // other lanes' columns are read below
    for (int r = 0; r < 64; r += 2) {
      const int rr = r + (lane >> 5), j = lane & 31;
      const int c = __shfl(wc, rr, 64);
      const int qr = __shfl(qid, rr, 64);
      if (j < c) __builtin_nontemporal_store(cols[64 * j + rr], a.inear + (size_t)qr * a.mmax + j);
    }
  } else if (!flag) {
    isort(out, cnt);
  }
  if (valid && !flag) {
    a.nnear[qid] = cnt;
    a.qflag[qid] = 0;
  }
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
    // rotated_dists (src/libkriging/rotation.f90:193-218): rotate (obs - query);
    // space-time adds (t_obs * at - t_query * at)^2 (kriging_st.F90:609-612)
    double df[4] = {0.0, 0.0, 0.0, 0.0}, r[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < DIM; ++k) df[k] = __dsub_rn(ocoord_orig[k + DIM * (size_t)i], q[k]);
    srot.apply<(DIM < 3 ? DIM : 3)>(df, r);
    if (DIM == 4) r[3] = df[3];
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

template <int DIM>
static void launch_sector(const SearchArgs& a, dim3 grd, dim3 blk, hipStream_t st) {
  switch (a.nmax) {
    case 1: hipLaunchKernelGGL((k_search_sector<DIM, 2>), grd, blk, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_search_sector<DIM, 3>), grd, blk, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_search_sector<DIM, 4>), grd, blk, 0, st, a); break;
    default: hipLaunchKernelGGL((k_search_sector<DIM, 5>), grd, blk, 0, st, a); break;
  }
}

// register-resident sector lists: up to 4 per sector
bool sector_kernel_fits(const SearchArgs& a) { return a.sort && a.ndim >= 2 && a.nmax >= 1 && a.nmax <= 4; }

hipError_t launch_search_wave(const SearchArgs& a, hipStream_t st) {
  const size_t lds = fast_search_lds(a.ndim, a.sort != 0, a.k, a.nmax);
  dim3 blk(64), grd((a.nq + 63) / 64);
  if (sector_kernel_fits(a)) {
    if (a.ndim == 2) launch_sector<2>(a, grd, blk, st);
    else launch_sector<3>(a, grd, blk, st);
    return hipGetLastError();
  }
  if (a.sort) {
    // sectored space-time searches go to the exact kernel
    if (a.ndim == 4) return hipErrorInvalidValue;
    if (a.ndim == 2) hipLaunchKernelGGL((k_search_wave<2, true>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_wave<3, true>), grd, blk, lds, st, a);
  } else {
    if (a.ndim == 1) hipLaunchKernelGGL((k_search_wave<1, false>), grd, blk, lds, st, a);
    else if (a.ndim == 2) hipLaunchKernelGGL((k_search_wave<2, false>), grd, blk, lds, st, a);
    else if (a.ndim == 3) hipLaunchKernelGGL((k_search_wave<3, false>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_wave<4, false>), grd, blk, lds, st, a);
  }
  return hipGetLastError();
}

hipError_t launch_search_fast(const SearchArgs& a, hipStream_t st) {
  const size_t lds = fast_search_lds(a.ndim, a.sort != 0, a.k, a.nmax);
  dim3 blk(64), grd((a.nq + 63) / 64);
  if (a.sort) {
    if (a.ndim == 4) return hipErrorInvalidValue;
    if (a.ndim == 2) hipLaunchKernelGGL((k_search_fast<2, true>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_fast<3, true>), grd, blk, lds, st, a);
  } else {
    if (a.ndim == 1) hipLaunchKernelGGL((k_search_fast<1, false>), grd, blk, lds, st, a);
    else if (a.ndim == 2) hipLaunchKernelGGL((k_search_fast<2, false>), grd, blk, lds, st, a);
    else if (a.ndim == 3) hipLaunchKernelGGL((k_search_fast<3, false>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search_fast<4, false>), grd, blk, lds, st, a);
  }
  return hipGetLastError();
}

bool exact_search_fits_lds(int k) { return (size_t)k * 12 * 32 <= 64 * 1024; }

hipError_t launch_search(const SearchArgs& a, hipStream_t st) {
  if (a.qlist && exact_search_fits_lds(a.k)) {
    // a few flagged queries: one per workgroup, so none waits on another's
    // divergent path; many: 16 per workgroup (measured on 3.35 M space-time
    // queries in Morton order: 38 ms, against 65 ms with 32 and 40 ms with 8)
    const int lanes = a.nq <= 4096 ? 1 : 16;
    const size_t lds = (size_t)a.k * 12 * lanes;
    dim3 blk(lanes), grd((a.nq + lanes - 1) / lanes);
    if (a.ndim == 1) hipLaunchKernelGGL((k_search<1, true>), grd, blk, lds, st, a);
    else if (a.ndim == 2) hipLaunchKernelGGL((k_search<2, true>), grd, blk, lds, st, a);
    else if (a.ndim == 3) hipLaunchKernelGGL((k_search<3, true>), grd, blk, lds, st, a);
    else hipLaunchKernelGGL((k_search<4, true>), grd, blk, lds, st, a);
    return hipGetLastError();
  }
  dim3 blk(256), grd((a.nq + 255) / 256);
  if (a.ndim == 1) hipLaunchKernelGGL((k_search<1, false>), grd, blk, 0, st, a);
  else if (a.ndim == 2) hipLaunchKernelGGL((k_search<2, false>), grd, blk, 0, st, a);
  else if (a.ndim == 3) hipLaunchKernelGGL((k_search<3, false>), grd, blk, 0, st, a);
  else hipLaunchKernelGGL((k_search<4, false>), grd, blk, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_search_dn(const SearchArgs& a, hipStream_t st) {
  if (!a.qlist || !a.dnq || !exact_search_fits_lds(a.k)) return hipErrorInvalidValue;
  // 16 lanes per workgroup; up to EXACT_ONE queries one per workgroup
  // (lane 0), more one per lane: a grid for a.nq queries either way
  if (a.nq <= 0) return hipSuccess;
  constexpr int lanes = 16;
  const size_t lds = (size_t)a.k * 12 * lanes;
  dim3 blk(lanes), grd(std::max((a.nq + lanes - 1) / lanes, std::min(a.nq, EXACT_ONE)));
  if (a.ndim == 1) hipLaunchKernelGGL((k_search<1, true>), grd, blk, lds, st, a);
  else if (a.ndim == 2) hipLaunchKernelGGL((k_search<2, true>), grd, blk, lds, st, a);
  else if (a.ndim == 3) hipLaunchKernelGGL((k_search<3, true>), grd, blk, lds, st, a);
  else hipLaunchKernelGGL((k_search<4, true>), grd, blk, lds, st, a);
  return hipGetLastError();
}

hipError_t launch_search_all(const SearchArgs& a, int nobs, const double* ocoord_orig, const Rot3& r,
                             hipStream_t st) {
  dim3 blk(256), grd((a.nq + 255) / 256);
  if (a.ndim == 1) hipLaunchKernelGGL(k_search_all<1>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  else if (a.ndim == 2) hipLaunchKernelGGL(k_search_all<2>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  else if (a.ndim == 3) hipLaunchKernelGGL(k_search_all<3>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  else hipLaunchKernelGGL(k_search_all<4>, grd, blk, 0, st, a, nobs, ocoord_orig, r);
  return hipGetLastError();
}

}  // namespace kk
