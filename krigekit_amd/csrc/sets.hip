// Neighbour-set grouping for factor reuse: blocks whose kriging matrices are
// identical -- the same neighbour count and ascending list, the same range
// scale and local nugget -- are solved from one factorisation, the reference's
// factor cache (fcache_matches / fhcache_lookup, kriging_base.F90:1746-1917,
// 2445-2485) done as a grouping pass over a batch of blocks:
//
//   k_set_init     clears the hash table and the per-block member counters;
//   k_set_hash     hashes each block's defining data (one coalesced row read
//                  per block);
//   k_set_find     finds each block's set in an open-addressing table keyed
//                  by the hash: the block whose insert wins a slot is the
//                  set's representative, blocks with the same hash join it;
//   k_set_verify   compares every member's data with its representative's:
//                  a block whose hash collided is a set of its own;
//   k_set_rank     ranks each block within its set;
//   k_set_items    (count pass, then fill pass) reserves each set's
//                  contiguous member range and its work items (at most
//                  min(item_max, item_cap(rows)) members each) in the bucket
//                  of its row count;
//   k_set_members  writes each block at its rank in its set's range, and
//                  the work item that starts at it.
// Counters are taken with one atomic per distinct key in a wave (set ranks)
// or per counter in a workgroup (the few counters of the item pass, summed
// in LDS first).
//
// Which block represents a set depends on the order of the inserts, but every
// member's matrix is identical, so the results do not.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kk_device.hpp"
#include "kk_types.hpp"

namespace kk {

constexpr unsigned long long SET_EMPTY = ~0ull;

__device__ inline unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// counters[key] += amount for every active lane, with one atomic per distinct
// key in the wave (blocks in Morton order share sets and buckets, so a few);
// each lane gets its exclusive offset: the counter's old value plus the
// amounts of the lanes below it with the same key
__device__ inline int wave_alloc(int* counters, int key, int amount, bool active) {
  const int lane = threadIdx.x & 63;
  int result = 0;
  unsigned long long pending = __ballot(active);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int k = __shfl(key, leader);
    const bool mine = active && key == k;
    const unsigned long long m = __ballot(mine);
    // inclusive prefix sum of the group's amounts over the wave, lane order
    int incl = mine ? amount : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const int tot = __shfl(incl, 63);
    int base = 0;
    if (lane == leader) base = atomicAdd(&counters[k], tot);
    base = __shfl(base, leader);
    if (mine) result = base + incl - amount;
    pending &= ~m;
  }
  return result;
}

__global__ void k_set_init(const int* sel, int nsel, unsigned long long* table, size_t tsize, int* cnt, int* ctr,
                           int nctr) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = t; i < tsize; i += stride) table[i] = SET_EMPTY;
  for (size_t i = t; i < (size_t)nsel; i += stride) cnt[sel ? sel[i] : (int)i] = 0;
  if (t < (size_t)nctr) ctr[t] = 0;
}

// A block's list row as seen by lane w of the L lanes that read it, V
// indices per lane (V = 4: 16-byte loads; rows then start 16-byte aligned)
template <int V>
__device__ inline void row_words(const int* inear, int mmax, int b, int w, int n, int (&v)[V]) {
  const int* row = inear + (size_t)b * mmax + V * w;
  if (V == 4 && V * w + 3 < n) {
    const int4 x = *reinterpret_cast<const int4*>(row);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
  } else {
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = V * w + i < n ? row[i] : 0;
  }
}

// Hash of each block's defining data, L lanes per block reading V indices
// each (L V >= mmax, so a block's list is one coalesced row read): the sum of
// a mix of (position, index) over the list, mixed with the count, range
// scale and local nugget.
template <int L, int V>
__global__ __launch_bounds__(256) void k_set_hash(const int* nnear, const int* inear, int mmax, const double* rs,
                                                  const double* ln, const int* sel, int nsel,
                                                  unsigned long long* hash) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int w = (int)(t % L);
  const size_t q = t / L;
  const int b = q < (size_t)nsel ? (sel ? sel[q] : (int)q) : -1;
  const int n = b >= 0 ? nnear[b] : 0;
  unsigned long long h = 0;
  if (b >= 0 && V * w < n) {
    int v[V];
    row_words<V>(inear, mmax, b, w, n, v);
#pragma unroll
    for (int i = 0; i < V; ++i)
      if (V * w + i < n) h += mix64(((unsigned long long)(V * w + i) << 32) | (unsigned int)v[i]);
  }
#pragma unroll
  for (int o = L / 2; o >= 1; o >>= 1) h += __shfl_xor(h, o, L);
  if (b < 0 || w != 0 || n <= 0) return;
  unsigned long long e = mix64((unsigned long long)(unsigned int)n);
  if (rs) e ^= mix64((unsigned long long)__double_as_longlong(rs[b]) + 1ull);
  if (ln) e ^= mix64((unsigned long long)__double_as_longlong(ln[b]) + 2ull);
  hash[b] = mix64(h ^ e);
}

// Each block's representative from the hash table: the block whose insert
// won the slot of its 64-bit hash (equal 32-bit tags of different hashes
// probe on).  k_set_verify then checks the data themselves.
__global__ __launch_bounds__(256) void k_set_find(const int* nnear, const int* sel, int nsel, const int* skip,
                                                  const unsigned long long* hash, unsigned long long* table,
                                                  unsigned int tmask, int* rep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsel) return;
  const int b = sel ? sel[t] : t;
  const int n = nnear[b];
  // no neighbours, more than the row-per-lane solver's 64 rows (solved
  // apart), or left to the exact search (whose lists may still be in
  // flight): no set
  if (n <= 0 || n > 64 || (skip && skip[b])) {
    rep[b] = -1;
    return;
  }
  const unsigned long long h = hash[b];
  const unsigned int tag = (unsigned int)(h >> 32);
  const unsigned long long mine = ((unsigned long long)tag << 32) | (unsigned int)b;
  unsigned int slot = (unsigned int)h & tmask;
  int r = b;
  // the table holds at least twice as many slots as blocks, so a free slot
  // is always found within tmask + 1 probes
  for (unsigned int probe = 0; probe <= tmask; ++probe) {
    unsigned long long cur = table[slot];
    if (cur == SET_EMPTY) {
      const unsigned long long old = atomicCAS(&table[slot], SET_EMPTY, mine);
      if (old == SET_EMPTY) break;
      cur = old;
    }
    const int c = (int)(unsigned int)cur;
    if ((unsigned int)(cur >> 32) == tag && hash[c] == h) {
      r = c;
      break;
    }
    slot = (slot + 1) & tmask;
  }
  rep[b] = r;
}

// Every member's data against its representative's, L lanes per block as in
// k_set_hash: a block whose hash collided becomes a set of its own.
template <int L, int V>
__global__ __launch_bounds__(256) void k_set_verify(const int* nnear, const int* inear, int mmax, const double* rs,
                                                    const double* ln, const int* sel, int nsel, int* rep) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int w = (int)(t % L);
  const size_t q = t / L;
  const int b = q < (size_t)nsel ? (sel ? sel[q] : (int)q) : -1;
  const int c = b >= 0 ? rep[b] : -1;
  const bool check = c >= 0 && c != b;
  const int n = check ? nnear[b] : 0;
  bool same = true;
  if (check) {
    if (w == 0) {
      same = nnear[c] == n;
      if (rs) same = same && __double_as_longlong(rs[c]) == __double_as_longlong(rs[b]);
      if (ln) same = same && __double_as_longlong(ln[c]) == __double_as_longlong(ln[b]);
    }
    if (V * w < n) {
      int vb[V], vc[V];
      row_words<V>(inear, mmax, b, w, n, vb);
      row_words<V>(inear, mmax, c, w, n, vc);
#pragma unroll
      for (int i = 0; i < V; ++i) same = same && vb[i] == vc[i];
    }
  }
  int ok = same ? 1 : 0;
#pragma unroll
  for (int o = L / 2; o >= 1; o >>= 1) ok &= __shfl_xor(ok, o, L);
  if (check && !ok && w == 0) rep[b] = b;
}

// ranks within the sets, a block per lane (one atomic per distinct set in a
// wave of 64 blocks)
__global__ __launch_bounds__(256) void k_set_rank(const int* sel, int nsel, const int* rep, int* rank, int* cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = t < nsel ? (sel ? sel[t] : t) : -1;
  const int r = b >= 0 ? rep[b] : -1;
  const int rk = wave_alloc(cnt, r >= 0 ? r : 0, 1, r >= 0);
  if (r >= 0) rank[b] = rk;
}

// count pass (fill == 0): per bucket the number of work items into ctr[0..8]
// (bucket 0: blocks without neighbours, NBUCKET: more than 64 rows), and the
// work items and blocks of single-neighbour sets into ctr[SET_N1_ITEMS],
// ctr[SET_N1_BLOCKS].  Fill pass: each set's member range from the pass-wide
// running total *total, its work items at ctr[SET_OFF + k] (bucket k's start
// in items) through the cursors ctr[SET_CURSOR + k]; blocks without
// neighbours into empties.
constexpr int ITEM_T = 1024;  // threads per workgroup of the item pass
__global__ __launch_bounds__(ITEM_T) void k_set_items(const int* nnear, const int* sel, int nsel, const int* skip,
                                                      const int* rep, const int* cnt, int item_max, int fill,
                                                      int* ctr, int* total, int* setbase, int* setitem,
                                                      int* empties) {
  // per-workgroup sums (count pass) or cursors (fill pass) in LDS, then one
  // global atomic per counter per workgroup
  __shared__ int lc[SET_NCTR];
  __shared__ int lb[NBUCKET + 1];
  __shared__ int ltot, lbase;
  if (threadIdx.x < SET_NCTR) lc[threadIdx.x] = 0;
  if (threadIdx.x == 0) ltot = 0;
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = t < nsel ? (sel ? sel[t] : t) : -1;
  const bool in = b >= 0 && !(skip && skip[b]);
  const int n = in ? nnear[b] : -1;
  const bool empty = n == 0;
  const bool isrep = n > 0 && rep[b] == b;
  const int c = isrep ? cnt[b] : 0;
  const int im = min(item_max, item_cap(n));
  const int ni = (c + im - 1) / im;
  const int k = n > 64 ? NBUCKET : n > 0 ? bucket_of(n) : 0;
  if (!fill) {
    if (empty) atomicAdd(&lc[0], 1);
    if (isrep) {
      atomicAdd(&lc[k], ni);
      if (n == 1) {
        atomicAdd(&lc[SET_N1_ITEMS], ni);
        atomicAdd(&lc[SET_N1_BLOCKS], c);
      }
    }
    __syncthreads();
    if (threadIdx.x < SET_NCTR && lc[threadIdx.x]) atomicAdd(&ctr[threadIdx.x], lc[threadIdx.x]);
    return;
  }
  int e = 0, s0 = 0, base = 0;
  if (empty) e = atomicAdd(&lc[0], 1);
  if (isrep) {
    s0 = atomicAdd(&lc[k], ni);
    base = atomicAdd(&ltot, c);
  }
  __syncthreads();
  if (threadIdx.x <= NBUCKET) lb[threadIdx.x] = lc[threadIdx.x] ? atomicAdd(&ctr[SET_CURSOR + threadIdx.x], lc[threadIdx.x]) : 0;
  if (threadIdx.x == 0) lbase = ltot ? atomicAdd(total, ltot) : 0;
  __syncthreads();
  if (empty) empties[lb[0] + e] = b;
  if (!isrep) return;
  setbase[b] = lbase + base;
  setitem[b] = ctr[SET_OFF + k] + lb[k] + s0;
}

// bucket starts in the item array: exclusive prefix over buckets 1..NBUCKET
__global__ void k_set_offsets(int* ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int acc = 0;
  ctr[SET_OFF] = 0;
  for (int k = 1; k <= NBUCKET; ++k) {
    ctr[SET_OFF + k] = acc;
    acc += ctr[k];
  }
}

// each block at its rank in its set's member range; the block at every
// item_max-th rank writes the work item that starts there
__global__ __launch_bounds__(256) void k_set_members(const int* nnear, const int* sel, int nsel, const int* rep,
                                                     const int* rank, const int* setbase, const int* setitem,
                                                     int item_max, int* members, int2* items) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsel) return;
  const int b = sel ? sel[t] : t;
  const int r = rep[b];
  if (r < 0) return;
  const int rk = rank[b], base = setbase[r];
  const int im = min(item_max, item_cap(nnear[r]));
  members[base + rk] = b;
  if (rk % im == 0) items[setitem[r] + rk / im] = make_int2(r, base + rk);
}

hipError_t launch_set_group(const SetGroupArgs& g, hipStream_t st) {
  if (g.nsel <= 0) return hipSuccess;
  const dim3 blk(256), grd((g.nsel + 255) / 256);
  const size_t work = g.tsize > (size_t)g.nsel ? g.tsize : (size_t)g.nsel;
  const unsigned int nblk = (unsigned int)((work + 255) / 256 < 8192 ? (work + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_set_init, dim3(nblk), blk, 0, st, g.sel, g.nsel, g.table, g.tsize, g.cnt, g.ctr,
                     SET_NCTR);
  // L lanes of V indices per list row
  auto coop = [&](auto lanes, auto words) -> hipError_t {
    constexpr int L = decltype(lanes)::value, V = decltype(words)::value;
    const dim3 cgrd((unsigned int)(((size_t)g.nsel * L + 255) / 256));
    hipLaunchKernelGGL((k_set_hash<L, V>), cgrd, blk, 0, st, g.nnear, g.inear, g.mmax, g.rs, g.ln, g.sel, g.nsel,
                       g.hash);
    hipLaunchKernelGGL(k_set_find, grd, blk, 0, st, g.nnear, g.sel, g.nsel, g.skip, g.hash, g.table,
                       (unsigned int)(g.tsize - 1), g.rep);
    hipLaunchKernelGGL((k_set_verify<L, V>), cgrd, blk, 0, st, g.nnear, g.inear, g.mmax, g.rs, g.ln, g.sel, g.nsel,
                       g.rep);
    return hipGetLastError();
  };
  using I1 = std::integral_constant<int, 1>;
  using I4 = std::integral_constant<int, 4>;
  hipError_t e;
  if (g.mmax % 4 == 0) {
    e = g.mmax <= 16   ? coop(std::integral_constant<int, 4>{}, I4{})
        : g.mmax <= 32 ? coop(std::integral_constant<int, 8>{}, I4{})
                       : coop(std::integral_constant<int, 16>{}, I4{});
  } else {
    e = g.mmax <= 8    ? coop(std::integral_constant<int, 8>{}, I1{})
        : g.mmax <= 16 ? coop(std::integral_constant<int, 16>{}, I1{})
        : g.mmax <= 32 ? coop(std::integral_constant<int, 32>{}, I1{})
                       : coop(std::integral_constant<int, 64>{}, I1{});
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_set_rank, grd, blk, 0, st, g.sel, g.nsel, g.rep, g.rank, g.cnt);
  const dim3 iblk(ITEM_T), igrd((g.nsel + ITEM_T - 1) / ITEM_T);
  hipLaunchKernelGGL(k_set_items, igrd, iblk, 0, st, g.nnear, g.sel, g.nsel, g.skip, g.rep, g.cnt, g.item_max, 0,
                     g.ctr, g.total, g.setbase, g.setitem, g.empties);
  hipLaunchKernelGGL(k_set_offsets, dim3(1), dim3(64), 0, st, g.ctr);
  hipLaunchKernelGGL(k_set_items, igrd, iblk, 0, st, g.nnear, g.sel, g.nsel, g.skip, g.rep, g.cnt, g.item_max, 1,
                     g.ctr, g.total, g.setbase, g.setitem, g.empties);
  hipLaunchKernelGGL(k_set_members, grd, blk, 0, st, g.nnear, g.sel, g.nsel, g.rep, g.rank, g.setbase, g.setitem,
                     g.item_max, g.members, g.items);
  return hipGetLastError();
}

}  // namespace kk
