// Kernel launchers shared between the host orchestration and the HIP files.
#pragma once
#include <hip/hip_runtime.h>

#include "kk_types.hpp"

namespace kk {

// 3x3 rotation applied like the reference's mat_mul (src/libkriging/rotation.f90:224-246)
struct Rot3 {
  double m[9];  // column-major
  template <int DIM>
  __host__ __device__ void apply(const double* x, double* y) const {
    if (DIM == 1) {
      y[0] = m[0] * x[0];
    } else if (DIM == 2) {
      y[0] = m[0] * x[0] + m[3] * x[1];
      y[1] = m[1] * x[0] + m[4] * x[1];
    } else {
      y[0] = m[0] * x[0] + m[3] * x[1] + m[6] * x[2];
      y[1] = m[1] * x[0] + m[4] * x[1] + m[7] * x[2];
      y[2] = m[2] * x[0] + m[5] * x[1] + m[8] * x[2];
    }
  }
};

hipError_t launch_search(const SearchArgs& a, hipStream_t st);
// exact search of *a.dnq listed queries (device-resident count, heap in LDS):
// a fixed grid that loops over them; a.nq must bound the count
hipError_t launch_search_dn(const SearchArgs& a, hipStream_t st);
bool exact_search_fits_lds(int k);
// two-pass LDS search; a.qflag marks queries the exact kernel must decide
hipError_t launch_search_fast(const SearchArgs& a, hipStream_t st);
// wave-cooperative traversal, same results and flags as launch_search_fast
hipError_t launch_search_wave(const SearchArgs& a, hipStream_t st);
size_t fast_search_lds(int ndim, bool sector, int k, int nmax);
hipError_t launch_search_all(const SearchArgs& a, int nobs, const double* ocoord_orig, const Rot3& r,
                             hipStream_t st);

// precision of the per-node solve
enum Precision : int { PREC_FP64 = 0, PREC_FP32R = 1 };

// neighbour-count buckets: 0 (no neighbours), then one per solve width
// (rows) below; larger counts go to the overflow bucket NBUCKET
constexpr int NBUCKET = 8;
constexpr int BUCKET_ROWS[NBUCKET - 1] = {8, 12, 16, 24, 32, 48, 64};
__host__ __device__ constexpr int bucket_of(int n) {  // 1 <= n <= 64
  return n <= 8 ? 1 : n <= 12 ? 2 : n <= 16 ? 3 : n <= 24 ? 4 : n <= 32 ? 5 : n <= 48 ? 6 : 7;
}

// bucket blocks by neighbour count: counts[NB] and lists (device), returns via counts
// blocks 0..nblock-1, or the nblock blocks listed in sel; blocks with a
// negative neighbour count are left out
// (dnsel: device-resident count of listed blocks; skip: per block, nonzero =
// leave out)
hipError_t launch_bucket(const int* nnear, const int* sel, int nblock, int* counts, int* lists, int cap,
                         hipStream_t st, const int* dnsel = nullptr, const int* skip = nullptr);
// *overflow = 1 when counts[c.idx[i]] > c.bound[i] for some i < c.n (<= 32)
hipError_t launch_check_bounds(const int* counts, const BoundCheck& c, int* overflow, hipStream_t st);
// p[0..n) = v by a kernel (no memset node in captured graphs)
hipError_t launch_fill(int* p, int n, int v, hipStream_t st);
int bucket_width(int b);  // solve width (matrix rows) of bucket b (0 for the empty bucket)
int solve_width(int n);   // smallest solve width holding n rows (BUCKET_ROWS)

// neighbour-set grouping of the blocks sel[0..nsel) (factor reuse), counters
// per round in ctr[SET_NCTR]: [0] blocks without neighbours, [1..NBUCKET]
// work items per bucket, [SET_N1_ITEMS], [SET_N1_BLOCKS] single-neighbour
// sets; bucket k's items start at items + ctr[SET_OFF + k]
constexpr int SET_CURSOR = NBUCKET + 1;
constexpr int SET_N1_ITEMS = SET_CURSOR + NBUCKET + 1;
constexpr int SET_N1_BLOCKS = SET_N1_ITEMS + 1;
constexpr int SET_OFF = SET_N1_BLOCKS + 1;
constexpr int SET_NCTR = 32;
static_assert(SET_OFF + NBUCKET + 1 <= SET_NCTR, "set counters fit");
struct SetGroupArgs {
  const int* nnear;
  const int* inear;
  int mmax;                    // list row stride, at most 64
  const double* rs;            // per-block range scale or null
  const double* ln;            // per-block local nugget or null
  const int* sel;              // blocks of this round (null: 0..nsel-1)
  int nsel;
  const int* skip;             // per block: nonzero = leave out (null: none)
  unsigned long long* table;   // tsize slots, a power of two >= 2 nsel
  size_t tsize;
  int item_max;                // members per work item
  int *rep, *rank, *cnt, *setbase, *setitem, *members;  // per block (block ids index them)
  unsigned long long* hash;    // per block
  int* ctr;                    // SET_NCTR counters of this round
  int* total;                  // members placed so far in this pass
  int2* items;                 // nsel work items (rep, first member slot)
  int* empties;                // blocks without neighbours
};
hipError_t launch_set_group(const SetGroupArgs& g, hipStream_t st);

hipError_t launch_solve(const SolveArgs& a, int width, int prec, bool tableB, hipStream_t st);
hipError_t launch_empty(const SolveArgs& a, hipStream_t st);
// Bunch-Kaufman fallback: scratch holds fallback_words(mmax) doubles per listed block
hipError_t launch_fallback(const SolveArgs& a, double* scratch, hipStream_t st);
__host__ __device__ constexpr size_t fallback_words(int mmax) {
  return 2 * (size_t)(mmax + 1) * (mmax + 1) + 3 * (size_t)(mmax + 1);
}
// leave-one-out estimates and their parameter derivatives, one block per workgroup
// Morton order of nb query points (ndim nd, the fourth coordinate scaled by
// tsc) into perm; work holds order_work_bytes(nb, nd) bytes
size_t order_work_bytes(int nb, int nd);
hipError_t launch_order(const double* bc, int nb, int nd, double tsc, int* perm, void* work, size_t work_bytes,
                        hipStream_t st);
hipError_t launch_loo_grad(const GradArgs& a, hipStream_t st);
// largest of nnear[0..nb) into the device int *out
hipError_t launch_max_count(const int* nnear, int nb, int* out, hipStream_t st);
// device reduction of the LOO squared error and gradient: work holds
// loo_reduce_words() doubles; the NGRAD + 2 sums (sum r^2, sum r dz_p, count)
// are left at its end, work + loo_reduce_words() - (NGRAD + 2)
int loo_reduce_words();
hipError_t launch_loo_reduce(const double* zhat, const double* dz, const double* val, int nb, double* work,
                             hipStream_t st);
// multi-variable block solve (one variable: single-variable systems above 64
// rows); rows + mv_extra_cols(...) must not exceed 128
int mv_extra_cols(int nvar, int std_ck, int unbias, int ndrift = 0);
hipError_t launch_solve_mv(const MVArgs& a, hipStream_t st);
// true when launch_solve_mv runs the FP32 + MFMA + refinement instance, which
// flags blocks for an FP64 re-solve (flag 1)
bool mv_fp32r_instance(const MVArgs& a);
// Bunch-Kaufman fallback of the blocks listed (flag 2 after launch_solve_mv):
// scratch holds mv_fallback_words(a.rows) doubles per listed block; flag 0
// when solved, 3 when the system stays singular (NaN estimates)
hipError_t launch_fallback_mv(const MVArgs& a, const int* list, int nlist, double* scratch, hipStream_t st);
__host__ __device__ constexpr size_t mv_fallback_words(int rows) {
  return 2 * (size_t)(rows + MAXP) * (rows + MAXP) + 2 * (size_t)(rows + MAXP) * MAXVAR + (rows + MAXP);
}
// estimates and variances from stored weights, G lanes per block
hipError_t launch_replay(const ReplayArgs& a, hipStream_t st);
hipError_t launch_copy(const double* src, double* dst, size_t n, hipStream_t st);
// blocks with more than limit neighbours into list (count: device counter)
hipError_t launch_collect_above(const int* nnear, int nblock, int limit, int* count, int* list, hipStream_t st);
// blocks whose flag equals code into list (count: device counter); with
// perm, taken in the order perm lists them
hipError_t launch_collect(const int* flag, int nblock, int code, int* count, int* list, hipStream_t st,
                          const int* perm = nullptr);

// sequential simulation: drop simulated neighbours that lie on one of the
// row's observations (combined-index rows, in place)
hipError_t launch_sgsim_lists(int* nnear, int* inear, int mmax, int nb, int nobs, int ndim, const double* coord,
                              hipStream_t st);
// one dependency level (order[l0..l1)) of every realisation
hipError_t launch_sgsim_level(const SgsimArgs& a, int l0, int l1, hipStream_t st);

}  // namespace kk
