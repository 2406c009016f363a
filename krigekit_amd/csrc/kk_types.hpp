// Shared host/device data layouts of the MI355X kriging backend.
//
// Everything the GPU reads is flattened into plain arrays: the kd-tree
// (node records + rearranged coordinates), the covariance model (structure
// parameters + lookup tables) and the per-solve block/observation arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kk {

constexpr int MAXS = 8;        // nested structures per variogram
constexpr int MAXDIM = 3;
constexpr int NODE_I = 6;      // ints per tree node: cut_dim, l, u, left, right, parent
constexpr int NODE_D = 11;     // doubles per tree node: cut_val, cut_left, cut_right, lo[4], hi[4]
constexpr int NODE_LO = 3;     // box lower bounds (up to 4 search dimensions: space-time)
constexpr int NODE_HI = 7;     // box upper bounds

// variogram shape ids (src/libkriging/vgmfunc.f90:6-16)
enum VType : int { NUG = 0, SPH = 1, EXP = 2, GAU = 3, HOL = 4, POW = 5, BSQ = 6,
                   CIR = 7, LIN = 8, CYC = 9, DCO = 10 };

// covariance evaluation mode
enum CovMode : int { COV_ANALYTIC = 0, COV_TABLE_B = 1, COV_TABLE_A = 2 };

// Covariance table zones: bounds h = 0, .1, .5, 3.5 with steps 1e-5, 1e-4,
// 1e-3; entries per zone max(1, nint(width/step)), 1/step = entries/width
// (src/libkriging/variogram.f90:223-255).  The host table builder and the
// device lookups both take the zone parameters from here.
constexpr double TAB_HB[4] = {0.0, 0.1, 0.5, 3.5};
constexpr double TAB_DH[3] = {1e-5, 1e-4, 1e-3};
constexpr int tab_zn(int k) {
  const int n = (int)((TAB_HB[k + 1] - TAB_HB[k]) / TAB_DH[k] + 0.5);
  return n < 1 ? 1 : n;
}
constexpr int tab_zoff(int k) { return k == 0 ? 0 : tab_zoff(k - 1) + tab_zn(k - 1); }
constexpr double tab_zinv(int k) { return (double)tab_zn(k) / (TAB_HB[k + 1] - TAB_HB[k]); }
constexpr int TAB_N = tab_zoff(3);  // entries 1..TAB_N follow C(0) at entry 0

// Covariance model of one variable pair, passed by value to kernels.
struct Model {
  int ndim;
  int mode;
  int nstruct;
  int b_tail;                 // Level-B table has an analytic tail
  int simple;                 // every structure has a simple shape (nugget, spherical,
                              // exponential, Gaussian, bounded-square, linear)
  int pad0;
  int vtype[MAXS];
  int is_product[MAXS];
  double sill[MAXS];
  double nugget[MAXS];
  double mat[MAXS][9];        // per-structure anisotropy matrix, column-major
  double bmat[9];             // Level-B table matrix (first non-nugget structure)
  double cov0;
  // piecewise-uniform table zones (h_bounds 0, .1, .5, 3.5)
  double zb[4];
  double zinv[3];
  int zoff[3];
  int zn[3];
  int ntab;                   // entries 0..ntab per table
  const double* tab;          // Level B: ntab+1; Level A: nstruct*(ntab+1)
};

// kd-tree as flattened arrays in device memory
struct Tree {
  int dim;
  int n;
  int nnodes;
  int root;
  const double* rdata;        // (dim, n) coordinates in tree order
  const int* ind;             // tree position -> 0-based observation index
  const int* ni;              // nnodes * NODE_I
  const double* nd;           // nnodes * NODE_D
};

struct SearchArgs {
  Tree tree;
  int ndim;
  int nq;                     // queries in this launch
  int q0;                     // first query id of this launch
  const double* qcoord;       // (ndim, nblock) query coordinates (already rotated)
  const double* ocoord;       // (ndim, nobs) original-order obs coordinates (rotated if aniso)
  int k;                      // candidates requested from the tree (heap size)
  int sort;                   // sector mode: heapsort then octant filter
  int nmax;                   // per-sector cap (sector) or list size (plain)
  int cv;                     // leave-one-out: skip query's own index (query id == obs id)
  int all_obs;                // nmax >= nobs: every observation, no tree
  double maxdist2;
  int mmax;                   // row stride of inear
  double* hdis;               // heap scratch: k * hstride
  int* hidx;
  int hstride;
  int* nnear;                 // [nblock]
  int* inear;                 // [nblock * mmax], 0-based obs ids, ascending
  const int* qlist;           // exact kernel: query ids to process (nullptr: q0..q0+nq)
  const int* dnq;             // exact kernel: device-resident query count (graph-captured passes)
  int* qflag;                 // fast kernel: 1 where the exact kernel must decide
  // sequential simulation (exact kernel): query qid sees only the points with
  // index < simlim + qid (the observations, then the blocks earlier on the
  // path) and takes min(k, simlim + qid) candidates; -1 = every point
  int simlim = -1;
};

// Sequential simulation, one dependency level at a time (blocks in path
// order p; combined neighbour indices: observations, then blocks by path
// position + nobs)
struct SgsimArgs {
  int nb, nobs, nsim, mmax;
  int unbias;
  double sk_mean;
  double bound_lo, bound_hi;
  const int* order;           // blocks grouped by level
  const int* nnear;           // [nb]
  const int* inear;           // [nb * mmax]
  const double* weights;      // [nb * mmax], matching inear
  const double* var;          // [nb] kriging variance
  const double* ovalue;       // [nobs]
  const double* sample;       // [nsim * nb], standard normal draw s of path position p at s + nsim p
  double* val;                // [nsim * nb], realisation s of path position p at s + nsim p
};

// Leave-one-out gradient: per block, the estimate and its derivatives with
// respect to the variogram parameters of the single structure.
constexpr int NGRAD = 8;  // nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge
struct GradArgs {
  const Model* model;
  int ndim;
  int nb;
  int mmax;
  int rows;                   // largest neighbour count of any block (<= mmax)
  const double* bcoord;       // (ndim, nb) targets (the observations themselves)
  const double* ocoord;       // (ndim, nobs)
  const double* ovalue;
  const double* ovar;
  const int* nnear;
  const int* inear;
  double dmat[6][9];          // d(anisotropy matrix)/d(a_major, a_minor1, a_minor2, azimuth, dip, plunge)
  double* zhat;               // [nb]
  double* dz;                 // [nb * NGRAD]
};

// Multi-variable (co)kriging, one workgroup per block.
constexpr int MAXVAR = 4;
constexpr int MAXDRIFT = 4;                // external drift functions
constexpr int MAXP = MAXVAR + MAXDRIFT;    // constraint rows: drift + unbiasedness
constexpr int MAXGRAD = 32;                // increment observations over all variables
struct MVArgs {
  const Model* models;        // vgm(a+1, b+1) at a + nvar * b
  int ndim;
  int nvar;
  int nb;
  int std_ck;                 // one unbiasedness row per variable (else one in total)
  int unbias;                 // 0: simple kriging, no constraint rows
  int nsub;                   // integration nodes per block (1 = point)
  int rows;                   // largest number of data rows (sum of the variables' mmax)
  int simple;                 // every structure of every model has a simple shape
  double sk_mean[MAXVAR];
  double bound_lo, bound_hi;
  const double* bcoord;       // (ndim, nb) block centres
  const double* scoord;       // (ndim, nb * nsub) integration nodes
  const double* sweight;      // (nb * nsub)
  const double* rangescale;   // [nb] or null
  const double* localnugget;  // [nb] or null
  const double* ocoord[MAXVAR];
  const double* ovalue[MAXVAR];
  const double* ovar[MAXVAR];
  const int* nnear[MAXVAR];
  const int* inear[MAXVAR];
  int mmax[MAXVAR];
  const int* list;            // blocks to solve (null: 0..nb-1)
  int nlist;
  int* ctr;                   // blocks sent to the fallback (or null)
  // kriging weights out, weights[k + mmax_w * (g + nvar * (q + nvar * b))]
  // for neighbour k of group g and target q (or null: not written);
  // weights_only leaves est and var untouched (k_fallback_mv also the flags)
  double* weights;
  int mmax_w;
  int weights_only;
  double* est;                // est[b + nb * q]
  double* var;                // var[b + nb * (i + nvar * j)]
  int* flag;                  // 1: FP32 pivot too small (FP64 re-solve); 2: K or S not positive
                              // definite (fallback); 3: fallback failed (NaN)
  int prec;                   // PREC_FP64 or PREC_FP32R (FP32 + MFMA factorisation, FP64 refinement)
  double pivot_tol;           // FP32 pivot ratio below which a block is re-solved in FP64
  // external drift (kriging.F90:675-849, 864-939): ndrift constraint rows
  // before the unbiasedness rows, with the drift values at the observations,
  // odrift[v][k + ndrift * i], and at the blocks, bdrift[k + ndrift * (q +
  // nvar * b)] for target variable q, as right-hand sides
  int ndrift;
  const double* odrift[MAXVAR];
  const double* bdrift;
  // weight correction (nvar 1, kriging_base.F90:3388-3391): negative weights
  // clipped to 0 and the rest rescaled to sum 1 before the estimate and
  // variance (needs weights, the weights instance)
  int wcorr;
  // increment (gradient-pair) observations z(x1) - z(x2) (kriging.F90:796-840,
  // 904-920): ngrad[v] of variable v join every block's system after the
  // neighbour groups, in variable order; coordinates (ndim, ngrad[v]),
  // drift values (ndrift, ngrad[v])
  int ngrad[MAXVAR];
  const double* gx1[MAXVAR];
  const double* gx2[MAXVAR];
  const double* gval[MAXVAR];
  const double* gvar[MAXVAR];
  const double* gdrift[MAXVAR];
};

// member blocks per work item of the single-variable solve (factor reuse):
// at most MAX_ITEM, fewer for wider systems, whose extra right-hand-side
// columns cost registers
constexpr int MAX_ITEM = 4;
__host__ __device__ constexpr int item_cap(int rows) { return rows <= 16 ? 4 : rows <= 32 ? 2 : 1; }

// counters and the bounds a graph-captured pass's launches were sized for
struct BoundCheck {
  int n;
  int idx[32];
  int bound[32];
};

struct SolveArgs {
  const Model* model;         // device copy of the covariance model (space-time: spatial part)
  const Model* tmodel;        // space-time only: temporal structures (1-D)
  double k_ps;                // space-time product-sum coefficient
  double st_cov0;             // space-time C(0, 0): k Cs0 Ct0 + Cs0 + Ct0 (product-sum),
                              // Cs0 + Ct0 + joint sills of non-nugget structures (sum-metric)
  // space-time sum-metric instead of product-sum: Cs + Ct + sum_k sill_st[k]
  // shape_k(sqrt(h_k^2 + f(dt)^2)), f(dt) = max(t_nug + t_sill (1 -
  // corefunc(t_vtype, |dt| / t_at)), 0), f(0) = 0
  int st_sum;
  int t_vtype;
  double t_nug, t_sill, t_at;
  double sill_st[MAXS];
  int ndim;
  int unbias;                 // 1 ordinary, 0 simple
  double sk_mean;
  double bound_lo, bound_hi;
  int nsub;                   // integration points per block (1 = point)
  const double* bcoord;       // (ndim, nblock) block centres
  const double* scoord;       // (ndim, nblock*nsub) sub-points or nullptr
  const double* sweight;      // nblock*nsub or nullptr
  const double* base_cov;     // per-block prior covariance (block kriging) or nullptr
  const double* rangescale;   // nblock or nullptr
  const double* localnugget;  // nblock or nullptr
  const double* ocoord;       // (ndim, nobs)
  const double* ovalue;       // nobs
  const double* ovar;         // nobs
  const int* nnear;
  const int* inear;
  int mmax;
  const int* list;            // block ids of this bucket (when items is null)
  int nlist;                  // blocks, or work items
  // device-resident count and offset (graph-captured passes): when dn is set
  // the list holds *dn entries and starts *doff entries in (doff may be
  // null); nlist then bounds the count (the launch is sized for it)
  const int* dn;
  const int* doff;
  int grid_cap;               // k_empty / k_fallback: at most this many workgroups, looping (0: one pass)
  // factor reuse: work items (representative block, first member slot), each
  // with min(item_max, setbase[rep] + setcnt[rep] - slot) members in members
  const int2* items;
  const int* setcnt;
  const int* setbase;
  const int* members;
  int item_max;
  int* ctr;                   // [0] flagged work items, [1] their blocks (or null)
  double pivot_tol;           // FP32 pivot ratio below which a system is re-solved in FP64
  double* est;
  double* var;
  double* weights;            // k_fallback: kriging weights out, nblock*mmax (or nullptr)
  int weights_only;           // k_fallback: write the weights only (est/var/flag untouched)
  int* flag;                  // per block: 0 ok, 1 needs FP64 re-solve, 2 Cholesky failed
};

// Estimates from stored weights (use_old_weight), the weight store's layouts:
// nnear[g + ng b], inear[k + nm (g + ng b)] (1-based), weight[k + nm (g + ng
// (q + nvar b))], var[i + nvar (j + nvar b)]; group g holds neighbours of
// variable g
struct ReplayArgs {
  int nb, nm, ng, nvar;
  const int* nnear;
  const int* inear;
  const double* weight;
  const double* wvar;
  const double* ovalue[MAXVAR];
  int nobs[MAXVAR];           // a stored id outside 1..nobs reads as NaN
  int std_ck, unbias;
  double sk_mean[MAXVAR];
  double bound_lo, bound_hi;
  double* est;                // est[b + nb * q]
  double* var;                // var[b + nb * (i + nvar * j)]
};

}  // namespace kk
