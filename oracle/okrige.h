/* CPU restatement of the reference kriging hot path -- TEST INFRASTRUCTURE.
 *
 * This is the parity oracle.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never does.
 * Every routine cites the reference file:line it restates (paths relative to
 * the reference repository root).
 */
#ifndef OKRIGE_H
#define OKRIGE_H

#ifdef __cplusplus
extern "C" {
#endif

/* one nested variogram structure, as passed to krige_set_vgm
 * (src/libkriging/kriging_capi.F90:245-270) */
typedef struct {
  int vtype;        /* 0 nug 1 sph 2 exp 3 gau 4 hol 5 pow 6 bsq 7 cir 8 lin 9 cyc 10 dco */
  int is_product;
  double nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge;
} okr_struct;

typedef struct {
  int ndim;                 /* 2 or 3 */
  int nobs;
  const double *obs_coord;  /* (ndim, nobs) column-major */
  const double *obs_value;  /* nobs */
  const double *obs_var;    /* nobs, may be NULL (zeros) */
  int nmax;                 /* as given to set_obs (INT_MAX = unlimited) */
  double maxdist;           /* as given to set_obs (DBL_MAX = unlimited) */
  double sk_mean;
  int unbias;               /* 1 ordinary, 0 simple */
  int sector;               /* sector (quadrant/octant) search */
  int cross_validation;     /* leave-one-out: blocks are the observations */
  int aniso_search;
  double s_anis1, s_anis2, s_azimuth, s_dip, s_plunge;
  int nblock;               /* ignored for cross-validation */
  const double *block_coord;/* (ndim, nblock) block centres */
  int nsub;                 /* sub-points per block (1 = point kriging) */
  const double *sub_coord;  /* (ndim, nblock*nsub) or NULL when nsub == 1 */
  const double *sub_weight; /* nblock*nsub or NULL */
  const double *rangescale; /* nblock or NULL */
  const double *localnugget;/* nblock or NULL */
  int nstruct;
  const okr_struct *structs;
  int table_mode;           /* 1: lookup table (reference default), 0: analytic */
  double bound_lo, bound_hi;
  int neglect_error;
  int nthread;
  /* space-time kriging (kriging_st.F90): obs/block coordinates are 4 rows
   * (x, y, z, t) with ndim = 3; the search runs on (x, y, z, t * time_at) and
   * the covariance is the product-sum k_ps Cs Ct + Cs + Ct
   * (variogram_st.f90:222-282) with the spatial structures above and the
   * isotropic 1-D temporal structures below; with st_sum the sum-metric
   * Cs + Ct + sum_k sill_st[k] shape_k(sqrt(h_k^2 + f(dt)^2)) instead
   * (variogram_st.f90:178-194, 241-261), f(dt) = max(t_nugget + t_sill
   * (1 - corefunc(t_vtype, |dt| / t_at)), 0) and f(0) = 0 (:126-141) */
  int st;
  double time_at;
  double k_ps;
  int nstruct_t;
  const okr_struct *structs_t;
  int st_sum;
  int t_vtype;
  double t_nugget, t_sill, t_at;
  const double *sill_st;    /* nstruct joint sills */
} okr_problem;

typedef struct {
  double *estimate;         /* nblock */
  double *variance;         /* nblock */
  int mmax;                 /* row stride of inear */
  int *nnear;               /* nblock */
  int *inear;               /* nblock * mmax, 1-based ascending */
  double *weights;          /* nblock * mmax, may be NULL */
  int stats[5];             /* fail, chol_fact, chol_reuse, ssytrf_fact, ssytrf_reuse */
  char errmsg[512];
} okr_result;

/* full solve: returns 0 on success, 1 on error (message in res->errmsg) */
int okr_solve(const okr_problem *P, okr_result *R);

/* size helper: the neighbour-list stride the solve will use */
int okr_mmax(const okr_problem *P);

/* covariance of the model at lag vectors (n lags, (ndim, n) column-major) */
int okr_cov(const okr_problem *P, int n, const double *lags, double *out);

/* kdtree2 single query (exact emulation): k nearest in heap order, or sorted
 * ascending by distance when sort != 0.  idx are 1-based. */
int okr_knn(int ndim, int n, const double *coord, const double *q, int k, int sort,
            int *idx, double *dis);

/* kdtree2_n_nearest_maxidx over one tree of n points, nq queries: query j
 * takes its k[j] nearest among the points with 1-based index <= max_idx[j]
 * (idx/dis rows of stride kmax, sorted ascending by distance when sort) */
int okr_knn_maxidx(int ndim, int n, const double *coord, int nq, const double *q, const int *k, int sort,
                   const int *max_idx, int kmax, int *idx, double *dis);

/* the reference's symmetric-indefinite fallback (SSYTRF/SSYTRS with one
 * jittered retry) on one column-major m x m augmented system K of n data rows
 * (K is modified by the jitter); B holds nrhs right-hand sides (ld m) and
 * receives the solutions.  Returns INFO (0 = solved). */
int okr_bk_fallback(double *K, int n, int m, double *B, int nrhs);

#ifdef __cplusplus
}
#endif
#endif
