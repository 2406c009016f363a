/*
 * C ABI of libkrigekit_amd.so, the MI355X (gfx950) kriging backend.
 *
 * The krige_* entry points have the names, argument meaning, array layouts
 * and error behaviour of the reference engine's C ABI, so krigekit's ctypes
 * wrapper can bind them unchanged.  Each declaration cites the reference
 * interface it replaces (paths relative to the reference repository root).
 *
 * Conventions (src/libkriging/kriging_capi_common.F90, kriging_err.f90):
 *   - handles are 1-based int64 slots of a process-global registry;
 *   - every entry clears the error slot first and returns 0 on success,
 *     1 on error; krige_get_last_error copies the first error message;
 *   - coordinates are (ndim, n) column-major; results are block-first;
 *   - input arrays are copied; outputs go to caller-allocated buffers.
 * Entry points for features outside this backend (indicator kriging, SGSIM,
 * drift, gradient constraints, varying variograms, factor cache) exist and
 * return 1 with a message.
 */
#ifndef KRIGEKIT_AMD_H
#define KRIGEKIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/libkriging/kriging_capi.F90:43 */
int krige_create(int64_t *handle);
/* src/libkriging/kriging_capi.F90 (krige_ind_create): returns 1, not provided */
int krige_ind_create(int64_t *handle);
/* src/libkriging/kriging_capi.F90 (krige_ind_set_ncat): returns 1, not provided */
int krige_ind_set_ncat(int64_t handle, int ncat);
/* src/libkriging/kriging_capi.F90 (krige_destroy): zeroes *handle */
int krige_destroy(int64_t *handle);

/* src/libkriging/kriging_capi.F90:143-192 */
int krige_initialize(int64_t handle, int ndim, int nvar, int ndrift, int unbias, int nsim,
                     int anisotropic_search, int weight_correction, int use_old_weight,
                     int store_weight, int cross_validation, int write_mat, int neglect_error,
                     int varying_vgm, int std_ck, int verbose, int pf_cache,
                     const char *weight_file, const double bounds[2], int seed);

/* src/libkriging/kriging_capi.F90:197-224; coord (ndim, nobs) column-major */
int krige_set_obs(int64_t handle, int ivar, int nobs, int ndim, const double *coord,
                  const double *value, const double *variance, int nmax, double maxdist,
                  double sk_mean);
/* src/libkriging/kriging_capi_common.F90:178: external drift values at the
 * observations, drift (ndrift, nobs) column-major */
int krige_set_obs_drift(int64_t handle, int ivar, int ndrift, int nobs, const double *drift);
/* src/libkriging/kriging_capi_common.F90:136 */
int krige_update_obs_value(int64_t handle, int ivar, int nobs, const double *value);
/* src/libkriging/kriging_capi_common.F90:156: returns 1, not provided */
int krige_set_nscore(int64_t handle, int ivar, double zmin, double zmax, int ltail, int utail,
                     double ltpar, double utpar, int nwt, const double *wt);

/* src/libkriging/kriging_capi.F90 (krige_reset_vgm) */
int krige_reset_vgm(int64_t handle, int ivar, int jvar);
/* src/libkriging/kriging_capi.F90:245-270 */
int krige_set_vgm(int64_t handle, int ivar, int jvar, const char *vtype, double nugget,
                  double sill, double a_major, double a_minor1, double a_minor2, double azimuth,
                  double dip, double plunge, int is_product);
/* src/libkriging/kriging_capi.F90 (krige_set_vgm_block): returns 1, not provided */
int krige_set_vgm_block(int64_t handle, int ivar, int jvar, int ib, const char *vtype,
                        double nugget, double sill, double a_major, double a_minor1,
                        double a_minor2, double azimuth, double dip, double plunge,
                        int is_product);

/* src/libkriging/kriging_capi.F90:311-332 */
int krige_set_grid(int64_t handle, int ngrid, int ndim, const double *coord,
                   const double *rangescale, const double *localnugget);
/* src/libkriging/kriging_capi.F90:334-374 (block_type -4: 4^ndim Gauss-Legendre
 * sub-points; block_type > 0: user sub-points with equal counts per block) */
int krige_set_grid_block(int64_t handle, int block_type, int ngrid, int ndim, const double *coord,
                         int nblock, const int *nblockpnt, const double *pointweight,
                         const double *blocksize, const double *rangescale,
                         const double *localnugget);
/* src/libkriging/kriging_capi_common.F90:195 */
int krige_set_grid_cv(int64_t handle);
/* src/libkriging/kriging_capi.F90 (krige_set_grid_drift): drift values at
 * the targets, (ndrift, nblocks) column-major; ivar < 1 for every variable */
int krige_set_grid_drift(int64_t handle, int ivar, int ndrift, int nblocks, const double *drift);
/* src/libkriging/kriging_capi.F90:397-443: sequential Gaussian simulation
 * path (int, 1-based permutation of the blocks) and standard normal draws
 * (double, (nsim, nvar, nblocks) column-major); NULL: drawn by the library's
 * own generator from the seed (the reference's Fortran generator is not
 * reproduced).  Simulation is provided for nvar = 1, point targets. */
int krige_set_sim(int64_t handle, int nblocks, const void *randpath, int nsim, int nvar,
                  const void *sample);
/* src/libkriging/kriging_capi.F90:448-467 (builds the kd-tree) */
int krige_set_search(int64_t handle, int ivar, double anis1, double anis2, double azimuth,
                     double dip, double plunge, int sector_search);
/* src/libkriging/kriging_capi.F90:472-506, kriging_base.F90:2591-2660: increment
 * observations z(coord1) - z(coord2) of variable ivar (1-based), with their
 * values, variances (may be NULL: 0) and, when the model has external drift,
 * ndrift drift values per pair ((ngrad, ndrift) column-major).  ngrad <= 0
 * clears the variable's increments.  Up to 32 increments in all; every
 * increment joins every block's system, after the neighbour groups.  Not
 * combined with space-time, stored weights, weight correction, cross-
 * validation or varying variograms (the solve reports the error). */
int krige_set_grad(int64_t handle, int ivar, int ngrad, int ndim, const double *coord1,
                   const double *coord2, const double *grad_val, const double *variance,
                   int ndrift, const double *drift_ext);

/* src/libkriging/kriging_capi_common.F90:212 (validation only) */
int krige_prepare(int64_t handle);
/* src/libkriging/kriging_capi_common.F90:229-239: the hot path, run on the GPU;
 * nthread and ncache are accepted and ignored */
int krige_solve(int64_t handle, int nthread, int ncache);

/* src/libkriging/kriging_capi_common.F90:246-392 */
int krige_get_nblocks(int64_t handle, int *n);
int krige_get_nsim(int64_t handle, int *n);
int krige_get_block_coord(int64_t handle, int ndim, int nblocks, double *out);
int krige_get_estimate(int64_t handle, int nsim, int nblocks, double *out);
int krige_get_estimate_all(int64_t handle, int nblocks, int nvar, int nsim, double *out);
int krige_get_variance(int64_t handle, int nblocks, double *out);
int krige_get_variance_all(int64_t handle, int nblocks, int nvar, double *out);

/* src/libkriging/kriging_capi_common.F90:417-484: return 1, not provided */
int krige_get_factor_info(int64_t handle, int *valid, int *npp, int *p);
int krige_get_factor_matrices(int64_t handle, int npp, int p, double *L, double *kinv_drift,
                              double *schur);
int krige_get_factor_system(int64_t handle, int matsize, int nvar, int which, double *matA,
                            double *rhsB);

/* src/libkriging/kriging_capi_common.F90:90-96 */
int krige_get_last_error(char *buffer, int nbuf);
/* src/libkriging/kriging_capi_common.F90:113-128: fail, chol_fact, chol_reuse,
 * ssytrf_fact, ssytrf_reuse */
void krige_solver_stats(int64_t handle, int out[5]);
/* src/libkriging/kriging_capi_common.F90:397: library-owned string */
intptr_t krige_to_str(int64_t handle);
/* src/krigekit/kriging.py:282-283 (OpenMP probes; this backend reports 1) */
void krige_get_max_threads(int *n);
void krige_get_num_threads(int *n);

/* The weight store, src/libkriging/kriging_capi_common.F90:491-619 (store
 * layouts kriging_base.F90:284-295): filled by a solve with store_weight
 * (and written to weight_file when one is named); krige_set_weights or a
 * weight_file with use_old_weight make the next solve replay the stored
 * weights against the current observation values, without search or solve.
 * Layouts (column-major): nnear(ngroups, nblock), inear(nmax, ngroups,
 * nblock) 1-based, weight(nmax, ngroups, nvar, nblock), var(nvar, nvar,
 * nblock), ngroups = nvar; entries past nnear are zero.  The getters copy
 * the leading part of the store of the requested dimensions. */
int krige_free_weight_store(int64_t handle);                                      /* :491-500 */
int krige_get_weight_dims(int64_t handle, int *nmax, int *ngroups, int *nblock);  /* :503-519 */
int krige_get_weight_nnear(int64_t handle, int ngroups, int nblock, int *out);    /* :522-537 */
int krige_get_weight_inear(int64_t handle, int nmax, int ngroups, int nblock, int *out); /* :540-555 */
int krige_get_weight_data(int64_t handle, int nmax, int ngroups, int nvar, int nblock,
                          double *out);                                           /* :558-573 */
int krige_get_weight_var(int64_t handle, int nvar, int nblock, double *out);      /* :576-596 */
int krige_set_weights(int64_t handle, int nmax, int ngroups, int nvar, int nblock, const int *nnear,
                      const int *inear, const double *weight, const int *order,
                      const double *var);                                         /* :600-619 */

/* ---- space-time kriging, src/libkriging/kriging_st_capi.f90 -----------------
 * Shares the handle registry and the solve/getter entry points above
 * (src/krigekit/kriging_st.py:105-212).  Coordinates are (x, y, z, t): the
 * spatial part has ndim <= 3 rows (zero-padded to 3), time is the last row.
 * Product-sum model only; sum-metric, gradients, blocks, drift and
 * simulation return 1 with a message. */
/* kriging_st_capi.f90:43 */
int krige_st_create(int64_t *handle);
/* kriging_st_capi.f90:61 */
int krige_st_destroy(int64_t *handle);
/* kriging_st_capi.f90:80-122 (ndim fixed at 3) */
int krige_st_initialize(int64_t handle, int nvar, int ndrift, int unbias, int nsim, int anisotropic_search,
                        int weight_correction, int use_old_weight, int store_weight, int cross_validation,
                        int write_mat, int neglect_error, int verbose, const char *weight_file,
                        const double bounds[2], int seed);
/* kriging_st_capi.f90:128-163: model "product_sum" (k_ps) or "sum_metric" */
int krige_st_set_st_model(int64_t handle, const char *model, const char *transform, double at,
                          double time_nugget, double time_sill, double k_ps);
/* kriging_st_capi.f90:165-192: coord (ndim+1, nobs), time in the last row; ndim must be 3 (the
   space-time object has 3 spatial rows + time, kriging_st.F90:121-123) */
int krige_st_set_obs(int64_t handle, int ivar, int nobs, int ndim, const double *coord, const double *value,
                     const double *variance, int nmax, double maxdist, double sk_mean);
/* kriging_st_capi.f90:201-239: returns 1, not provided */
int krige_st_set_grad(int64_t handle, int ivar, int ngrad, int ndim, const double *coord1,
                      const double *coord2, const double *grad_val, const double *variance, int ndrift_c,
                      const double *drift_ext);
/* kriging_st_capi.f90:244-258 */
int krige_st_reset_vgm(int64_t handle, int ivar, int jvar);
/* kriging_st_capi.f90:260-286: spatial structure */
int krige_st_set_vgm(int64_t handle, int ivar, int jvar, const char *vtype, double nugget, double sill,
                     double a_major, double a_minor1, double a_minor2, double azimuth, double dip, double plunge);
/* kriging_st_capi.f90:288-307: temporal structure, isotropic range at_k */
int krige_st_set_vgm_temporal(int64_t handle, int ivar, int jvar, const char *vtype, double nugget,
                              double sill, double at_k);
/* kriging_st_capi.f90:309-327: sum-metric joint sills, one per spatial
 * structure of the pair added so far (nvar = 1) */
int krige_st_set_vgm_joint_sills(int64_t handle, int ivar, int jvar, int n, const double *sills);
/* kriging_st_capi.f90:330-353: coord (3, ngrid), time (ngrid) */
int krige_st_set_grid(int64_t handle, int ngrid, const double *coord, const double *time,
                      const double *rangescale, const double *localnugget);
/* kriging_st_capi.f90:355-388: returns 1, not provided */
int krige_st_set_grid_block(int64_t handle, int nblocks, const double *coord, const double *time,
                            const int *nblockpnt, int npnts_total, const double *blockcoord,
                            const double *blocktime, const double *pointweight, const double *rangescale,
                            const double *localnugget);
/* kriging_st_capi.f90:390-406: returns 1, not provided */
int krige_st_set_grid_drift(int64_t handle, int ndrift_c, int nblocks, const double *drift);
/* kriging_st_capi.f90:414-453: returns 1, not provided */
int krige_st_set_sim(int64_t handle, int nblocks, const void *randpath, int nsim_c, const void *sample);
/* kriging_st_capi.f90:455-490: time_at scales time in the search metric */
int krige_st_set_search(int64_t handle, int ivar, double time_at, double anis1, double anis2, double azimuth,
                        double dip, double plunge, int sector_search);
/* kriging_st_capi.f90:496-520 (this backend reports 1) */
void krige_st_get_max_threads(int *n);
void krige_st_get_num_threads(int *n);

/* ---- extensions: no counterpart in the reference -------------------------- */
/* precision 0 = FP64 solve, 1 = FP32 factorisation + FP64 refinement;
 * pivot_tol: FP32 pivot ratio below which a node is re-solved in FP64;
 * table_mode: 1 = lookup-table covariance (the reference default), 0 = analytic */
int kk_set_options(int64_t handle, int precision, double pivot_tol, int table_mode);
/* factor reuse over identical neighbour sets (the reference's factor cache,
 * src/libkriging/kriging_base.F90:1746-1917,2445-2485, done as a grouping
 * pass): blocks with the same neighbour list, range scale and local nugget
 * share one factorisation, item_max (1..4, default 4) member blocks per
 * factorisation.  mode -1
 * (default): on unless cross_validation (whose sets are all distinct); 0 off;
 * 1 on.  Results are identical either way; krige_solver_stats counts the
 * factorisations and the reuses. */
int kk_set_factor_reuse(int64_t handle, int mode, int item_max);
/* mode 2 (default): wave-cooperative two-pass LDS search, mode 1: per-lane
 * two-pass LDS search (both run the exact kernel only for the queries they
 * flag), mode 0: exact-emulation search kernel for every query.
 * reorder: 0 queries in the given order; 1 (default) Morton order, searched
 * in two chunks, the first chunk's blocks solved while the second chunk is
 * searched; 2 Morton order in one search launch, so that search and solve
 * are timed apart. */
int kk_set_search_mode(int64_t handle, int mode, int reorder);
/* solve with inputs and results left resident in HBM (no result copy-back) */
int kk_solve_resident(int64_t handle);
/* copy the results of the last kk_solve_resident to the host */
int kk_fetch(int64_t handle);
/* capture the device pass of a solved single-variable handle into a HIP
 * graph: search, set grouping or bucketing, one solve per bucket, exact
 * search of flagged queries (second stream), FP64 re-solve and fallback, with
 * every count consumed on the device (no host synchronisation inside).
 * Needs the wave-cooperative search, at most 64 rows per system and an
 * exact-search heap that fits LDS; any setter invalidates the graph. */
int kk_graph_capture(int64_t handle);
/* replay the captured pass over the resident inputs; results stay in HBM
 * (kk_fetch copies them back), timing total = the graph's device time */
int kk_graph_launch(int64_t handle);
/* per-bucket solve schedules (k_solve instance width: lanes and rows per system),
 * timed on a solved single-variable handle's own work lists (reps launches
 * per candidate); the fastest go into the library's table, used by later
 * passes and captures, and into the text file at path when path is not
 * NULL or empty.  The library reads such a file when it is loaded
 * (KRIGEKIT_AMD_TUNING, else tuning.txt beside the library). */
int kk_autotune(int64_t handle, const char *path, int reps);
/* replace the library's schedule table by the file at path (NULL or empty:
 * every bucket on its own width with a full grid); drops captured graphs */
int kk_set_tuning(const char *path);
/* out[15]: search, bucket, solve, re-solve/fallback, total (ms), FP64 re-solved
 * nodes, Bunch-Kaufman fallback nodes, queries decided by the exact search kernel, fast
 * search launches (0: none), fast search kernel ms (all launches), exact
 * search kernel ms, LOO-gradient kernel ms, weight-store pass (store_weight)
 * or stored-weight replay (use_old_weight) ms, sequential-simulation
 * recurrence ms and its dependency levels (nsim > 0; the other entries then
 * describe the pass that solved every block's system).
 * Times come from HIP events on the stream the kernels run on.  When the fast
 * search leaves queries to the exact kernel, that kernel runs on a second
 * stream while the other blocks are solved: it is outside "search" and its
 * wait, if any, is inside "solve".  With more than one fast search launch
 * (Morton chunks), the blocks of the first chunk are solved on a third stream
 * while the second is searched, so part of the solve lies inside "search". */
int kk_get_timing(int64_t handle, double *out);
/* Leave-one-out cross-validation MSE and its gradient with respect to the
 * variogram parameters (nugget, sill, a_major, a_minor1, a_minor2, azimuth,
 * dip, plunge) of a single-structure ordinary-kriging model, by an adjoint
 * kernel over the per-block systems (no counterpart in the reference; the
 * target of its LOO workflow, tests/test_cross_validation.py).  Needs
 * cross_validation=1 and set_grid_cv.  grad[8]; zhat[nblock] and
 * dz[nblock*8] (block-major) may be NULL. */
int kk_loo_grad(int64_t handle, double *mse, double *grad, double *zhat, double *dz);
/* No counterpart in the reference: kk_solve_resident for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree: the
 * pass's neighbour lists, flags, bucket lists, search heaps and fallback
 * matrices live in a fixed 64 MiB device array (one per GPU, held until the
 * process ends); a problem whose pass needs more (about 4*mmax + 90 bytes per
 * block plus the search-heap and fallback areas) is rejected with an error.
 * est and var are caller-owned device buffers of nblock doubles.  The handle
 * must hold a finished single-variable solve of its current inputs (its
 * inputs, streams and query order stay resident); its own results are left
 * unchanged.  All calls on one GPU share the array: calls that overlap on one
 * GPU race. */
int kk_solve_resident_noalloc(int64_t handle, double *est, double *var);
/* hipDeviceSynchronize */
int kk_device_synchronize(void);
/* plain device copy of nbytes, reps times; *ms = mean kernel time.  A kernel
 * of known size for calibrating HBM traffic counters. */
int kk_copy_probe(double nbytes, int reps, double *ms);
/* neighbour lists of the last solve (0-based ids ascending, row stride mmax) */
int kk_get_mmax(int64_t handle, int *mmax);
int kk_get_neighbours(int64_t handle, int nblock, int mmax, int *nnear, int *inear);

#ifdef __cplusplus
}
#endif
#endif /* KRIGEKIT_AMD_H */
