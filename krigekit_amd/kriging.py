"""Python front end of the MI355X kriging backend.

Mirrors the public surface of krigekit's ``Kriging`` class
(``src/krigekit/kriging.py:403-1985``): same constructor flags and defaults
(``neglect_error=True``, ``std_ck=False``, ``pf_cache=False``,
``anisotropic_search=False``), same setters, the same (n, ndim) ->
(ndim, n) column-major transposes, the same ``RuntimeError(last_error)``
convention and the same result shapes and DataFrame columns.  ``solve()``
runs the hot path (neighbour search, assembly, factorisation, estimate and
variance) on the GPU through the C ABI of ``libkrigekit_amd.so``.

Additional keyword arguments (no counterpart in the reference):
``precision`` ("fp64" or "fp32r": FP32 factorisation with one FP64
refinement step), ``table_mode`` (True: the reference's default lookup-table
covariance; False: analytic covariance) and ``pivot_tol``.
"""
import ctypes
import os
import random
import sys
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkrigekit_amd.so")

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_i64 = ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

_lib = None


def _load_lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"HIP kriging library not found at {LIB_PATH!r}; build it with "
            "`python -m krigekit_amd.build`.")
    lib = ctypes.CDLL(LIB_PATH)

    def sig(name, args, res=_c_int):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res

    sig("krige_create", [ctypes.POINTER(_i64)])
    sig("krige_destroy", [ctypes.POINTER(_i64)])
    sig("krige_initialize", [_i64] + [_c_int] * 16 + [ctypes.c_char_p, _pd, _c_int])
    sig("krige_set_obs", [_i64, _c_int, _c_int, _c_int, _pd, _pd, _pd, _c_int, _c_double, _c_double])
    sig("krige_update_obs_value", [_i64, _c_int, _c_int, _pd])
    sig("krige_set_obs_drift", [_i64, _c_int, _c_int, _c_int, _pd])
    sig("krige_set_grid_drift", [_i64, _c_int, _c_int, _c_int, _pd])
    sig("krige_reset_vgm", [_i64, _c_int, _c_int])
    sig("krige_set_vgm", [_i64, _c_int, _c_int, ctypes.c_char_p] + [_c_double] * 8 + [_c_int])
    sig("krige_set_grid", [_i64, _c_int, _c_int, _pd, _pd, _pd])
    sig("krige_set_grid_block", [_i64, _c_int, _c_int, _c_int, _pd, _c_int, _pi, _pd, _pd, _pd, _pd])
    sig("krige_set_grid_cv", [_i64])
    sig("krige_set_sim", [_i64, _c_int, ctypes.c_void_p, _c_int, _c_int, ctypes.c_void_p])
    sig("krige_set_grad", [_i64, _c_int, _c_int, _c_int, _pd, _pd, _pd, _pd, _c_int, _pd])
    sig("krige_set_search", [_i64, _c_int] + [_c_double] * 5 + [_c_int])
    sig("krige_solve", [_i64, _c_int, _c_int])
    sig("krige_get_nblocks", [_i64, _pi])
    sig("krige_get_nsim", [_i64, _pi])
    sig("krige_get_block_coord", [_i64, _c_int, _c_int, _pd])
    sig("krige_get_estimate_all", [_i64, _c_int, _c_int, _c_int, _pd])
    sig("krige_get_variance_all", [_i64, _c_int, _c_int, _pd])
    sig("krige_get_last_error", [ctypes.POINTER(ctypes.c_char), _c_int])
    sig("krige_solver_stats", [_i64, _pi], None)
    sig("krige_to_str", [_i64], ctypes.c_void_p)
    sig("krige_get_weight_dims", [_i64, _pi, _pi, _pi])
    sig("krige_get_weight_nnear", [_i64, _c_int, _c_int, _pi])
    sig("krige_get_weight_inear", [_i64, _c_int, _c_int, _c_int, _pi])
    sig("krige_get_weight_data", [_i64, _c_int, _c_int, _c_int, _c_int, _pd])
    sig("krige_get_weight_var", [_i64, _c_int, _c_int, _pd])
    sig("krige_set_weights", [_i64, _c_int, _c_int, _c_int, _c_int, _pi, _pi, _pd, _pi, _pd])
    sig("krige_free_weight_store", [_i64])
    sig("kk_set_options", [_i64, _c_int, _c_double, _c_int])
    sig("kk_set_search_mode", [_i64, _c_int, _c_int])
    sig("kk_set_factor_reuse", [_i64, _c_int, _c_int])
    sig("kk_graph_capture", [_i64])
    sig("kk_graph_launch", [_i64])
    sig("kk_autotune", [_i64, ctypes.c_char_p, _c_int])
    sig("kk_set_tuning", [ctypes.c_char_p])
    sig("kk_solve_resident_noalloc", [_i64, ctypes.c_void_p, ctypes.c_void_p])
    sig("kk_solve_resident", [_i64])
    sig("kk_fetch", [_i64])
    sig("kk_get_timing", [_i64, _pd])
    sig("kk_get_neighbours", [_i64, _c_int, _c_int, _pi, _pi])
    sig("kk_get_mmax", [_i64, _pi])
    sig("kk_device_synchronize", [])
    sig("kk_copy_probe", [_c_double, _c_int, _pd])
    sig("kk_loo_grad", [_i64, _pd, _pd, _pd, _pd])
    # space-time (krige_st_*)
    sig("krige_st_create", [ctypes.POINTER(_i64)])
    sig("krige_st_destroy", [ctypes.POINTER(_i64)])
    sig("krige_st_initialize", [_i64] + [_c_int] * 12 + [ctypes.c_char_p, _pd, _c_int])
    sig("krige_st_set_st_model", [_i64, ctypes.c_char_p, ctypes.c_char_p] + [_c_double] * 4)
    sig("krige_st_set_obs", [_i64, _c_int, _c_int, _c_int, _pd, _pd, _pd, _c_int, _c_double, _c_double])
    sig("krige_st_reset_vgm", [_i64, _c_int, _c_int])
    sig("krige_st_set_vgm", [_i64, _c_int, _c_int, ctypes.c_char_p] + [_c_double] * 8)
    sig("krige_st_set_vgm_temporal", [_i64, _c_int, _c_int, ctypes.c_char_p] + [_c_double] * 3)
    sig("krige_st_set_vgm_joint_sills", [_i64, _c_int, _c_int, _c_int, _pd])
    sig("krige_st_set_grid", [_i64, _c_int, _pd, _pd, _pd, _pd])
    sig("krige_st_set_search", [_i64, _c_int] + [_c_double] * 6 + [_c_int])
    _lib = lib
    return lib


def _last_error() -> str:
    buf = ctypes.create_string_buffer(4096)
    _load_lib().krige_get_last_error(buf, 4096)
    return buf.value.decode("utf-8", errors="replace").strip()


def _check(ierr, name):
    if int(ierr) != 0:
        raise RuntimeError(_last_error() or f"{name} failed with ierr={int(ierr)}")


def _fcoord(coord):
    """(n, ndim) -> (ndim, n) column-major, as kriging.py:342-360."""
    a = np.asarray(coord, dtype=np.float64)
    if a.ndim == 1:
        return np.asfortranarray(a.reshape(-1, 1))
    return np.asfortranarray(a.T)


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


_PRECISIONS = {"fp64": 0, "fp32r": 1}


class Kriging:
    """GPU drop-in for krigekit's ``Kriging`` (same methods and argument meaning)."""

    def __init__(self, ndim: int = 2, nvar: int = 1, ndrift: int = 0, unbias: int = 1, nsim: int = 0,
                 anisotropic_search: bool = False, weight_correction: bool = False,
                 use_old_weight: bool = False, store_weight: bool = False,
                 cross_validation: bool = False, write_mat: bool = False, neglect_error: bool = True,
                 varying_vgm: bool = False, std_ck: bool = False, verbose: bool = False,
                 pf_cache: bool = False, weight_file: str = "", bounds: Optional[tuple] = None,
                 seed: Optional[int] = None, precision: str = "fp64", table_mode: bool = True,
                 pivot_tol: float = 1.0e-3):
        lib = _load_lib()
        h = _i64(0)
        _check(lib.krige_create(ctypes.byref(h)), "krige_create")
        self._handle = h.value
        huge = sys.float_info.max * 1e3
        self.bounds = np.asfortranarray(bounds if bounds is not None else [-huge, huge], dtype=np.float64)
        seed = seed or random.randint(0, 2 ** 31 - 1)
        self.store_weight = store_weight
        self.use_old_weight = use_old_weight
        self.weight_file = weight_file
        # store_weight without a weight_file keeps the weights in memory only
        # (the reference's own check at kriging.py:572-573 compares the str
        # with b"" and so never fires; its tests rely on that)
        if store_weight and use_old_weight:
            raise ValueError("store_weight and use_old_weight are mutually exclusive")
        if cross_validation and nsim > 0:
            raise ValueError("nsim>0 and cross_validation are mutually exclusive")
        _check(lib.krige_initialize(
            self._handle, ndim, nvar, ndrift, unbias, nsim, int(anisotropic_search),
            int(weight_correction), int(use_old_weight), int(store_weight), int(cross_validation),
            int(write_mat), int(neglect_error), int(varying_vgm), int(std_ck), int(verbose),
            int(pf_cache), weight_file.encode("utf-8") if weight_file else b"", _dp(self.bounds),
            int(seed) & 0x7FFFFFFF), "krige_initialize")
        self.ndim, self.nvar, self.ndrift, self.nsim = ndim, nvar, ndrift, nsim
        self.unbias = unbias
        self.cross_validation = cross_validation
        self.verbose = verbose
        self.seed = seed
        self._nblock = 0
        self._nobs = np.zeros(nvar, dtype=np.int64)
        self.set_options(precision=precision, table_mode=table_mode, pivot_tol=pivot_tol)

    # ------------------------------------------------------------------
    def set_options(self, precision: str = "fp64", table_mode: bool = True, pivot_tol: float = 1.0e-3):
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        _check(_load_lib().kk_set_options(self._handle, _PRECISIONS[precision], pivot_tol,
                                          int(table_mode)), "kk_set_options")
        self.precision, self.table_mode, self.pivot_tol = precision, table_mode, pivot_tol

    def set_obs(self, ivar: int, coord, value, variance=None, nmax: Optional[int] = None,
                maxdist: Optional[float] = None, sk_mean: float = 0.0):
        cf = _fcoord(coord)
        v = np.asfortranarray(np.asarray(value, dtype=np.float64).ravel())
        nobs = cf.shape[1]
        if v.size != nobs:
            raise ValueError(f"value length ({v.size}) must match nobs ({nobs})")
        if variance is not None:
            var = np.asfortranarray(np.asarray(variance, dtype=np.float64).ravel())
            if var.size != nobs:
                raise ValueError(f"variance length ({var.size}) must match nobs ({nobs})")
        else:
            var = np.zeros(nobs)
        nm = nmax if nmax is not None else np.iinfo(np.int32).max
        md = maxdist if maxdist is not None else sys.float_info.max
        _check(_load_lib().krige_set_obs(self._handle, ivar, nobs, cf.shape[0], _dp(cf), _dp(v),
                                         _dp(var), nm, md, sk_mean), "krige_set_obs")
        self._nobs[ivar - 1] = nobs

    def set_obs_drift(self, ivar: int, drift):
        """external drift values at variable ``ivar``'s observations, shape
        (nobs, ndrift), passed to the engine as (ndrift, nobs)
        (src/krigekit/kriging.py:665-687)"""
        d = np.asfortranarray(np.asarray(drift, dtype=np.float64).T)
        if d.ndim == 1:
            d = d.reshape(1, -1)
        _check(_load_lib().krige_set_obs_drift(self._handle, ivar, d.shape[0], d.shape[1], _dp(d)),
               "krige_set_obs_drift")

    def set_grid_drift(self, drift, ivar: Optional[int] = None):
        """external drift values at the blocks, shape (nblocks, ndrift), for
        target variable ``ivar`` (None: every target)
        (src/krigekit/kriging.py:1096-1132)"""
        d = np.asfortranarray(np.asarray(drift, dtype=np.float64).T)
        if d.ndim == 1:
            d = d.reshape(1, -1)
        _check(_load_lib().krige_set_grid_drift(self._handle, ivar if ivar is not None else -1, d.shape[0],
                                                d.shape[1], _dp(d)), "krige_set_grid_drift")

    def update_obs_value(self, ivar: int, value):
        v = np.asfortranarray(np.asarray(value, dtype=np.float64).ravel())
        if v.size != self._nobs[ivar - 1]:
            raise ValueError(f"value length ({v.size}) must match nobs ({self._nobs[ivar - 1]}) "
                             f"set by the previous set_obs call for ivar={ivar}")
        _check(_load_lib().krige_update_obs_value(self._handle, ivar, v.size, _dp(v)),
               "krige_update_obs_value")

    def set_vgm(self, ivar: int, jvar: int, vtype: str, nugget: float = 0.0, sill: float = 1.0,
                a_major: float = 1.0, a_minor1: Optional[float] = None,
                a_minor2: Optional[float] = None, azimuth: float = 0.0, dip: float = 0.0,
                plunge: float = 0.0, append: bool = True, product: bool = False):
        if a_minor1 is None:
            a_minor1 = a_major
        if a_minor2 is None:
            a_minor2 = a_minor1
        lib = _load_lib()
        if not append:
            _check(lib.krige_reset_vgm(self._handle, ivar, jvar), "krige_reset_vgm")
        if jvar < ivar:  # same pair collapse as kriging.py:878-881
            jvar = ivar
        _check(lib.krige_set_vgm(self._handle, ivar, jvar, vtype.lower()[:3].encode("utf-8"),
                                 nugget, sill, a_major, a_minor1, a_minor2, azimuth, dip, plunge,
                                 int(product)), "krige_set_vgm")

    def set_grid(self, coord=None, rangescale=None, localnugget=None):
        if coord is None:
            self.set_grid_cv()
            return
        cf = _fcoord(coord)
        n = cf.shape[1]
        rs = np.asfortranarray(rangescale if rangescale is not None else np.ones(n), dtype=np.float64)
        ln = np.asfortranarray(localnugget if localnugget is not None else np.zeros(n), dtype=np.float64)
        _check(_load_lib().krige_set_grid(self._handle, n, cf.shape[0], _dp(cf), _dp(rs), _dp(ln)),
               "krige_set_grid")
        self._nblock = n

    def set_grid_block(self, coord, block_type: int, nblockpnt=None, pointweight=None,
                       blocksize=None, rangescale=None, localnugget=None):
        cf = _fcoord(coord)
        ngrid = cf.shape[1]
        if block_type == -4:
            nblock = ngrid
            if blocksize is None:
                raise ValueError("blocksize must be specified for Gaussian quadrature blocks.")
            bs = np.asarray(blocksize, dtype=np.float64)
            if bs.ndim == 1:
                bs = np.tile(bs, (nblock, 1))
            bsf = _fcoord(bs)
            nbp = np.ones(nblock, dtype=np.int32)
            pw = np.ones(nblock)
        else:
            nbp = np.ascontiguousarray(nblockpnt, dtype=np.int32)
            nblock = len(nbp)
            npoint = int(nbp.sum())
            if np.any(nbp <= 0):
                raise ValueError("nblockpnt must contain positive counts")
            if npoint != ngrid:
                raise ValueError(f"sum(nblockpnt) ({npoint}) must match coord rows ({ngrid})")
            bsf = _fcoord(np.zeros((nblock, self.ndim)))
            pw = (np.asfortranarray(pointweight, dtype=np.float64) if pointweight is not None
                  else np.repeat(1.0 / nbp, nbp))
        rs = np.asfortranarray(rangescale if rangescale is not None else np.ones(nblock), dtype=np.float64)
        ln = np.asfortranarray(localnugget if localnugget is not None else np.zeros(nblock),
                               dtype=np.float64)
        _check(_load_lib().krige_set_grid_block(self._handle, block_type, ngrid, cf.shape[0], _dp(cf),
                                                nblock, _ip(nbp), _dp(pw), _dp(bsf), _dp(rs), _dp(ln)),
               "krige_set_grid_block")
        self._nblock = nblock

    def set_grid_cv(self):
        _check(_load_lib().krige_set_grid_cv(self._handle), "krige_set_grid_cv")
        self._nblock = int(self._nobs[0])

    def set_grad(self, coord1, coord2, grad_value, ivar: int = 1, variance=None, drift_ext=None):
        """Increment (gradient-pair) observations Z(coord1[i]) - Z(coord2[i]) =
        grad_value[i] of variable ivar (reference kriging.py set_grad): every
        pair joins every block's system; variance relaxes the constraint;
        drift_ext (ngrad, ndrift) holds the drift differences when ndrift > 0.
        An empty set clears the variable's pairs."""
        c1, c2 = _fcoord(coord1), _fcoord(coord2)
        ng = c1.shape[1] if np.asarray(coord1).size else 0
        gv = np.asfortranarray(np.asarray(grad_value, dtype=np.float64).ravel())
        if gv.size != ng or c2.shape != c1.shape:
            raise ValueError("coord1, coord2 and grad_value must describe the same pairs")
        var = np.zeros(max(ng, 1)) if variance is None else np.asfortranarray(variance, dtype=np.float64).ravel()
        if variance is not None and var.size != ng:
            raise ValueError(f"variance length ({var.size}) must match ngrad ({ng})")
        if drift_ext is not None:
            de = np.asarray(drift_ext, dtype=np.float64)
            de = de.reshape(ng, -1) if de.ndim < 2 else de
            df, nd = np.asfortranarray(de.T), de.shape[1]
        else:
            df, nd = np.zeros((1, max(ng, 1)), order="F"), 0
        if ng == 0:
            c1 = c2 = np.zeros((self.ndim, 1), order="F")
            gv = np.zeros(1)
        _check(_load_lib().krige_set_grad(self._handle, ivar, ng, c1.shape[0], _dp(c1), _dp(c2), _dp(gv), _dp(var),
                                          nd, _dp(df)), "krige_set_grad")

    def set_sim(self, randpath: Optional[np.ndarray] = None, sample: Optional[np.ndarray] = None):
        """Sequential Gaussian simulation path and draws (reference
        kriging.py:1134-1205): randpath a 1-based permutation of the blocks,
        sample standard normal draws (nblocks,), (nblocks, nsim) or (nblocks,
        nvar, nsim), the one of path position p at row p.  Either left out is
        drawn by the library from the seed."""
        if self.nsim <= 0:
            raise ValueError("nsim must be > 0 when setting SGSIM parameters.")
        nb = self._nblock
        rp_ptr = s_ptr = None
        keep = []
        if randpath is not None:
            rp = np.ascontiguousarray(np.asarray(randpath, dtype=np.int32).ravel())
            if rp.size != nb:
                raise ValueError(f"randpath length ({rp.size}) must match nblocks ({nb})")
            if not np.array_equal(np.sort(rp), np.arange(1, nb + 1, dtype=np.int32)):
                raise ValueError("randpath must be a 1-based permutation of 1..nblocks")
            keep.append(rp)
            rp_ptr = rp.ctypes.data
        nsim_c, nvar_c = 0, 0
        if sample is not None:
            sa = np.asarray(sample, dtype=np.float64)
            if sa.ndim == 1:
                sa = sa[:, None, None]
            elif sa.ndim == 2:
                sa = sa[:, None, :]
            elif sa.ndim != 3:
                raise ValueError(f"sample must be 2-D (nblocks, nsim) or 3-D (nblocks, nvar, nsim), got {sa.shape}")
            sf = np.asfortranarray(sa.transpose(2, 1, 0))  # (nsim, nvar, nblocks)
            nsim_c, nvar_c = sf.shape[0], sf.shape[1]
            if nsim_c != self.nsim or nvar_c != self.nvar or sf.shape[2] != nb:
                raise ValueError(f"sample shape {sf.shape} must be ({self.nsim}, {self.nvar}, {nb})")
            keep.append(sf)
            s_ptr = sf.ctypes.data
        _check(_load_lib().krige_set_sim(self._handle, nb, rp_ptr, nsim_c, nvar_c, s_ptr), "krige_set_sim")

    def set_search(self, ivar: int = 1, anis1: float = 1.0, anis2: float = 1.0, azimuth: float = 0.0,
                   dip: float = 0.0, plunge: float = 0.0, sector_search: bool = False):
        _check(_load_lib().krige_set_search(self._handle, ivar, anis1, anis2, azimuth, dip, plunge,
                                            int(sector_search)), "krige_set_search")

    def solve(self, nthread: int = 0, ncache: Optional[int] = None):
        _check(_load_lib().krige_solve(self._handle, nthread, -1 if ncache is None else int(ncache)),
               "krige_solve")

    # ---- resident-data entry points (inputs and results stay in HBM) ----
    def solve_resident(self):
        _check(_load_lib().kk_solve_resident(self._handle), "kk_solve_resident")

    def fetch(self):
        _check(_load_lib().kk_fetch(self._handle), "kk_fetch")

    @property
    def timing(self) -> dict:
        out = np.zeros(15)
        _check(_load_lib().kk_get_timing(self._handle, _dp(out)), "kk_get_timing")
        return dict(search_ms=out[0], bucket_ms=out[1], solve_ms=out[2], fallback_ms=out[3],
                    total_ms=out[4], fp64_resolved=int(out[5]), bk_fallback=int(out[6]),
                    exact_search_queries=int(out[7]), fast_search_launches=int(out[8]),
                    fast_search_kernel_ms=out[9],
                    exact_search_kernel_ms=out[10], loo_grad_kernel_ms=out[11],
                    weights_ms=out[12], sim_ms=out[13], sim_levels=int(out[14]))

    GRAD_PARAMS = ("nugget", "sill", "a_major", "a_minor1", "a_minor2", "azimuth", "dip", "plunge")

    def loo_grad(self, per_block: bool = False):
        """Leave-one-out MSE and its gradient with respect to the parameters of
        the single variogram structure, from the adjoint kernel (FP64).

        Needs ``cross_validation=True`` and ``set_grid_cv()``.  Returns
        ``(mse, grad)`` with ``grad`` a dict over ``GRAD_PARAMS`` (angles in
        degrees, as ``set_vgm`` takes them); with ``per_block`` also the
        leave-one-out estimates and their (nblock, 8) derivatives."""
        nbc = ctypes.c_int(0)
        _check(_load_lib().krige_get_nblocks(self._handle, ctypes.byref(nbc)), "krige_get_nblocks")
        nb = nbc.value
        mse = ctypes.c_double(0.0)
        grad = np.zeros(8)
        zhat = np.zeros(nb) if per_block else None
        dz = np.zeros((nb, 8)) if per_block else None
        _check(_load_lib().kk_loo_grad(self._handle, ctypes.byref(mse), _dp(grad),
                                        _dp(zhat) if per_block else None,
                                        _dp(dz) if per_block else None), "kk_loo_grad")
        g = dict(zip(self.GRAD_PARAMS, grad.tolist()))
        if per_block:
            return mse.value, g, zhat, dz
        return mse.value, g

    def graph_capture(self):
        """capture the solved handle's device pass into a HIP graph"""
        _check(_load_lib().kk_graph_capture(self._handle), "kk_graph_capture")

    def graph_launch(self):
        """replay the captured pass; results stay on the device (``fetch``)"""
        _check(_load_lib().kk_graph_launch(self._handle), "kk_graph_launch")

    def autotune(self, path: Optional[str] = None, reps: int = 5):
        """time candidate per-bucket solve schedules on this solved handle's
        work lists, keep the fastest (and write them to ``path``)"""
        _check(_load_lib().kk_autotune(self._handle, (path or "").encode("utf-8"), int(reps)), "kk_autotune")

    def set_factor_reuse(self, mode: int = -1, item_max: int = 4):
        """Factor reuse over identical neighbour sets: -1 auto (on unless
        cross-validation), 0 off, 1 on; item_max (1..4) member blocks per
        factorisation.  Results are identical either way."""
        _check(_load_lib().kk_set_factor_reuse(self._handle, int(mode), int(item_max)), "kk_set_factor_reuse")

    def set_search_mode(self, fast=True, reorder=True):
        """fast: True/2 wave-cooperative two-pass search (default), 1 per-lane
        two-pass search, False/0 exact-emulation kernel for every query.
        reorder: True/1 Morton order in two overlapped chunks (default), 2
        Morton order in one launch (phases timed apart), False/0 as given."""
        mode = 2 if fast is True else 0 if fast is False else int(fast)
        _check(_load_lib().kk_set_search_mode(self._handle, mode, int(reorder)), "kk_set_search_mode")

    def neighbours(self):
        """(nnear[nblock], inear[nblock, mmax]) of the last solve; 1-based ids, ascending."""
        lib = _load_lib()
        mm = ctypes.c_int(0)
        _check(lib.kk_get_mmax(self._handle, ctypes.byref(mm)), "kk_get_mmax")
        nb = self._nblock
        nnear = np.zeros(nb, dtype=np.int32)
        inear = np.zeros((nb, mm.value), dtype=np.int32)
        _check(lib.kk_get_neighbours(self._handle, nb, mm.value, _ip(nnear), _ip(inear)),
               "kk_get_neighbours")
        mask = np.arange(mm.value)[None, :] < nnear[:, None]
        return nnear, np.where(mask, inear + 1, 0)

    # ------------------------------------------------------------------
    # weight store (reference kriging.py:1747-1960)
    def free_weight_store(self):
        """Release the in-memory weight store."""
        _check(_load_lib().krige_free_weight_store(self._handle), "krige_free_weight_store")

    def set_weights(self, weights: dict) -> None:
        """Load kriging weights into the store so that solve() replays them
        against the current observation values (no search, no solve) and
        restores the stored variance.  ``weights`` as returned by
        :meth:`get_weights`: ``nnear`` (nblock, ngroups), ``inear`` (nblock,
        ngroups, nmax) 1-based, ``weight`` (nblock, ngroups, nmax) for nvar 1
        or (nblock, nvar, ngroups, nmax), optional ``order`` (nblock,) and
        ``variance`` (nblock,) or (nblock, nvar, nvar), zero when absent."""
        if self.store_weight:
            raise ValueError(
                "set_weights() is incompatible with store_weight=True. "
                "Construct Kriging without store_weight=True when reusing weights.")
        nnear = np.ascontiguousarray(weights["nnear"], dtype=np.int32)
        inear = np.ascontiguousarray(weights["inear"], dtype=np.int32)
        weight = np.asarray(weights["weight"], dtype=np.float64)
        if nnear.ndim != 2:
            raise ValueError("nnear must be 2-D (nblock, ngroups)")
        if inear.ndim != 3:
            raise ValueError("inear must be 3-D (nblock, ngroups, nmax)")
        nb, ng = nnear.shape
        nm = inear.shape[2]
        nv = self.nvar
        order_raw = weights.get("order", None)
        if order_raw is None or self.nsim == 0:
            order_f = np.arange(1, nb + 1, dtype=np.int32)
        else:
            order_f = np.ascontiguousarray(order_raw, dtype=np.int32)
        # the store's column-major layouts are the block-major arrays reversed
        nnear_f = np.asfortranarray(nnear.T)
        inear_f = np.asfortranarray(inear.T)
        if weight.ndim == 3:
            weight_f = np.asfortranarray(weight.T[:, :, np.newaxis, :])
        elif weight.ndim == 4:
            weight_f = np.asfortranarray(np.transpose(weight, (3, 2, 1, 0)))
        else:
            raise ValueError("weight must be 3-D (nblock, ngroups, nmax) or "
                             "4-D (nblock, nvar, ngroups, nmax)")
        if weights.get("variance", None) is not None:
            var = np.asarray(weights["variance"], dtype=np.float64)
            if var.ndim == 1:
                var = var[:, np.newaxis, np.newaxis]
            var_f = np.asfortranarray(np.transpose(var, (1, 2, 0)))
        else:
            var_f = np.zeros((nv, nv, nb), order="F")
        if weight_f.shape != (nm, ng, nv, nb) or var_f.shape != (nv, nv, nb):
            raise ValueError("weight/variance shapes do not match nnear and inear")
        _check(_load_lib().krige_set_weights(
            self._handle, nm, ng, nv, nb, _ip(nnear_f), _ip(inear_f), _dp(weight_f),
            _ip(order_f), _dp(var_f)), "krige_set_weights")
        self.use_old_weight = True

    def get_weights(self) -> dict:
        """The stored weights: ``nnear`` (nblock, ngroups) int32, ``inear``
        (nblock, ngroups, nmax) int32 1-based, ``weight`` (nblock, nvar,
        ngroups, nmax) -- (nblock, ngroups, nmax) for nvar 1 -- and
        ``variance`` (nblock,) for nvar 1, else (nblock, nvar, nvar).
        Entries past nnear are zero."""
        lib = _load_lib()
        nm, ng, nb = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.krige_get_weight_dims(self._handle, ctypes.byref(nm), ctypes.byref(ng),
                                         ctypes.byref(nb)), "krige_get_weight_dims")
        nm, ng, nb, nv = nm.value, ng.value, nb.value, self.nvar
        nnear_f = np.zeros((ng, nb), dtype=np.int32, order="F")
        inear_f = np.zeros((nm, ng, nb), dtype=np.int32, order="F")
        weight_f = np.zeros((nm, ng, nv, nb), dtype=np.float64, order="F")
        var_f = np.zeros((nv, nv, nb), dtype=np.float64, order="F")
        _check(lib.krige_get_weight_nnear(self._handle, ng, nb, _ip(nnear_f)), "krige_get_weight_nnear")
        _check(lib.krige_get_weight_inear(self._handle, nm, ng, nb, _ip(inear_f)), "krige_get_weight_inear")
        _check(lib.krige_get_weight_data(self._handle, nm, ng, nv, nb, _dp(weight_f)),
               "krige_get_weight_data")
        _check(lib.krige_get_weight_var(self._handle, nv, nb, _dp(var_f)), "krige_get_weight_var")
        weight = np.ascontiguousarray(weight_f.T)
        var = np.ascontiguousarray(var_f.T)
        if nv == 1:
            weight = weight[:, 0, :, :]
            var = var[:, 0, 0]
        return {"nnear": np.ascontiguousarray(nnear_f.T), "inear": np.ascontiguousarray(inear_f.T),
                "weight": weight, "variance": var}

    @property
    def solver_stats(self) -> dict:
        buf = (ctypes.c_int * 5)()
        _load_lib().krige_solver_stats(self._handle, buf)
        return dict(zip(["fail", "chol_fact", "chol_reuse", "ssytrf_fact", "ssytrf_reuse"], list(buf)))

    def get_results(self, copy: bool = False, squeeze: bool = True):
        lib = _load_lib()
        nb, ns = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.krige_get_nblocks(self._handle, ctypes.byref(nb)), "krige_get_nblocks")
        _check(lib.krige_get_nsim(self._handle, ctypes.byref(ns)), "krige_get_nsim")
        nb, nv, ns = nb.value, self.nvar, max(1, ns.value)
        est = np.empty((nb, nv, ns), order="F")
        var = np.empty((nb, nv, nv), order="F")
        _check(lib.krige_get_estimate_all(self._handle, nb, nv, ns, _dp(est)), "krige_get_estimate_all")
        _check(lib.krige_get_variance_all(self._handle, nb, nv, _dp(var)), "krige_get_variance_all")
        if squeeze:
            if nv == 1:
                est = est[:, 0]
                var = var[:, 0, 0]
            if ns <= 1:
                est = est[..., 0]
        if copy:
            est = np.array(est, order="C", copy=True)
            var = np.array(var, order="C", copy=True)
        return est, var

    def get_result_array(self) -> np.ndarray:
        lib = _load_lib()
        nb = ctypes.c_int(0)
        _check(lib.krige_get_nblocks(self._handle, ctypes.byref(nb)), "krige_get_nblocks")
        nb = nb.value
        nd, nv = self.ndim, self.nvar
        cf = np.empty((nd, nb), order="F")
        _check(lib.krige_get_block_coord(self._handle, nd, nb, _dp(cf)), "krige_get_block_coord")
        est, var = self.get_results(squeeze=False)
        names = ["x", "y", "z"][:nd]
        fields = [(n, np.float64) for n in names]
        fields += [("estimate", np.float64)] if nv == 1 else [(f"est_v{i + 1}", np.float64) for i in range(nv)]
        fields += [("variance", np.float64)] if nv == 1 else [(f"var_v{i + 1}", np.float64) for i in range(nv)]
        out = np.empty(nb, dtype=fields)
        for i, n in enumerate(names):
            out[n] = cf[i]
        if nv == 1:
            out["estimate"] = est[:, 0, 0]
            out["variance"] = var[:, 0, 0]
        else:
            for i in range(nv):
                out[f"est_v{i + 1}"] = est[:, i, 0]
                out[f"var_v{i + 1}"] = var[:, i, i]
        return out

    def get_result_df(self):
        import pandas as pd
        return pd.DataFrame(self.get_result_array())

    def to_str(self) -> str:
        p = _load_lib().krige_to_str(self._handle)
        return ctypes.string_at(p).decode() if p else ""

    def close(self):
        if getattr(self, "_handle", 0):
            h = _i64(self._handle)
            _load_lib().krige_destroy(ctypes.byref(h))
            self._handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_probe(nbytes, reps=10):
    """mean time (ms) of a plain device copy of nbytes"""
    ms = ctypes.c_double(0.0)
    _check(_load_lib().kk_copy_probe(float(nbytes), int(reps), ctypes.byref(ms)), "kk_copy_probe")
    return ms.value


def set_tuning(path: Optional[str] = None):
    """load a schedule table written by ``Kriging.autotune`` (None: defaults)"""
    _check(_load_lib().kk_set_tuning((path or "").encode("utf-8")), "kk_set_tuning")


def device_synchronize():
    _check(_load_lib().kk_device_synchronize(), "kk_device_synchronize")


def from_spec(spec, **kw):
    """Builds a ``Kriging`` object from a workload description
    (``krigekit_amd.synth``) and returns it ready to ``solve()``."""
    nvar = spec.get("nvar", 1)
    if nvar > 1:
        kw.setdefault("std_ck", spec.get("std_ck", True))
    k = Kriging(ndim=spec["ndim"], nvar=nvar, unbias=spec.get("unbias", 1),
                cross_validation=bool(spec.get("cross_validation", False)), **kw)
    for iv, o in enumerate(spec["obs"], start=1):
        k.set_obs(iv, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"),
                  maxdist=o.get("maxdist"), sk_mean=o.get("sk_mean", 0.0))
    if spec.get("grid") is None:
        k.set_grid_cv()
    elif spec.get("blocksize") is not None:
        # Gauss-Legendre integration nodes, 4 per axis
        k.set_grid_block(spec["grid"], -4, blocksize=np.asarray(spec["blocksize"]),
                         rangescale=spec.get("rangescale"), localnugget=spec.get("localnugget"))
    else:
        k.set_grid(spec["grid"], rangescale=spec.get("rangescale"), localnugget=spec.get("localnugget"))
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    return k
