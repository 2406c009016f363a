"""ctypes driver for the compiled reference engine (test infrastructure only).

Loads ``oracle/_ref/libkriging_ref.so`` (table-mode covariance, the reference's
shipped default) or ``libkriging_ref_analytic.so`` built by ``oracle/Makefile``
and drives its C ABI directly, without importing any reference Python.  The
call sequence and argument conventions follow ``src/krigekit/kriging.py``
(coordinates transposed to (ndim, n) column-major, nmax/maxdist "unlimited"
sentinels at ``kriging.py:649-650``, vtype truncated to 3 lower-case chars at
``kriging.py:884``).

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use this module; the product library never does.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_i64 = ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

_LIBS = {}


def ref_available(mode="table"):
    return os.path.exists(_libpath(mode))


def _libpath(mode):
    name = "libkriging_ref.so" if mode == "table" else "libkriging_ref_analytic.so"
    return os.path.join(REF_DIR, name)


def load(mode="table"):
    if mode not in _LIBS:
        path = _libpath(mode)
        if not os.path.exists(path):
            raise FileNotFoundError(f"reference library not built: {path}")
        _LIBS[mode] = ctypes.CDLL(path)
    return _LIBS[mode]


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


def _fcoord(c):
    a = np.asarray(c, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a.T)


class RefError(RuntimeError):
    pass


class RefKriging:
    """Minimal mirror of the reference Kriging call sequence over its C ABI."""

    def __init__(self, mode="table", ndim=2, nvar=1, ndrift=0, unbias=1, nsim=0,
                 anisotropic_search=False, weight_correction=False,
                 store_weight=False, cross_validation=False, neglect_error=True,
                 std_ck=False, bounds=None, seed=12345):
        self.lib = load(mode)
        self.ndim, self.nvar = ndim, nvar
        h = _i64(0)
        self._chk(self.lib.krige_create(ctypes.byref(h)), "krige_create")
        self.h = h.value
        huge = sys.float_info.max * 1e3
        self._bounds = np.asfortranarray(bounds if bounds is not None else [-huge, huge],
                                         dtype=np.float64)
        f = self.lib.krige_initialize
        f.restype = _c_int
        f.argtypes = [_i64] + [_c_int] * 16 + [ctypes.c_char_p, _pd, _c_int]
        self._chk(f(self.h, ndim, nvar, ndrift, unbias, nsim,
                    int(anisotropic_search), int(weight_correction), 0,
                    int(store_weight), int(cross_validation), 0, int(neglect_error),
                    0, int(std_ck), 0, 0, b"", _dp(self._bounds), seed),
                  "krige_initialize")
        self.nobs = {}

    def _chk(self, ierr, name):
        if int(ierr) != 0:
            buf = ctypes.create_string_buffer(4096)
            self.lib.krige_get_last_error(buf, 4096)
            raise RefError(buf.value.decode(errors="replace").strip() or name)

    def set_obs(self, ivar, coord, value, variance=None, nmax=None, maxdist=None, sk_mean=0.0):
        cf = _fcoord(coord)
        v = np.ascontiguousarray(value, dtype=np.float64).ravel()
        var = (np.zeros(v.size) if variance is None
               else np.ascontiguousarray(variance, dtype=np.float64).ravel())
        f = self.lib.krige_set_obs
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd, _pd, _pd, _c_int, _c_double, _c_double]
        nm = nmax if nmax is not None else np.iinfo(np.int32).max
        md = maxdist if maxdist is not None else sys.float_info.max
        self._chk(f(self.h, ivar, v.size, cf.shape[0], _dp(cf), _dp(v), _dp(var), nm, md, sk_mean),
                  "krige_set_obs")
        self.nobs[ivar] = v.size

    def set_obs_drift(self, ivar, drift):
        """drift (nobs, ndrift), passed as (ndrift, nobs) column-major (kriging.py:665-685)"""
        d = np.asfortranarray(np.asarray(drift, dtype=np.float64).T)
        f = self.lib.krige_set_obs_drift
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd]
        self._chk(f(self.h, ivar, d.shape[0], d.shape[1], _dp(d)), "krige_set_obs_drift")

    def set_grid_drift(self, drift, ivar=None):
        """drift (nblocks, ndrift); ivar None broadcasts to every target (kriging.py:1096-1132)"""
        d = np.asfortranarray(np.asarray(drift, dtype=np.float64).T)
        f = self.lib.krige_set_grid_drift
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd]
        self._chk(f(self.h, -1 if ivar is None else ivar, d.shape[0], d.shape[1], _dp(d)),
                  "krige_set_grid_drift")

    def set_vgm(self, ivar, jvar, vtype, nugget=0.0, sill=1.0, a_major=1.0, a_minor1=None,
                a_minor2=None, azimuth=0.0, dip=0.0, plunge=0.0, product=False):
        if a_minor1 is None:
            a_minor1 = a_major
        if a_minor2 is None:
            a_minor2 = a_minor1
        if jvar < ivar:
            ivar, jvar = jvar, ivar
        f = self.lib.krige_set_vgm
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, ctypes.c_char_p] + [_c_double] * 8 + [_c_int]
        self._chk(f(self.h, ivar, jvar, vtype.lower()[:3].encode(), nugget, sill, a_major,
                    a_minor1, a_minor2, azimuth, dip, plunge, int(product)), "krige_set_vgm")

    def set_grid(self, coord, rangescale=None, localnugget=None):
        cf = _fcoord(coord)
        n = cf.shape[1]
        rs = np.ones(n) if rangescale is None else np.ascontiguousarray(rangescale, dtype=np.float64)
        ln = np.zeros(n) if localnugget is None else np.ascontiguousarray(localnugget, dtype=np.float64)
        f = self.lib.krige_set_grid
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _pd, _pd, _pd]
        self._chk(f(self.h, n, cf.shape[0], _dp(cf), _dp(rs), _dp(ln)), "krige_set_grid")
        self.nblock = n

    def set_grid_block_gq(self, coord, blocksize, rangescale=None, localnugget=None):
        """block_type=-4 (4^ndim Gauss-Legendre sub-nodes), as kriging.py:1040-1052."""
        cf = _fcoord(coord)
        nb = cf.shape[1]
        bs = np.asarray(blocksize, dtype=np.float64)
        if bs.ndim == 1:
            bs = np.tile(bs, (nb, 1))
        bsf = _fcoord(bs)
        nbp = np.ones(nb, dtype=np.int32)
        pw = np.ones(nb)
        rs = np.ones(nb) if rangescale is None else np.ascontiguousarray(rangescale, dtype=np.float64)
        ln = np.zeros(nb) if localnugget is None else np.ascontiguousarray(localnugget, dtype=np.float64)
        f = self.lib.krige_set_grid_block
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd, _c_int, _pi, _pd, _pd, _pd, _pd]
        self._chk(f(self.h, -4, nb, cf.shape[0], _dp(cf), nb, _ip(nbp), _dp(pw), _dp(bsf),
                    _dp(rs), _dp(ln)), "krige_set_grid_block")
        self.nblock = nb

    def set_grid_cv(self):
        f = self.lib.krige_set_grid_cv
        f.restype = _c_int
        f.argtypes = [_i64]
        self._chk(f(self.h), "krige_set_grid_cv")
        self.nblock = self.nobs[1]

    def set_search(self, ivar=1, anis1=1.0, anis2=1.0, azimuth=0.0, dip=0.0, plunge=0.0,
                   sector_search=False):
        f = self.lib.krige_set_search
        f.restype = _c_int
        f.argtypes = [_i64, _c_int] + [_c_double] * 5 + [_c_int]
        self._chk(f(self.h, ivar, anis1, anis2, azimuth, dip, plunge, int(sector_search)),
                  "krige_set_search")

    def set_grad(self, ivar, coord1, coord2, value, variance=None, drift=None):
        """increment observations z(coord1) - z(coord2) of variable ivar
        (kriging_capi.F90:472-506); drift (n, ndrift) when ndrift > 0"""
        c1, c2 = _fcoord(coord1), _fcoord(coord2)
        v = np.ascontiguousarray(value, dtype=np.float64).ravel()
        var = np.zeros(v.size) if variance is None else np.ascontiguousarray(variance, dtype=np.float64).ravel()
        if drift is None:
            d, nd = np.zeros((1, max(v.size, 1)), order="F"), 0
        else:
            d = np.asfortranarray(np.asarray(drift, dtype=np.float64).reshape(v.size, -1).T)
            nd = d.shape[0]
        f = self.lib.krige_set_grad
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd, _pd, _pd, _pd, _c_int, _pd]
        self._chk(f(self.h, ivar, v.size, c1.shape[0], _dp(c1), _dp(c2), _dp(v), _dp(var), nd, _dp(d)),
                  "krige_set_grad")

    def set_sim(self, randpath, sample):
        """SGSIM path (1-based permutation of the blocks) and standard normal
        draws (nblock, nsim) for nvar 1, passed as (nsim, nvar, nblock)
        column-major (kriging_capi.F90:397-443)"""
        rp = np.ascontiguousarray(randpath, dtype=np.int32).ravel()
        smp = np.asarray(sample, dtype=np.float64)
        nb, ns = smp.shape
        sf = np.asfortranarray(smp.T.reshape(ns, 1, nb))
        f = self.lib.krige_set_sim
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, ctypes.c_void_p, _c_int, _c_int, ctypes.c_void_p]
        self._chk(f(self.h, nb, rp.ctypes.data, ns, 1, sf.ctypes.data), "krige_set_sim")
        self.nsim = ns

    def solve(self, nthread=1, ncache=-1):
        f = self.lib.krige_solve
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int]
        self._chk(f(self.h, nthread, ncache), "krige_solve")

    def results(self):
        """estimate (nblock, nvar), or (nblock, nvar, nsim) after set_sim, and
        variance (nblock, nvar, nvar)"""
        nb, nv = self.nblock, self.nvar
        ns = getattr(self, "nsim", 0)
        if ns:
            est = np.empty((nb, nv, ns), order="F")
            f = self.lib.krige_get_estimate_all
            f.restype = _c_int
            f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd]
            self._chk(f(self.h, nb, nv, ns, _dp(est)), "krige_get_estimate_all")
            var = np.empty((nb, nv, nv), order="F")
            g = self.lib.krige_get_variance_all
            g.restype = _c_int
            g.argtypes = [_i64, _c_int, _c_int, _pd]
            self._chk(g(self.h, nb, nv, _dp(var)), "krige_get_variance_all")
            return np.ascontiguousarray(est), np.ascontiguousarray(var)
        est = np.empty((nb, nv, 1), order="F")
        var = np.empty((nb, nv, nv), order="F")
        f = self.lib.krige_get_estimate_all
        f.restype = _c_int
        f.argtypes = [_i64, _c_int, _c_int, _c_int, _pd]
        self._chk(f(self.h, nb, nv, 1, _dp(est)), "krige_get_estimate_all")
        g = self.lib.krige_get_variance_all
        g.restype = _c_int
        g.argtypes = [_i64, _c_int, _c_int, _pd]
        self._chk(g(self.h, nb, nv, _dp(var)), "krige_get_variance_all")
        return np.ascontiguousarray(est[:, :, 0]), np.ascontiguousarray(var)

    def solver_stats(self):
        buf = (ctypes.c_int * 5)()
        self.lib.krige_solver_stats(_i64(self.h), buf)
        return dict(zip(["fail", "chol_fact", "chol_reuse", "ssytrf_fact", "ssytrf_reuse"], list(buf)))

    def neighbours(self):
        """Per-block neighbour index lists (1-based, ascending) from the
        in-memory weight store (requires store_weight=True at construction)."""
        nm, ng, nb = _c_int(), _c_int(), _c_int()
        f = self.lib.krige_get_weight_dims
        f.restype = _c_int
        f.argtypes = [_i64, _pi, _pi, _pi]
        self._chk(f(self.h, ctypes.byref(nm), ctypes.byref(ng), ctypes.byref(nb)),
                  "krige_get_weight_dims")
        nnear = np.zeros((nb.value, ng.value), dtype=np.int32)  # F-order (ng, nb)
        g = self.lib.krige_get_weight_nnear
        g.restype = _c_int
        g.argtypes = [_i64, _c_int, _c_int, _pi]
        self._chk(g(self.h, ng.value, nb.value, _ip(nnear)), "krige_get_weight_nnear")
        inear = np.zeros((nb.value, ng.value, nm.value), dtype=np.int32)  # F-order (nm, ng, nb)
        k = self.lib.krige_get_weight_inear
        k.restype = _c_int
        k.argtypes = [_i64, _c_int, _c_int, _c_int, _pi]
        self._chk(k(self.h, nm.value, ng.value, nb.value, _ip(inear)), "krige_get_weight_inear")
        return nnear, inear

    def weights(self):
        nm, ng, nb = _c_int(), _c_int(), _c_int()
        f = self.lib.krige_get_weight_dims
        f.restype = _c_int
        f.argtypes = [_i64, _pi, _pi, _pi]
        self._chk(f(self.h, ctypes.byref(nm), ctypes.byref(ng), ctypes.byref(nb)),
                  "krige_get_weight_dims")
        w = np.zeros((nb.value, self.nvar, ng.value, nm.value))  # F-order (nm, ng, nvar, nb)
        g = self.lib.krige_get_weight_data
        g.restype = _c_int
        g.argtypes = [_i64, _c_int, _c_int, _c_int, _c_int, _pd]
        self._chk(g(self.h, nm.value, ng.value, self.nvar, nb.value, _dp(w)), "krige_get_weight_data")
        return w

    def close(self):
        if getattr(self, "h", 0):
            h = _i64(self.h)
            self.lib.krige_destroy(ctypes.byref(h))
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
