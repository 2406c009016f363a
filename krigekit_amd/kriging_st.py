"""GPU drop-in for krigekit's ``SpaceTimeKriging`` (src/krigekit/kriging_st.py:275-1000).

Same call sequence as the reference class -- ``set_st_model``, ``set_obs``
(combined ``(nobs, ndim+1)`` coordinates or split ``coord`` + ``time``),
``set_vgm`` (spatial structures), ``set_vgm_temporal``, ``set_grid``,
``set_search(ivar, time_at, ...)``, ``solve``, ``get_results`` -- over the
``krige_st_*`` entry points of ``libkrigekit_amd.so``.  Both space-time
models are provided: the sum-metric (the default; joint sills through
``set_vgm_joint_sills``) and the product-sum ``C = k Cs Ct + Cs + Ct``;
gradients, blocks, drift and simulation raise ``RuntimeError`` from the
library.
"""
import ctypes
import random
import sys
from typing import Optional

import numpy as np

from .kriging import _PRECISIONS, Kriging, _check, _dp, _i64, _load_lib


_TRANSFORM_ALIASES = {"linear": "lin", "bounded": "exp", "power": "pow"}


def _split(coord, time):
    a = np.asarray(coord, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError(f"coord must be 2-D, got shape {a.shape}")
    if time is None:
        if a.shape[1] not in (2, 3, 4):
            raise ValueError(f"Combined coord must be (n, ndim+1) with ndim=1,2,3, got {a.shape}")
        return a[:, :-1], a[:, -1]
    if a.shape[1] not in (1, 2, 3):
        raise ValueError(f"Split coord must be (n, ndim) with ndim=1,2,3, got {a.shape}")
    t = np.asarray(time, dtype=np.float64).ravel()
    if len(t) != len(a):
        raise ValueError(f"time length ({len(t)}) != n ({len(a)})")
    return a, t


class SpaceTimeKriging(Kriging):
    """Space-time kriging on the GPU; solve/results/timing as ``Kriging``."""

    def __init__(self, nvar: int = 1, ndrift: int = 0, unbias: int = 1, nsim: int = 0,
                 anisotropic_search: bool = False, weight_correction: bool = False,
                 use_old_weight: bool = False, store_weight: bool = False,
                 cross_validation: bool = False, write_mat: bool = False, neglect_error: bool = True,
                 verbose: bool = False, weight_file: str = "", bounds: Optional[tuple] = None,
                 seed: Optional[int] = None, precision: str = "fp64", table_mode: bool = True,
                 pivot_tol: float = 1.0e-3):
        lib = _load_lib()
        h = _i64(0)
        _check(lib.krige_st_create(ctypes.byref(h)), "krige_st_create")
        self._handle = h.value
        huge = sys.float_info.max
        self.bounds = np.asfortranarray(bounds if bounds is not None else [-huge, huge], dtype=np.float64)
        seed = seed or random.randint(0, 2 ** 31 - 1)
        _check(lib.krige_st_initialize(
            self._handle, nvar, ndrift, unbias, nsim, int(anisotropic_search), int(weight_correction),
            int(use_old_weight), int(store_weight), int(cross_validation), int(write_mat),
            int(neglect_error), int(verbose), weight_file.encode("utf-8") if weight_file else b"",
            _dp(self.bounds), int(seed) & 0x7FFFFFFF), "krige_st_initialize")
        self.ndim = None  # spatial ndim, set by the first set_obs
        self.nvar, self.ndrift, self.nsim, self.unbias = nvar, ndrift, nsim, unbias
        self.store_weight = store_weight
        self.use_old_weight = use_old_weight
        self.weight_file = weight_file
        self.cross_validation = cross_validation
        self.verbose = verbose
        self.seed = seed
        self._nblock = 0
        self._nobs = np.zeros(nvar, dtype=np.int64)
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        self.set_options(precision=precision, table_mode=table_mode, pivot_tol=pivot_tol)

    def set_st_model(self, model: str = "sum_metric", transform: str = "linear", at: float = 1.0,
                     time_nugget: float = 0.0, time_sill: float = 1.0, k_ps: float = 0.0):
        """Global space-time model (reference kriging_st.py:356-391): the
        sum-metric Cs + Ct + sum_k sill_st_k shape_k(sqrt(h_k^2 + f(dt)^2)),
        f(dt) = time_nugget + time_sill (1 - corefunc(transform, |dt| / at)),
        or the product-sum k_ps Cs Ct + Cs + Ct.  transform aliases: linear
        -> lin, bounded -> exp, power -> pow."""
        model = model.lower()
        if model not in ("sum_metric", "product_sum"):
            raise ValueError("model must be 'sum_metric' or 'product_sum'")
        t = _TRANSFORM_ALIASES.get(str(transform).lower(), str(transform).lower())
        _check(_load_lib().krige_st_set_st_model(self._handle, model.encode(), t.encode(),
                                                 at, time_nugget, time_sill, k_ps), "krige_st_set_st_model")

    def set_obs(self, ivar: int, coord, value, time=None, variance=None, nmax: Optional[int] = None,
                maxdist: Optional[float] = None, sk_mean: float = 0.0):
        spatial, t = _split(coord, time)
        nd = spatial.shape[1]
        if self.ndim is None:
            self.ndim = nd
        elif self.ndim != nd:
            raise ValueError(f"ndim mismatch: {self.ndim} vs {nd}")
        n = len(spatial)
        cf = np.empty((nd + 1, n), dtype=np.float64, order="F")
        cf[:nd] = spatial.T
        cf[nd] = t
        v = np.asfortranarray(np.asarray(value, dtype=np.float64).ravel())
        if v.size != n:
            raise ValueError(f"value length ({v.size}) != nobs ({n})")
        var = np.zeros(n) if variance is None else np.asfortranarray(np.asarray(variance, dtype=np.float64))
        nm = nmax if nmax is not None else np.iinfo(np.int32).max
        md = maxdist if maxdist is not None else sys.float_info.max
        _check(_load_lib().krige_st_set_obs(self._handle, ivar, n, nd, _dp(cf), _dp(v), _dp(var), nm, md,
                                            sk_mean), "krige_st_set_obs")
        self._nobs[ivar - 1] = n

    def set_vgm(self, ivar: int, jvar: int, vtype: str, nugget: float = 0.0, sill: float = 1.0,
                a_major: float = 1.0, a_minor1: Optional[float] = None, a_minor2: Optional[float] = None,
                azimuth: float = 0.0, dip: float = 0.0, plunge: float = 0.0, append: bool = True):
        a_minor1 = a_major if a_minor1 is None else a_minor1
        a_minor2 = a_minor1 if a_minor2 is None else a_minor2
        if not append:
            _check(_load_lib().krige_st_reset_vgm(self._handle, ivar, jvar), "krige_st_reset_vgm")
        _check(_load_lib().krige_st_set_vgm(self._handle, ivar, jvar, vtype.lower()[:3].encode(), nugget, sill,
                                            a_major, a_minor1, a_minor2, azimuth, dip, plunge), "krige_st_set_vgm")

    def set_vgm_temporal(self, ivar: int, jvar: int, vtype: str = "sph", nugget: float = 0.0,
                         sill: float = 1.0, at_k: float = 1.0):
        _check(_load_lib().krige_st_set_vgm_temporal(self._handle, ivar, jvar, vtype.lower()[:3].encode(),
                                                     nugget, sill, at_k), "krige_st_set_vgm_temporal")

    def set_vgm_joint_sills(self, ivar: int, jvar: int, *sills: float):
        s = np.asarray(sills, dtype=np.float64)
        _check(_load_lib().krige_st_set_vgm_joint_sills(self._handle, ivar, jvar, len(s), _dp(s)),
               "krige_st_set_vgm_joint_sills")

    def set_grid(self, coord=None, time=None, rangescale=None, localnugget=None):
        spatial, t = _split(coord, time)
        n, nd = spatial.shape
        c3 = np.zeros((3, n), dtype=np.float64, order="F")
        c3[:nd] = spatial.T
        tf = np.asfortranarray(t)
        rs = np.asfortranarray(rangescale if rangescale is not None else np.ones(n), dtype=np.float64)
        ln = np.asfortranarray(localnugget if localnugget is not None else np.zeros(n), dtype=np.float64)
        _check(_load_lib().krige_st_set_grid(self._handle, n, _dp(c3), _dp(tf), _dp(rs), _dp(ln)),
               "krige_st_set_grid")
        self._nblock = n

    def set_search(self, ivar: int = 1, time_at: float = 1.0, anis1: float = 1.0, anis2: float = 1.0,
                   azimuth: float = 0.0, dip: float = 0.0, plunge: float = 0.0, sector_search: bool = False):
        _check(_load_lib().krige_st_set_search(self._handle, ivar, time_at, anis1, anis2, azimuth, dip, plunge,
                                               int(sector_search)), "krige_st_set_search")


def from_st_spec(spec, **kw):
    """``SpaceTimeKriging`` ready to solve from a space-time workload
    description (``krigekit_amd.synth.st``)."""
    k = SpaceTimeKriging(unbias=spec.get("unbias", 1), cross_validation=bool(spec.get("cross_validation", False)),
                         **kw)
    if spec.get("st_model", "product_sum") == "sum_metric":
        k.set_st_model("sum_metric", spec.get("transform", "lin"), at=spec.get("at", 1.0),
                       time_nugget=spec.get("time_nugget", 0.0), time_sill=spec.get("time_sill", 1.0))
    else:
        k.set_st_model("product_sum", k_ps=spec["k_ps"])
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"),
              maxdist=o.get("maxdist"), sk_mean=o.get("sk_mean", 0.0))
    if spec.get("grid") is None:
        k.set_grid_cv()
    else:
        k.set_grid(spec["grid"])
    for v in spec["vgm"]:
        k.set_vgm(**{key: val for key, val in v.items() if key != "product"})
    for v in spec["vgm_t"]:
        k.set_vgm_temporal(1, 1, v["vtype"], nugget=v.get("nugget", 0.0), sill=v.get("sill", 1.0), at_k=v["at_k"])
    if spec.get("sill_st") is not None:
        k.set_vgm_joint_sills(1, 1, *spec["sill_st"])
    for s in spec["search"]:
        k.set_search(**s)
    return k
