"""Run one problem (tests/problems.py format) through the C oracle, the compiled
reference engine or the GPU library and return comparable results:
estimate, variance, nnear and inear (1-based ids, ascending, zero padded)."""
import numpy as np

from oracle import okrige


def _mask(nnear, inear):
    m = np.arange(inear.shape[1])[None, :] < nnear[:, None]
    return np.where(m, inear, 0)


def run_oracle(spec, table_mode=True, nthread=8, weights=False):
    o = okrige.solve(spec, table_mode=table_mode, nthread=nthread, want_weights=weights)
    o["inear"] = _mask(o["nnear"], o["inear"])
    return o


def run_ref(spec, mode="table", nthread=4, want_weights=False):
    from oracle.refdrv import RefKriging
    k = RefKriging(mode=mode, ndim=spec["ndim"], store_weight=True, unbias=spec.get("unbias", 1),
                   cross_validation=bool(spec.get("cross_validation", False)), bounds=spec.get("bounds"),
                   anisotropic_search=bool(spec.get("anisotropic_search", False)))
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"),
              maxdist=o.get("maxdist"), sk_mean=o.get("sk_mean", 0.0))
    if spec.get("grid") is None:
        k.set_grid_cv()
    elif spec.get("blocksize") is not None:
        k.set_grid_block_gq(spec["grid"], spec["blocksize"], spec.get("rangescale"), spec.get("localnugget"))
    else:
        k.set_grid(spec["grid"], spec.get("rangescale"), spec.get("localnugget"))
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    k.solve(nthread=nthread)
    e, v = k.results()
    nn, inn = k.neighbours()
    nnear = nn[:, 0].astype(np.int32)
    out = dict(estimate=e[:, 0], variance=v[:, 0, 0], nnear=nnear, inear=_mask(nnear, inn[:, 0, :]),
               stats=k.solver_stats())
    if want_weights:
        out["weight"] = k.weights()[:, 0, 0, :]
    k.close()
    return out


def make_gpu(spec, **kw):
    from krigekit_amd.kriging import Kriging
    kw.setdefault("bounds", spec.get("bounds"))
    kw.setdefault("anisotropic_search", bool(spec.get("anisotropic_search", False)))
    k = Kriging(ndim=spec["ndim"], unbias=spec.get("unbias", 1),
                cross_validation=bool(spec.get("cross_validation", False)), **kw)
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"),
              maxdist=o.get("maxdist"), sk_mean=o.get("sk_mean", 0.0))
    if spec.get("grid") is None:
        k.set_grid_cv()
    elif spec.get("blocksize") is not None:
        k.set_grid_block(spec["grid"], -4, blocksize=np.asarray(spec["blocksize"]),
                         rangescale=spec.get("rangescale"), localnugget=spec.get("localnugget"))
    else:
        k.set_grid(spec["grid"], rangescale=spec.get("rangescale"), localnugget=spec.get("localnugget"))
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    return k


def run_gpu(spec, fast=True, **kw):
    k = make_gpu(spec, **kw)
    k.set_search_mode(fast)
    k.solve()
    e, v = k.get_results()
    nn, inn = k.neighbours()
    out = dict(estimate=np.asarray(e), variance=np.asarray(v), nnear=nn, inear=inn,
               stats=k.solver_stats, timing=k.timing)
    k.close()
    return out


def rel_err(a, b, floor=1e-4):
    """|a-b| / max(|b|, floor); NaN where both are NaN counts as 0.  The floor
    (1e-4, against covariance sills of order 1) keeps values that vanish at data
    locations (variance 0, exact interpolation) from turning one ulp into a large
    relative error."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    e[both_nan] = 0.0
    return e


def _store_view(d):
    """(nnear, inear) as the reference's weight store reports them: a block
    whose solve failed (NaN estimate, variance max(NaN, 0) = 0) keeps one
    neighbour (ctx%nnear = 1, kriging_base.F90:3369-3383)"""
    nn, inn = np.asarray(d["nnear"]), np.asarray(d["inear"])
    if "estimate" not in d or "variance" not in d:
        return nn, inn
    failed = np.isnan(d["estimate"]) & (np.asarray(d["variance"]) == 0.0)
    if not failed.any():
        return nn, inn
    nn = np.where(failed, 1, nn)
    inn = inn.copy()
    inn[failed, 1:] = 0
    return nn, inn


def same_lists(x, y):
    """neighbour counts and ascending id lists identical (failed blocks as the
    reference's weight store reports them)"""
    xn, xi = _store_view(x)
    yn, yi = _store_view(y)
    x = dict(nnear=xn, inear=xi)
    y = dict(nnear=yn, inear=yi)
    if not np.array_equal(x["nnear"], y["nnear"]):
        return False
    w = min(x["inear"].shape[1], y["inear"].shape[1])
    if np.any(x["inear"][:, w:]) or np.any(y["inear"][:, w:]):
        return False
    return np.array_equal(x["inear"][:, :w], y["inear"][:, :w])


def max_errors(x, ref):
    """(max relative estimate error, max relative variance error) of x against
    ref.  Blocks with exactly one neighbour are compared on the variance only:
    the reference skips their solve (kriging_base.F90:2127) and its estimate
    reuses weights of another block.  NaN positions must agree."""
    one = ref["nnear"] == 1
    e = rel_err(x["estimate"][~one], ref["estimate"][~one])
    v = rel_err(x["variance"], ref["variance"])
    nan_ok = np.array_equal(np.isnan(x["variance"]), np.isnan(ref["variance"])) and \
        np.array_equal(np.isnan(x["estimate"][~one]), np.isnan(ref["estimate"][~one]))
    if not nan_ok:
        return np.inf, np.inf
    return float(np.max(e, initial=0.0)), float(np.max(v, initial=0.0))


def run_ref_ck(spec, mode="table", nthread=4, want_weights=False):
    """the compiled reference engine on a multi-variable problem"""
    from oracle.refdrv import RefKriging
    k = RefKriging(mode=mode, ndim=spec["ndim"], nvar=spec["nvar"], std_ck=spec.get("std_ck", True),
                   unbias=spec.get("unbias", 1), store_weight=want_weights, ndrift=spec.get("ndrift", 0),
                   weight_correction=spec.get("weight_correction", False))
    for v, o in enumerate(spec["obs"], 1):
        k.set_obs(v, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"),
                  maxdist=o.get("maxdist"), sk_mean=o.get("sk_mean", 0.0))
        if o.get("drift") is not None:
            k.set_obs_drift(v, o["drift"])
    if spec.get("sub_coord") is not None:
        k.set_grid_block_gq(spec["grid"], spec["blocksize"])
    else:
        k.set_grid(spec["grid"])
    if spec.get("grid_drift") is not None:
        k.set_grid_drift(spec["grid_drift"])
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    for g in spec.get("grad", []):
        k.set_grad(g["ivar"], g["coord1"], g["coord2"], g["value"], variance=g.get("variance"), drift=g.get("drift"))
    k.solve(nthread=nthread)
    e, v = k.results()
    out = dict(estimate=e, variance=v, stats=k.solver_stats())
    if want_weights:
        out["nnear"], out["inear"] = k.neighbours()
        out["weight"] = k.weights()
    k.close()
    return out


def make_gpu_ck(spec, **kw):
    from krigekit_amd.kriging import Kriging
    kw.setdefault("ndrift", spec.get("ndrift", 0))
    kw.setdefault("weight_correction", spec.get("weight_correction", False))
    k = Kriging(ndim=spec["ndim"], nvar=spec["nvar"], std_ck=spec.get("std_ck", True),
                unbias=spec.get("unbias", 1), **kw)
    for v, o in enumerate(spec["obs"], 1):
        k.set_obs(v, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"),
                  maxdist=o.get("maxdist"), sk_mean=o.get("sk_mean", 0.0))
        if o.get("drift") is not None:
            k.set_obs_drift(v, o["drift"])
    rl = dict(rangescale=spec.get("rangescale"), localnugget=spec.get("localnugget"))
    if spec.get("sub_coord") is not None:
        k.set_grid_block(spec["grid"], -4, blocksize=np.asarray(spec["blocksize"]), **rl)
    else:
        k.set_grid(spec["grid"], **rl)
    if spec.get("grid_drift") is not None:
        k.set_grid_drift(spec["grid_drift"])
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    for g in spec.get("grad", []):
        k.set_grad(g["coord1"], g["coord2"], g["value"], ivar=g["ivar"], variance=g.get("variance"),
                   drift_ext=g.get("drift"))
    return k


def run_gpu_ck(spec, **kw):
    k = make_gpu_ck(spec, **kw)
    k.solve()
    e, v = k.get_results(copy=True)
    out = dict(estimate=e, variance=v, timing=k.timing, stats=k.solver_stats)
    k.close()
    return out


def ck_errors(x, ref):
    """max relative errors (floor 1e-4) of estimates and variances; NaN
    positions must agree"""
    e = rel_err(x["estimate"], ref["estimate"])
    v = rel_err(x["variance"], ref["variance"])
    if not (np.array_equal(np.isnan(x["estimate"]), np.isnan(ref["estimate"])) and
            np.array_equal(np.isnan(x["variance"]), np.isnan(ref["variance"]))):
        return np.inf, np.inf
    return float(np.max(e, initial=0.0)), float(np.max(v, initial=0.0))


def run_gpu_st(spec, fast=True, **kw):
    from krigekit_amd.kriging_st import from_st_spec
    k = from_st_spec(spec, **kw)
    k.set_search_mode(fast)
    k.solve()
    e, v = k.get_results()
    nn, inn = k.neighbours()
    out = dict(estimate=np.asarray(e), variance=np.asarray(v), nnear=nn, inear=inn, timing=k.timing)
    k.close()
    return out


def run_ref_sgsim(spec, mode="table"):
    """the compiled reference's sequential simulation with the spec's path
    and draws: realisations (nblock, nsim) and variances (nblock,)"""
    from oracle.refdrv import RefKriging
    k = RefKriging(mode=mode, ndim=spec["ndim"], nsim=spec["nsim"], unbias=spec.get("unbias", 1),
                   bounds=spec.get("bounds"))
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"), maxdist=o.get("maxdist"),
              sk_mean=o.get("sk_mean", 0.0))
    k.set_grid(spec["grid"])
    k.set_sim(spec["path"], spec["sample"])
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    k.solve()
    e, v = k.results()
    k.close()
    return e[:, 0, :], v[:, 0, 0]


def run_gpu_sgsim(spec, **kw):
    """the library's sequential simulation with the spec's path and draws"""
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=spec["ndim"], nsim=spec["nsim"], unbias=spec.get("unbias", 1), bounds=spec.get("bounds"),
                **kw)
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], variance=o.get("variance"), nmax=o.get("nmax"), maxdist=o.get("maxdist"),
              sk_mean=o.get("sk_mean", 0.0))
    k.set_grid(spec["grid"])
    k.set_sim(spec["path"], spec["sample"])
    for v in spec["vgm"]:
        k.set_vgm(**v)
    for s in spec["search"]:
        k.set_search(**s)
    k.solve()
    e, v = k.get_results(copy=True, squeeze=False)
    t = k.timing
    k.close()
    return e[:, 0, :], v[:, 0, 0], t
