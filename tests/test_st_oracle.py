"""Pins the oracle's space-time product-sum path (CPU only).

The compiled reference cannot run its own space-time path here: every
krige_st_set_vgm / krige_st_set_vgm_temporal call reaches
`present(product) .and. product` (variogram.f90:462) with `product` absent,
and amdflang evaluates both operands (a null dereference).  So the oracle is
pinned by
  * two exact reductions to the reference's SPATIAL path, which does build:
    all times equal -> 3-D OK with Cs, variance x (1 + k Ct(0));
    all points at one place -> 1-D OK on the times with Ct, variance x
    (1 + k Cs(0)) (ordinary kriging weights do not change when every
    covariance is scaled by a > 0 and shifted by b; the variance scales by a);
  * the covariance composition k Cs Ct + Cs + Ct against the oracle's own
    spatial and 1-D covariances (themselves pinned against the reference);
  * an independent numpy brute force (distance sort in the search metric,
    dense bordered solve) in general position;
  * the reference's own space-time known-answer tests
    (tests/test_spacetime_kriging.py:168-183, 243-282), which hold for any
    positive-definite model and so are run with the product-sum model.
"""
import numpy as np
import pytest

from oracle import okrige, refdrv
from runners import rel_err, run_oracle, run_ref

needs_ref = pytest.mark.skipif(not refdrv.ref_available("table"), reason="compiled reference not built")


def _st_spec(coord4, value, grid4, vgm, vgm_t, k_ps, nmax, maxdist=None, time_at=1.0, sector=False):
    return dict(name="st", st=True, ndim=3, k_ps=k_ps,
                obs=[dict(coord=coord4, value=value, nmax=nmax, maxdist=maxdist)], grid=grid4,
                vgm=[vgm], vgm_t=[vgm_t], search=[dict(ivar=1, time_at=time_at, sector_search=sector)])


@needs_ref
@pytest.mark.parametrize("table", [True, False])
@pytest.mark.parametrize("sector", [False, True])
def test_equal_times_reduce_to_spatial_reference(table, sector):
    rng = np.random.default_rng(7)
    n = 400
    xyz = np.column_stack([rng.uniform(0, 200, (n, 2)), rng.uniform(0, 20, n)])
    val = np.sin(xyz[:, 0] / 30) + np.cos(xyz[:, 1] / 45) + rng.normal(0, .1, n)
    g = np.column_stack([rng.uniform(0, 200, (50, 2)), rng.uniform(0, 20, 50)])
    nm = 3 if sector else 16
    vg = dict(ivar=1, jvar=1, vtype="exp", nugget=0.05, sill=1.0, a_major=50.0, a_minor1=30.0, a_minor2=10.0,
              azimuth=20.0)
    vt = dict(vtype="sph", nugget=0.1, sill=0.7, at_k=5.0)
    k = 0.5
    sp = dict(name="s", ndim=3, obs=[dict(coord=xyz, value=val, nmax=nm, maxdist=60.0)], grid=g, vgm=[vg],
              search=[dict(ivar=1, sector_search=sector)])
    st = _st_spec(np.column_stack([xyz, np.full(n, 3.0)]), val, np.column_stack([g, np.full(len(g), 3.0)]), vg,
                  vt, k, nm, maxdist=60.0, time_at=10.0, sector=sector)
    r = run_ref(sp, "table" if table else "analytic")
    o = run_oracle(st, table_mode=table)
    a = 1.0 + k * (0.7 + 0.1)  # 1 + k Ct(0), Ct(0) = sill + nugget
    assert np.array_equal(o["nnear"], r["nnear"]) and np.array_equal(o["inear"], r["inear"])
    assert np.max(rel_err(o["estimate"], r["estimate"])) < 1e-10
    assert np.max(rel_err(o["variance"], a * r["variance"])) < 1e-10


@needs_ref
@pytest.mark.parametrize("table", [True, False])
def test_colocated_reduces_to_1d_reference(table):
    rng = np.random.default_rng(8)
    nt = 300
    t = rng.uniform(0, 100, nt)
    v = np.sin(t / 7) + rng.normal(0, .1, nt)
    tg = rng.uniform(0, 100, 40)
    k, time_at = 0.3, 2.0
    vs = dict(ivar=1, jvar=1, vtype="sph", nugget=0.05, sill=0.8, a_major=50.0)
    vt = dict(vtype="exp", nugget=0.02, sill=0.6, at_k=8.0)
    st = _st_spec(np.column_stack([np.zeros((nt, 3)), t]), v, np.column_stack([np.zeros((40, 3)), tg]), vs, vt,
                  k, 12, maxdist=40.0, time_at=time_at)
    sp = dict(name="s", ndim=1, obs=[dict(coord=t[:, None], value=v, nmax=12, maxdist=40.0 / time_at)],
              grid=tg[:, None], vgm=[dict(ivar=1, jvar=1, vtype="exp", nugget=0.02, sill=0.6, a_major=8.0)],
              search=[dict(ivar=1)])
    r = run_ref(sp, "table" if table else "analytic")
    o = run_oracle(st, table_mode=table)
    a = 1.0 + k * (0.8 + 0.05)  # 1 + k Cs(0), Cs(0) = sill + nugget
    assert np.array_equal(o["nnear"], r["nnear"]) and np.array_equal(o["inear"], r["inear"])
    assert np.max(rel_err(o["estimate"], r["estimate"])) < 1e-10
    assert np.max(rel_err(o["variance"], a * r["variance"])) < 1e-10


VS = dict(ivar=1, jvar=1, vtype="exp", nugget=0.05, sill=1.0, a_major=50.0, a_minor1=20.0, a_minor2=8.0,
          azimuth=35.0, dip=10.0)
VT = dict(vtype="gau", nugget=0.03, sill=0.5, at_k=5.0)


@pytest.mark.parametrize("table", [True, False])
def test_product_sum_composition(table):
    rng = np.random.default_rng(9)
    lags = np.column_stack([rng.uniform(-60, 60, (500, 3)), rng.uniform(-12, 12, 500)])
    lags[:50, :3] = 0.0   # same place, different times
    lags[50:100, 3] = 0.0  # same time
    k = 0.4
    st = _st_spec(np.zeros((2, 4)), np.zeros(2), np.zeros((1, 4)), VS, VT, k, 4)
    c = okrige.cov(st, lags, table_mode=table)
    sp = dict(ndim=3, obs=[dict(coord=np.zeros((2, 3)), value=np.zeros(2))], grid=np.zeros((1, 3)), vgm=[VS],
              search=[dict(ivar=1)])
    tp = dict(ndim=1, obs=[dict(coord=np.zeros((2, 1)), value=np.zeros(2))], grid=np.zeros((1, 1)),
              vgm=[dict(ivar=1, jvar=1, vtype="gau", nugget=0.03, sill=0.5, a_major=5.0)], search=[dict(ivar=1)])
    cs = okrige.cov(sp, lags[:, :3], table_mode=table)
    ct = okrige.cov(tp, np.abs(lags[:, 3:]), table_mode=table)
    np.testing.assert_allclose(c, k * cs * ct + cs + ct, rtol=1e-14, atol=1e-15)
    # the temporal covariance is even in the time lag
    flip = lags.copy()
    flip[:, 3] *= -1
    np.testing.assert_array_equal(okrige.cov(st, flip, table_mode=table), c)


def _brute_st(coord4, val, grid4, k, time_at, nmax, maxdist, table):
    """dense product-sum OK per target, neighbours by a full distance sort in
    the search metric (x, y, z, t time_at); covariances from the oracle's
    spatial and 1-D functions"""
    sp = dict(ndim=3, obs=[dict(coord=np.zeros((2, 3)), value=np.zeros(2))], grid=np.zeros((1, 3)), vgm=[VS],
              search=[dict(ivar=1)])
    tp = dict(ndim=1, obs=[dict(coord=np.zeros((2, 1)), value=np.zeros(2))], grid=np.zeros((1, 1)),
              vgm=[dict(ivar=1, jvar=1, vtype="gau", nugget=0.03, sill=0.5, a_major=5.0)], search=[dict(ivar=1)])

    def C(a, b):
        lag = a[:, None, :] - b[None, :, :]
        s = okrige.cov(sp, lag[..., :3].reshape(-1, 3), table).reshape(lag.shape[:2])
        t = okrige.cov(tp, np.abs(lag[..., 3]).reshape(-1, 1), table).reshape(lag.shape[:2])
        return k * s * t + s + t

    cs0 = VS["sill"] + VS["nugget"]
    ct0 = VT["sill"] + VT["nugget"]
    cov0 = k * cs0 * ct0 + cs0 + ct0
    sc = coord4 * np.array([1, 1, 1, time_at])
    est, var, lists = [], [], []
    for g in grid4:
        d2 = np.sum((sc - g * np.array([1, 1, 1, time_at])) ** 2, axis=1)
        idx = np.argsort(d2, kind="stable")[:nmax]
        idx = np.sort(idx[d2[idx] <= maxdist ** 2])
        n = len(idx)
        K = np.ones((n + 1, n + 1))
        K[n, n] = 0.0
        K[:n, :n] = C(coord4[idx], coord4[idx])
        K[np.arange(n), np.arange(n)] = cov0
        r = np.ones(n + 1)
        r[:n] = C(coord4[idx], g[None, :])[:, 0]
        x = np.linalg.solve(K, r)
        est.append(x[:n] @ val[idx])
        var.append(max(cov0 - x @ r, 0.0))
        lists.append(idx + 1)
    return np.array(est), np.array(var), lists


@pytest.mark.parametrize("table", [True, False])
def test_general_position_against_brute_force(table):
    rng = np.random.default_rng(10)
    ns, nd = 25, 30
    st_xy = rng.uniform(0, 120, (ns, 2))
    coord4 = np.column_stack([np.repeat(st_xy, nd, axis=0), np.repeat(rng.uniform(0, 5, ns), nd),
                              np.tile(np.arange(nd, dtype=float), ns) + np.repeat(rng.uniform(0, .5, ns), nd)])
    val = np.sin(coord4[:, 0] / 25) + 0.3 * np.cos(coord4[:, 3] / 4) + rng.normal(0, .05, len(coord4))
    grid4 = np.column_stack([rng.uniform(0, 120, (60, 2)), rng.uniform(0, 5, 60), rng.uniform(0, 29, 60)])
    k, time_at, nmax, md = 0.5, 4.0, 20, 45.0
    spec = _st_spec(coord4, val, grid4, VS, VT, k, nmax, maxdist=md, time_at=time_at)
    o = run_oracle(spec, table_mode=table)
    e, v, lists = _brute_st(coord4, val, grid4, k, time_at, nmax, md, table)
    for b in range(len(grid4)):
        assert o["nnear"][b] == len(lists[b])
        np.testing.assert_array_equal(o["inear"][b, :o["nnear"][b]], lists[b])
    assert np.max(rel_err(o["estimate"], e)) < 1e-10
    assert np.max(rel_err(o["variance"], v)) < 1e-10


def _kat_spec(obs4, value, grid4, nmax, maxdist=None, time_at=1.0):
    vs = dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.8, a_major=500.0, a_minor1=300.0, a_minor2=100.0)
    vt = dict(vtype="exp", nugget=0.0, sill=0.5, at_k=10.0)
    return _st_spec(np.asarray(obs4, float), np.asarray(value, float), np.asarray(grid4, float), vs, vt, 0.5, nmax,
                    maxdist=maxdist, time_at=time_at)


@pytest.mark.parametrize("time_at,expect", [(0.1, 10.0), (100.0, 20.0)])
def test_time_at_selects_nearest_pair(time_at, expect):
    """tests/test_spacetime_kriging.py:243-255: a co-located pair ten time
    units away against a contemporaneous pair 3 units away in space"""
    obs = [[0, 0, 0, 10], [0, 0, 0, 11], [3, 0, 0, 0], [3.1, 0, 0, 0], [100, 0, 0, 0], [0, 100, 0, 0],
           [0, 0, 100, 0]]
    o = run_oracle(_kat_spec(obs, [10, 10, 20, 20, -1, -2, -3], [[0, 0, 0, 0]], nmax=2, time_at=time_at))
    assert o["estimate"][0] == pytest.approx(expect)


@pytest.mark.parametrize("time_at,inside", [(50.0, True), (200.0, False)])
def test_time_at_scales_search_radius(time_at, inside):
    """tests/test_spacetime_kriging.py:257-282: maxdist applies in the search
    metric, where a time lag counts dt * time_at"""
    o = run_oracle(_kat_spec([[0, 0, 0, .04], [0, 0, 0, .05]], [10, 10], [[0, 0, 0, 0]], nmax=2, maxdist=5.0,
                             time_at=time_at))
    if inside:
        assert o["estimate"][0] == pytest.approx(10.0)
    else:
        assert np.isnan(o["estimate"][0])


def test_exact_at_an_observation():
    """tests/test_spacetime_kriging.py:168-183: a target on an observation
    returns its value"""
    rng = np.random.default_rng(42)
    c = rng.uniform([0, 0, 0, 0], [1000, 1000, 100, 50], size=(200, 4))
    v = np.sin(c[:, 0] / 200) + np.cos(c[:, 1] / 300) + 0.01 * c[:, 3] + rng.normal(0, 0.1, 200)
    o = run_oracle(_kat_spec(c, v, c[[5]], nmax=20))
    assert abs(o["estimate"][0] - v[5]) < 1e-9
