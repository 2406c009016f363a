"""Pins the oracle's space-time sum-metric path (CPU only).

C = Cs(h_s) + Ct(|dt|) + sum_k sill_st_k shape_k(sqrt(h_k^2 + f(dt)^2)),
f(dt) = max(time_nugget + time_sill (1 - corefunc(transform, |dt| / at)), 0),
f(0) = 0 (variogram_st.f90:126-141, 178-194, 241-261); C(0, 0) = Cs(0) +
Ct(0) + the joint sills of the non-nugget structures (:344-375).

The compiled reference's space-time entry points cannot run here (the
null-dereference described in test_st_oracle.py), so the oracle is pinned by
  * two exact reductions to the reference's SPATIAL path in analytic mode:
    all times equal (f = 0): a spatial model whose structures are those of Cs
    followed by a copy of each non-nugget one with sill sill_st_k and no
    nugget, shifted by the constant Ct(0); all points at one place (h_k = 0)
    with the linear transform below saturation: a 1-D model on the times with
    the structures of Ct followed by one per joint sill (shape of Cs structure
    k, range at), shifted by Cs(0).  Ordinary kriging weights and variances do
    not change under a constant shift of every covariance;
  * the composition against the oracle's own spatial, 1-D and per-structure
    tables in both table layouts (one composite table: analytic joint term;
    per-structure tables: the structure's table over its sill);
  * an independent numpy brute force (distance sort in the search metric,
    dense bordered solve) in general position;
  * the reference's own sum-metric known-answer tests
    (tests/test_spacetime_kriging.py:168-189, 207-240, 243-282).
"""
import numpy as np
import pytest

from oracle import okrige, refdrv
from runners import rel_err, run_oracle, run_ref

needs_ref = pytest.mark.skipif(not refdrv.ref_available("table"), reason="compiled reference not built")

CORE = {"sph": lambda r: np.where(r < 1, 1 - 1.5 * r + 0.5 * r ** 3, 0.0), "exp": lambda r: np.exp(-3 * r),
        "gau": lambda r: np.exp(-3.0625 * r ** 2), "lin": lambda r: np.where(r < 1, 1 - r, 0.0),
        "nug": lambda r: np.zeros_like(r)}


def _sm_spec(coord4, value, grid4, vgm, vgm_t, sill_st, nmax, maxdist=None, time_at=1.0, sector=False,
             transform="lin", at=1.0, time_nugget=0.0, time_sill=1.0):
    return dict(name="stsm", st=True, ndim=3, st_model="sum_metric", transform=transform, at=at,
                time_nugget=time_nugget, time_sill=time_sill, sill_st=list(sill_st),
                obs=[dict(coord=coord4, value=value, nmax=nmax, maxdist=maxdist)], grid=grid4,
                vgm=list(vgm), vgm_t=list(vgm_t), search=[dict(ivar=1, time_at=time_at, sector_search=sector)])


def _f_time(dt, transform, at, tn, ts):
    adt = np.abs(dt)
    dw = np.maximum(tn + ts * (1.0 - CORE[transform](adt / at)), 0.0)
    return np.where(adt <= np.finfo(float).eps, 0.0, dw)


@needs_ref
@pytest.mark.parametrize("sector", [False, True])
def test_equal_times_reduce_to_spatial_reference(sector):
    rng = np.random.default_rng(17)
    n = 400
    xyz = np.column_stack([rng.uniform(0, 200, (n, 2)), rng.uniform(0, 20, n)])
    val = np.sin(xyz[:, 0] / 30) + np.cos(xyz[:, 1] / 45) + rng.normal(0, .1, n)
    g = np.column_stack([rng.uniform(0, 200, (50, 2)), rng.uniform(0, 20, 50)])
    nm = 3 if sector else 16
    cs = [dict(ivar=1, jvar=1, vtype="nug", nugget=0.05, sill=0.0, a_major=1.0),
          dict(ivar=1, jvar=1, vtype="exp", nugget=0.0, sill=0.7, a_major=50.0, a_minor1=30.0, a_minor2=10.0,
               azimuth=20.0),
          dict(ivar=1, jvar=1, vtype="sph", nugget=0.02, sill=0.2, a_major=90.0)]
    joint = [0.2, 0.3, 0.15]
    vt = dict(vtype="sph", nugget=0.1, sill=0.7, at_k=5.0)
    # the spatial equivalent: Cs, then each non-nugget structure again with its joint sill
    eq = cs + [dict(v, sill=j, nugget=0.0) for v, j in zip(cs, joint) if v["vtype"] != "nug"]
    sp = dict(name="s", ndim=3, obs=[dict(coord=xyz, value=val, nmax=nm, maxdist=60.0)], grid=g, vgm=eq,
              search=[dict(ivar=1, sector_search=sector)])
    st = _sm_spec(np.column_stack([xyz, np.full(n, 3.0)]), val, np.column_stack([g, np.full(len(g), 3.0)]), cs,
                  [vt], joint, nm, maxdist=60.0, time_at=10.0, sector=sector, transform="exp", at=4.0,
                  time_nugget=0.2, time_sill=0.8)
    r = run_ref(sp, "analytic")
    o = run_oracle(st, table_mode=False)
    assert np.array_equal(o["nnear"], r["nnear"]) and np.array_equal(o["inear"], r["inear"])
    assert np.max(rel_err(o["estimate"], r["estimate"])) < 1e-10
    assert np.max(rel_err(o["variance"], r["variance"])) < 1e-10


@needs_ref
def test_colocated_reduces_to_1d_reference():
    rng = np.random.default_rng(18)
    nt = 300
    t = rng.uniform(0, 100, nt)
    v = np.sin(t / 7) + rng.normal(0, .1, nt)
    tg = rng.uniform(0, 100, 40)
    at = 500.0  # linear transform below saturation: f(dt) = |dt| / at
    cs = [dict(ivar=1, jvar=1, vtype="sph", nugget=0.05, sill=0.8, a_major=50.0, a_minor1=20.0)]
    vt = dict(vtype="exp", nugget=0.02, sill=0.6, at_k=8.0)
    st = _sm_spec(np.column_stack([np.zeros((nt, 3)), t]), v, np.column_stack([np.zeros((40, 3)), tg]), cs, [vt],
                  [0.4], 12, maxdist=40.0, time_at=2.0, transform="linear", at=at)
    sp = dict(name="s", ndim=1, obs=[dict(coord=t[:, None], value=v, nmax=12, maxdist=40.0 / 2.0)],
              grid=tg[:, None], search=[dict(ivar=1)],
              vgm=[dict(ivar=1, jvar=1, vtype="exp", nugget=0.02, sill=0.6, a_major=8.0),
                   dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.4, a_major=at)])
    r = run_ref(sp, "analytic")
    o = run_oracle(st, table_mode=False)
    assert np.array_equal(o["nnear"], r["nnear"]) and np.array_equal(o["inear"], r["inear"])
    assert np.max(rel_err(o["estimate"], r["estimate"])) < 1e-10
    assert np.max(rel_err(o["variance"], r["variance"])) < 1e-10


def _spatial(vgm, ndim=3):
    z = np.zeros((2, ndim))
    return dict(ndim=ndim, obs=[dict(coord=z, value=np.zeros(2))], grid=z[:1], vgm=vgm, search=[dict(ivar=1)])


ISO2 = [dict(ivar=1, jvar=1, vtype="exp", nugget=0.05, sill=0.6, a_major=50.0),
        dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.3, a_major=120.0)]
ISO1 = ISO2[:1]


@pytest.mark.parametrize("table", [True, False])
@pytest.mark.parametrize("cs", [ISO1, ISO2], ids=["one_table", "per_structure_tables"])
def test_sum_metric_composition(cs, table):
    rng = np.random.default_rng(19)
    lags = np.column_stack([rng.uniform(-150, 150, (600, 3)), rng.uniform(-15, 15, 600)])
    lags[:60, :3] = 0.0   # same place, different times
    lags[60:120, 3] = 0.0  # same time
    lags[120:130] = 0.0    # same place and time
    vt = dict(vtype="gau", nugget=0.03, sill=0.5, at_k=5.0)
    joint = [0.25, 0.15][:len(cs)]
    tr, at, tn, ts = "exp", 7.0, 0.1, 0.9
    st = _sm_spec(np.zeros((2, 4)), np.zeros(2), np.zeros((1, 4)), cs, [vt], joint, 4, transform="bounded", at=at,
                  time_nugget=tn, time_sill=ts)
    c = okrige.cov(st, lags, table_mode=table)
    cs_v = okrige.cov(_spatial(cs), lags[:, :3], table_mode=table)
    ct_v = okrige.cov(_spatial([dict(ivar=1, jvar=1, vtype="gau", nugget=0.03, sill=0.5, a_major=5.0)], 1),
                      np.abs(lags[:, 3:]), table_mode=table)
    dw = _f_time(lags[:, 3], tr, at, tn, ts)
    expect = cs_v + ct_v
    per_structure = table and len(cs) > 1
    for k, v in enumerate(cs):
        hst = np.sqrt((np.linalg.norm(lags[:, :3], axis=1) / v["a_major"]) ** 2 + dw ** 2)
        if per_structure:
            # the structure's own table at h_st (a 1-D unit-range model has h = |lag|), over its sill
            unit = _spatial([dict(v, a_major=1.0, a_minor1=1.0, a_minor2=1.0)], 1)
            expect = expect + joint[k] * okrige.cov(unit, hst[:, None], table_mode=True) / v["sill"]
        else:
            expect = expect + joint[k] * CORE[v["vtype"]](hst)
    np.testing.assert_allclose(c, expect, rtol=1e-12, atol=1e-14)
    # even in the time lag, and the value at the origin is C(0, 0) minus nothing:
    # cov at a zero lag is Cs(0) + Ct(0) + the joint sills (all shapes are 1 at 0)
    flip = lags.copy()
    flip[:, 3] *= -1
    np.testing.assert_array_equal(okrige.cov(st, flip, table_mode=table), c)


def _brute_sm(coord4, val, grid4, cs, vt, joint, tr, at, tn, ts, time_at, nmax, maxdist):
    """dense sum-metric OK per target (analytic shapes, isotropic structures),
    neighbours by a full distance sort in the search metric (x, y, z, t time_at)"""
    def C(a, b):
        lag = a[:, None, :] - b[None, :, :]
        h = np.linalg.norm(lag[..., :3], axis=2)
        dt = lag[..., 3]
        out = np.zeros(h.shape)
        for v in cs:
            out += np.where(h / v["a_major"] > 2.2250738585072014e-305, v["sill"] * CORE[v["vtype"]](h / v["a_major"]),
                            v["sill"] + v["nugget"])
        adt = np.abs(dt) / vt["at_k"]
        out += np.where(adt > 2.2250738585072014e-305, vt["sill"] * CORE[vt["vtype"]](adt), vt["sill"] + vt["nugget"])
        dw = _f_time(dt, tr, at, tn, ts)
        for v, j in zip(cs, joint):
            out += j * CORE[v["vtype"]](np.sqrt((h / v["a_major"]) ** 2 + dw ** 2))
        return out

    cov0 = sum(v["sill"] + v["nugget"] for v in cs) + vt["sill"] + vt["nugget"] + \
        sum(j for v, j in zip(cs, joint) if v["vtype"] != "nug")
    sc = coord4 * np.array([1, 1, 1, time_at])
    est, var, lists = [], [], []
    for g in grid4:
        d2 = np.sum((sc - g * np.array([1, 1, 1, time_at])) ** 2, axis=1)
        idx = np.argsort(d2, kind="stable")[:nmax]
        idx = np.sort(idx[d2[idx] <= maxdist ** 2])
        n = len(idx)
        lists.append(idx + 1)
        if n < 2:  # no system (a block with one neighbour is not solved)
            est.append(np.nan)
            var.append(np.nan)
            continue
        K = np.ones((n + 1, n + 1))
        K[n, n] = 0.0
        K[:n, :n] = C(coord4[idx], coord4[idx])
        K[np.arange(n), np.arange(n)] = cov0
        r = np.ones(n + 1)
        r[:n] = C(coord4[idx], g[None, :])[:, 0]
        x = np.linalg.solve(K, r)
        est.append(x[:n] @ val[idx])
        var.append(max(cov0 - x @ r, 0.0))
    return np.array(est), np.array(var), lists


def test_general_position_against_brute_force():
    rng = np.random.default_rng(20)
    ns, nd = 25, 30
    st_xy = rng.uniform(0, 120, (ns, 2))
    coord4 = np.column_stack([np.repeat(st_xy, nd, axis=0), np.repeat(rng.uniform(0, 5, ns), nd),
                              np.tile(np.arange(nd, dtype=float), ns) + np.repeat(rng.uniform(0, .5, ns), nd)])
    val = np.sin(coord4[:, 0] / 25) + 0.3 * np.cos(coord4[:, 3] / 4) + rng.normal(0, .05, len(coord4))
    grid4 = np.column_stack([rng.uniform(0, 120, (60, 2)), rng.uniform(0, 5, 60), rng.uniform(0, 29, 60)])
    cs = [dict(ivar=1, jvar=1, vtype="nug", nugget=0.05, sill=0.0, a_major=1.0), ISO2[0], ISO2[1]]
    vt = dict(vtype="exp", nugget=0.03, sill=0.5, at_k=6.0)
    joint = [0.5, 0.25, 0.15]
    tr, at, tn, ts = "sph", 12.0, 0.05, 1.5
    time_at, nmax, md = 4.0, 20, 45.0
    spec = _sm_spec(coord4, val, grid4, cs, [vt], joint, nmax, maxdist=md, time_at=time_at, transform=tr, at=at,
                    time_nugget=tn, time_sill=ts)
    o = run_oracle(spec, table_mode=False)
    e, v, lists = _brute_sm(coord4, val, grid4, cs, vt, joint, tr, at, tn, ts, time_at, nmax, md)
    for b in range(len(grid4)):
        assert o["nnear"][b] == len(lists[b])
        np.testing.assert_array_equal(o["inear"][b, :o["nnear"][b]], lists[b])
    ok = o["nnear"] >= 2
    assert ok.sum() > 40
    assert np.max(rel_err(o["estimate"][ok], e[ok])) < 1e-10
    assert np.max(rel_err(o["variance"][ok], v[ok])) < 1e-10


def test_terms_matter():
    """the brute force tells the terms apart: dropping the joint term, the
    temporal transform's nugget or the joint sill of the non-nugget structures
    in C(0, 0) each moves the result"""
    rng = np.random.default_rng(21)
    coord4 = np.column_stack([rng.uniform(0, 100, (200, 3)), rng.uniform(0, 20, 200)])
    val = rng.normal(size=200)
    grid4 = np.column_stack([rng.uniform(0, 100, (20, 3)), rng.uniform(0, 20, 20)])
    vt = dict(vtype="exp", nugget=0.03, sill=0.5, at_k=6.0)
    spec = _sm_spec(coord4, val, grid4, ISO2, [vt], [0.25, 0.15], 12, time_at=3.0, transform="exp", at=5.0,
                    time_nugget=0.1, time_sill=0.9)
    o = run_oracle(spec, table_mode=False)
    base = _brute_sm(coord4, val, grid4, ISO2, vt, [0.25, 0.15], "exp", 5.0, 0.1, 0.9, 3.0, 12, 1e150)
    assert np.max(rel_err(o["estimate"], base[0])) < 1e-10
    for alt in (dict(joint=[0.0, 0.0]), dict(tn=0.0), dict(joint=[0.25, 0.3])):
        a = dict(joint=[0.25, 0.15], tn=0.1)
        a.update(alt)
        e, v, _ = _brute_sm(coord4, val, grid4, ISO2, vt, a["joint"], "exp", 5.0, a["tn"], 0.9, 3.0, 12, 1e150)
        assert np.max(rel_err(o["estimate"], e)) > 1e-6 or np.max(rel_err(o["variance"], v)) > 1e-6


def _kat(obs4, value, grid4, nmax, maxdist=None, time_at=1.0, transform="linear", at=100.0, time_sill=10.0):
    """the reference's probe (tests/test_spacetime_kriging.py:66-84)"""
    vs = dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.8, a_major=500.0, a_minor1=300.0, a_minor2=100.0)
    vt = dict(vtype="exp", nugget=0.0, sill=0.5, at_k=10.0)
    return _sm_spec(np.asarray(obs4, float), np.asarray(value, float), np.asarray(grid4, float), [vs], [vt], [0.3],
                    nmax, maxdist=maxdist, time_at=time_at, transform=transform, at=at, time_sill=time_sill)


@pytest.mark.parametrize("time_at,expect", [(0.1, 10.0), (100.0, 20.0)])
def test_time_at_selects_nearest_pair(time_at, expect):
    """tests/test_spacetime_kriging.py:243-255"""
    obs = [[0, 0, 0, 10], [0, 0, 0, 11], [3, 0, 0, 0], [3.1, 0, 0, 0], [100, 0, 0, 0], [0, 100, 0, 0], [0, 0, 100, 0]]
    o = run_oracle(_kat(obs, [10, 10, 20, 20, -1, -2, -3], [[0, 0, 0, 0]], nmax=2, time_at=time_at))
    assert o["estimate"][0] == pytest.approx(expect)


@pytest.mark.parametrize("time_at,inside", [(50.0, True), (200.0, False)])
def test_time_at_scales_search_radius(time_at, inside):
    """tests/test_spacetime_kriging.py:257-282"""
    o = run_oracle(_kat([[0, 0, 0, .04], [0, 0, 0, .05]], [10, 10], [[0, 0, 0, 0]], nmax=2, maxdist=5.0,
                        time_at=time_at))
    if inside:
        assert o["estimate"][0] == pytest.approx(10.0)
    else:
        assert np.isnan(o["estimate"][0])


def _obs_1var():
    """the reference's fixture shape (tests/test_spacetime_kriging.py:33-44)"""
    rng = np.random.default_rng(42)
    coord = rng.uniform([0, 0, 0], [1000, 1000, 100], size=(200, 3))
    time = rng.uniform(0, 50, size=200)
    value = np.sin(coord[:, 0] / 200) + np.cos(coord[:, 1] / 300) + 0.01 * time + rng.normal(0, 0.1, 200)
    return np.column_stack([coord, time]), value


def test_exact_at_an_observation():
    """tests/test_spacetime_kriging.py:168-189: a target on an observation
    (space and time) returns its value (no spatial nugget)"""
    coord, value = _obs_1var()
    o = run_oracle(_kat(coord, value, coord[[5]], nmax=20, transform="linear", at=10.0, time_sill=1.0))
    assert abs(o["estimate"][0] - value[5]) < 1e-3


@pytest.mark.parametrize("table", [True, False])
def test_bounded_transform_and_maxdist(table):
    """tests/test_spacetime_kriging.py:191-240"""
    coord, value = _obs_1var()
    rng = np.random.default_rng(43)
    grid = np.column_stack([rng.uniform([0, 0, 0], [1000, 1000, 100], size=(50, 3)), rng.uniform(0, 50, 50)])
    wide = run_oracle(_kat(coord, value, grid, nmax=20, transform="bounded", at=10.0, time_sill=1.0),
                      table_mode=table)
    tight = run_oracle(_kat(coord, value, grid, nmax=20, maxdist=1.0, transform="bounded", at=10.0, time_sill=1.0),
                       table_mode=table)
    assert np.all(wide["variance"] >= 0) and np.all(np.isfinite(wide["estimate"]))
    assert np.sum(np.isnan(tight["estimate"])) >= np.sum(np.isnan(wide["estimate"]))
