"""The C oracle (oracle/okrige.c) pinned against the reference: its own
known-answer tests and fixtures, textbook identities, brute force, and the
compiled reference engine (oracle/_ref, built from the reference sources)."""
import numpy as np
import pytest

import problems
from oracle import okrige, refdrv
from runners import max_errors, rel_err, run_oracle, run_ref, same_lists

need_ref = pytest.mark.skipif(not refdrv.ref_available("table"),
                              reason="compiled reference oracle/_ref not built (make -C oracle ref)")


# ---- closed forms: reference tests/test_cokriging.py:737-970 and test_api.py:39-70 ----
def _one(spec, **kw):
    o = run_oracle(spec, weights=True, **kw)
    return o["estimate"][0], o["variance"][0], o


@pytest.mark.parametrize("table", [True, False])
def test_pure_nugget_ok(table):
    spec = problems.onedim([0.0, 1.0, 2.0], [1.0, 3.0, 5.0], 10.0,
                           dict(vtype="sph", nugget=2.0, sill=0.0, a_major=1.0))
    est, var, _ = _one(spec, table_mode=table)
    assert est == pytest.approx(3.0, rel=1e-12)
    assert var == pytest.approx(2.0 * (1.0 + 1.0 / 3.0), rel=1e-12)


@pytest.mark.parametrize("table", [True, False])
def test_symmetric_spherical_system(table):
    # K = [[4,0],[0,4]], F = 1, b = [1.25, 1.25] -> w = [0.5, 0.5], mu = -0.75
    spec = problems.onedim([0.0, 2.0], [1.0, 3.0], 1.0,
                           dict(vtype="sph", nugget=0.0, sill=4.0, a_major=2.0))
    est, var, o = _one(spec, table_mode=table)
    assert est == pytest.approx(2.0, rel=1e-12)
    assert var == pytest.approx(3.5, rel=1e-12)
    np.testing.assert_allclose(o["weights"][0, :2], [0.5, 0.5], rtol=1e-12)


def test_exact_interpolation():
    x, z = [0.0, 1.5, 4.0], [2.0, 7.0, 4.0]
    spec = problems.onedim(x, z, x, dict(vtype="sph", nugget=0.0, sill=1.0, a_major=6.0))
    o = run_oracle(spec, table_mode=False)
    np.testing.assert_allclose(o["estimate"], z, rtol=1e-10)
    np.testing.assert_allclose(o["variance"], 0.0, atol=1e-10)


def test_constant_field_and_weights_sum_to_one():
    spec = problems.random_problem(nq=400, seed=3)
    spec["obs"][0]["value"] = np.full(len(spec["obs"][0]["value"]), 2.71828)
    o = run_oracle(spec, weights=True)
    np.testing.assert_allclose(o["estimate"], 2.71828, rtol=1e-11)
    s = np.array([o["weights"][b, :o["nnear"][b]].sum() for b in range(len(o["nnear"]))])
    np.testing.assert_allclose(s, 1.0, rtol=1e-11)
    assert np.all(o["variance"] >= 0.0)


# ---- reference fixtures ----------------------------------------------------
def test_pc2d_grid_regression():
    """tests/test_ordinary_kriging.py:138-170: corr(est, grid2d.csv) > 0.99"""
    spec = problems.pc2d()
    o = run_oracle(spec)
    ref = problems.read_csv("grid2d.csv")
    assert np.corrcoef(o["estimate"], ref["estimate"].values)[0, 1] > 0.99
    assert np.all(o["variance"] >= 0.0)
    assert np.corrcoef(o["variance"], ref["variance"].values)[0, 1] > 0.99


def test_pc2d_loo_reference_column():
    """tests/test_cross_validation.py:132-141: corr > 0.999, RMSE < 0.01"""
    o = run_oracle(problems.pc2d_loo())
    ref = problems.read_csv("pc2d.csv")["loo-cv"].values
    assert np.corrcoef(o["estimate"], ref)[0, 1] > 0.999
    assert np.sqrt(np.mean((o["estimate"] - ref) ** 2)) < 0.01
    assert np.all(o["variance"] > 0)


def test_gq_layout_matches_fixture():
    """tests/test_bk.py:185-212: block_type -4 nodes/weights == gridblockpnt2d.csv"""
    pnt = problems.read_csv("gridblockpnt2d.csv")
    blk = problems.read_csv("gridblock2d.csv")
    blk.columns = [c.strip().lstrip("#").strip() for c in blk.columns]
    centre = np.array([[blk["x"].iloc[0], blk["y"].iloc[0]]])
    sub, w = problems.gq_subpoints(centre, np.array([2000.0, 2000.0]))
    fx = pnt[["x", "y"]].values
    fw = pnt["weight"].values
    for (x, y), ww in zip(sub, w):
        d = np.hypot(fx[:, 0] - x, fx[:, 1] - y)
        j = int(np.argmin(d))
        assert d[j] < 0.01  # fixture coordinates are printed to 1 mm
        assert fw[j] == pytest.approx(ww, abs=1e-6)  # printed to 6 decimals


@pytest.mark.parametrize("ndim", [2, 3])
def test_sector_exact_sets(ndim):
    """tests/test_sector_search.py:20-110: one neighbour per quadrant/octant"""
    if ndim == 2:
        obs = np.array([[0.1, 0.1], [0.2, 0.2], [-5.0, 5.0], [-5.0, -5.0], [5.0, -5.0]])
        expect = [1, 3, 4, 5]
    else:
        signs = [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, 1, -1), (-1, 1, -1),
                 (1, -1, -1), (-1, -1, -1)]
        obs = [[0.1, 0.1, 0.1], [0.2, 0.2, 0.2]] + [[5.0 * a, 5.0 * b, 5.0 * c] for a, b, c in signs[1:]]
        obs = np.array(obs)
        expect = [1] + list(range(3, 10))
    spec = dict(name="sector", ndim=ndim, obs=[dict(coord=obs, value=np.arange(len(obs), dtype=float), nmax=1)],
                grid=np.zeros((1, ndim)), vgm=[dict(ivar=1, jvar=1, vtype="sph", sill=1.0, a_major=100.0)],
                search=[dict(ivar=1, sector_search=True)])
    o = run_oracle(spec)
    assert list(o["inear"][0, :o["nnear"][0]]) == expect


# ---- kd-tree against brute force (distinct distances) -----------------------
@pytest.mark.parametrize("ndim,k", [(2, 1), (2, 16), (3, 40), (3, 128)])
def test_knn_brute_force(ndim, k):
    rng = np.random.default_rng(ndim * 100 + k)
    pts = rng.uniform(0, 1000, (3000, ndim))
    for q in rng.uniform(-100, 1100, (25, ndim)):
        idx, dis = okrige.knn(pts, q, k, sort=True)
        d = ((pts - q) ** 2).sum(axis=1)
        order = np.argsort(d, kind="stable")[:k]
        assert set(idx - 1) == set(order)
        np.testing.assert_array_equal(dis, d[order])
        assert np.all(np.diff(dis) >= 0)


# ---- covariance: closed forms and the table's documented behaviour -----------
SHAPES = {
    "sph": lambda r: np.where(r < 1, 1 - 1.5 * r + 0.5 * r ** 3, 0.0),
    "exp": lambda r: np.exp(-3 * r),
    "gau": lambda r: np.exp(-3.0625 * r ** 2),
    "lin": lambda r: np.where(r < 1, 1 - r, 0.0),
    "bsq": lambda r: np.where(r < 1, (1 - r ** 2) ** 2, 0.0),
    "dco": lambda r: np.exp(-3 * r) * np.cos(np.pi * r),
}


@pytest.mark.parametrize("vt", sorted(SHAPES))
def test_covariance_shapes(vt):
    spec = dict(ndim=2, obs=[dict(coord=np.zeros((1, 2)), value=[0.0])], grid=np.zeros((1, 2)),
                vgm=[dict(ivar=1, jvar=1, vtype=vt, nugget=0.3, sill=2.0, a_major=100.0, a_minor1=50.0,
                          azimuth=30.0)], search=[dict(ivar=1)])
    rng = np.random.default_rng(7)
    lags = rng.uniform(-250, 250, (4000, 2))
    # major axis along azimuth 30 deg clockwise from north (variogram.f90:185, rotation.f90:51)
    az = np.deg2rad(30.0)
    major = lags @ np.array([np.sin(az), np.cos(az)])
    minor = lags @ np.array([np.cos(az), -np.sin(az)])
    r = np.sqrt((major / 100.0) ** 2 + (minor / 50.0) ** 2)
    exact = 2.0 * SHAPES[vt](r)
    ana = okrige.cov(spec, lags, table_mode=False)
    np.testing.assert_allclose(ana, exact, rtol=1e-12, atol=1e-14)
    tab = okrige.cov(spec, lags, table_mode=True)
    # table mode = piecewise-linear interpolation on the zone grid
    # h_bounds (0, .1, .5, 3.5), steps (1e-5, 1e-4, 1e-3) (kriging.F90:416-417),
    # entry 0 = C(0) with the nugget (variogram.f90:618); numpy.interp on that grid
    grid = np.concatenate([[0.0], 0.0 + np.arange(1, 10001) * (0.1 / 10000),
                           0.1 + np.arange(1, 4001) * (0.4 / 4000), 0.5 + np.arange(1, 3001) * (3.0 / 3000)])
    vals = 2.0 * SHAPES[vt](grid)
    vals[0] = 2.3
    tail = r >= 3.5
    np.testing.assert_allclose(tab[~tail], np.interp(r[~tail], grid, vals), rtol=1e-9, atol=1e-12)
    if vt in ("exp", "gau", "dco"):  # analytic tail beyond h = 3.5
        np.testing.assert_allclose(tab[tail], exact[tail], rtol=1e-12)
    else:
        assert np.all(tab[tail] == 0.0)
    # C(0) includes the nugget; lags below the first table step interpolate towards it
    z = okrige.cov(spec, np.array([[0.0, 0.0], [1e-4, 0.0]]), table_mode=True)
    assert z[0] == pytest.approx(2.3)
    assert 2.0 * SHAPES[vt](np.array([1e-6]))[0] < z[1] < 2.3


# ---- against the compiled reference engine ----------------------------------
@need_ref
@pytest.mark.parametrize("spec", problems.parity_cases() + problems.all_shapes() + [problems.gq_blocks()],
                         ids=lambda s: s["name"])
def test_oracle_matches_compiled_reference(spec):
    r = run_ref(spec)
    o = run_oracle(spec)
    assert same_lists(o, r), "neighbour lists differ from the reference"
    # one-neighbour blocks: the reference's estimate reuses another block's
    # weights (solve skipped, kriging_base.F90:2127); compared: variance only
    one = r["nnear"] == 1
    assert np.max(rel_err(o["estimate"][~one], r["estimate"][~one]), initial=0) < 1e-10
    assert np.max(rel_err(o["variance"], r["variance"])) < 1e-10
    assert np.array_equal(np.isnan(o["variance"]), np.isnan(r["variance"]))
    assert np.all(np.isnan(o["estimate"][one]))


@need_ref
def test_oracle_matches_reference_analytic_mode():
    spec = problems.random_problem(seed=61)
    r = run_ref(spec, mode="analytic")
    o = run_oracle(spec, table_mode=False)
    assert same_lists(o, r)
    assert np.max(rel_err(o["estimate"], r["estimate"])) < 1e-10
    assert np.max(rel_err(o["variance"], r["variance"])) < 1e-10


def test_committed_reference_vectors():
    """the committed reference vectors agree with the oracle (the same
    comparison the GPU tests make against both)"""
    vec = np.load(problems.GOLDEN + "/ref_vectors.npz")
    for spec in problems.parity_cases():
        key = spec["name"]
        o = run_oracle(spec)
        nn = vec[f"{key}/nnear"]
        inn = vec[f"{key}/inear"].astype(np.int32)
        ref = dict(estimate=vec[f"{key}/estimate"], variance=vec[f"{key}/variance"], nnear=nn, inear=inn)
        assert same_lists(o, ref), key
        ee, ev = max_errors(o, ref)
        assert ee < 1e-10 and ev < 1e-10, key
        # failures, Cholesky solves and symmetric-indefinite solves
        st = o["stats"]
        got = [st["fail"], st["chol_fact"] + st["chol_reuse"], st["ssytrf_fact"] + st["ssytrf_reuse"]]
        assert got == list(vec[f"{key}/stats"]), (key, got, list(vec[f"{key}/stats"]))


def test_committed_reference_weights():
    """the reference's stored weights (krige_get_weight_data after a solve
    with store_weight, sampled blocks of two or more neighbours) agree with
    the oracle's weights"""
    vec = np.load(problems.GOLDEN + "/ref_vectors.npz")
    n = 0
    for spec in problems.parity_cases():
        key = spec["name"]
        if f"{key}/weight" not in vec.files:
            continue
        o = run_oracle(spec, weights=True)
        sel = vec[f"{key}/wblocks"]
        wr = vec[f"{key}/weight"]
        wo = np.where(np.arange(o["weights"].shape[1])[None, :] < o["nnear"][sel, None],
                      o["weights"][sel], 0.0)[:, :wr.shape[1]]
        d = np.abs(wo - wr) / np.maximum(1.0, np.abs(wr))
        assert d.max() < 1e-9, (key, d.max())
        n += 1
    assert n >= 10


def test_failure_ladder_is_reached():
    """the ladder problem reaches every step of the reference's fallback
    (kriging_base.F90:3244-3398): Cholesky, the symmetric-indefinite
    factorisation, its jittered retry, and failure"""
    spec = problems.ladder()
    o = run_oracle(spec)
    v = spec["obs"][0]["variance"]
    npos = np.array([(v[o["inear"][b, :o["nnear"][b]] - 1] > 0).sum() for b in range(len(o["nnear"]))])
    assert set(np.unique(npos)) >= {0, 1, 3, 4}
    assert np.all(np.isnan(o["estimate"][npos == 0])) and np.all(o["variance"][npos == 0] == 0.0)
    assert np.all(np.isfinite(o["estimate"][npos > 0]))
    assert o["stats"]["fail"] == (npos == 0).sum()
    assert o["stats"]["ssytrf_fact"] == ((npos > 0) & (npos < 4)).sum()
    with pytest.raises(RuntimeError):
        okrige.solve(spec, neglect_error=False)


def test_bounds_clip_estimates():
    spec = problems.bounds_case()
    o = run_oracle(spec)
    lo, hi = spec["bounds"]
    assert (o["estimate"] == lo).sum() > 50 and (o["estimate"] == hi).sum() > 50
    free = dict(spec)
    free.pop("bounds")
    f = run_oracle(free)
    np.testing.assert_array_equal(o["estimate"], np.clip(f["estimate"], lo, hi))


def test_planted_mistake_is_caught():
    """a wrong sign in the Lagrange term shifts the variance: the closed form catches it"""
    spec = problems.onedim([0.0, 2.0], [1.0, 3.0], 1.0,
                           dict(vtype="sph", nugget=0.0, sill=4.0, a_major=2.0))
    o = run_oracle(spec, weights=True)
    w = o["weights"][0, :2]
    b = np.array([1.25, 1.25])
    mu = -0.75
    wrong = 4.0 - (w @ b) + mu  # sign of the Lagrange term flipped
    assert o["variance"][0] == pytest.approx(3.5)
    assert abs(wrong - 3.5) > 1.0
