"""The numpy cokriging restatement (oracle/ckrige.py) pinned on the CPU
against the compiled reference's results: the committed vectors
(tests/golden/ck_vectors.npz) always, the live compiled reference when
oracle/_ref is built."""
import numpy as np
import pytest

import problems
from oracle import ckrige
from runners import ck_errors

CASES = problems.ck_cases()
IDS = [c["name"] for c in CASES]


@pytest.fixture(scope="module")
def vectors():
    return np.load(problems.GOLDEN + "/ck_vectors.npz")


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_oracle_matches_reference_vectors(spec, mode, vectors):
    o = ckrige.solve(spec, table_mode=(mode == "table"))
    ref = dict(estimate=vectors[f"{spec['name']}/{mode}/estimate"],
               variance=vectors[f"{spec['name']}/{mode}/variance"])
    ee, ev = ck_errors(o, ref)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


def test_oracle_matches_live_reference():
    from oracle import refdrv
    if not refdrv.ref_available("table"):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    from runners import run_ref_ck
    spec = problems.ck_small(True, True, nb=60, seed=3)
    ee, ev = ck_errors(ckrige.solve(spec), run_ref_ck(spec))
    assert ee < 1e-10 and ev < 1e-10


def test_table_rule_matters():
    """cross blocks of K are analytic in table mode; evaluating them from the
    tables instead (a planted mistake) must break the agreement"""
    from oracle import okrige
    spec = problems.ck_small(True, False, nb=40, seed=5)
    v = np.load(problems.GOLDEN + "/ck_vectors.npz")
    orig = okrige.cov

    def all_tables(s, lags, table_mode=True):
        return orig(s, lags, table_mode=True)

    okrige.cov = all_tables
    try:
        o = ckrige.solve(spec, table_mode=True)
    finally:
        okrige.cov = orig
    ref = ckrige.solve(spec, table_mode=True)
    assert np.max(np.abs(o["estimate"] - ref["estimate"])) > 1e-9


def test_point_cokriging_closed_form():
    """one neighbour per variable, std_ck: the constraints fix both weights,
    so the primary estimate is the primary value and the secondary weight 0"""
    rng = np.random.default_rng(0)
    spec = dict(name="ck_closed", ndim=2, nvar=2, std_ck=True, unbias=1,
                obs=[dict(coord=np.array([[0.0, 0.0], [500.0, 0.0]]), value=np.array([2.0, 5.0]), nmax=1),
                     dict(coord=np.array([[10.0, 0.0], [480.0, 0.0]]), value=np.array([-1.0, 3.0]), nmax=1)],
                grid=np.array([[5.0, 1.0]]),
                vgm=[dict(ivar=1, jvar=1, vtype="sph", sill=1.0, a_major=300.0, nugget=0.1),
                     dict(ivar=1, jvar=2, vtype="sph", sill=0.5, a_major=300.0),
                     dict(ivar=2, jvar=2, vtype="sph", sill=1.0, a_major=300.0, nugget=0.1)],
                search=[dict(ivar=1, sector_search=False), dict(ivar=2, sector_search=False)])
    o = ckrige.solve(spec, table_mode=False)
    assert o["estimate"][0, 0] == pytest.approx(2.0, rel=1e-12)
    assert o["estimate"][0, 1] == pytest.approx(-1.0, rel=1e-12)
    del rng


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_oracle_weights_match_reference_store(spec, vectors):
    """the weights of the reference's store (krige_get_weight_data after a
    solve with store_weight) on sampled blocks, and its neighbour counts,
    against the oracle's solution of the bordered system"""
    key = spec["name"]
    o = ckrige.solve(spec, want_weights=True)
    # a block the fallback cannot solve keeps one neighbour in every group
    # (ctx%nnear = 1, kriging_base.F90:3369-3383)
    nn = np.array(o["nnear"])
    failed = np.isnan(o["estimate"]).any(axis=1) & (nn.sum(axis=1) >= 2)
    nn[failed] = 1
    np.testing.assert_array_equal(nn, vectors[f"{key}/nnear"])
    sel = vectors[f"{key}/wblocks"]
    wr = vectors[f"{key}/weight"]
    wo = o["weights"][sel]
    k = min(wr.shape[-1], wo.shape[-1])
    fin = np.isfinite(wr).all(axis=(1, 2, 3)) & np.isfinite(o["estimate"][sel]).all(axis=1)
    assert fin.sum() > 0
    d = np.abs(wo[fin][..., :k] - wr[fin][..., :k]) / np.maximum(1.0, np.abs(wr[fin][..., :k]))
    assert d.max() < 1e-9, d.max()



DRIFT = problems.drift_cases()


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", DRIFT, ids=[c["name"] for c in DRIFT])
def test_drift_oracle_matches_compiled_reference(spec, mode):
    """external drift (KED, UK on blocks, drift without unbiasedness, drift
    in cokriging) and weight correction: the numpy oracle against the
    compiled reference's committed vectors"""
    v = np.load(problems.GOLDEN + "/drift_vectors.npz")
    o = ckrige.solve(spec, table_mode=(mode == "table"))
    ref = dict(estimate=v[f"{spec['name']}/{mode}/estimate"], variance=v[f"{spec['name']}/{mode}/variance"])
    got = dict(estimate=o["estimate"].reshape(ref["estimate"].shape),
               variance=o["variance"].reshape(ref["variance"].shape))
    ee, ev = ck_errors(got, ref)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("name", ["wcorr_ok", "drift_ked"])
def test_drift_and_correction_terms_matter(name):
    """the committed vectors exercise what they pin: without the weight
    correction (or without the drift rows) the oracle misses them by far more
    than the bar"""
    spec = {c["name"]: c for c in DRIFT}[name]
    v = np.load(problems.GOLDEN + "/drift_vectors.npz")
    off = dict(spec)
    if name == "wcorr_ok":
        off["weight_correction"] = False
    else:
        off["ndrift"] = 0
        off["grid_drift"] = None
        off["obs"] = [dict(o, drift=None) for o in spec["obs"]]
    o = ckrige.solve(off)
    ref = dict(estimate=v[f"{name}/table/estimate"], variance=v[f"{name}/table/variance"])
    got = dict(estimate=o["estimate"].reshape(ref["estimate"].shape),
               variance=o["variance"].reshape(ref["variance"].shape))
    ee, ev = ck_errors(got, ref)
    assert max(ee, ev) > 1e-3, (ee, ev)


GRAD = problems.grad_cases()


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", GRAD, ids=[c["name"] for c in GRAD])
def test_grad_oracle_matches_compiled_reference(spec, mode):
    """increment (gradient-pair) constraints: the numpy oracle against the
    compiled reference's committed vectors"""
    v = np.load(problems.GOLDEN + "/grad_vectors.npz")
    o = ckrige.solve(spec, table_mode=(mode == "table"))
    ref = dict(estimate=v[f"{spec['name']}/{mode}/estimate"], variance=v[f"{spec['name']}/{mode}/variance"])
    got = dict(estimate=o["estimate"].reshape(ref["estimate"].shape),
               variance=o["variance"].reshape(ref["variance"].shape))
    ee, ev = ck_errors(got, ref)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("change", ["drop", "cross_zero", "sign", "no_var"])
def test_grad_terms_matter(change):
    """the vectors tell the increment terms apart: dropping the increments,
    filling the increment-increment block across variables, flipping the
    pair order of the covariance, or dropping the increment variance each
    misses them by far more than the bar"""
    name = "grad_ck" if change == "cross_zero" else "grad_ok"
    spec = {c["name"]: c for c in GRAD}[name]
    v = np.load(problems.GOLDEN + "/grad_vectors.npz")
    if change == "drop":
        off = dict(spec)
        off.pop("grad")
        o = ckrige.solve(off)
    elif change == "no_var":
        off = dict(spec)
        off["grad"] = [dict(g, variance=None) for g in spec["grad"]]
        o = ckrige.solve(off)
    else:
        src = open(ckrige.__file__).read()
        if change == "cross_zero":
            # the block of one variable's increments, filled across variables too
            src = src.replace("K[np.ix_(no + ga, no + ga)] = blk", "K[np.ix_(no + ga, no + ga)] = blk\n"
                              "            K[no:n, no:n] = np.where(K[no:n, no:n] == 0.0, 0.05, K[no:n, no:n])")
        else:
            src = src.replace("c = (cov(g, i, l1) - cov(g, i, l2))", "c = (cov(g, i, l1) + cov(g, i, l2))")
        ns = {}
        exec(compile(src, "oracle/ckrige.py (changed)", "exec"), ns)
        o = ns["solve"](spec)
    ref = dict(estimate=v[f"{name}/table/estimate"], variance=v[f"{name}/table/variance"])
    got = dict(estimate=o["estimate"].reshape(ref["estimate"].shape),
               variance=o["variance"].reshape(ref["variance"].shape))
    ee, ev = ck_errors(got, ref)
    assert max(ee, ev) > 1e-4, (ee, ev)
