"""GPU parity: the HIP library (through its C ABI) against the CPU oracle and
the committed reference vectors.

Bars: neighbour lists bit-identical (counts and ascending ids); FP64 path
estimates and variances within 1e-10 relative (floor 1e-4, see
runners.rel_err); FP32-factorisation + FP64-refinement path within 1e-6.
"""
import numpy as np
import pytest

import problems
from runners import make_gpu, max_errors, run_gpu, run_oracle, same_lists

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32R_TOL = 1e-6

CASES = problems.parity_cases()
IDS = [c["name"] for c in CASES]


@pytest.fixture(scope="module")
def ref_vectors():
    return np.load(problems.GOLDEN + "/ref_vectors.npz")


def _ref(vec, key):
    nn = vec[f"{key}/nnear"]
    inn = vec[f"{key}/inear"].astype(np.int32)
    return dict(estimate=vec[f"{key}/estimate"], variance=vec[f"{key}/variance"], nnear=nn, inear=inn,
                stats=vec[f"{key}/stats"])


def _check_counters(g, ref):
    """failures and symmetric-indefinite solves as the reference counts them;
    blocks solved through a Cholesky factorisation (made for them or reused)
    too when no block has a single neighbour (the reference skips those
    solves, kriging_base.F90:2127)"""
    st = g["stats"]
    f, c, s = (int(x) for x in ref["stats"])
    assert st["fail"] == f and st["ssytrf_fact"] == s, (st, ref["stats"])
    if not np.any(ref["nnear"] == 1):
        assert st["chol_fact"] + st["chol_reuse"] == c, (st, ref["stats"])


@pytest.mark.parametrize("spec", CASES, ids=IDS)
@pytest.mark.parametrize("fast", [2, 1, 0], ids=["wave_search", "lane_search", "exact_search"])
def test_fp64_matches_reference_and_oracle(spec, fast, ref_vectors):
    print("case", spec["name"], "fast", fast, flush=True)
    g = run_gpu(spec, fast=fast)
    ref = _ref(ref_vectors, spec["name"])
    assert same_lists(g, ref), "neighbour lists differ from the compiled reference"
    ee, ev = max_errors(g, ref)
    assert ee < FP64_TOL and ev < FP64_TOL, (ee, ev)
    _check_counters(g, ref)
    o = run_oracle(spec)
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < FP64_TOL and ev < FP64_TOL, (ee, ev)


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_fp32_refined_within_1e6(spec, ref_vectors):
    """FP32 factorisation + FP64 refinement within 1e-6 of the reference and
    the oracle; the near-duplicate problem's systems fall below the FP32
    pivot-ratio threshold (1e-3) and are re-solved in FP64"""
    print("case", spec["name"], "fp32r", flush=True)
    g = run_gpu(spec, precision="fp32r")
    ref = _ref(ref_vectors, spec["name"])
    assert same_lists(g, ref)
    ee, ev = max_errors(g, ref)
    assert ee < FP32R_TOL and ev < FP32R_TOL, (ee, ev)
    _check_counters(g, ref)
    o = run_oracle(spec)
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < FP32R_TOL and ev < FP32R_TOL, (ee, ev)
    if spec["name"] == "near_duplicates":
        assert g["timing"]["fp64_resolved"] > 0


def test_failure_without_neglect_error():
    """a system that stays singular after the jittered retry is an error
    unless neglect_error (kriging_base.F90:3369-3383)"""
    k = make_gpu(problems.ladder(), neglect_error=False)
    with pytest.raises(RuntimeError, match="Singular matrix"):
        k.solve()
    k.close()


@pytest.mark.parametrize("spec", problems.all_shapes(), ids=lambda s: s["name"])
@pytest.mark.parametrize("table", [True, False], ids=["table", "analytic"])
def test_variogram_shapes(spec, table):
    g = run_gpu(spec, table_mode=table)
    o = run_oracle(spec, table_mode=table)
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < FP64_TOL and ev < FP64_TOL, (ee, ev)


@pytest.mark.parametrize("ndim", [2, 3])
def test_gauss_legendre_blocks(ndim):
    spec = problems.gq_blocks(ndim=ndim, nb=200 if ndim == 3 else 300)
    g = run_gpu(spec)
    o = run_oracle(spec)
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < FP64_TOL and ev < FP64_TOL, (ee, ev)


def test_closed_forms_1d():
    nug = problems.onedim([0.0, 1.0, 2.0], [1.0, 3.0, 5.0], 10.0,
                          dict(vtype="sph", nugget=2.0, sill=0.0, a_major=1.0))
    g = run_gpu(nug)
    assert g["estimate"][0] == pytest.approx(3.0, rel=1e-12)
    assert g["variance"][0] == pytest.approx(8.0 / 3.0, rel=1e-12)
    sph = problems.onedim([0.0, 2.0], [1.0, 3.0], 1.0, dict(vtype="sph", nugget=0.0, sill=4.0, a_major=2.0))
    g = run_gpu(sph)
    assert g["estimate"][0] == pytest.approx(2.0, rel=1e-12)
    assert g["variance"][0] == pytest.approx(3.5, rel=1e-12)
    x, z = [0.0, 1.5, 4.0], [2.0, 7.0, 4.0]
    ex = problems.onedim(x, z, x, dict(vtype="sph", nugget=0.0, sill=1.0, a_major=6.0))
    g = run_gpu(ex, table_mode=False)
    np.testing.assert_allclose(g["estimate"], z, rtol=1e-10)
    np.testing.assert_allclose(g["variance"], 0.0, atol=1e-10)


@pytest.mark.parametrize("ndim,nmax,maxdist", [(2, 10, None), (3, 20, None), (2, 24, None), (2, 40, None),
                                               (2, 48, None), (2, 60, 600.0), (3, 44, None), (3, 56, None)])
@pytest.mark.parametrize("precision", ["fp64", "fp32r"])
@pytest.mark.parametrize("table", [True, False], ids=["table", "analytic"])
def test_solve_widths(ndim, nmax, maxdist, precision, table):
    """every solve width: the 12-, 24- and 48-row instances (16, 32 and 64
    lanes) and the full-width ones; the maxdist case spreads the neighbour
    counts over every bucket in one solve"""
    kw = dict(maxdist=maxdist) if maxdist else {}
    spec = problems.random_problem(ndim, nobs=800, nq=400, nmax=nmax, seed=40 + nmax, **kw)
    print("case widths", ndim, nmax, maxdist, precision, table, flush=True)
    g = run_gpu(spec, precision=precision, table_mode=table)
    o = run_oracle(spec, table_mode=table)
    assert same_lists(g, o)
    nn = g["nnear"]
    if maxdist:
        assert nn.min() <= 8 and nn.max() > 48
    else:
        assert (nn == nmax).all()
    tol = FP64_TOL if precision == "fp64" else FP32R_TOL
    ee, ev = max_errors(g, o)
    assert ee < tol and ev < tol, (ee, ev)


def test_reference_fixtures_pc2d():
    """tests/test_ordinary_kriging.py:138-170 and test_cross_validation.py:132-141"""
    spec = problems.pc2d()
    g = run_gpu(spec)
    ref = problems.read_csv("grid2d.csv")
    assert np.corrcoef(g["estimate"], ref["estimate"].values)[0, 1] > 0.99
    assert np.all(g["variance"] >= 0)
    ee, ev = max_errors(g, run_oracle(spec))
    assert ee < FP64_TOL and ev < FP64_TOL
    loo = run_gpu(problems.pc2d_loo())
    col = problems.read_csv("pc2d.csv")["loo-cv"].values
    assert np.corrcoef(loo["estimate"], col)[0, 1] > 0.999
    assert np.sqrt(np.mean((loo["estimate"] - col) ** 2)) < 0.01
    assert np.all(loo["variance"] > 0)


def test_no_neighbour_nan_and_error():
    """maxdist below every distance: NaN with neglect_error, error otherwise
    (kriging_base.F90:3198-3207; tests/test_ordinary_kriging.py:236-250)"""
    spec = problems._base("far", 2, np.array([[0.0, 0.0], [1.0, 0.0]]), np.array([1.0, 2.0]),
                          np.array([[100.0, 0.0], [0.5, 0.0]]),
                          [dict(ivar=1, jvar=1, vtype="sph", sill=1.0, a_major=5.0)], nmax=2, maxdist=1.0)
    g = run_gpu(spec)
    assert np.isnan(g["estimate"][0]) and np.isnan(g["variance"][0])
    assert np.isfinite(g["estimate"][1])
    k = make_gpu(spec, neglect_error=False)
    with pytest.raises(RuntimeError, match="not enough neighbors"):
        k.solve()
    k.close()


def test_resident_solve_and_determinism():
    spec = problems.ok3d_slice(step=512)
    k = make_gpu(spec)
    k.solve()
    e1, v1 = k.get_results(copy=True)
    k.solve_resident()
    k.solve_resident()
    k.fetch()
    e2, v2 = k.get_results(copy=True)
    assert np.array_equal(e1, e2) and np.array_equal(v1, v2)
    k.close()


def test_pipeline_orders_agree():
    """the default pipeline (Morton order, two search launches, the first
    chunk's blocks solved while the second is searched), one Morton-ordered
    launch, and the given query order: identical lists and bitwise identical
    results (every block is solved on its own)"""
    from krigekit_amd import synth
    spec = synth.ok3d(scale=1.0 / 16)  # 262,144 nodes, above the two-chunk threshold
    runs = []
    for reorder in (1, 2, 0):
        print("case pipeline order", reorder, flush=True)
        k = make_gpu(spec)
        k.set_search_mode(True, reorder)
        k.solve()
        e, v = k.get_results(copy=True)
        nn, inn = k.neighbours()
        k.close()
        inn = np.where(np.arange(inn.shape[1])[None, :] < nn[:, None], inn, 0)  # valid entries only
        runs.append((e, v, nn, inn))
    e0, v0, nn0, inn0 = runs[0]
    for e, v, nn, inn in runs[1:]:
        assert np.array_equal(nn, nn0) and np.array_equal(inn, inn0)
        assert np.array_equal(e, e0, equal_nan=True) and np.array_equal(v, v0, equal_nan=True)


def test_full_size_3d_octant():
    """4,194,304-node 3-D octant workload: a strided subset of nodes against
    the oracle (lists bit-exact, 1e-10), the FP32+refine path against FP64 at
    every node (1e-6), and size-independent properties (variance >= 0, finite
    estimates, neighbour counts within the octant caps)."""
    from krigekit_amd import synth
    spec = synth.ok3d()
    k64 = make_gpu(spec)
    k64.solve()
    e64, v64 = k64.get_results(copy=True)
    nn, inn = k64.neighbours()
    k64.close()
    assert np.all(np.isfinite(e64)) and np.all(v64 >= 0.0)
    assert nn.min() >= 1 and nn.max() <= 32
    k32 = make_gpu(spec, precision="fp32r")
    k32.solve()
    e32, v32 = k32.get_results(copy=True)
    k32.close()
    ref = dict(estimate=e64, variance=v64, nnear=nn)
    ee, ev = max_errors(dict(estimate=e32, variance=v32), ref)
    assert ee < FP32R_TOL and ev < FP32R_TOL, (ee, ev)
    sub = dict(spec)
    idx = np.arange(0, len(spec["grid"]), 997)
    sub["grid"] = spec["grid"][idx]
    o = run_oracle(sub, nthread=16)
    assert np.array_equal(nn[idx], o["nnear"])
    w = o["inear"].shape[1]
    assert np.array_equal(inn[idx][:, :w], o["inear"])
    ee, ev = max_errors(dict(estimate=e64[idx], variance=v64[idx]), o)
    assert ee < FP64_TOL and ev < FP64_TOL, (ee, ev)


@pytest.mark.parametrize("case", ["ok3d_chunks", "gq2d", "local_params", "maxdist", "ladder", "near_duplicates"])
@pytest.mark.parametrize("precision", ["fp64", "fp32r"])
def test_factor_reuse_is_bitwise_neutral(case, precision):
    """factor reuse over identical neighbour sets (the reference's factor
    cache, kriging_base.F90:1746-1917,2445-2485) gives bitwise the results of
    solving every block on its own; the counters split the Cholesky solves
    into factorisations and reuses"""
    from krigekit_amd import synth
    spec = {"ok3d_chunks": lambda: synth.ok3d(scale=1.0 / 16), "gq2d": lambda: problems.gq_blocks(),
            "local_params": problems.local_params, "maxdist": problems.maxdist_problem,
            "ladder": problems.ladder, "near_duplicates": problems.near_duplicates}[case]()
    out = []
    for mode in (0, 1):
        print("case reuse", case, precision, mode, flush=True)
        k = make_gpu(spec, precision=precision)
        k.set_factor_reuse(mode, 4)
        k.solve()
        e, v = k.get_results(copy=True)
        nn, _ = k.neighbours()
        out.append((e, v, k.solver_stats, nn))
        k.close()
    (e0, v0, s0, nn), (e1, v1, s1, _) = out
    assert np.array_equal(e0, e1, equal_nan=True) and np.array_equal(v0, v1, equal_nan=True)
    assert s1["fail"] == s0["fail"] and s1["ssytrf_fact"] == s0["ssytrf_fact"]
    assert s0["chol_reuse"] == 0
    if not np.any(nn == 1):
        assert s1["chol_fact"] + s1["chol_reuse"] == s0["chol_fact"]
    if case == "ok3d_chunks":
        assert s1["chol_reuse"] > s1["chol_fact"]


def test_factor_reuse_auto_mode_turns_off_for_distinct_sets():
    """auto mode groups the first pass; when fewer than a quarter of the solved
    blocks shared a factorisation, the next pass over the same inputs skips the
    grouping (chol_reuse 0), with bitwise the same results; new inputs group
    again"""
    spec = problems.random_problem(ndim=2, nobs=1500, nq=3000, nmax=12, seed=31)
    k = make_gpu(spec)
    k.solve()
    s0 = k.solver_stats
    e0, v0 = k.get_results(copy=True)
    assert 0 < 4 * s0["chol_reuse"] < s0["chol_fact"] + s0["chol_reuse"], s0
    k.solve()
    s1 = k.solver_stats
    e1, v1 = k.get_results(copy=True)
    assert s1["chol_reuse"] == 0 and s1["chol_fact"] == s0["chol_fact"] + s0["chol_reuse"], (s0, s1)
    assert np.array_equal(e0, e1, equal_nan=True) and np.array_equal(v0, v1, equal_nan=True)
    k.set_grid(spec["grid"][::-1].copy())
    k.solve()
    assert k.solver_stats["chol_reuse"] > 0
    k.close()
