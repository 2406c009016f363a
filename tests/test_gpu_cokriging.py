"""GPU multi-variable (co)kriging (one workgroup per block) against the
compiled reference's committed vectors and the numpy oracle: the FP64 path at
1e-10, the FP32 + MFMA + FP64-refinement path at 1e-6."""
import numpy as np
import pytest

import problems
from oracle import ckrige
from runners import ck_errors, run_gpu_ck

pytestmark = pytest.mark.gpu

CASES = problems.ck_cases()
IDS = [c["name"] for c in CASES]


@pytest.fixture(scope="module")
def vectors():
    return np.load(problems.GOLDEN + "/ck_vectors.npz")


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_cokriging_matches_reference(spec, mode, vectors):
    print("case", spec["name"], mode, flush=True)
    g = run_gpu_ck(spec, table_mode=(mode == "table"))
    ref = dict(estimate=vectors[f"{spec['name']}/{mode}/estimate"],
               variance=vectors[f"{spec['name']}/{mode}/variance"])
    ee, ev = ck_errors(g, ref)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)
    # failures and symmetric-indefinite solves as the reference counts them
    f, c, sy = (int(x) for x in vectors[f"{spec['name']}/{mode}/stats"])
    st = g["stats"]
    assert (st["fail"], st["chol_fact"], st["ssytrf_fact"]) == (f, c, sy), (st, (f, c, sy))


def test_cokriging_failure_without_neglect_error():
    """a variable without neighbours under std_ck leaves the Schur complement
    singular and the fallback fails: an error unless neglect_error
    (kriging_base.F90:3369-3383)"""
    from runners import make_gpu_ck
    k = make_gpu_ck(problems.ck_empty_var(True), neglect_error=False)
    with pytest.raises(RuntimeError, match="Singular matrix"):
        k.solve()
    k.close()


def test_full_size_cokriging():
    """2 variables, 20,000 + 200,000 observations, 10^6 blocks of 4 x 4
    Gauss-Legendre nodes, 24 + 48 neighbours, std_ck: a strided subset of
    blocks against the oracle (Cholesky and Schur complement, as the
    reference), and size-independent properties."""
    from krigekit_amd import synth
    spec = synth.ck2d()
    print("ck2d full", len(spec["grid"]), flush=True)
    g = run_gpu_ck(spec)
    e, v = g["estimate"], g["variance"]
    assert e.shape == (len(spec["grid"]), 2) and v.shape == (len(spec["grid"]), 2, 2)
    assert np.all(np.isfinite(e)) and np.all(v >= 0.0)
    np.testing.assert_array_equal(v[:, 0, 1], v[:, 1, 0])
    idx = np.arange(0, len(spec["grid"]), 4999)
    sub = dict(spec)
    sub["grid"] = spec["grid"][idx]
    sub["sub_coord"] = spec["sub_coord"].reshape(-1, 16, 2)[idx].reshape(-1, 2)
    sub["sub_weight"] = spec["sub_weight"].reshape(-1, 16)[idx].ravel()
    o = ckrige.solve(sub)
    ee, ev = ck_errors(dict(estimate=e[idx], variance=v[idx]), o)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("nmax", [(23, 46), (22, 44), (21, 41)], ids=["69rows", "66rows", "62rows"])
def test_cokriging_row_counts(nmax):
    """data row counts that are not multiples of the four-pivot panel (the
    last panel holds 1-3 pivots and the extra columns start inside it),
    against the oracle"""
    spec = problems.ck_small(True, True, nb=120, seed=11)
    for o, m in zip(spec["obs"], nmax):
        o["nmax"] = m
    print("case ck rows", nmax, flush=True)
    g = run_gpu_ck(spec)
    o = ckrige.solve(spec)
    ee, ev = ck_errors(g, o)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("gq", [True, False], ids=["blocks", "points"])
def test_cokriging_range_scale_and_local_nugget(gq):
    """per-block range scale (lags divided by it in the matrix and the
    right-hand sides, not in the block base covariance) and local nugget
    (added to every data diagonal), against the oracle"""
    spec = problems.ck_small(True, gq, nb=150, seed=13)
    rng = np.random.default_rng(14)
    spec["rangescale"] = rng.uniform(0.7, 1.4, len(spec["grid"]))
    spec["localnugget"] = rng.uniform(0.0, 0.1, len(spec["grid"]))
    print("case ck rangescale", gq, flush=True)
    g = run_gpu_ck(spec)
    o = ckrige.solve(spec)
    ee, ev = ck_errors(g, o)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_cokriging_fp32r_matches_reference(spec, mode, vectors):
    """FP32 factorisation with the MFMA trailing update and FP64 refinement
    (systems up to 80 columns; the 128-column instance solves in FP64) against
    the reference vectors at 1e-6, with the reference's failure and
    symmetric-indefinite counts"""
    print("case fp32r", spec["name"], mode, flush=True)
    g = run_gpu_ck(spec, table_mode=(mode == "table"), precision="fp32r")
    ref = dict(estimate=vectors[f"{spec['name']}/{mode}/estimate"],
               variance=vectors[f"{spec['name']}/{mode}/variance"])
    ee, ev = ck_errors(g, ref)
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)
    f, c, sy = (int(x) for x in vectors[f"{spec['name']}/{mode}/stats"])
    st = g["stats"]
    assert (st["fail"], st["chol_fact"], st["ssytrf_fact"]) == (f, c, sy), (st, (f, c, sy))


def test_cokriging_fp32r_resolves_small_pivots():
    """near-duplicate observations without a nugget: blocks whose FP32 pivot
    ratio falls below pivot_tol (1e-3) are solved again in FP64; every block
    within 1e-6 of the FP64 path, which stays within 1e-8 of the oracle"""
    spec = problems.ck_near_duplicates()
    print("case ck near duplicates", flush=True)
    g = run_gpu_ck(spec, precision="fp32r")
    f = run_gpu_ck(spec)
    assert g["timing"]["fp64_resolved"] > 0
    ee, ev = ck_errors(g, f)
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)
    o = ckrige.solve(spec)
    ee, ev = ck_errors(f, o)
    assert ee < 1e-8 and ev < 1e-8, (ee, ev)


def test_full_size_cokriging_fp32r():
    """the 10^6-block workload in FP32 + MFMA + refinement: within 1e-6 of the
    FP64 path everywhere"""
    from krigekit_amd import synth
    spec = synth.ck2d()
    print("ck2d full fp32r", len(spec["grid"]), flush=True)
    g = run_gpu_ck(spec, precision="fp32r")
    f = run_gpu_ck(spec)
    assert np.all(np.isfinite(g["estimate"])) and np.all(g["variance"] >= 0.0)
    ee, ev = ck_errors(g, f)
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)
