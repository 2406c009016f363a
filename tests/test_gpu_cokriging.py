"""GPU multi-variable (co)kriging (one workgroup per block) against the
compiled reference's committed vectors and the numpy oracle."""
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


def test_full_size_cokriging():
    """2 variables, 20,000 + 200,000 observations, 10^6 blocks of 4 x 4
    Gauss-Legendre nodes, 24 + 48 neighbours, std_ck: a strided subset of
    blocks against the oracle, and size-independent properties.  The oracle
    solves the 74-unknown bordered systems by LU (numpy), the library by
    Cholesky and Schur complement, as the reference does: their rounding
    differs at the 1e-14 level, which the relative metric (floor 1e-4) turns
    into up to ~2e-10 on estimates near zero, hence 1e-9 here (the reference's
    own results are matched at 1e-10 above)."""
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
    assert ee < 1e-9 and ev < 1e-9, (ee, ev)


def test_indefinite_coregionalisation_fails_the_block():
    """a cross sill above the geometric mean of the sills, with collocated
    data, makes the block's system indefinite: the Cholesky meets a negative
    pivot (the reference's spotrf returns info > 0, solver.f90:224-243), the
    block's results are NaN, and without neglect_error the solve reports a
    singular matrix.  A block elsewhere, whose variables' data lie beyond the
    range of each other, is solved as usual."""
    a = np.array([[0.0, 0.0], [30.0, 0.0], [0.0, 30.0]])
    b1 = a + [10000.0, 0.0]
    b2 = b1 + [0.0, 500.0]
    x1, x2 = np.vstack([a, b1]), np.vstack([a, b2])
    spec = dict(ndim=2, nvar=2, std_ck=True, unbias=1,
                obs=[dict(coord=x1, value=np.arange(6.0), nmax=3), dict(coord=x2, value=np.arange(6.0)[::-1], nmax=3)],
                grid=np.array([[10.0, 10.0], [10010.0, 250.0]]),
                vgm=[dict(ivar=i, jvar=j, vtype="sph", sill=s, a_major=100.0) for (i, j), s in
                     (((1, 1), 1.0), ((1, 2), 2.0), ((2, 2), 1.0))],
                search=[dict(ivar=1, sector_search=False), dict(ivar=2, sector_search=False)])
    print("case indefinite LMC", flush=True)
    g = run_gpu_ck(spec)
    assert np.all(np.isnan(g["estimate"][0])) and np.all(np.isnan(g["variance"][0]))
    assert np.all(np.isfinite(g["estimate"][1])) and np.all(g["variance"][1] >= 0.0)
    from runners import make_gpu_ck
    k = make_gpu_ck(spec, neglect_error=False)
    with pytest.raises(RuntimeError, match="Singular matrix"):
        k.solve()
    k.close()


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
    assert ee < 1e-9 and ev < 1e-9, (ee, ev)


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
    assert ee < 1e-9 and ev < 1e-9, (ee, ev)
