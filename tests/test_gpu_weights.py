"""The weight store on the GPU (store_weight, get_weights, set_weights,
use_old_weight, update_obs_value, the text factor file) against the compiled
reference's stored weights (tests/golden/ref_vectors.npz, ck_vectors.npz:
krige_get_weight_data after a solve with store_weight) and the reference's
own round-trip properties (tests/test_weight_store.py and
tests/test_update_obs_value.py of the reference)."""
import os

import numpy as np
import pytest

import problems
from runners import make_gpu, make_gpu_ck, rel_err

pytestmark = pytest.mark.gpu

CASES = [c for c in problems.parity_cases()
         if c["name"] in ("ok2d", "lattice2d_s1_cv0", "random3d_octant", "simple_kriging",
                          "rangescale_localnugget", "maxdist", "ladder", "near_duplicates", "big80_2d",
                          "loo2d", "bounds")]
IDS = [c["name"] for c in CASES]
CK = [c for c in problems.ck_cases() if c["name"] != "ck_indefinite"]
CK_IDS = [c["name"] for c in CK]


@pytest.fixture(scope="module")
def vec():
    return np.load(problems.GOLDEN + "/ref_vectors.npz")


@pytest.fixture(scope="module")
def ckvec():
    return np.load(problems.GOLDEN + "/ck_vectors.npz")


def _estimate_from_store(w, values):
    """sum_k w(k) z(inear(k)) per block (single variable, unbiased)"""
    nb, _, nm = w["inear"].shape
    idx = np.maximum(w["inear"][:, 0, :] - 1, 0)
    m = np.arange(nm)[None, :] < w["nnear"][:, 0][:, None]
    return np.where(m, w["weight"][:, 0, :] * values[idx], 0.0).sum(axis=1)


def _solve_store(spec, **kw):
    k = make_gpu(spec, store_weight=True, **kw)
    k.solve()
    e, v = k.get_results(copy=True)
    w = k.get_weights()
    return k, e, v, w


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_store_matches_reference(spec, vec):
    """neighbour counts and lists bitwise, weights within 1e-9 (relative,
    floor 1) of the reference's store; the stored variance is the solve's;
    the weights reproduce the solve's estimates"""
    key = spec["name"]
    print("case", key, flush=True)
    k, e, v, w = _solve_store(spec)
    k.close()
    nb = len(e)
    assert w["nnear"].shape == (nb, 1) and w["weight"].shape == w["inear"].shape
    nn = vec[f"{key}/nnear"]
    np.testing.assert_array_equal(w["nnear"][:, 0], nn)
    ref_in = vec[f"{key}/inear"].astype(np.int32)
    np.testing.assert_array_equal(w["inear"][:, 0, :ref_in.shape[1]], ref_in)
    assert not np.any(w["inear"][:, 0, ref_in.shape[1]:])
    # entries past nnear are zero
    m = np.arange(w["inear"].shape[2])[None, :] >= w["nnear"][:, 0][:, None]
    assert not np.any(w["inear"][:, 0, :][m]) and not np.any(w["weight"][:, 0, :][m])
    sel = vec[f"{key}/wblocks"]
    wr = vec[f"{key}/weight"]
    wg = w["weight"][sel, 0, :wr.shape[1]]
    d = np.abs(wg - wr) / np.maximum(1.0, np.abs(wr))
    assert d.max() < 1e-9, d.max()
    np.testing.assert_array_equal(w["variance"], v)
    # estimates from the stored weights (blocks of two or more neighbours,
    # not failed, unbiased kriging without bounds)
    if spec.get("unbias", 1) and spec.get("bounds") is None:
        ok = (nn >= 2) & np.isfinite(e)
        rec = _estimate_from_store(w, np.asarray(spec["obs"][0]["value"]))
        assert np.max(rel_err(rec[ok], e[ok])) < 1e-10
    # failed blocks: one neighbour, a NaN weight; single neighbours: NaN
    bad = (nn == 1)
    assert np.all(np.isnan(w["weight"][bad, 0, 0]))


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_set_weights_replays(spec):
    """set_weights into a fresh object, then solve: the stored weights
    against the current values, the stored variance; with the values
    transformed z -> 2 z + 3 the estimate becomes 2 e + 3 sum_k w_k (the
    replay is linear in the values)"""
    key = spec["name"]
    print("case", key, flush=True)
    k1, e1, v1, w = _solve_store(spec)
    k1.close()
    k2 = make_gpu(spec, use_old_weight=True)
    k2.set_weights(w)
    k2.solve()
    e2, v2 = k2.get_results(copy=True)
    t = k2.timing
    assert t["weights_ms"] > 0 and t["search_ms"] == 0
    assert k2.solver_stats["chol_fact"] == 0
    assert np.array_equal(np.isnan(e2), np.isnan(e1))
    assert np.max(rel_err(e2, e1)) < 1e-10
    np.testing.assert_array_equal(v2, v1)
    if spec.get("unbias", 1) and spec.get("bounds") is None:
        z = np.asarray(spec["obs"][0]["value"])
        k2.update_obs_value(1, 2.0 * z + 3.0)
        k2.solve()
        e3, _ = k2.get_results(copy=True)
        fin = np.isfinite(e1) & (w["nnear"][:, 0] >= 2)
        ws = w["weight"][:, 0, :].sum(axis=1)
        assert np.max(rel_err(e3[fin], 2.0 * e2[fin] + 3.0 * ws[fin])) < 1e-10
    k2.close()


def test_update_obs_value_switches_to_replay():
    """after a solve with store_weight, update_obs_value makes the next solve
    replay the store (kriging_base.F90:669-678); it equals a fresh solve with
    the new values"""
    spec = problems.ok2d_small()
    k, e, v, w = _solve_store(spec)
    z2 = np.asarray(spec["obs"][0]["value"]) * 1.5 + 0.1
    k.update_obs_value(1, z2)
    k.solve()
    e2, v2 = k.get_results(copy=True)
    assert k.timing["search_ms"] == 0 and k.timing["weights_ms"] > 0
    k.close()
    fresh = dict(spec)
    fresh["obs"] = [dict(spec["obs"][0], value=z2)]
    kf = make_gpu(fresh)
    kf.solve()
    ef, vf = kf.get_results(copy=True)
    kf.close()
    assert np.max(rel_err(e2, ef)) < 1e-10
    assert np.max(rel_err(v2, vf)) < 1e-10


def test_weight_file_roundtrip(tmp_path):
    """store_weight with a weight_file writes the factor file; a new object
    with use_old_weight and that file needs only observations and grid
    (weights carry ten decimals, variances ten digits)"""
    spec = problems.ok2d_small()
    fac = str(tmp_path / "weights.fac")
    k1, e1, v1, w = _solve_store(spec, weight_file=fac)
    k1.close()
    assert os.path.isfile(fac)
    lines = open(fac).read().splitlines()
    nb = len(e1)
    nm = w["inear"].shape[2]
    assert [int(x) for x in lines[0].split()] == [nb, 1, nm]
    assert len(lines) == 1 + 4 * nb
    # block records: order nnear / inear / variance / weights
    b = 7
    assert [int(x) for x in lines[1 + 4 * b].split()] == [b + 1, int(w["nnear"][b, 0])]
    assert [int(x) for x in lines[2 + 4 * b].split()] == list(w["inear"][b, 0, :w["nnear"][b, 0]])
    from krigekit_amd.kriging import Kriging
    o = spec["obs"][0]
    k2 = Kriging(ndim=spec["ndim"], use_old_weight=True, weight_file=fac)
    k2.set_obs(1, o["coord"], o["value"], nmax=o.get("nmax"))
    k2.set_grid(spec["grid"])
    k2.solve()
    e2, v2 = k2.get_results(copy=True)
    k2.close()
    assert np.max(np.abs(e2 - e1)) < 1e-8
    assert np.max(rel_err(v2, v1)) < 1e-9


@pytest.mark.parametrize("spec", CK, ids=CK_IDS)
def test_cokriging_store_matches_reference(spec, ckvec):
    """cokriging: per-group neighbour counts and lists, weights of sampled
    blocks within 1e-9 of the reference's store; set_weights replays the
    estimates and variances"""
    key = spec["name"]
    print("case", key, flush=True)
    k = make_gpu_ck(spec, store_weight=True)
    k.solve()
    e, v = k.get_results(copy=True)
    w = k.get_weights()
    k.close()
    nn = ckvec[f"{key}/nnear"].astype(np.int32)
    np.testing.assert_array_equal(w["nnear"], nn)
    ref_in = ckvec[f"{key}/inear"].astype(np.int32)
    # a failed block's empty group holds a left-over id in the reference
    # (ctx%inear of an earlier block); compared where the group has one
    failed = np.isnan(e).any(axis=1) & (nn.sum(axis=1) >= 2)
    keep = np.ones_like(ref_in, dtype=bool)
    keep[failed] = False
    np.testing.assert_array_equal(np.where(keep, w["inear"][:, :, :ref_in.shape[2]], 0),
                                  np.where(keep, ref_in, 0))
    sel = ckvec[f"{key}/wblocks"]
    wr = ckvec[f"{key}/weight"]
    wg = w["weight"][sel][..., :wr.shape[3]]
    fin = np.isfinite(wr).all(axis=(1, 2, 3))
    d = np.abs(wg[fin] - wr[fin]) / np.maximum(1.0, np.abs(wr[fin]))
    assert d.max() < 1e-9, d.max()
    # replay
    k2 = make_gpu_ck(spec, use_old_weight=True)
    k2.set_weights(w)
    k2.solve()
    e2, v2 = k2.get_results(copy=True)
    k2.close()
    assert np.array_equal(np.isnan(e2), np.isnan(e))
    assert np.max(rel_err(e2, e)) < 1e-10
    np.testing.assert_array_equal(v2, v)


def test_space_time_store_replays():
    """space-time blocks take their weights from the Bunch-Kaufman path of
    each block; the replay reproduces the solve and the weights sum to one"""
    from krigekit_amd.kriging_st import from_st_spec
    spec = problems.st_small()
    k = from_st_spec(spec, store_weight=True)
    k.solve()
    e, v = k.get_results(copy=True)
    w = k.get_weights()
    k.close()
    nn = w["nnear"][:, 0]
    ok = nn >= 2
    s = w["weight"][:, 0, :].sum(axis=1)
    assert np.max(np.abs(s[ok] - 1.0)) < 1e-10
    rec = _estimate_from_store(w, np.asarray(spec["obs"][0]["value"]))
    assert np.max(rel_err(rec[ok], e[ok])) < 1e-10
    k2 = from_st_spec(spec, use_old_weight=True)
    k2.set_weights(w)
    k2.solve()
    e2, v2 = k2.get_results(copy=True)
    k2.close()
    assert np.max(rel_err(e2, e)) < 1e-10
    np.testing.assert_array_equal(v2, v)


def test_fp32r_store_holds_fp64_weights():
    """the weights pass is FP64 whatever the solve precision: the stores of
    an FP64 and an FP32-refined solve are identical, and the replay matches
    the FP32-refined estimates to its 1e-6 bar"""
    spec = problems.ok2d_small()
    k64, e64, _, w64 = _solve_store(spec)
    k64.close()
    k32, e32, _, w32 = _solve_store(spec, precision="fp32r")
    k32.close()
    np.testing.assert_array_equal(w32["weight"], w64["weight"])
    np.testing.assert_array_equal(w32["inear"], w64["inear"])
    rec = _estimate_from_store(w32, np.asarray(spec["obs"][0]["value"]))
    ok = w32["nnear"][:, 0] >= 2
    assert np.max(rel_err(rec[ok], e32[ok])) < 1e-6


def test_replay_out_of_range_id_gives_nan():
    """a stored id outside the observations reads as NaN for its block only
    (the reference would read outside its array)"""
    spec = problems.ok2d_small()
    k1, e1, v1, w = _solve_store(spec)
    k1.close()
    w["inear"][5, 0, 0] = len(spec["obs"][0]["value"]) + 1
    w["inear"][9, 0, 1] = 0
    k2 = make_gpu(spec, use_old_weight=True)
    k2.set_weights(w)
    k2.solve()
    e2, _ = k2.get_results(copy=True)
    k2.close()
    assert np.isnan(e2[5]) and np.isnan(e2[9])
    keep = np.ones(len(e2), dtype=bool)
    keep[[5, 9]] = False
    assert np.max(rel_err(e2[keep], e1[keep])) < 1e-10

