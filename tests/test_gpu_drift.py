"""External drift (kriging with external drift, universal kriging on blocks,
drift without the unbiasedness row, drift in cokriging) and weight correction
on the GPU, solved by the bordered solver, against the compiled reference's
committed vectors (tests/golden/drift_vectors.npz): FP64 at 1e-10, FP32 +
MFMA + refinement at 1e-6, stored weights at 1e-9."""
import numpy as np
import pytest

import problems
from runners import ck_errors, make_gpu_ck, run_gpu_ck

pytestmark = pytest.mark.gpu

CASES = problems.drift_cases()
IDS = [c["name"] for c in CASES]


@pytest.fixture(scope="module")
def vec():
    return np.load(problems.GOLDEN + "/drift_vectors.npz")


def _ref(vec, spec, mode):
    return dict(estimate=vec[f"{spec['name']}/{mode}/estimate"], variance=vec[f"{spec['name']}/{mode}/variance"])


def _shaped(g, ref):
    return dict(estimate=np.asarray(g["estimate"]).reshape(ref["estimate"].shape),
                variance=np.asarray(g["variance"]).reshape(ref["variance"].shape))


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_drift_matches_reference(spec, mode, vec):
    print("case drift", spec["name"], mode, flush=True)
    g = run_gpu_ck(spec, table_mode=(mode == "table"))
    ref = _ref(vec, spec, mode)
    ee, ev = ck_errors(_shaped(g, ref), ref)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)
    f, c, sy = (int(x) for x in vec[f"{spec['name']}/{mode}/stats"])
    st = g["stats"]
    assert (st["fail"], st["chol_fact"], st["ssytrf_fact"]) == (f, c, sy), (st, (f, c, sy))


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_drift_fp32r(spec, vec):
    print("case drift fp32r", spec["name"], flush=True)
    g = run_gpu_ck(spec, precision="fp32r")
    ref = _ref(vec, spec, "table")
    ee, ev = ck_errors(_shaped(g, ref), ref)
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_drift_weight_store(spec, vec):
    """the stored weights (corrected ones under weight correction) of
    sampled blocks against the reference's store"""
    key = spec["name"]
    print("case drift store", key, flush=True)
    k = make_gpu_ck(spec, store_weight=True)
    k.solve()
    w = k.get_weights()
    k.close()
    sel = vec[f"{key}/wblocks"]
    wr = vec[f"{key}/weight"]
    # one variable: the store is (block, 1, k); the reference's (block, 1, 1, k)
    wg = np.asarray(w["weight"])[sel][..., :wr.shape[3]].reshape(wr.shape)
    d = np.abs(wg - wr) / np.maximum(1.0, np.abs(wr))
    assert d.max() < 1e-9, d.max()


def test_drift_setter_errors():
    """the reference's argument checks (kriging_base.F90:756-791, 1173-1220)"""
    spec = problems.drift_ked(nb=10)
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=2, ndrift=2, unbias=0)
    with pytest.raises(RuntimeError, match="Observation needs to be set before adding drift"):
        k.set_obs_drift(1, np.ones((5, 2)))
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], nmax=10)
    with pytest.raises(RuntimeError, match="size\\(drift, 1\\) /= ndrift"):
        k.set_obs_drift(1, np.ones((len(o["value"]), 3)))
    with pytest.raises(RuntimeError, match="Grid needs to be set before adding drift"):
        k.set_grid_drift(np.ones((3, 2)))
    k.set_grid(spec["grid"])
    with pytest.raises(RuntimeError, match="one drift value per block"):
        k.set_grid_drift(np.ones((len(spec["grid"]) + 1, 2)))
    k.close()
    k = Kriging(ndim=2)
    k.set_obs(1, o["coord"], o["value"], nmax=10)
    with pytest.raises(RuntimeError, match="ndrift==0"):
        k.set_obs_drift(1, np.ones((len(o["value"]), 1)))
    k.close()
