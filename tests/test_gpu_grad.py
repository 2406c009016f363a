"""GPU increment (gradient-pair) constraints against the compiled reference's
vectors (tests/golden/grad_vectors.npz; the oracle restating them is pinned
in test_ckrige.py): FP64 within 1e-10 in table and analytic mode, FP32 +
refinement within 1e-6; setup errors."""
import numpy as np
import pytest

import problems
from runners import ck_errors, make_gpu_ck, run_gpu_ck

pytestmark = pytest.mark.gpu

CASES = problems.grad_cases()
IDS = [c["name"] for c in CASES]


@pytest.fixture(scope="module")
def vec():
    return np.load(problems.GOLDEN + "/grad_vectors.npz")


def _cmp(g, vec, key, mode):
    ref = dict(estimate=vec[f"{key}/{mode}/estimate"], variance=vec[f"{key}/{mode}/variance"])
    got = dict(estimate=np.asarray(g["estimate"]).reshape(ref["estimate"].shape),
               variance=np.asarray(g["variance"]).reshape(ref["variance"].shape))
    return ck_errors(got, ref)


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_grad_matches_reference(spec, mode, vec):
    print("case", spec["name"], mode, flush=True)
    g = run_gpu_ck(spec, table_mode=(mode == "table"))
    ee, ev = _cmp(g, vec, spec["name"], mode)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_grad_fp32r(spec, vec):
    print("case", spec["name"], "fp32r", flush=True)
    g = run_gpu_ck(spec, precision="fp32r")
    ee, ev = _cmp(g, vec, spec["name"], "table")
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)


def test_grad_setup_errors():
    spec = CASES[0]
    g = spec["grad"][0]
    k = make_gpu_ck(spec)
    with pytest.raises(RuntimeError, match="nlag"):
        k.set_grad(g["coord1"][:, :1], g["coord2"][:, :1], g["value"])
    n = 40
    with pytest.raises(RuntimeError, match="increment observations"):
        k.set_grad(np.zeros((n, 2)), np.ones((n, 2)), np.zeros(n))
    k.close()
    k = make_gpu_ck(spec, store_weight=True)
    with pytest.raises(RuntimeError, match="increment"):
        k.solve()
    k.close()
    # an empty set clears the pairs: the result is the plain problem's
    plain = dict(spec)
    plain.pop("grad")
    k = make_gpu_ck(spec)
    k.set_grad(np.zeros((0, 2)), np.zeros((0, 2)), np.zeros(0))
    k.solve()
    e0, _ = k.get_results(copy=True)
    k.close()
    e1 = run_gpu_ck(plain)["estimate"]
    assert np.array_equal(np.asarray(e0).ravel(), np.asarray(e1).ravel())
