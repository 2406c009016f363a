"""GPU sequential Gaussian simulation against the compiled reference's
realisations for the same paths and draws (tests/golden/sgsim_vectors.npz;
the oracle that restates it is pinned in test_sgsim_oracle.py): within 1e-10
(relative, floor 1e-4) in table and analytic mode; draws and paths from the
library's own generator; setup errors."""
import os

import numpy as np
import pytest

import problems
from runners import rel_err, run_gpu_sgsim

pytestmark = pytest.mark.gpu

VEC = np.load(os.path.join(os.path.dirname(__file__), "golden", "sgsim_vectors.npz"))
CASES = problems.sgsim_cases()
IDS = [c["name"] for c in CASES]


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_sgsim_matches_reference_vectors(spec, mode):
    print("case", spec["name"], mode, flush=True)
    e, v, t = run_gpu_sgsim(spec, table_mode=(mode == "table"))
    er, vr = VEC[f"{spec['name']}/{mode}/sim"], VEC[f"{spec['name']}/{mode}/variance"]
    assert e.shape == er.shape
    assert np.max(rel_err(e, er)) < 1e-10, np.max(rel_err(e, er))
    assert np.max(rel_err(v, vr)) < 1e-10
    assert 1 <= t["sim_levels"] < len(spec["grid"])


def _setup(spec, **kw):
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=spec["ndim"], nsim=spec["nsim"], **kw)
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"], nmax=o["nmax"])
    k.set_grid(spec["grid"])
    for v in spec["vgm"]:
        k.set_vgm(**v)
    return k


def test_sgsim_own_generator():
    """path and draws left to the library: reproducible for a seed, and the
    realisations of one run differ from each other"""
    spec = CASES[0]
    out = []
    for seed in (11, 11, 12):
        k = _setup(spec, seed=seed)
        k.set_sim()
        k.set_search(1)
        k.solve()
        e, v = k.get_results(copy=True)
        k.close()
        assert e.shape == (len(spec["grid"]), spec["nsim"]) and np.all(np.isfinite(e)) and np.all(v >= 0)
        out.append(e)
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])
    assert not np.allclose(out[0][:, 0], out[0][:, 1])


def test_sgsim_setup_errors():
    spec = CASES[0]
    k = _setup(spec, seed=3)
    k.set_search(1)
    with pytest.raises(RuntimeError, match="set_sim"):
        k.solve()
    with pytest.raises(ValueError):
        k.set_sim(randpath=np.arange(len(spec["grid"])))  # 0-based
    with pytest.raises(ValueError):
        k.set_sim(sample=np.zeros((len(spec["grid"]), spec["nsim"] + 1)))
    k.close()
    from krigekit_amd.kriging import Kriging
    k = Kriging(ndim=2, nsim=2, seed=1)
    o = spec["obs"][0]
    k.set_obs(1, o["coord"], o["value"])  # no nmax
    k.set_grid(spec["grid"])
    with pytest.raises(RuntimeError, match="nmax"):
        k.set_sim()
    k.close()
