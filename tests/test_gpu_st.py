"""GPU space-time product-sum kriging against the oracle (whose space-time
path is pinned in test_st_oracle.py): neighbour lists identical, FP64 within
1e-10 (relative, floor 1e-4), FP32 factorisation + FP64 refinement within 1e-6."""
import numpy as np
import pytest

import problems
from runners import max_errors, run_gpu_st, run_oracle, same_lists

pytestmark = pytest.mark.gpu

CASES = problems.st_cases()
IDS = [c["name"] for c in CASES]


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_st_matches_oracle(spec, mode):
    print("case", spec["name"], mode, flush=True)
    o = run_oracle(spec, table_mode=(mode == "table"))
    g = run_gpu_st(spec, table_mode=(mode == "table"))
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("spec", CASES, ids=IDS)
def test_st_fp32r(spec):
    print("case", spec["name"], "fp32r", flush=True)
    o = run_oracle(spec)
    g = run_gpu_st(spec, precision="fp32r")
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)


@pytest.mark.parametrize("precision", ["fp64", "fp32r"])
def test_st_48_rows(precision):
    """the full-size workload's system size: 33-48 rows (64 lanes, 48-row instance)"""
    spec = problems.st_small(nmax=44, name="st_rows44")
    print("case st_rows44", precision, flush=True)
    o = run_oracle(spec)
    g = run_gpu_st(spec, precision=precision)
    assert same_lists(g, o)
    assert 32 < g["nnear"].max() <= 48
    tol = 1e-10 if precision == "fp64" else 1e-6
    ee, ev = max_errors(g, o)
    assert ee < tol and ev < tol, (ee, ev)


@pytest.mark.parametrize("fast", [2, 1, 0])
def test_st_search_modes_agree(fast):
    spec = problems.st_small(nstation=80, ndays=40, side=200.0, grid_side=12, ntimes=4, name="st_modes")
    print("case st_modes", fast, flush=True)
    o = run_oracle(spec)
    g = run_gpu_st(spec, fast=fast)
    assert same_lists(g, o)


def _kat(obs4, value, grid4, nmax, maxdist=None, time_at=1.0):
    vs = dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.8, a_major=500.0, a_minor1=300.0, a_minor2=100.0)
    return dict(name="kat", st=True, ndim=3, k_ps=0.5,
                obs=[dict(coord=np.asarray(obs4, float), value=np.asarray(value, float), nmax=nmax, maxdist=maxdist)],
                grid=np.asarray(grid4, float), vgm=[vs], vgm_t=[dict(vtype="exp", sill=0.5, at_k=10.0)],
                search=[dict(ivar=1, time_at=time_at)])


@pytest.mark.parametrize("time_at,expect", [(0.1, 10.0), (100.0, 20.0)])
def test_st_time_at_selects_nearest_pair(time_at, expect):
    """tests/test_spacetime_kriging.py:243-255 of the reference"""
    obs = [[0, 0, 0, 10], [0, 0, 0, 11], [3, 0, 0, 0], [3.1, 0, 0, 0], [100, 0, 0, 0], [0, 100, 0, 0],
           [0, 0, 100, 0]]
    g = run_gpu_st(_kat(obs, [10, 10, 20, 20, -1, -2, -3], [[0, 0, 0, 0]], nmax=2, time_at=time_at))
    assert g["estimate"][0] == pytest.approx(expect)


@pytest.mark.parametrize("time_at,inside", [(50.0, True), (200.0, False)])
def test_st_time_at_scales_search_radius(time_at, inside):
    """tests/test_spacetime_kriging.py:257-282 of the reference"""
    g = run_gpu_st(_kat([[0, 0, 0, .04], [0, 0, 0, .05]], [10, 10], [[0, 0, 0, 0]], nmax=2, maxdist=5.0,
                        time_at=time_at))
    if inside:
        assert g["estimate"][0] == pytest.approx(10.0)
    else:
        assert np.isnan(g["estimate"][0])


def test_st_full_size():
    """500 stations x 365 days (182,500 observations), 500 x 500 x 30 nodes
    (7.5 million), nmax 48: every 7,499th node against the oracle, and
    size-independent properties of the whole result"""
    from krigekit_amd import synth
    spec = synth.st()
    nb = len(spec["grid"])
    print("st full", nb, flush=True)
    g = run_gpu_st(spec)
    e, v, nn = g["estimate"], g["variance"], g["nnear"]
    assert e.shape == (nb,) and np.all(np.isfinite(e)) and np.all(v >= 0.0)
    assert nn.min() >= 2 and nn.max() == 48
    idx = np.arange(0, nb, 7499)
    sub = dict(spec)
    sub["grid"] = spec["grid"][idx]
    o = run_oracle(sub)
    sel = dict(estimate=e[idx], variance=v[idx], nnear=nn[idx], inear=g["inear"][idx])
    assert same_lists(sel, o)
    ee, ev = max_errors(sel, o)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


SM_CASES = problems.st_sum_cases()
SM_IDS = [c["name"] for c in SM_CASES]


@pytest.mark.parametrize("mode", ["table", "analytic"])
@pytest.mark.parametrize("spec", SM_CASES, ids=SM_IDS)
def test_st_sum_metric_matches_oracle(spec, mode):
    """the sum-metric model (oracle pinned in test_st_sum_metric.py)"""
    print("case", spec["name"], mode, flush=True)
    o = run_oracle(spec, table_mode=(mode == "table"))
    g = run_gpu_st(spec, table_mode=(mode == "table"))
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < 1e-10 and ev < 1e-10, (ee, ev)


@pytest.mark.parametrize("spec", SM_CASES[:2], ids=SM_IDS[:2])
def test_st_sum_metric_fp32r(spec):
    print("case", spec["name"], "fp32r", flush=True)
    o = run_oracle(spec)
    g = run_gpu_st(spec, precision="fp32r")
    assert same_lists(g, o)
    ee, ev = max_errors(g, o)
    assert ee < 1e-6 and ev < 1e-6, (ee, ev)


def test_st_sum_metric_setup_errors():
    """set_st_model / set_vgm_joint_sills / solve checks of the reference
    (kriging_st.F90:143-165, variogram_st.f90:325-337, 388-402)"""
    from krigekit_amd.kriging_st import SpaceTimeKriging
    spec = problems.st_small(nstation=10, ndays=10, grid_side=2, ntimes=1, name="st_err")
    o = spec["obs"][0]
    k = SpaceTimeKriging()
    for bad in (dict(transform="hol"), dict(transform="cyc"), dict(at=0.0), dict(time_sill=-1.0),
                dict(transform="xyz")):
        args = dict(model="sum_metric", transform="linear", at=5.0)
        args.update(bad)
        with pytest.raises(RuntimeError):
            k.set_st_model(**args)
    k.set_st_model("sum_metric", "bounded", at=5.0)
    k.set_obs(1, o["coord"], o["value"], nmax=o["nmax"])
    k.set_vgm(**spec["vgm"][0])
    k.set_vgm_temporal(1, 1, "exp", sill=0.5, at_k=5.0)
    with pytest.raises(RuntimeError, match="cs%nstruct"):
        k.set_vgm_joint_sills(1, 1, 0.3, 0.2)
    k.set_grid(spec["grid"])
    k.set_search(1, time_at=10.0)
    with pytest.raises(RuntimeError, match="sill_st not set"):
        k.solve()
    k.set_vgm_joint_sills(1, 1, 0.3)
    k.solve()
    e, v = k.get_results()
    assert np.all(np.isfinite(e)) and np.all(v >= 0)
    k.close()


@pytest.mark.parametrize("time_at,expect", [(0.1, 10.0), (100.0, 20.0)])
def test_st_sum_metric_time_at_selects_nearest_pair(time_at, expect):
    """tests/test_spacetime_kriging.py:66-84, 243-255 of the reference (its
    probe is sum-metric, linear transform, at 100, time_sill 10)"""
    obs = [[0, 0, 0, 10], [0, 0, 0, 11], [3, 0, 0, 0], [3.1, 0, 0, 0], [100, 0, 0, 0], [0, 100, 0, 0],
           [0, 0, 100, 0]]
    spec = problems.st_sum(_kat(obs, [10, 10, 20, 20, -1, -2, -3], [[0, 0, 0, 0]], nmax=2, time_at=time_at), [0.3],
                           transform="linear", at=100.0, time_sill=10.0, name="kat_sm")
    g = run_gpu_st(spec)
    assert g["estimate"][0] == pytest.approx(expect)
