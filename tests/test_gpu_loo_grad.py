"""GPU leave-one-out gradient (kk_loo_grad, the adjoint kernel) against the
numpy oracle (oracle/loograd.py) and against central finite differences of
the GPU's own leave-one-out MSE."""
import numpy as np
import pytest

import problems
from oracle import loograd as lg
from runners import make_gpu

pytestmark = pytest.mark.gpu

NAMES = ["nugget", "sill", "a_major", "a_minor1", "azimuth"]


def gpu_grad(spec, table, per_block=False):
    k = make_gpu(spec, table_mode=table)
    out = k.loo_grad(per_block=per_block)
    k.close()
    return out


def gpu_mse(spec, table):
    k = make_gpu(spec, table_mode=table)
    k.solve()
    e, _ = k.get_results(copy=True)
    nn, _ = k.neighbours()
    k.close()
    z = np.asarray(spec["obs"][0]["value"])
    ok = np.isfinite(e) & (nn >= 2)
    return float(np.mean((e[ok] - z[ok]) ** 2))


@pytest.mark.parametrize("nmax", [16, 6, 24, 40])
@pytest.mark.parametrize("table", [False, True], ids=["analytic", "table"])
def test_gradient_matches_oracle(table, nmax):
    """nmax 6, 16 and 24 run 8, 16 and 32 lanes per block (several blocks per
    wave); nmax 40 one block per workgroup"""
    spec = problems.loo_grad_problem(nmax=nmax)
    print("loo_grad_problem table", table, "nmax", nmax, flush=True)
    mse, g, zh, dz = gpu_grad(spec, table, per_block=True)
    omse, og, ozh, odz = lg.loo_grad(spec, table_mode=table)
    assert mse == pytest.approx(omse, rel=1e-12)
    ok = np.isfinite(ozh)
    np.testing.assert_array_equal(np.isfinite(zh), ok)
    np.testing.assert_allclose(zh[ok], ozh[ok], rtol=0, atol=1e-11)
    scale = np.max(np.abs(odz[ok]), axis=0) + 1e-300
    assert np.max(np.abs(dz[ok] - odz[ok]) / scale) < 1e-8
    for n in lg.PARAMS:
        assert abs(g[n] - og[n]) <= 1e-9 * max(abs(og[n]), 1e-12), (n, g[n], og[n])


def test_gradient_matches_finite_differences():
    """adjoint gradient against central differences (step 1e-4 relative) of the
    GPU forward MSE, analytic covariance: agreement to 1e-5"""
    spec = problems.loo_grad_problem(n_uniform=4500, n_dup=500)
    _, g = gpu_grad(spec, False)
    fd = lg.fd_grad(spec, NAMES, lambda s: gpu_mse(s, False))
    for n in NAMES:
        assert abs(g[n] - fd[n]) <= 1e-5 * abs(fd[n]), (n, g[n], fd[n])


def test_full_size_loo_gradient():
    """100,000 observations (10% jittered duplicates), nmax 16, anisotropic
    spherical (0.1, 1.0, 1500, 600, 45): adjoint vs central differences"""
    from krigekit_amd import synth
    spec = synth.loo2d()
    print("loo2d full", flush=True)
    mse, g = gpu_grad(spec, False)
    assert np.isfinite(mse)
    fd = lg.fd_grad(spec, NAMES, lambda s: gpu_mse(s, False))
    for n in NAMES:
        assert abs(g[n] - fd[n]) <= 1e-5 * abs(fd[n]), (n, g[n], fd[n])


def test_autograd_function():
    import torch
    from krigekit_amd.autograd import loo_mse
    spec = problems.loo_grad_problem()
    k = make_gpu(spec, table_mode=False)
    v = spec["vgm"][0]
    base = dict(vtype=v["vtype"], nugget=v["nugget"], sill=v["sill"], a_major=v["a_major"],
                a_minor1=v["a_minor1"], azimuth=v["azimuth"])
    theta = torch.tensor([base[n] for n in NAMES], dtype=torch.float64, requires_grad=True)
    loss = loo_mse(k, theta, base)
    loss.backward()
    _, og, _, _ = lg.loo_grad(spec, table_mode=False)
    np.testing.assert_allclose(theta.grad.numpy(), [og[n] for n in NAMES], rtol=1e-9)
    assert float(loss.detach()) == pytest.approx(lg.loo_mse(spec, False), rel=1e-12)
    k.close()


def _same(a, b):
    assert a[0] == b[0]
    assert a[1] == b[1]
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[3], b[3])


@pytest.mark.parametrize("table", [False, True], ids=["analytic", "table"])
def test_new_variogram_keeps_lists(table):
    """after set_vgm the handle keeps its neighbour lists and re-runs only the
    adjoint kernel: the result equals a fresh handle's at the new variogram,
    bit for bit"""
    spec = problems.loo_grad_problem()
    print("loo keep lists table", table, flush=True)
    k = make_gpu(spec, table_mode=table)
    k.loo_grad(per_block=True)
    v = dict(spec["vgm"][0], sill=1.3, a_major=1700.0, azimuth=30.0)
    k.set_vgm(append=False, **v)
    again = k.loo_grad(per_block=True)
    k.close()
    s2 = dict(spec, vgm=[v])
    _same(again, gpu_grad(s2, table, per_block=True))


def test_new_inputs_rebuild_lists():
    """new observation values or a new search (here nmax) are searched again:
    the result equals a fresh handle's"""
    spec = problems.loo_grad_problem()
    print("loo rebuild lists", flush=True)
    k = make_gpu(spec, table_mode=False)
    k.loo_grad(per_block=True)
    z2 = np.asarray(spec["obs"][0]["value"]) * 1.5 + 0.25
    k.update_obs_value(1, z2)
    got = k.loo_grad(per_block=True)
    o2 = dict(spec["obs"][0], value=z2)
    _same(got, gpu_grad(dict(spec, obs=[o2]), False, per_block=True))
    o3 = dict(spec["obs"][0], nmax=8)
    k.set_obs(1, o3["coord"], o3["value"], nmax=8)
    k.set_grid_cv()
    for sr in spec["search"]:
        k.set_search(**sr)
    got = k.loo_grad(per_block=True)
    k.close()
    _same(got, gpu_grad(dict(spec, obs=[o3]), False, per_block=True))
