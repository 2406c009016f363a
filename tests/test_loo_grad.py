"""The numpy restatement of the leave-one-out gradient (oracle/loograd.py)
pinned on the CPU: covariances against the C oracle, estimates against the C
oracle's leave-one-out solve, the gradient against central finite
differences of the C oracle's MSE, and the scale invariance of ordinary
kriging weights."""
import numpy as np
import pytest

import problems
from oracle import loograd as lg
from oracle import okrige

NAMES = ["nugget", "sill", "a_major", "a_minor1", "azimuth"]


@pytest.fixture(scope="module")
def spec():
    return problems.loo_grad_problem()


@pytest.mark.parametrize("table", [False, True], ids=["analytic", "table"])
def test_covariance_restatement(spec, table):
    rng = np.random.default_rng(3)
    lags = rng.normal(0.0, 600.0, (3000, 2))
    lags[:5] = 0.0
    c_np, _ = lg.cov_and_grad("sph", lg.full_params(spec["vgm"][0]), lags, table)
    np.testing.assert_allclose(c_np, okrige.cov(spec, lags, table_mode=table), rtol=0, atol=1e-14)


@pytest.mark.parametrize("vt", ["sph", "exp", "gau", "hol", "pow", "bsq", "cir", "lin", "cyc", "dco"])
def test_shape_derivatives(vt):
    r = np.linspace(0.013, 2.7, 400)
    r = r[np.abs(r - 1.0) > 1e-3]  # kinks of the compact shapes at r = 1
    fd = (lg.rho(vt, r + 1e-7) - lg.rho(vt, r - 1e-7)) / 2e-7
    np.testing.assert_allclose(lg.drho(vt, r), fd, rtol=0, atol=1e-6)


@pytest.mark.parametrize("table", [False, True], ids=["analytic", "table"])
def test_estimates_match_oracle(spec, table):
    _, _, zhat, _ = lg.loo_grad(spec, table_mode=table)
    o = okrige.solve(spec, table_mode=table)
    ok = o["nnear"] >= 2
    np.testing.assert_allclose(zhat[ok], o["estimate"][ok], rtol=0, atol=1e-12)


def test_gradient_matches_finite_differences_analytic(spec):
    _, g, _, _ = lg.loo_grad(spec, table_mode=False)
    fd = lg.fd_grad(spec, NAMES, lambda s: lg.loo_mse(s, False))
    for n in NAMES:
        assert abs(g[n] - fd[n]) <= 1e-6 * abs(fd[n]), (n, g[n], fd[n])


def test_gradient_matches_finite_differences_table(spec):
    """the table covariance is piecewise linear in h: central differences
    average the slopes of the segments they cross, hence the looser bound"""
    _, g, _, _ = lg.loo_grad(spec, table_mode=True)
    fd = lg.fd_grad(spec, NAMES, lambda s: lg.loo_mse(s, True))
    for n in NAMES:
        assert abs(g[n] - fd[n]) <= 1e-3 * abs(fd[n]), (n, g[n], fd[n])


@pytest.mark.parametrize("table", [False, True], ids=["analytic", "table"])
def test_scale_invariance(spec, table):
    """OK weights do not change when every covariance is scaled, so
    nugget * dMSE/dnugget + sill * dMSE/dsill = 0"""
    _, g, _, _ = lg.loo_grad(spec, table_mode=table)
    v = lg.full_params(spec["vgm"][0])
    s = v["nugget"] * g["nugget"] + v["sill"] * g["sill"]
    assert abs(s) <= 1e-9 * abs(v["sill"] * g["sill"])


def test_planted_mistake_fails(spec):
    """dropping the y' dK w term of the adjoint must break the agreement with
    finite differences"""
    _, g, _, _ = lg.loo_grad(spec, table_mode=False, plant="no_dK")
    fd = lg.fd_grad(spec, ["a_major", "sill"], lambda s: lg.loo_mse(s, False))
    assert any(abs(g[n] - fd[n]) > 1e-3 * abs(fd[n]) for n in ("a_major", "sill"))
