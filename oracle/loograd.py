"""Leave-one-out MSE and its variogram-parameter gradient restated with numpy
-- TEST INFRASTRUCTURE (parity oracle for the library's adjoint kernel; only
tests/ use it).

The reference has no gradient; what it fixes is the function being
differentiated: its leave-one-out ordinary-kriging estimates
(kriging.F90:490-666 search with the query's own index removed, :714-939
assembly, solver.f90:224-300 solve), with the covariance of one structure in
analytic form (variogram.f90:327-337, vgmfunc.f90:48-87) or from the lookup
table (variogram.f90:223-325, 682-710).  Here:

  * neighbour lists and the estimates come from the C oracle (okrige.solve);
  * K, k and their parameter derivatives are rebuilt in numpy, with
    dM/dt of the anisotropy matrix (rotation.f90:51-95, scaled by 1/a_major)
    taken by central differences of a numpy restatement of calc_rotmat;
  * the gradient is d zhat/dt = y_w' dk - y_w' dK w for the bordered system
    A = [K 1; 1' 0], x = A^-1 [k; 1], y = A^-1 [z; 0].
The tests pin it against central finite differences of the C oracle's MSE.
"""
import numpy as np

from oracle import okrige

PARAMS = ("nugget", "sill", "a_major", "a_minor1", "a_minor2", "azimuth", "dip", "plunge")
GEOM = PARAMS[2:]
HB = np.array([0.0, 0.1, 0.5, 3.5])   # variogram.f90 h_bounds
DH = np.array([1e-5, 1e-4, 1e-3])     # per-zone step
TAIL = {"exp", "gau", "hol", "cyc", "dco"}


def rho(vt, r):
    """correlation shapes, vgmfunc.f90:48-87"""
    r = np.asarray(r, dtype=np.float64)
    pi = np.pi
    inside = r < 1.0
    if vt == "sph":
        return np.where(inside, (1.0 - 1.5 * r) + 0.5 * r ** 3, 0.0)
    if vt == "exp":
        return np.exp(-3.0 * r)
    if vt == "gau":
        return np.exp(-3.0625 * r * r)
    if vt == "hol":
        return np.cos(pi * r)
    if vt == "pow":
        return np.where(inside, 1.0 - np.abs(r) ** 1.5, 0.0)
    if vt == "bsq":
        return np.where(inside, (1.0 - r * r) ** 2, 0.0)
    if vt == "cir":
        rr = np.minimum(r, 1.0)
        return np.where(inside, 1.0 - (2.0 * rr * np.sqrt(1.0 - rr * rr) + 2.0 * np.arcsin(rr)) / pi, 0.0)
    if vt == "lin":
        return np.where(inside, 1.0 - r, 0.0)
    if vt == "cyc":
        return np.exp(-2.0 * np.sin(pi * r) ** 2)
    if vt == "dco":
        return np.exp(-3.0 * r) * np.cos(pi * r)
    raise ValueError(vt)


def drho(vt, r):
    """d rho / dr, by hand from the shapes above"""
    r = np.asarray(r, dtype=np.float64)
    pi = np.pi
    inside = r < 1.0
    if vt == "sph":
        return np.where(inside, -1.5 + 1.5 * r * r, 0.0)
    if vt == "exp":
        return -3.0 * np.exp(-3.0 * r)
    if vt == "gau":
        return -6.125 * r * np.exp(-3.0625 * r * r)
    if vt == "hol":
        return -pi * np.sin(pi * r)
    if vt == "pow":
        return np.where(inside, -1.5 * np.sqrt(np.abs(r)), 0.0)
    if vt == "bsq":
        return np.where(inside, -4.0 * r * (1.0 - r * r), 0.0)
    if vt == "cir":
        rr = np.minimum(r, 1.0)
        return np.where(inside, -4.0 * np.sqrt(1.0 - rr * rr) / pi, 0.0)
    if vt == "lin":
        return np.where(inside, -1.0, 0.0)
    if vt == "cyc":
        return -2.0 * pi * np.sin(2.0 * pi * r) * np.exp(-2.0 * np.sin(pi * r) ** 2)
    if vt == "dco":
        return -np.exp(-3.0 * r) * (3.0 * np.cos(pi * r) + pi * np.sin(pi * r))
    raise ValueError(vt)


def rotmat(az, dip, pl, an1, an2):
    """calc_rotmat, rotation.f90:51-95 (row-major numpy; u = R @ lag)"""
    d2r = np.pi / 180.0
    R = np.diag([1.0 / an1, 1.0, 1.0 / an2])
    if abs(pl) > 1e-10:
        a = -pl * d2r
        c, s = np.cos(a), np.sin(a)
        R = R @ np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])
    if abs(dip) > 1e-10:
        a = -dip * d2r
        c, s = np.cos(a), np.sin(a)
        R = R @ np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
    if abs(az) > 1e-10:
        a = az * d2r
        c, s = np.cos(a), np.sin(a)
        R = R @ np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    return R


def full_params(v):
    p = dict(nugget=0.0, sill=1.0, a_minor1=None, a_minor2=None, azimuth=0.0, dip=0.0, plunge=0.0)
    p.update({k: v[k] for k in v if k in PARAMS})
    if p["a_minor1"] is None:
        p["a_minor1"] = p["a_major"]
    if p["a_minor2"] is None:
        p["a_minor2"] = p["a_minor1"]
    return p


def aniso(p):
    return rotmat(p["azimuth"], p["dip"], p["plunge"], p["a_minor1"] / p["a_major"],
                  p["a_minor2"] / p["a_major"]) / p["a_major"]


def aniso_derivs(p):
    """dM/dt for the geometric parameters, central differences (the entries
    are smooth; the truncation error is O(step^2) ~ 1e-12 relative)"""
    out = {}
    for g in GEOM:
        step = 1e-6 * max(abs(p[g]), 1.0)
        hi, lo = dict(p), dict(p)
        hi[g] = p[g] + step
        lo[g] = p[g] - step
        out[g] = (aniso(hi) - aniso(lo)) / (2.0 * step)
    return out


def _table(vt, p):
    """composite table of one structure: entry 0 = cov0, then sill*rho(h_i)"""
    n = np.rint((HB[1:] - HB[:-1]) / DH).astype(int)
    zinv = n / (HB[1:] - HB[:-1])
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    tab = np.zeros(int(n.sum()) + 1)
    grid = np.zeros_like(tab)
    for k in range(3):
        i = np.arange(1, n[k] + 1)
        grid[off[k] + i] = HB[k] + i * (1.0 / zinv[k])
    tab[1:] = p["sill"] * rho(vt, grid[1:])
    tab[0] = p["sill"] + p["nugget"]
    return tab, grid, n, zinv, off


def cov_and_grad(vt, p, lags, table_mode, dmat=None):
    """C(lag) and dC/dt for t in PARAMS; lags (m, ndim)"""
    lags = np.atleast_2d(np.asarray(lags, dtype=np.float64))
    nd = lags.shape[1]
    M = aniso(p)[:nd, :nd]
    u = lags @ M.T
    h = np.sqrt(np.sum(u * u, axis=1))
    m = len(h)
    c = np.zeros(m)
    dc = np.zeros((m, len(PARAMS)))
    dcdh = np.zeros(m)
    eps = 2.2250738585072014e-308 * 1.0e3
    if table_mode:
        tab, grid, n, zinv, off = _table(vt, p)
        inn = (h > 0.0) & (h < HB[3])
        k = (h >= HB[1]).astype(int) + (h >= HB[2]).astype(int)
        fi = (h - HB[k]) * zinv[k]
        ifi = np.floor(fi).astype(int)
        il = np.minimum(ifi, n[k] - 1)
        i = off[k] + il
        i = np.where(inn, i, 0)
        f = fi - ifi
        t0, t1 = tab[i], tab[np.minimum(i + 1, len(tab) - 1)]
        r0 = np.where(i == 0, 1.0, rho(vt, grid[i]))
        r1 = rho(vt, grid[np.minimum(i + 1, len(tab) - 1)])
        c = np.where(inn, t0 + f * (t1 - t0), c)
        dcdh = np.where(inn, (t1 - t0) * zinv[k], 0.0)
        dc[:, 1] = np.where(inn, r0 + f * (r1 - r0), 0.0)
        dc[:, 0] = np.where(inn & (i == 0), 1.0 - f, 0.0)
        zero = h <= 0.0
        c = np.where(zero, tab[0], c)
        dc[zero, 0] = 1.0
        dc[zero, 1] = 1.0
        far = h >= HB[3]
        if vt in TAIL:
            c = np.where(far, p["sill"] * rho(vt, h), c)
            dc[far, 1] = rho(vt, h[far])
            dcdh = np.where(far, p["sill"] * drho(vt, h), dcdh)
    else:
        pos = h > eps
        c = np.where(pos, p["sill"] * rho(vt, h), p["sill"] + p["nugget"])
        dc[:, 1] = np.where(pos, rho(vt, h), 1.0)
        dc[:, 0] = np.where(pos, 0.0, 1.0)
        dcdh = np.where(pos, p["sill"] * drho(vt, h), 0.0)
    if dmat is None:
        dmat = aniso_derivs(p)
    hs = np.where(h > 0.0, h, 1.0)
    for gi, g in enumerate(GEOM):
        du = lags @ dmat[g][:nd, :nd].T
        dc[:, 2 + gi] = np.where(h > 0.0, dcdh * np.sum(u * du, axis=1) / hs, 0.0)
    return c, dc


def loo_grad(spec, table_mode=False, nthread=8, plant=None):
    """(mse, grad over PARAMS, zhat, dzhat) of a one-structure LOO problem.
    plant="no_dK" drops the y' dK w term (a planted mistake for the tests)."""
    assert spec.get("cross_validation") and len(spec["vgm"]) == 1
    v = spec["vgm"][0]
    vt = v["vtype"].lower()[:3]
    p = full_params(v)
    o = okrige.solve(spec, table_mode=table_mode, nthread=nthread)
    X = np.asarray(spec["obs"][0]["coord"], dtype=np.float64)
    z = np.asarray(spec["obs"][0]["value"], dtype=np.float64)
    ovar = spec["obs"][0].get("variance")
    ovar = np.zeros(len(z)) if ovar is None else np.asarray(ovar, dtype=np.float64)
    cov0 = p["sill"] + p["nugget"]
    dmat = aniso_derivs(p)
    nb = len(z)
    zhat = np.full(nb, np.nan)
    dz = np.zeros((nb, len(PARAMS)))
    nn = o["nnear"]
    for n in np.unique(nn):
        if n < 2:
            continue
        sel = np.nonzero(nn == n)[0]
        idx = o["inear"][sel, :n] - 1                      # (B, n), 0-based
        d = X[idx] - X[sel][:, None, :]                    # (B, n, nd) node-local
        B = len(sel)
        lag = d[:, :, None, :] - d[:, None, :, :]          # (B, n, n, nd)
        cK, dK = cov_and_grad(vt, p, lag.reshape(-1, X.shape[1]), table_mode, dmat)
        K = cK.reshape(B, n, n)
        dK = dK.reshape(B, n, n, len(PARAMS))
        ii = np.arange(n)
        K[:, ii, ii] = cov0 + ovar[idx]
        dK[:, ii, ii, :] = 0.0
        dK[:, ii, ii, 0] = 1.0
        dK[:, ii, ii, 1] = 1.0
        ck, dk = cov_and_grad(vt, p, d.reshape(-1, X.shape[1]), table_mode, dmat)
        ck = ck.reshape(B, n)
        dk = dk.reshape(B, n, len(PARAMS))
        A = np.zeros((B, n + 1, n + 1))
        A[:, :n, :n] = K
        A[:, :n, n] = 1.0
        A[:, n, :n] = 1.0
        rhs = np.zeros((B, n + 1, 2))
        rhs[:, :n, 0] = ck
        rhs[:, n, 0] = 1.0
        rhs[:, :n, 1] = z[idx]
        sol = np.linalg.solve(A, rhs)
        w = sol[:, :n, 0]
        y = sol[:, :n, 1]
        zhat[sel] = np.sum(w * z[idx], axis=1)
        dz[sel] = np.einsum("bj,bjp->bp", y, dk)
        if plant != "no_dK":
            dz[sel] -= np.einsum("bj,bjkp,bk->bp", y, dK, w)
    ok = np.isfinite(zhat)
    r = zhat[ok] - z[ok]
    mse = float(np.mean(r * r))
    grad = 2.0 * (r @ dz[ok]) / ok.sum()
    return mse, dict(zip(PARAMS, grad.tolist())), zhat, dz


def loo_mse(spec, table_mode=False, nthread=8):
    """MSE of the C oracle's leave-one-out estimates (the function itself)"""
    o = okrige.solve(spec, table_mode=table_mode, nthread=nthread)
    z = np.asarray(spec["obs"][0]["value"], dtype=np.float64)
    ok = np.isfinite(o["estimate"]) & (o["nnear"] >= 2)
    r = o["estimate"][ok] - z[ok]
    return float(np.mean(r * r))


def with_param(spec, name, value):
    s = dict(spec)
    v = full_params(spec["vgm"][0])
    v[name] = value
    vg = dict(spec["vgm"][0])
    vg.update(v)
    s["vgm"] = [vg]
    return s


def fd_grad(spec, names, mse_fn, rel=1e-4):
    """central differences with step rel*|t| (rel for zero parameters)"""
    v = full_params(spec["vgm"][0])
    out = {}
    for n in names:
        step = rel * (abs(v[n]) if v[n] != 0.0 else 1.0)
        out[n] = (mse_fn(with_param(spec, n, v[n] + step)) - mse_fn(with_param(spec, n, v[n] - step))) / (2 * step)
    return out
