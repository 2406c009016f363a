"""Multi-variable (co)kriging restated with numpy -- TEST INFRASTRUCTURE
(parity oracle for the library's cokriging path; only tests/ use it).

Follows the reference's per-block computation for nvar > 1:
  * neighbours: one search per variable, each on its own observation set
    (kriging_base.F90:3172-3175), here taken from the C oracle's search;
  * rows: the variables' neighbour groups in order (istart, :3177-3190),
    each group sorted by index;
  * K(r, c) = C_{v(r) v(c)}((x_r - x_c) / rs), diagonal cov0 + obs variance +
    local nugget (kriging.F90:714-849), covariances from the C oracle
    (okrige.cov on the variable pair's structures);
  * table mode: the reference builds lookup tables only for vgm(a, b) with
    a <= b (kriging.F90:410-420) and evaluates the others analytically
    (variogram.f90:365); K's cross blocks use vgm(col var, row var) with the
    column group after the row group, so they are analytic, and the
    right-hand side uses vgm(group var, target var);
  * constraint rows (kriging_base.F90:614-624, 940-950): first the ndrift
    external drifts (the drift values at the row's observation,
    kriging.F90:675-849), then unbiasedness: std_ck -> one per variable (row
    v(r) of F is 1), otherwise one row of ones; none for simple kriging;
  * right-hand side per target variable q: sum over the block's integration
    nodes of w_k C_{v(r) q}((x_r - g_k) / rs), constraint part the block's
    drift values, then e_q (std_ck) or 1 (kriging.F90:864-939);
  * gradient (increment) rows (kriging.F90:796-840, 904-920; kriging_base.F90:
    2591-2660): every increment observation z(x1) - z(x2) of a variable joins
    every block's system after the neighbour groups, with covariances
    C(x1 - x) - C(x2 - x) against data rows (vgm(increment variable, data
    variable)), c11 - c12 - c21 + c22 between increments of one variable
    (zero across variables), diagonal 2 C(0) - 2 C(x1 - x2) + variance (no
    local nugget), its drift values in the drift rows and 0 in the
    unbiasedness rows; right-hand side sum_k w_k C(x1 - g_k) - sum_k w_k
    C(x2 - g_k); the estimate adds the increments' weights times their
    values; a block without data neighbours is NaN however many increments;
  * weight correction (nvar 1, kriging_base.F90:3388-3391, 1995-2001): the
    data weights clipped at 0 and divided by their sum when it is positive,
    before the estimate and variance;
  * estimate (kriging_base.F90:3407-3503): sum of weights x values, the
    Isaaks-Srivastava mean shift for non-standard cokriging, the simple
    kriging mean term, bounds;
  * variance (kriging.F90:968-1028): base_ij - x_i' b_j over weights and
    multipliers, base = cov0 (points) or the double sum over integration
    nodes (lag not range-scaled); symmetrised; diagonal and off-diagonal
    clamped at 0.
The bordered system is solved as the reference solves it (solver.f90:224-300,
kriging_base.F90:3244-3398): Cholesky K = U'U, K^-1 F, the Schur complement
S = F'K^-1 F and its Cholesky factor, mu = S^-1 (F'K^-1 b - d),
w = K^-1 b - K^-1 F mu; when either factorisation meets a pivot that is not
positive, the Bunch-Kaufman factorisation of the full augmented system with
one jittered retry (okrige.bk_fallback, the C restatement of SSYTRF/SSYTRS),
and NaN weights when that fails too (estimate NaN, variance max(NaN, 0) = 0).
"""
import numpy as np
import scipy.linalg as sla

from oracle import okrige


def _pair_vgm(spec, i, j):
    a, b = min(i, j), max(i, j)
    return [v for v in spec["vgm"] if (min(v["ivar"], v["jvar"]), max(v["ivar"], v["jvar"])) == (a, b)]


def _cov0(structs):
    tot, i = 0.0, 0
    while i < len(structs):
        g = structs[i].get("sill", 1.0) + structs[i].get("nugget", 0.0)
        while i + 1 < len(structs) and structs[i + 1].get("product", False):
            i += 1
            g *= structs[i].get("sill", 1.0) + structs[i].get("nugget", 0.0)
        tot += g
        i += 1
    return tot


def neighbour_lists(spec, nthread=8):
    """per variable: (nnear, inear 0-based) from the C oracle's search"""
    out = []
    for v in range(spec["nvar"]):
        sv = dict(name="ckv", ndim=spec["ndim"], obs=[spec["obs"][v]], grid=spec["grid"],
                  vgm=_pair_vgm(spec, v + 1, v + 1), search=[spec["search"][v]],
                  unbias=spec.get("unbias", 1))
        o = okrige.solve(sv, table_mode=True, nthread=nthread)
        out.append((o["nnear"], o["inear"] - 1))
    return out


def _solve_like_reference(K, R, n, p):
    """solution of the augmented system [K F; F' 0] x = R as the reference
    computes it, or None when it fails (NaN weights)"""
    try:
        U = sla.cholesky(K[:n, :n], lower=False, check_finite=False)
        xb = sla.cho_solve((U, False), R[:n], check_finite=False)
        if p:
            F = K[:n, n:]
            kf = sla.cho_solve((U, False), F, check_finite=False)
            Us = sla.cholesky(F.T @ kf, lower=False, check_finite=False)
            lam = sla.cho_solve((Us, False), kf.T @ R[:n] - R[n:], check_finite=False)
            return np.vstack([xb - kf @ lam, lam])
        return xb
    except np.linalg.LinAlgError:
        info, x = okrige.bk_fallback(K, n, R)
        return None if info else x


def solve(spec, table_mode=True, nthread=8, want_weights=False):
    nvar, nd = spec["nvar"], spec["ndim"]
    std = spec.get("std_ck", True)
    unbias = spec.get("unbias", 1)
    centres = np.asarray(spec["grid"], dtype=np.float64)
    nb = len(centres)
    nsub = int(spec.get("nsub", 1))
    if spec.get("sub_coord") is not None:
        sub = np.asarray(spec["sub_coord"], dtype=np.float64).reshape(nb, nsub, nd)
        sw = np.asarray(spec["sub_weight"], dtype=np.float64).reshape(nb, nsub)
    else:
        sub = centres[:, None, :]
        sw = np.ones((nb, 1))
    rs = np.ones(nb) if spec.get("rangescale") is None else np.asarray(spec["rangescale"])
    ln = np.zeros(nb) if spec.get("localnugget") is None else np.asarray(spec["localnugget"])
    huge = 1.7976931348623157e308
    blo, bhi = spec.get("bounds", (-huge, huge))
    lists = neighbour_lists(spec, nthread)
    X = [np.asarray(o["coord"], dtype=np.float64) for o in spec["obs"]]
    Z = [np.asarray(o["value"], dtype=np.float64) for o in spec["obs"]]
    V = [np.zeros(len(z)) if o.get("variance") is None else np.asarray(o["variance"], dtype=np.float64)
         for o, z in zip(spec["obs"], Z)]
    skm = [o.get("sk_mean", 0.0) for o in spec["obs"]]
    pairs = {(i, j): _pair_vgm(spec, i + 1, j + 1) for i in range(nvar) for j in range(nvar)}
    cov0 = {k: _cov0(v) for k, v in pairs.items()}
    base_spec = dict(ndim=nd, obs=[spec["obs"][0]], grid=centres[:1], search=[spec["search"][0]])

    def cov(a, b, lags):
        """C from vgm(a+1, b+1): table only where the reference built one"""
        s = dict(base_spec)
        s["vgm"] = pairs[(a, b)]
        return okrige.cov(s, np.asarray(lags, dtype=np.float64).reshape(-1, nd),
                          table_mode=table_mode and a <= b)

    ndr = int(spec.get("ndrift", 0))
    D = [np.zeros((len(z), ndr)) if o.get("drift") is None else np.asarray(o["drift"], dtype=np.float64).reshape(len(z), ndr)
         for o, z in zip(spec["obs"], Z)]
    gd = np.zeros((nb, ndr)) if spec.get("grid_drift") is None else np.asarray(spec["grid_drift"],
                                                                                dtype=np.float64).reshape(nb, ndr)
    wcorr = bool(spec.get("weight_correction", False)) and nvar == 1
    # increments per variable: (x1, x2, value, variance, drift)
    G = []
    for v in range(nvar):
        gs = [g for g in spec.get("grad", []) if g["ivar"] == v + 1]
        if gs:
            g = gs[-1]
            c1 = np.asarray(g["coord1"], dtype=np.float64).reshape(-1, nd)
            ng = len(c1)
            G.append((c1, np.asarray(g["coord2"], dtype=np.float64).reshape(-1, nd),
                      np.asarray(g["value"], dtype=np.float64).ravel(),
                      np.zeros(ng) if g.get("variance") is None else np.asarray(g["variance"], dtype=np.float64),
                      np.zeros((ng, ndr)) if g.get("drift") is None else np.asarray(g["drift"], dtype=np.float64).reshape(ng, ndr)))
        else:
            G.append(None)
    gv = np.concatenate([np.full(len(g[0]), v) for v, g in enumerate(G) if g is not None]).astype(int) \
        if any(g is not None for g in G) else np.zeros(0, dtype=int)
    gx1 = np.concatenate([g[0] for g in G if g is not None]) if len(gv) else np.zeros((0, nd))
    gx2 = np.concatenate([g[1] for g in G if g is not None]) if len(gv) else np.zeros((0, nd))
    gz = np.concatenate([g[2] for g in G if g is not None]) if len(gv) else np.zeros(0)
    gvr = np.concatenate([g[3] for g in G if g is not None]) if len(gv) else np.zeros(0)
    gdr = np.concatenate([g[4] for g in G if g is not None]) if len(gv) else np.zeros((0, ndr))
    nge = len(gv)
    pu = (nvar if std else 1) if unbias else 0
    p = ndr + pu
    est = np.full((nb, nvar), np.nan)
    var = np.full((nb, nvar, nvar), np.nan)
    nnear = np.stack([l[0] for l in lists], axis=1)
    # weight store layout: weights[b, q, g, k] for neighbour k of group g
    wmax = max(int(nnear.max()), 1) if nb else 1
    weights = np.zeros((nb, nvar, nvar, wmax)) if want_weights else None
    for b in range(nb):
        groups = [l[1][b, :l[0][b]] for l in lists]
        rv = np.concatenate([np.full(len(g), v) for v, g in enumerate(groups)]).astype(int)
        no = len(rv)
        if no == 0:
            continue
        n = no + nge
        xr = np.concatenate([X[v][g] for v, g in enumerate(groups)])
        zr = np.concatenate([Z[v][g] for v, g in enumerate(groups)])
        vr = np.concatenate([V[v][g] for v, g in enumerate(groups)])
        dr = np.concatenate([D[v][g] for v, g in enumerate(groups)])
        K = np.zeros((n + p, n + p))
        for i in range(nvar):
            for j in range(nvar):
                ri, rj = np.nonzero(rv == i)[0], np.nonzero(rv == j)[0]
                if len(ri) == 0 or len(rj) == 0:
                    continue
                lag = (xr[ri][:, None, :] - xr[rj][None, :, :]) / rs[b]
                K[np.ix_(ri, rj)] = cov(max(i, j), min(i, j), lag).reshape(len(ri), len(rj))
        K[np.arange(no), np.arange(no)] = np.array([cov0[(v, v)] for v in rv]) + vr + ln[b]
        for g in range(nvar):
            ga = np.nonzero(gv == g)[0]
            if len(ga) == 0:
                continue
            # increment - data: C(x1 - x) - C(x2 - x), vgm(increment var, data var)
            for i in range(nvar):
                ri = np.nonzero(rv == i)[0]
                if len(ri) == 0:
                    continue
                l1 = (gx1[ga][:, None, :] - xr[ri][None, :, :]) / rs[b]
                l2 = (gx2[ga][:, None, :] - xr[ri][None, :, :]) / rs[b]
                c = (cov(g, i, l1) - cov(g, i, l2)).reshape(len(ga), len(ri))
                K[np.ix_(no + ga, ri)] = c
                K[np.ix_(ri, no + ga)] = c.T
            # increments of one variable: c11 - c12 - c21 + c22 (pair i before j),
            # diagonal 2 C(0) - 2 C(x1 - x2) + variance
            ll = lambda u, w: (u[:, None, :] - w[None, :, :]) / rs[b]  # noqa: E731
            m = len(ga)
            c11 = cov(g, g, ll(gx1[ga], gx1[ga])).reshape(m, m)
            c12 = cov(g, g, ll(gx1[ga], gx2[ga])).reshape(m, m)
            c21 = cov(g, g, ll(gx2[ga], gx1[ga])).reshape(m, m)
            c22 = cov(g, g, ll(gx2[ga], gx2[ga])).reshape(m, m)
            gg = c11 - c12 - c21 + c22
            iu = np.triu_indices(m, 1)
            blk = np.zeros((m, m))
            blk[iu] = gg[iu]
            blk = blk + blk.T
            dg = 2.0 * cov0[(g, g)] - 2.0 * cov(g, g, (gx1[ga] - gx2[ga]) / rs[b]) + gvr[ga]
            blk[np.arange(m), np.arange(m)] = dg
            K[np.ix_(no + ga, no + ga)] = blk
        if p:
            F = np.zeros((n, p))
            F[:no, :ndr] = dr
            F[no:, :ndr] = gdr
            if pu:
                if std:
                    F[np.arange(no), ndr + rv] = 1.0
                else:
                    F[:no, ndr] = 1.0
            K[:n, n:] = F
            K[n:, :n] = F.T
        R = np.zeros((n + p, nvar))
        for q in range(nvar):
            for i in range(nvar):
                ri = np.nonzero(rv == i)[0]
                if len(ri) == 0:
                    continue
                lag = (xr[ri][:, None, :] - sub[b][None, :, :]) / rs[b]
                R[ri, q] = cov(i, q, lag).reshape(len(ri), -1) @ sw[b]
            for g in range(nvar):
                ga = np.nonzero(gv == g)[0]
                if len(ga) == 0:
                    continue
                t1 = cov(g, q, (gx1[ga][:, None, :] - sub[b][None, :, :]) / rs[b]).reshape(len(ga), -1) @ sw[b]
                t2 = cov(g, q, (gx2[ga][:, None, :] - sub[b][None, :, :]) / rs[b]).reshape(len(ga), -1) @ sw[b]
                R[no + ga, q] = t1 - t2
            if p:
                R[n:n + ndr, q] = gd[b]
            if pu:
                R[n + ndr + (q if std else 0), q] = 1.0
        if n == 1:
            # the reference skips the solve of a one-neighbour block
            # (kriging_base.F90:2127): NaN, as the library reports
            continue
        x = _solve_like_reference(K, R, n, p)
        if x is None:
            est[b] = np.nan
            var[b] = 0.0
            continue
        if wcorr:
            wc = np.where(x[:n, 0] > 0.0, x[:n, 0], 0.0)
            tot = 0.0
            for v in wc:
                tot += v
            if tot > 0.0:
                wc = wc / tot
            x = x.copy()
            x[:n, 0] = wc
        w = x[:no]
        if want_weights:
            for g in range(nvar):
                rows = np.nonzero(rv == g)[0]
                weights[b, :, g, :len(rows)] = w[rows, :].T
        for q in range(nvar):
            e = w[:, q] @ zr
            if nge:
                e += x[no:n, q] @ gz
            if not std and unbias and nvar > 1:
                means = [zr[rv == i].mean() if np.any(rv == i) else 0.0 for i in range(nvar)]
                for i in range(nvar):
                    if i != q and np.any(rv == i):
                        e += (means[q] - means[i]) * w[rv == i, q].sum()
            if unbias == 0 and skm[q] != 0.0:
                e += (1.0 - w[:, q].sum()) * skm[q]
            est[b, q] = min(max(e, blo), bhi)
        for i in range(nvar):
            for j in range(i, nvar):
                if nsub == 1:
                    base = cov0[(i, j)]
                else:
                    g = sub[b]
                    lag = g[:, None, :] - g[None, :, :]
                    C = cov(i, j, lag).reshape(nsub, nsub)
                    iu = np.triu_indices(nsub, 1)
                    base = cov0[(i, j)] * np.sum(sw[b] ** 2) + 2.0 * np.sum(C[iu] * sw[b][iu[0]] * sw[b][iu[1]])
                s = base - x[:, i] @ R[:, j]
                var[b, i, j] = var[b, j, i] = s
        for i in range(nvar):
            var[b, i, i] = max(var[b, i, i], 0.0)
            for j in range(i + 1, nvar):
                var[b, j, i] = var[b, i, j] = max(var[b, j, i], 0.0)
    return dict(estimate=est, variance=var, nnear=nnear, lists=lists, weights=weights)
