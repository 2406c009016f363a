"""Sequential Gaussian simulation restated with numpy -- TEST INFRASTRUCTURE
(parity oracle for the library's simulation path; only tests/ use it).

Follows the reference for one variable and point targets:
  * set_sim (kriging_base.F90:1226-1316): the blocks are visited along the
    path; the draw of path position p is sample[p]; the search pool is the
    observations followed by the blocks in path order;
  * search (kriging.F90:535-612): block p (0-based) searches the pool points
    with 1-based index <= nobs + p (kdtree2_n_nearest_maxidx, here the C
    oracle's emulation of the tree): nmax nearest, or every such point when
    they are no more than nmax; with sector search the 2^ndim nmax 4 nearest
    (fewer when the pool is smaller) through the sector filter
    (:1093-1154) over both groups; the simulated group within maxdist; a
    simulated block on one of the observations of the row (every coordinate
    within 1e-6 (1 + |x_obs|)) dropped; the observation group within maxdist;
  * the system (kriging.F90:714-849): one ordinary-kriging system over both
    groups, diagonal C(0) + observation variance + local nugget (simulated
    rows: no observation variance), lags over the block's range scale;
    solved as the reference solves it (oracle.ckrige._solve_like_reference);
  * estimate_block_base (kriging_base.F90:3405-3505): the observation part
    (simple kriging with a mean: + (1 - sum of the observation weights)
    mean), plus per realisation the simulated neighbours' values, plus
    sqrt(variance) times the draw (the 1x1 Cholesky factor, 0 when the
    variance is not positive), clipped to the bounds;
  * reorder_sgsim (kriging_base.F90:2268-2282): back to block order.
A block with fewer than two neighbours has no solved system: NaN (the
reference's estimate for a single neighbour comes from an unsolved system).
"""
import numpy as np

from oracle import ckrige, okrige


def simulate(spec, table_mode=True):
    """realisations (nblock, nsim) and variances (nblock,)"""
    nd = spec["ndim"]
    o = spec["obs"][0]
    X = np.asarray(o["coord"], dtype=np.float64).reshape(-1, nd)
    Z = np.asarray(o["value"], dtype=np.float64)
    V = np.zeros(len(Z)) if o.get("variance") is None else np.asarray(o["variance"], dtype=np.float64)
    G = np.asarray(spec["grid"], dtype=np.float64).reshape(-1, nd)
    path = np.asarray(spec["path"], dtype=np.int64) - 1
    smp = np.asarray(spec["sample"], dtype=np.float64)
    nobs, nb, ns = len(X), len(G), smp.shape[1]
    sector = bool(spec["search"][0].get("sector_search", False))
    nmax = int(o["nmax"])
    md = o.get("maxdist")
    md2 = np.inf if md is None else md * md
    skm = o.get("sk_mean", 0.0)
    unbias = spec.get("unbias", 1)
    huge = 1.7976931348623157e308
    blo, bhi = spec.get("bounds", (-huge, huge))
    rs = np.ones(nb) if spec.get("rangescale") is None else np.asarray(spec["rangescale"])[path]
    ln = np.zeros(nb) if spec.get("localnugget") is None else np.asarray(spec["localnugget"])[path]
    P = np.vstack([X, G[path]])
    Q = G[path]
    nmax_eff = min(nmax, nobs + nb)
    lim = nobs + np.arange(nb)
    k = np.minimum(lim, (2 ** nd) * nmax * 4) if sector else np.minimum(nmax_eff, lim)
    idx, dis = okrige.knn_maxidx(P, Q, k, lim, sector)
    cov0 = ckrige._cov0(spec["vgm"])
    base = dict(ndim=nd, obs=[dict(coord=X[:2], value=Z[:2])], grid=G[:1], vgm=spec["vgm"],
                search=[dict(ivar=1)])

    def cov(lags):
        return okrige.cov(base, np.asarray(lags, dtype=np.float64).reshape(-1, nd), table_mode=table_mode)

    val = np.full((ns, nb), np.nan)
    var = np.full(nb, np.nan)
    for p in range(nb):
        c, d = idx[p, :k[p]], dis[p, :k[p]]
        if sector:
            cnt = np.zeros(2 ** nd, dtype=int)
            og, sg = [], []
            for i in range(k[p]):
                if d[i] > md2:
                    continue
                df = P[c[i] - 1] - Q[p]
                sec = int(df[0] < 0.0) + 2 * int(df[1] < 0.0) + (4 * int(df[2] < 0.0) if nd == 3 else 0)
                if cnt[sec] < nmax:
                    cnt[sec] += 1
                    (og if c[i] <= nobs else sg).append((c[i], d[i]))
        else:
            og = [(c[i], d[i]) for i in range(k[p]) if c[i] <= nobs]
            sg = [(c[i], d[i]) for i in range(k[p]) if c[i] > nobs and d[i] <= md2]
        keep = []
        for b, db in sg:
            dup = any(np.all(np.abs(P[b - 1] - X[i - 1]) <= 1.0e-6 * (1.0 + np.abs(X[i - 1]))) for i, _ in og)
            if not dup:
                keep.append(b)
        obs = np.sort([i for i, di in og if di <= md2]).astype(int)
        sim = np.sort(keep).astype(int)
        rows = np.concatenate([obs, sim])
        n = len(rows)
        if n < 2:
            continue
        xr = P[rows - 1]
        pu = 1 if unbias else 0
        K = np.zeros((n + pu, n + pu))
        K[:n, :n] = cov(((xr[:, None, :] - xr[None, :, :]) / rs[p]).reshape(-1, nd)).reshape(n, n)
        K[np.arange(n), np.arange(n)] = np.concatenate([cov0 + V[obs - 1], np.full(len(sim), cov0)]) + ln[p]
        R = np.zeros((n + pu, 1))
        R[:n, 0] = cov((xr - Q[p]) / rs[p])
        if pu:
            K[:n, n] = K[n, :n] = 1.0
            R[n, 0] = 1.0
        x = ckrige._solve_like_reference(K, R, n, pu)
        if x is None:
            var[p] = 0.0
            continue
        x = x[:, 0]
        var[p] = max(cov0 - x @ R[:, 0], 0.0)
        w = x[:n]
        wo, wsim = w[:len(obs)], w[len(obs):]
        avg = wo @ Z[obs - 1]
        if unbias == 0 and skm != 0.0:
            avg += (1.0 - wo.sum()) * skm
        L = np.sqrt(var[p]) if var[p] > 0.0 else 0.0
        for s in range(ns):
            v = avg + wsim @ val[s, sim - 1 - nobs] + L * smp[p, s]
            val[s, p] = min(max(v, blo), bhi)
    out = np.full((nb, ns), np.nan)
    vb = np.full(nb, np.nan)
    out[path] = val.T
    vb[path] = var
    return out, vb
