"""Seeded kriging problems shared by the tests (dict format of krigekit_amd.synth).

Each builder returns a problem small enough for the CPU oracle to finish in
well under a second, chosen to cover one behaviour of the reference's hot
path: plain k-nearest search, octant/quadrant search, leave-one-out,
lattices whose equal distances exercise the kd-tree tie rules, maxdist,
per-block range scale and local nugget, simple kriging with a mean, nested
and product structures (table Level A and Level B), analytic covariance,
observation error variance and Gauss-Legendre block kriging.
"""
import os

import numpy as np

from krigekit_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Gauss-Legendre constants of the reference (gaussian_quadrature.f90:3-4)
GQ_FRAC = np.array([0.4305681558, 0.1699905218, -0.1699905218, -0.4305681558])
GQ_W = np.array([0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226])


def read_csv(name):
    import pandas as pd
    return pd.read_csv(os.path.join(GOLDEN, name))


def _base(name, ndim, obs, val, grid, vgm, nmax=None, sector=False, cv=False, **kw):
    o = dict(coord=obs, value=val, nmax=nmax)
    for key in ("maxdist", "variance", "sk_mean"):
        if key in kw:
            o[key] = kw.pop(key)
    d = dict(name=name, ndim=ndim, cross_validation=cv, obs=[o], grid=None if cv else grid,
             vgm=vgm, search=[dict(ivar=1, sector_search=sector)])
    d.update(kw)
    return d


def lattice(ndim=2, nside=30, qoff=5.0, qstep=5.0, sector=False, nmax=8, cv=False, seed=5):
    """Observations on a 10 m lattice, queries on a lattice offset by qoff:
    many queries have equal distances at the k-th / per-octant boundary."""
    rng = np.random.default_rng(seed)
    ax = np.arange(nside) * 10.0
    obs = np.stack([m.ravel() for m in np.meshgrid(*([ax] * ndim), indexing="ij")], 1)
    val = rng.normal(size=len(obs))
    qa = np.arange(qoff, 10.0 * nside, qstep)
    grid = np.stack([m.ravel() for m in np.meshgrid(*([qa] * ndim), indexing="ij")], 1)
    vgm = [dict(ivar=1, jvar=1, vtype="exp", nugget=0.1, sill=1.0, a_major=60.0)]
    return _base(f"lattice{ndim}d_s{int(sector)}_cv{int(cv)}", ndim, obs, val, grid, vgm, nmax=nmax,
                 sector=sector, cv=cv)


def random_problem(ndim=2, nobs=1500, nq=3000, nmax=12, sector=False, seed=11, vgm=None, **kw):
    rng = np.random.default_rng(seed)
    ext = 5000.0
    obs = rng.uniform(0, ext, (nobs, ndim))
    if ndim == 3:
        obs[:, 2] *= 0.05
    val = np.sin(obs[:, 0] / 700.0) + np.cos(obs[:, 1] / 900.0) + rng.normal(0, 0.1, nobs)
    grid = rng.uniform(0, ext, (nq, ndim))
    if ndim == 3:
        grid[:, 2] *= 0.05
    if vgm is None:
        vgm = [dict(ivar=1, jvar=1, vtype="sph", nugget=0.05, sill=1.0, a_major=1200.0,
                    a_minor1=700.0, a_minor2=100.0, azimuth=25.0)]
    return _base(f"random{ndim}d", ndim, obs, val, grid, vgm, nmax=nmax, sector=sector, **kw)


def ok2d_small(scale=0.1):
    return synth.ok2d(scale=scale)


def ok3d_slice(step=256):
    return synth.ok3d(scale=1.0 / step)


def loo_small(n_uniform=4500, n_dup=500, nmax=16):
    return synth.loo2d(n_uniform=n_uniform, n_dup=n_dup, nmax=nmax)


def nested_levelA():
    """two structures with different anisotropy: per-structure tables"""
    vgm = [dict(ivar=1, jvar=1, vtype="nug", nugget=0.05, sill=0.0, a_major=1.0),
           dict(ivar=1, jvar=1, vtype="sph", sill=0.6, a_major=1500.0, a_minor1=500.0, azimuth=30.0),
           dict(ivar=1, jvar=1, vtype="gau", sill=0.4, a_major=800.0, a_minor1=800.0)]
    p = random_problem(vgm=vgm, seed=21)
    p["name"] = "nested_levelA"
    return p


def product_structure():
    """exponential decay times hole effect (product group), one matrix: Level B"""
    vgm = [dict(ivar=1, jvar=1, vtype="exp", sill=1.0, a_major=2000.0, nugget=0.02),
           dict(ivar=1, jvar=1, vtype="hol", sill=1.0, a_major=900.0, product=True)]
    p = random_problem(vgm=vgm, seed=23)
    p["name"] = "product_levelB"
    return p


def all_shapes():
    out = []
    for vt in ("sph", "exp", "gau", "pow", "bsq", "cir", "lin", "cyc", "dco"):
        vgm = [dict(ivar=1, jvar=1, vtype=vt, nugget=0.1, sill=1.0, a_major=1500.0, a_minor1=900.0,
                    azimuth=10.0)]
        p = random_problem(vgm=vgm, nobs=600, nq=500, nmax=10, seed=31)
        p["name"] = f"shape_{vt}"
        out.append(p)
    return out


def local_params():
    p = random_problem(seed=41)
    rng = np.random.default_rng(42)
    n = len(p["grid"])
    p["rangescale"] = rng.uniform(0.7, 1.5, n)
    p["localnugget"] = rng.uniform(0.0, 0.2, n)
    p["name"] = "rangescale_localnugget"
    return p


def obs_error():
    p = random_problem(seed=43)
    rng = np.random.default_rng(44)
    p["obs"][0]["variance"] = rng.uniform(0.0, 0.05, len(p["obs"][0]["value"]))
    p["name"] = "obs_variance"
    return p


def simple_kriging():
    p = random_problem(seed=45)
    p["unbias"] = 0
    p["obs"][0]["sk_mean"] = 0.3
    p["name"] = "simple_kriging"
    return p


def maxdist_problem():
    p = random_problem(seed=47, nq=2000)
    p["obs"][0]["maxdist"] = 180.0
    p["name"] = "maxdist"
    return p


def gq_blocks(ndim=2, nb=300, seed=51):
    """block_type=-4: 4^ndim Gauss-Legendre sub-points per block"""
    p = random_problem(ndim=ndim, nq=nb, seed=seed)
    bs = np.array([120.0, 80.0, 10.0][:ndim])
    p["blocksize"] = bs
    sub, w = gq_subpoints(p["grid"], bs)
    p["nsub"] = 4 ** ndim
    p["sub_coord"] = sub
    p["sub_weight"] = w
    p["name"] = f"gq{ndim}d"
    return p


def gq_subpoints(centres, blocksize):
    """sub-points and weights as set_gaussian_quadrature builds them
    (gaussian_quadrature.f90:10-41): index i fastest, then j, then k."""
    centres = np.asarray(centres, dtype=np.float64)
    nd = centres.shape[1]
    ngq = 4 ** nd
    ii = np.array([[g % 4, (g // 4) % 4, g // 16] for g in range(ngq)])[:, :nd]
    off = GQ_FRAC[ii] * np.asarray(blocksize)[None, :]
    w = np.prod(GQ_W[ii], axis=1)
    sub = (centres[:, None, :] + off[None, :, :]).reshape(-1, nd)
    return sub, np.tile(w, len(centres))


def pc2d(nmax=62):
    d = read_csv("pc2d.csv")
    g = read_csv("grid2d.csv")
    vgm = [dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.12, a_major=5000.0)]
    return _base("pc2d", 2, d[["x", "y"]].values, d["pc"].values, g[["x", "y"]].values, vgm,
                 nmax=nmax)


def pc2d_loo():
    d = read_csv("pc2d.csv")
    vgm = [dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.12, a_major=5000.0)]
    return _base("pc2d_loo", 2, d[["x", "y"]].values, d["pc"].values, None, vgm, nmax=None, cv=True)


def onedim(coords, values, target, vgm, nmax=50):
    obs = np.asarray(coords, dtype=float)[:, None]
    grid = np.asarray([[float(t)] for t in np.atleast_1d(target)])
    return _base("1d", 1, obs, np.asarray(values, dtype=float), grid,
                 [dict(ivar=1, jvar=1, **vgm)], nmax=nmax)


def loo_grad_problem(n_uniform=900, n_dup=100, nmax=16, shrink=0.15):
    """C4-shaped leave-one-out problem (anisotropic spherical, jittered
    duplicates), shrunk so that neighbour lags span the variogram ranges"""
    p = synth.loo2d(n_uniform=n_uniform, n_dup=n_dup, nmax=nmax)
    p["obs"][0]["coord"] = p["obs"][0]["coord"] * shrink
    p["name"] = "loo_grad"
    return p


def ck_small(std=True, gq=True, nb=200, seed=7, unbias=1, sk=(0.0, 0.0)):
    """two-variable LMC cokriging (krigekit_amd.synth.ck2d at 0.64% of the
    observations, same densities), a random subset of its 20 m blocks"""
    s = synth.ck2d(scale=0.0064, std_ck=std)
    rng = np.random.default_rng(seed)
    sel = rng.choice(len(s["grid"]), nb, replace=False)
    s["grid"] = s["grid"][sel]
    if gq:
        s["sub_coord"] = s["sub_coord"].reshape(-1, 16, 2)[sel].reshape(-1, 2)
        s["sub_weight"] = s["sub_weight"].reshape(-1, 16)[sel].ravel()
    else:
        for k in ("sub_coord", "sub_weight", "blocksize"):
            s.pop(k)
        s["nsub"] = 1
    s["unbias"] = unbias
    for o, m in zip(s["obs"], sk):
        o["sk_mean"] = m
    s["name"] = f"ck_std{int(std)}_gq{int(gq)}_ub{unbias}"
    return s


def ck3_small(nb=150, seed=9):
    """three variables (third = noisy mix of the first two), point support"""
    s = ck_small(std=True, gq=False, nb=nb, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x3 = rng.uniform(0.0, 1600.0, (600, 2))
    z3 = np.sin(x3[:, 0] / 300.0) + rng.normal(0.0, 0.2, len(x3))
    s["nvar"] = 3
    s["obs"].append(dict(coord=x3, value=z3, nmax=12))
    s["search"].append(dict(ivar=3, sector_search=False))
    sills = {(1, 1): 1.0, (1, 2): 0.6, (1, 3): 0.3, (2, 2): 0.8, (2, 3): 0.2, (3, 3): 0.7}
    s["vgm"] = [dict(ivar=i, jvar=j, vtype="exp", sill=v, a_major=700.0, nugget=0.05 if i == j else 0.0)
                for (i, j), v in sills.items()]
    s["name"] = "ck3_point"
    return s


def ck_empty_var(std=True):
    """a maxdist on the secondary variable leaves some blocks without
    secondary neighbours: under std_ck its constraint row is empty, the Schur
    complement singular, and the reference's fallback ends in NaN estimates,
    variances 0 and n_fail; non-standard cokriging solves those blocks"""
    s = ck_small(std=std, gq=False, nb=80, seed=15)
    s["obs"][1]["maxdist"] = 25.0
    s["name"] = f"ck_empty_var_std{int(std)}"
    return s


def ck_indefinite():
    """a cross sill above the geometric mean of the sills with collocated data
    makes the first block's K indefinite (negative Cholesky pivot): the
    reference solves it with the symmetric-indefinite factorisation of the
    augmented system; the second block's variables lie beyond each other's
    range and factor by Cholesky"""
    a = np.array([[0.0, 0.0], [30.0, 0.0], [0.0, 30.0]])
    b1 = a + [10000.0, 0.0]
    b2 = b1 + [0.0, 500.0]
    x1, x2 = np.vstack([a, b1]), np.vstack([a, b2])
    return dict(name="ck_indefinite", ndim=2, nvar=2, std_ck=True, unbias=1,
                obs=[dict(coord=x1, value=np.arange(6.0), nmax=3),
                     dict(coord=x2, value=np.arange(6.0)[::-1], nmax=3)],
                grid=np.array([[10.0, 10.0], [10010.0, 250.0]]),
                vgm=[dict(ivar=i, jvar=j, vtype="sph", sill=s, a_major=100.0) for (i, j), s in
                     (((1, 1), 1.0), ((1, 2), 2.0), ((2, 2), 1.0))],
                search=[dict(ivar=1, sector_search=False), dict(ivar=2, sector_search=False)])


def ck_wide():
    """two variables with 40 + 60 neighbours: 100 data rows, the 128-column
    instance of the block solver"""
    s = ck_small(std=True, gq=True, nb=60, seed=17)
    s["obs"][0]["nmax"] = 40
    s["obs"][1]["nmax"] = 60
    s["name"] = "ck_wide_100rows"
    return s


def ck4_small(nb=100, seed=19):
    """four variables (the third and fourth noisy mixes of the first two),
    point support, standard cokriging"""
    s = ck3_small(nb=nb, seed=seed)
    rng = np.random.default_rng(seed + 2)
    x4 = rng.uniform(0.0, 1600.0, (500, 2))
    z4 = np.cos(x4[:, 1] / 250.0) + rng.normal(0.0, 0.2, len(x4))
    s["nvar"] = 4
    s["obs"].append(dict(coord=x4, value=z4, nmax=10))
    s["search"].append(dict(ivar=4, sector_search=False))
    sills = {(1, 1): 1.0, (1, 2): 0.6, (1, 3): 0.3, (1, 4): 0.2, (2, 2): 0.8, (2, 3): 0.2, (2, 4): 0.1,
             (3, 3): 0.7, (3, 4): 0.15, (4, 4): 0.6}
    s["vgm"] = [dict(ivar=i, jvar=j, vtype="exp", sill=v, a_major=700.0, nugget=0.05 if i == j else 0.0)
                for (i, j), v in sills.items()]
    s["name"] = "ck4_point"
    return s


def ck_near_duplicates(nb=120, seed=21):
    """two-variable point cokriging with 40 primary observations duplicated
    0.01 m away and no nugget: the FP32 pivot ratio of such a pair (about
    3e-5) falls below 1e-3, so the FP32+refine path re-solves those blocks in
    FP64"""
    s = ck_small(std=True, gq=False, nb=nb, seed=seed)
    rng = np.random.default_rng(seed + 1)
    o = s["obs"][0]
    pick = rng.choice(len(o["coord"]), 40, replace=False)
    ang = rng.uniform(0, 2 * np.pi, len(pick))
    dup = o["coord"][pick] + 0.01 * np.stack([np.cos(ang), np.sin(ang)], 1)
    o["coord"] = np.concatenate([o["coord"], dup])
    o["value"] = np.concatenate([o["value"], o["value"][pick] + rng.normal(0, 0.01, len(pick))])
    # blocks centred on the duplicated points, so their pairs enter the systems
    s["grid"] = np.concatenate([s["grid"], o["coord"][pick[:60]] + 3.0])
    s["name"] = "ck_near_duplicates"
    return s


def _one_var(s):
    """variable 1 of a cokriging problem on its own"""
    s["nvar"] = 1
    s["obs"] = s["obs"][:1]
    s["search"] = s["search"][:1]
    s["vgm"] = [v for v in s["vgm"] if v["ivar"] == 1 and v["jvar"] == 1]
    return s


def drift_ked(nb=150, seed=23):
    """external-drift kriging: one drift function (a smooth trend surface),
    ordinary unbiasedness (p = 2 constraint rows)"""
    s = _one_var(ck_small(std=True, gq=False, nb=nb, seed=seed))
    trend = lambda x: 0.8 * np.sin(x[:, 0] / 450.0) + 0.5 * np.cos(x[:, 1] / 620.0)  # noqa: E731
    o = s["obs"][0]
    o["drift"] = trend(o["coord"])[:, None]
    o["value"] = o["value"] + 2.0 * o["drift"][:, 0]
    s["grid_drift"] = trend(s["grid"])[:, None]
    s["ndrift"] = 1
    s["name"] = "drift_ked"
    return s


def drift_uk(nb=120, seed=25):
    """universal kriging: the coordinates (km) as two drift functions with
    ordinary unbiasedness, on 4 x 4 Gauss-Legendre blocks (block drift at
    the centre)"""
    s = _one_var(ck_small(std=True, gq=True, nb=nb, seed=seed))
    o = s["obs"][0]
    o["drift"] = o["coord"] / 1000.0
    o["value"] = o["value"] + 0.5 * o["drift"][:, 0] - 0.3 * o["drift"][:, 1]
    s["grid_drift"] = s["grid"] / 1000.0
    s["ndrift"] = 2
    s["name"] = "drift_uk_blocks"
    return s


def drift_sk(nb=100, seed=27):
    """two drift functions without the unbiasedness row (unbias=0), as the
    reference's own drift test (tests/test_api.py:290-306)"""
    s = drift_uk(nb=nb, seed=seed)
    for k in ("sub_coord", "sub_weight", "blocksize"):
        s.pop(k, None)
    s["nsub"] = 1
    s["unbias"] = 0
    s["name"] = "drift_sk"
    return s


def drift_ck(nb=100, seed=29):
    """two-variable standard cokriging with one drift function on both
    variables (p = 1 + 2 constraint rows), the block drift broadcast to both
    targets"""
    s = ck_small(std=True, gq=False, nb=nb, seed=seed)
    trend = lambda x: x[:, 0] / 1000.0 - 0.5 * x[:, 1] / 1000.0  # noqa: E731
    for o in s["obs"]:
        o["drift"] = trend(o["coord"])[:, None]
        o["value"] = o["value"] + o["drift"][:, 0]
    s["grid_drift"] = trend(s["grid"])[:, None]
    s["ndrift"] = 1
    s["name"] = "drift_ck"
    return s


def wcorr_ok(nb=150, seed=31):
    """weight correction (negative weights clipped, the rest rescaled to sum 1)
    on ordinary kriging with clustered data, where screening gives negative
    weights"""
    s = _one_var(ck_small(std=True, gq=False, nb=nb, seed=seed))
    s["weight_correction"] = True
    s["name"] = "wcorr_ok"
    return s


def wcorr_drift(nb=120, seed=33):
    """weight correction with one external drift function"""
    s = drift_ked(nb=nb, seed=seed)
    s["weight_correction"] = True
    s["name"] = "wcorr_drift"
    return s


def _add_grads(s, ivar, n, seed, sep=40.0, var=0.01, trend=None):
    """n increment observations z(x1) - z(x2) of variable ivar: pairs sep m
    apart inside the domain, values from a smooth field plus noise; with
    trend, their drift values trend(x1) - trend(x2)"""
    rng = np.random.default_rng(seed)
    x1 = rng.uniform(100.0, 1500.0, (n, 2))
    ang = rng.uniform(0.0, 2.0 * np.pi, n)
    x2 = x1 + sep * np.stack([np.cos(ang), np.sin(ang)], 1)
    f = lambda x: np.sin(x[:, 0] / 300.0) + 0.5 * np.cos(x[:, 1] / 410.0)  # noqa: E731
    g = dict(ivar=ivar, coord1=x1, coord2=x2, value=f(x1) - f(x2) + rng.normal(0.0, 0.02, n),
             variance=np.full(n, var))
    if trend is not None:
        g["drift"] = (trend(x1) - trend(x2))[:, None]
    s.setdefault("grad", []).append(g)
    return s


def grad_cases():
    """increment (gradient-pair) constraints (kriging.F90:796-840, 904-920):
    ordinary kriging; with an external drift; standard cokriging with
    increments of both variables; non-standard cokriging on Gauss-Legendre
    blocks; simple kriging with a mean"""
    a = _add_grads(_one_var(ck_small(std=True, gq=False, nb=120, seed=41)), 1, 8, 42)
    a["name"] = "grad_ok"
    b = drift_ked(nb=100, seed=43)
    trend = lambda x: 0.8 * np.sin(x[:, 0] / 450.0) + 0.5 * np.cos(x[:, 1] / 620.0)  # noqa: E731
    b = _add_grads(b, 1, 6, 44, trend=trend)
    b["name"] = "grad_ked"
    c = _add_grads(_add_grads(ck_small(std=True, gq=False, nb=100, seed=45), 2, 6, 46), 1, 4, 47, var=0.0)
    c["name"] = "grad_ck"
    d = _add_grads(ck_small(std=False, gq=True, nb=80, seed=48), 1, 5, 49)
    d["name"] = "grad_ck_nonstd_blocks"
    e = _add_grads(_one_var(ck_small(True, False, nb=100, seed=50, unbias=0, sk=(0.2, -0.1))), 1, 6, 51)
    e["name"] = "grad_sk"
    return [a, b, c, d, e]


def drift_cases():
    return [drift_ked(), drift_uk(), drift_sk(), drift_ck(), wcorr_ok(), wcorr_drift()]


def ck_cases():
    return [ck_small(True, True), ck_small(True, False), ck_small(False, True), ck_small(False, False),
            ck_small(True, False, unbias=0, sk=(0.2, -0.1)), ck3_small(), ck_empty_var(True), ck_empty_var(False),
            ck_indefinite(), ck_wide(), ck4_small()]


def st_small(nstation=60, ndays=60, side=170.0, grid_side=8, ntimes=3, nmax=24, maxdist=40.0, sector=False,
             cv=False, name="st_small", seed=1005):
    """space-time product-sum problem at C5 densities (60 stations x 60 days)"""
    p = synth.st(nstation=nstation, ndays=ndays, side=side, grid_side=grid_side, ntimes=ntimes, nmax=nmax,
                 maxdist=maxdist, seed=seed)
    p["search"][0]["sector_search"] = sector
    if sector:
        p["obs"][0]["nmax"] = 3
    if cv:
        p["grid"] = None
        p["cross_validation"] = True
    p["name"] = name
    return p


def st_cases():
    return [st_small(), st_small(sector=True, name="st_octant"), st_small(nmax=8, maxdist=25.0, name="st_maxdist"),
            st_small(nstation=6, ndays=9, nmax=200, maxdist=500.0, name="st_allobs"),
            st_small(nstation=30, ndays=20, cv=True, nmax=12, name="st_loo")]


def st_sum(base, sill_st, transform="lin", at=20.0, time_nugget=0.0, time_sill=1.0, vgm=None, name="st_sum"):
    """a space-time problem switched to the sum-metric model
    (variogram_st.f90:178-194): joint sills per spatial structure and the
    temporal transform f(dt)"""
    p = dict(base)
    p.update(st_model="sum_metric", transform=transform, at=at, time_nugget=time_nugget, time_sill=time_sill,
             sill_st=list(sill_st), name=name)
    if vgm is not None:
        p["vgm"] = vgm
    return p


def st_sum_cases():
    """sum-metric variants: linear transform that saturates within the data's
    time span (one spatial structure: a composite table, analytic joint term);
    bounded transform with a temporal nugget over two spatial structures of
    different anisotropy (per-structure tables, joint term from them); a
    nugget structure with joint sill 0 beside a spherical one (the nested
    case of tests/test_spacetime_kriging.py:474-492); octant search"""
    two = [dict(ivar=1, jvar=1, vtype="exp", nugget=0.05, sill=0.6, a_major=50.0, a_minor1=30.0, azimuth=30.0),
           dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.3, a_major=90.0)]
    nug = [dict(ivar=1, jvar=1, vtype="nug", nugget=0.1, sill=0.0, a_major=1.0),
           dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=0.7, a_major=80.0, a_minor1=60.0)]
    return [st_sum(st_small(), [0.3], name="stsm_linear"),
            st_sum(st_small(nmax=20), [0.25, 0.15], transform="bounded", at=10.0, time_nugget=0.1, time_sill=0.9,
                   vgm=two, name="stsm_bounded_tabA"),
            st_sum(st_small(nmax=16), [0.0, 0.3], transform="sph", at=30.0, vgm=nug, name="stsm_nested_nug"),
            st_sum(st_small(sector=True), [0.4], transform="gau", at=15.0, time_sill=2.0, name="stsm_octant")]


def sgsim_case(name, ndim=2, nobs=250, side=1000.0, ngrid=12, nmax=12, nsim=3, sector=False, maxdist=None,
               unbias=1, sk_mean=0.0, bounds=None, on_obs=0, obs_var=None, vgm=None, seed=2024):
    """sequential simulation with a given path and draws: observations
    uniform in a square (cube 10x flatter in z), values a sinusoid field plus
    noise; targets an ngrid^ndim lattice of cell centres, the first on_obs of
    them moved onto observations (the co-location rule).  Targets on error-
    free observations have a variance that is zero up to rounding, whose
    square root (about 1e-8) scales the draw: such cases give the
    observations a measurement variance."""
    rng = np.random.default_rng(seed)
    lo = np.zeros(ndim)
    hi = np.full(ndim, side)
    if ndim == 3:
        hi[2] = side / 10.0
    x = rng.uniform(lo, hi, (nobs, ndim))
    v = np.sin(x[:, 0] / 150.0) + np.cos(x[:, 1] / 230.0) + rng.normal(0.0, 0.1, nobs)
    axes = [(np.arange(ngrid) + 0.5) * (hi[k] / ngrid) for k in range(ndim)]
    g = np.stack([m.ravel() for m in np.meshgrid(*axes, indexing="ij")], 1)
    if on_obs:
        g[:on_obs] = x[:on_obs]
    nb = len(g)
    vg = vgm or [dict(ivar=1, jvar=1, vtype="sph", nugget=0.05, sill=1.0, a_major=400.0, a_minor1=250.0,
                      azimuth=30.0)]
    spec = _base(name, ndim, x, v, g, vg, nmax=nmax, maxdist=maxdist)
    spec["obs"][0]["sk_mean"] = sk_mean
    if obs_var is not None:
        spec["obs"][0]["variance"] = np.full(nobs, obs_var)
    spec["unbias"] = unbias
    spec["search"][0]["sector_search"] = sector
    if bounds is not None:
        spec["bounds"] = bounds
    spec["nsim"] = nsim
    spec["path"] = rng.permutation(nb) + 1
    spec["sample"] = rng.normal(size=(nb, nsim))
    return spec


def sgsim_cases():
    return [sgsim_case("sgs_2d"),
            sgsim_case("sgs_octant", nmax=3, sector=True, seed=2025),
            sgsim_case("sgs_3d_maxdist", ndim=3, ngrid=6, nobs=400, nmax=16, maxdist=300.0, seed=2026,
                       vgm=[dict(ivar=1, jvar=1, vtype="exp", nugget=0.02, sill=0.8, a_major=500.0, a_minor1=300.0,
                                 a_minor2=60.0, azimuth=20.0)]),
            sgsim_case("sgs_sk_bounds", unbias=0, sk_mean=0.4, bounds=(-0.8, 1.2), nsim=4, seed=2027),
            sgsim_case("sgs_on_obs", on_obs=20, nmax=10, obs_var=0.02, seed=2028)]


def ladder():
    """the reference's failure ladder (kriging_base.F90:3244-3398) on exact
    zeros: observations on a 10 m lattice, a spherical model whose range
    (4 m) is below the spacing, so every off-diagonal covariance is exactly 0,
    and a local nugget of -1 cancelling C(0) = 1, so the diagonal is the
    observation variance (powers of two, so the factorisations are exact).
    Every block takes the 4 corners of one lattice cell; by quadrant of the
    lattice, a cell has 4 positive variances (Cholesky), one zero (Cholesky
    fails, the symmetric-indefinite factorisation of the bordered system
    succeeds), one positive (that factorisation is exactly singular; the
    1e-6 mean-diagonal jitter repairs it) or none (no retry when the diagonal
    sums to 0: NaN estimate, variance max(NaN, 0) = 0, n_fail)."""
    ax = np.arange(8) * 10.0
    obs = np.stack([m.ravel() for m in np.meshgrid(ax, ax, indexing="ij")], 1)
    rng = np.random.default_rng(71)
    val = rng.normal(size=len(obs))
    i, j = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    i, j = i.ravel(), j.ravel()
    q = (i >= 4).astype(int) + 2 * (j >= 4).astype(int)
    odd = (i % 2 == 1) & (j % 2 == 1)
    even = (i % 2 == 0) & (j % 2 == 0)
    pos = np.where(q == 0, True, np.where(q == 1, ~odd, np.where(q == 2, even, False)))
    var = np.where(pos, rng.choice([0.25, 0.5], len(obs)), 0.0)
    # queries at cell centres (zero right-hand sides) and 2.9 m from a lattice
    # node (one non-zero right-hand side), both seeing that cell's 4 corners
    c = np.arange(7) * 10.0 + 5.0
    g1 = np.stack([m.ravel() for m in np.meshgrid(c, c, indexing="ij")], 1)
    g2 = np.stack([m.ravel() for m in np.meshgrid(c - 3.5, c + 2.5, indexing="ij")], 1)
    grid = np.concatenate([g1, g2])
    n = len(grid)
    return _base("ladder", 2, obs, val, grid, [dict(ivar=1, jvar=1, vtype="sph", nugget=0.0, sill=1.0, a_major=4.0)],
                 nmax=4, variance=var, localnugget=np.full(n, -1.0))


def bounds_case():
    """estimates clipped to the bounds (kriging_base.F90:3497-3498)"""
    p = random_problem(seed=81, nq=1500)
    p["bounds"] = (-0.4, 0.9)
    p["name"] = "bounds"
    return p


def aniso_search_case(ndim=2, sector=False):
    """anisotropic_search: the kd-tree on rotated, axis-scaled coordinates and
    rotated query lags (kriging.F90:354-357, 371-381, 518-522)"""
    p = random_problem(ndim=ndim, seed=83 + ndim, nq=1200, nmax=3 if sector else 10, sector=sector)
    p["anisotropic_search"] = True
    if ndim == 2:
        p["search"][0].update(anis1=0.4, azimuth=35.0)
    else:
        p["search"][0].update(anis1=0.5, anis2=0.1, azimuth=35.0, dip=10.0, plunge=5.0)
    p["name"] = f"aniso_search{ndim}d_s{int(sector)}"
    return p


def near_duplicates(nmax=12):
    """jittered duplicate observations (0.01 m apart) and a 1e-4 nugget: the
    FP32 pivot ratio of such a pair falls far below 1e-3, so the
    FP32+refine path re-solves those systems in FP64"""
    rng = np.random.default_rng(87)
    base = rng.uniform(0, 3000.0, (700, 2))
    dup = base[rng.choice(len(base), 40, replace=False)]
    ang = rng.uniform(0, 2 * np.pi, len(dup))
    dup = dup + 0.01 * np.stack([np.cos(ang), np.sin(ang)], 1)
    obs = np.concatenate([base, dup])
    val = np.sin(obs[:, 0] / 500.0) + np.cos(obs[:, 1] / 700.0) + rng.normal(0, 0.05, len(obs))
    grid = rng.uniform(0, 3000.0, (1500, 2))
    vgm = [dict(ivar=1, jvar=1, vtype="sph", nugget=1e-4, sill=1.0, a_major=1500.0, a_minor1=900.0, azimuth=20.0)]
    return _base("near_duplicates", 2, obs, val, grid, vgm, nmax=nmax)


def big_systems():
    """neighbourhoods above 64 rows (solved by the one-workgroup bordered
    solver): plain 80 nearest, 3-D octants of 12 (96 rows), and a maxdist
    that spreads the counts over every solve width and beyond 64"""
    return [_named(random_problem(2, nobs=1500, nq=400, nmax=80, seed=91), "big80_2d"),
            _named(random_problem(3, nobs=3000, nq=300, nmax=12, sector=True, seed=93), "big96_3d_octant"),
            _named(random_problem(2, nobs=1500, nq=500, nmax=110, seed=95, maxdist=640.0), "big_mixed_2d")]


def _named(p, name):
    p["name"] = name
    return p


def parity_cases():
    """small problems compared against the oracle and the compiled reference"""
    cases = [ok2d_small(), ok3d_slice(), loo_small(),
             lattice(2), lattice(2, sector=True, nmax=3), lattice(3, nside=12, qoff=2.5),
             lattice(3, nside=12, qoff=2.5, sector=True, nmax=2), lattice(2, qoff=0, qstep=1, cv=True, nmax=7),
             lattice(3, nside=10, qoff=0, qstep=1, cv=True, sector=True, nmax=2),
             _named(random_problem(3, sector=True, nmax=3), "random3d_octant"), nested_levelA(), product_structure(),
             local_params(), obs_error(), simple_kriging(), maxdist_problem(),
             ladder(), bounds_case(), aniso_search_case(2), aniso_search_case(3),
             aniso_search_case(3, sector=True), near_duplicates()] + big_systems()
    return cases
