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


def ck_cases():
    return [ck_small(True, True), ck_small(True, False), ck_small(False, True), ck_small(False, False),
            ck_small(True, False, unbias=0, sk=(0.2, -0.1)), ck3_small()]


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
             local_params(), obs_error(), simple_kriging(), maxdist_problem()]
    return cases
