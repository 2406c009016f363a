"""Seeded synthetic workloads for the kriging hot path.

Each generator returns a dict describing one kriging problem: observations,
target nodes, variogram structures and search settings, in the argument
conventions of the ``Kriging`` class (coordinates (n, ndim), 1-based ivar).
All randomness comes from ``numpy.random.default_rng(seed)`` (PCG64).

Workloads (sizes are the full ones; ``scale`` shrinks the target grid for
parity tests while keeping observations, model and search identical):

- ``ok2d``   2-D ordinary kriging: 2,000 obs uniform in [0, 10000]^2 m, value =
  8-term random-phase sinusoid sum (wavelengths 500-4000 m) + N(0, 0.1^2);
  200 x 200 nodes at 50 m; spherical, sill 1.0, nugget 0.05, range 1500 m;
  nmax = 16, no sectors.
- ``ok3d``   3-D ordinary kriging: 50,000 obs from 20 Gaussian clusters
  (sigma 300/300/10 m, reflected into a 12.8 km x 12.8 km x 320 m box);
  256 x 256 x 64 = 4,194,304 nodes at (50, 50, 5) m; exponential, sill 1.0,
  nugget 0.02, ranges 2000/800/50 m, azimuth 30 deg; octant search with
  at most 4 per octant (<= 32 neighbours from 128 candidates).
- ``loo2d``  leave-one-out cross-validation: 90,000 uniform obs in
  [0, 20000]^2 + 10,000 near-duplicates displaced 0.01 m; anisotropic
  spherical (nugget 0.1, sill 1.0, ranges 1500/600, azimuth 45 deg); nmax 16.
"""
import numpy as np


def _sinusoid_field_2d(rng, xy, nterm=8, lmin=500.0, lmax=4000.0):
    lam = rng.uniform(lmin, lmax, nterm)
    ang = rng.uniform(0.0, 2 * np.pi, nterm)
    phi = rng.uniform(0.0, 2 * np.pi, nterm)
    u = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    proj = xy @ u.T  # (n, nterm)
    return np.sin(2 * np.pi * proj / lam + phi).sum(axis=1)


def _sinusoid_field_3d(rng, xyz, nterm=8):
    lam_h = rng.uniform(500.0, 4000.0, nterm)
    lam_v = rng.uniform(40.0, 320.0, nterm)
    ang = rng.uniform(0.0, 2 * np.pi, nterm)
    phi = rng.uniform(0.0, 2 * np.pi, nterm)
    u = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    proj = xyz[:, :2] @ u.T
    return np.sin(2 * np.pi * (proj / lam_h + xyz[:, 2:3] / lam_v) + phi).sum(axis=1)


def _grid(shape, spacing, origin):
    axes = [origin[d] + spacing[d] * np.arange(shape[d]) for d in range(len(shape))]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.ravel() for m in mesh], axis=1)


def ok2d(scale=1.0, seed=1001):
    rng = np.random.default_rng(seed)
    n = 2000
    xy = rng.uniform(0.0, 10000.0, (n, 2))
    val = _sinusoid_field_2d(rng, xy) + rng.normal(0.0, 0.1, n)
    ng = max(2, int(round(200 * np.sqrt(scale))))
    grid = _grid((ng, ng), (50.0 * 200 / ng,) * 2, (25.0, 25.0))
    return dict(name="ok2d", ndim=2, obs=[dict(coord=xy, value=val, nmax=16)], grid=grid,
                vgm=[dict(ivar=1, jvar=1, vtype="sph", nugget=0.05, sill=1.0, a_major=1500.0)],
                search=[dict(ivar=1, sector_search=False)])


def ok3d_obs(seed=1002):
    rng = np.random.default_rng(seed)
    box = np.array([12800.0, 12800.0, 320.0])
    sig = np.array([300.0, 300.0, 10.0])
    ncl, n = 20, 50000
    centres = rng.uniform(0.0, 1.0, (ncl, 3)) * box
    lab = rng.integers(0, ncl, n)
    pts = centres[lab] + rng.normal(0.0, 1.0, (n, 3)) * sig
    # reflect into the box
    pts = np.abs(pts)
    pts = np.where(pts > box, 2 * box - pts, pts)
    pts = np.clip(pts, 0.0, box)
    val = _sinusoid_field_3d(rng, pts) + rng.normal(0.0, 0.1, n)
    return rng, pts, val


def ok3d(scale=1.0, seed=1002, grid_shape=(256, 256, 64)):
    """Full size is 256 x 256 x 64 nodes; ``scale`` < 1 keeps every
    ``1/scale``-th node of the same lattice (a strided subset), so parity
    runs see the same geometry at a fraction of the nodes."""
    _, pts, val = ok3d_obs(seed)
    grid = _grid(grid_shape, (50.0, 50.0, 5.0), (25.0, 25.0, 2.5))
    if scale < 1.0:
        step = max(1, int(round(1.0 / scale)))
        grid = grid[::step]
    return dict(name="ok3d", ndim=3, obs=[dict(coord=pts, value=val, nmax=4)], grid=grid,
                vgm=[dict(ivar=1, jvar=1, vtype="exp", nugget=0.02, sill=1.0, a_major=2000.0,
                          a_minor1=800.0, a_minor2=50.0, azimuth=30.0)],
                search=[dict(ivar=1, sector_search=True)])


def loo2d(n_uniform=90000, n_dup=10000, seed=1004, nmax=16):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.0, 20000.0, (n_uniform, 2))
    src = xy[rng.choice(n_uniform, n_dup, replace=False)] if n_dup else xy[:0]
    ang = rng.uniform(0.0, 2 * np.pi, n_dup)
    dup = src + 0.01 * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    pts = np.concatenate([xy, dup])
    val = _sinusoid_field_2d(rng, pts) + rng.normal(0.0, 0.1, len(pts))
    return dict(name="loo2d", ndim=2, cross_validation=True,
                obs=[dict(coord=pts, value=val, nmax=nmax)], grid=None,
                vgm=[dict(ivar=1, jvar=1, vtype="sph", nugget=0.1, sill=1.0, a_major=1500.0,
                          a_minor1=600.0, azimuth=45.0)],
                search=[dict(ivar=1, sector_search=False)])


def _sinusoid_field_2d_fixed(seed, xy, nterm=8, lmin=500.0, lmax=4000.0):
    """a sinusoid-sum field whose wavelengths/phases come from their own seed,
    so that observation sets of different sizes sample the same field"""
    return _sinusoid_field_2d(np.random.default_rng(seed), xy, nterm, lmin, lmax)


def gq_nodes(centres, blocksize):
    """Gauss-Legendre 4-point nodes and weights per block (block_type=-4):
    index i fastest, then j, then k; fractions +-0.4305681558, +-0.1699905218
    of the block size, weights 0.1739274226 and 0.3260725774 per axis"""
    frac = np.array([0.4305681558, 0.1699905218, -0.1699905218, -0.4305681558])
    wgt = np.array([0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226])
    centres = np.asarray(centres, dtype=np.float64)
    nd = centres.shape[1]
    ngq = 4 ** nd
    ii = np.array([[g % 4, (g // 4) % 4, g // 16] for g in range(ngq)])[:, :nd]
    off = frac[ii] * np.asarray(blocksize, dtype=np.float64)[None, :]
    w = np.prod(wgt[ii], axis=1)
    sub = (centres[:, None, :] + off[None, :, :]).reshape(-1, nd)
    return sub, np.tile(w, len(centres))


def ck2d(scale=1.0, seed=1003, grid_side=None, std_ck=True):
    """Two-variable cokriging under a linear model of coregionalisation.

    20,000 primary and 200,000 secondary observations (times `scale`) uniform
    in a square of side 20 km * sqrt(scale); z1 = f + N(0, 0.1^2),
    z2 = 0.7 f + 0.3 g + N(0, 0.1^2) with f, g seeded sinusoid-sum fields.
    LMC: spherical range 1200 m, sills [[1.0, 0.6], [0.6, 0.8]], plus
    Gaussian range 400 m, sills [[0.3, 0.1], [0.1, 0.2]].  Neighbours: 24
    primary + 48 secondary.  Targets: blocks at 20 m spacing (centres
    10 + 20 i), 20 m x 20 m with 4 x 4 Gauss-Legendre nodes."""
    rng = np.random.default_rng(seed)
    side = 20000.0 * np.sqrt(scale)
    n1, n2 = max(int(20000 * scale), 30), max(int(200000 * scale), 60)
    x1 = rng.uniform(0.0, side, (n1, 2))
    x2 = rng.uniform(0.0, side, (n2, 2))
    f1, f2 = _sinusoid_field_2d_fixed(seed + 1, x1), _sinusoid_field_2d_fixed(seed + 1, x2)
    g2 = _sinusoid_field_2d_fixed(seed + 2, x2)
    z1 = f1 + rng.normal(0.0, 0.1, n1)
    z2 = 0.7 * f2 + 0.3 * g2 + rng.normal(0.0, 0.1, n2)
    ns = int(round(side / 20.0)) if grid_side is None else grid_side
    ax = 10.0 + 20.0 * np.arange(ns)
    gx, gy = np.meshgrid(ax, ax, indexing="ij")
    centres = np.stack([gx.ravel(), gy.ravel()], axis=1)
    bs = np.array([20.0, 20.0])
    sub, w = gq_nodes(centres, bs)
    vgm = []
    for (i, j), s1, s2 in (((1, 1), 1.0, 0.3), ((1, 2), 0.6, 0.1), ((2, 2), 0.8, 0.2)):
        vgm.append(dict(ivar=i, jvar=j, vtype="sph", sill=s1, a_major=1200.0))
        vgm.append(dict(ivar=i, jvar=j, vtype="gau", sill=s2, a_major=400.0))
    return dict(name="ck2d", ndim=2, nvar=2, std_ck=std_ck, unbias=1,
                obs=[dict(coord=x1, value=z1, nmax=24), dict(coord=x2, value=z2, nmax=48)],
                grid=centres, blocksize=bs, nsub=16, sub_coord=sub, sub_weight=w, vgm=vgm,
                search=[dict(ivar=1, sector_search=False), dict(ivar=2, sector_search=False)])


def st(nstation=500, ndays=365, side=500.0, grid_side=500, ntimes=30, nmax=48, maxdist=40.0, seed=1005):
    """Space-time product-sum kriging (km and days).

    `nstation` stations uniform in a `side` km square (z = 0) observed daily
    for `ndays` days: value = sin(2 pi t / 365) + a seeded spatial sinusoid-sum
    field (wavelengths 50-400 km) + a per-station AR(1) series (phi 0.7,
    innovations N(0, 0.1^2)).  Targets: a grid_side x grid_side grid at
    side/grid_side km spacing (cell centres) x `ntimes` times 6 + 12 k (integer
    days, so time ties occur).  Model: Cs exponential, nugget 0.05, sill 1.0,
    range 50 km; Ct exponential, sill 0.5, range 5 days; k_ps 0.5.  Search:
    time_at 10 km/day, nmax 48, maxdist 40 km-equivalent (interior nodes get
    48 neighbours, edges and corners fewer)."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.0, side, (nstation, 2))
    field = _sinusoid_field_2d(np.random.default_rng(seed + 1), xy, nterm=8, lmin=50.0, lmax=400.0)
    t = np.arange(ndays, dtype=np.float64)
    ar = np.zeros((nstation, ndays))
    for d in range(1, ndays):
        ar[:, d] = 0.7 * ar[:, d - 1] + rng.normal(0.0, 0.1, nstation)
    val = (np.sin(2 * np.pi * t / 365.0)[None, :] + field[:, None] + ar).ravel()
    coord = np.zeros((nstation * ndays, 4))
    coord[:, 0] = np.repeat(xy[:, 0], ndays)
    coord[:, 1] = np.repeat(xy[:, 1], ndays)
    coord[:, 3] = np.tile(t, nstation)
    h = side / grid_side
    ax = h * (0.5 + np.arange(grid_side))
    times = 6.0 + 12.0 * np.arange(ntimes)
    gx, gy, gt = np.meshgrid(ax, ax, times, indexing="ij")
    grid = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size), gt.ravel()], axis=1)
    return dict(name="st", st=True, ndim=3, k_ps=0.5,
                obs=[dict(coord=coord, value=val, nmax=nmax, maxdist=maxdist)], grid=grid,
                vgm=[dict(ivar=1, jvar=1, vtype="exp", nugget=0.05, sill=1.0, a_major=50.0)],
                vgm_t=[dict(vtype="exp", sill=0.5, at_k=5.0)],
                search=[dict(ivar=1, time_at=10.0, sector_search=False)])
