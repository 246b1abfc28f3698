"""Seeded synthetic meshes and cameras for the benchmark workloads.

Only ``Generator.random()`` uniforms are drawn (the PCG64 bit stream is
stable across numpy versions, the distribution methods are not); every other
distribution is built here from them.  Workloads:

* ``uniform_mesh``   N points uniform in [0, L)^3, rho = exp(Z), Z ~ N(0, 1).
* ``clustered_mesh`` 70% of the points in 200 clumps (power-law counts of
  slope -2, r^-2 profiles with scale radii log-uniform in [0.05, 5],
  isotropic), 30% uniform background, wrapped into the periodic box;
  generators closer than ``d_min`` to an earlier one are dropped so cell
  sizes stay within about three decades; rho = 32 / (4/3 pi d32^3) from the
  periodic 32-NN distance.
* ``rgba_transfer``  R, G, B, alpha from three Gaussian bumps in log10 rho
  (volume rendering), alpha = 1 - exp(-kappa rho).
* ``camera_rays``    isotropic viewing directions through the box centre.
"""
from __future__ import annotations

import hashlib

import numpy as np


def _u(rng, n):
    return rng.random(n)


def normal(rng, n):
    """Standard normals by Box-Muller from uniforms."""
    m = (n + 1) // 2
    u1 = 1.0 - _u(rng, m)  # (0, 1]
    u2 = _u(rng, m)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def _isotropic(rng, n):
    mu = 2.0 * _u(rng, n) - 1.0
    phi = 2.0 * np.pi * _u(rng, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    return np.stack([s * np.cos(phi), s * np.sin(phi), mu], axis=1)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def uniform_mesh(n: int = 100_000, box: float = 100.0, seed: int = 1):
    rng = np.random.default_rng(seed)
    pos = box * _u(rng, 3 * n).reshape(n, 3)
    rho = np.exp(normal(rng, n))
    return pos, rho


def _largest_remainder(weights, total):
    q = weights / weights.sum() * total
    base = np.floor(q).astype(np.int64)
    short = total - base.sum()
    order = np.argsort(-(q - base), kind="stable")
    base[order[:short]] += 1
    return base


def clustered_mesh(n: int = 2_000_000, box: float = 100.0, seed: int = 2,
                   n_clumps: int = 200, frac_clumped: float = 0.7,
                   u_min: float = 0.01, d_min: float = 1e-3,
                   with_rho: bool = True, workers: int = -1):
    """Clustered periodic mesh; returns (pos (n, 3), rho (n,) or None)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    n_gen = int(n * 1.15)  # candidates; close pairs are dropped below
    n_cl = int(round(frac_clumped * n_gen))
    centres = box * _u(rng, 3 * n_clumps).reshape(n_clumps, 3)
    w = 1.0 / (1.0 - _u(rng, n_clumps))  # Pareto(1): N(>m) ~ 1/m, slope -2
    w = np.minimum(w, w.sum() / 20.0)
    counts = _largest_remainder(w, n_cl)
    r_s = 0.05 * (100.0 ** _u(rng, n_clumps))  # log-uniform in [0.05, 5]
    which = np.repeat(np.arange(n_clumps), counts)
    rad = r_s[which] * (u_min + (1.0 - u_min) * _u(rng, n_cl))  # M(<r) ~ r
    clump = centres[which] + rad[:, None] * _isotropic(rng, n_cl)
    bg = box * _u(rng, 3 * (n_gen - n_cl)).reshape(-1, 3)
    cand = np.mod(np.concatenate([clump, bg]), box)
    # shuffle deterministically so the kept prefix mixes clumps and background
    perm = np.argsort(_u(rng, n_gen), kind="stable")
    cand = cand[perm]
    cand[cand >= box] = 0.0  # mod can round up to exactly box
    tree = cKDTree(cand, boxsize=box)
    pairs = tree.query_pairs(d_min, output_type="ndarray")
    drop = np.zeros(n_gen, dtype=bool)
    if len(pairs):
        drop[np.maximum(pairs[:, 0], pairs[:, 1])] = True
    kept = cand[~drop]
    if len(kept) < n:
        raise RuntimeError("clustered_mesh: too many close pairs; lower density")
    pos = np.ascontiguousarray(kept[:n])
    rho = None
    if with_rho:
        t2 = cKDTree(pos, boxsize=box)
        d32, _ = t2.query(pos, k=33, workers=workers)
        d = d32[:, 32]
        rho = 32.0 / ((4.0 * np.pi / 3.0) * d ** 3)
    return pos, rho


def rgba_transfer(rho, kappa: float):
    """(R, G, B, alpha) per cell: three Gaussian bumps in x = log10 rho at the
    50th, 90th and 99th percentiles (0.15 dex) with fixed colours."""
    x = np.log10(rho)
    cents = np.percentile(x, [50.0, 90.0, 99.0])
    cols = np.array([[0.2, 0.4, 1.0], [0.2, 1.0, 0.3], [1.0, 0.3, 0.1]])
    g = np.exp(-0.5 * ((x[:, None] - cents[None, :]) / 0.15) ** 2)
    rgb = np.clip(g @ cols, 0.0, 1.0)
    alpha = 1.0 - np.exp(-kappa * rho)
    return np.ascontiguousarray(np.column_stack([rgb, alpha]))


def kappa_for_median_opacity(rho, cell_ids, ds, offsets, target=0.5):
    """kappa such that the median over rays of 1 - exp(-sum alpha ds) is
    `target`, given traced segments (cell ids into rho, ds, offsets)."""
    ray = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))

    def med(k):
        a = 1.0 - np.exp(-k * rho[cell_ids])
        tau = np.bincount(ray, weights=a * ds, minlength=len(offsets) - 1)
        return np.median(1.0 - np.exp(-tau))

    lo, hi = 1e-12, 1e12
    for _ in range(200):
        mid = np.sqrt(lo * hi)
        if med(mid) < target:
            lo = mid
        else:
            hi = mid
    return np.sqrt(lo * hi)


def camera_views(rng_seed: int, n_views: int, box: float = 100.0,
                 sizes=(64, 128, 256, 512, 1024), extent_range=(10.0, 100.0)):
    """Per-view grid.generate_projection_grid arguments: views through the
    box centre with isotropic directions d (yaw = atan2(d_y, d_x),
    pitch = arccos(d_z), roll = 0, so that R z = d), extents uniform in
    extent_range and image sizes drawn from `sizes`."""
    rng = np.random.default_rng(rng_seed)
    dirs = _isotropic(rng, n_views)
    ext = extent_range[0] + (extent_range[1] - extent_range[0]) * _u(rng, n_views)
    pick = np.minimum((len(sizes) * _u(rng, n_views)).astype(int), len(sizes) - 1)
    c = [box / 2] * 3
    views = []
    for v in range(n_views):
        d = dirs[v]
        e = ext[v]
        views.append(dict(extent=[box / 2 - e / 2, box / 2 + e / 2],
                          nres=int(sizes[pick[v]]), bounds=[0.0, box], center=c,
                          yaw=float(np.arctan2(d[1], d[0])),
                          pitch=float(np.arccos(np.clip(d[2], -1.0, 1.0))), roll=0.0))
    return views


def camera_rays(rng_seed: int, n_views: int, box: float = 100.0,
                sizes=(64, 128, 256, 512, 1024), extent_range=(10.0, 100.0)):
    """(start, end) host rays of camera_views(...), each (n_i, 3), as the
    reference's grid recipe generates them."""
    from . import grid

    views = []
    for v in camera_views(rng_seed, n_views, box, sizes, extent_range):
        s, t = grid.generate_projection_grid(**v)
        views.append((s.reshape(-1, 3), t.reshape(-1, 3)))
    return views
