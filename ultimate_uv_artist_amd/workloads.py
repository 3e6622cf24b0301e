"""Seeded synthetic meshes for the parameterization workloads.

Every generator returns charts as lists of ``(P (k,3) float64, T (t,3) int64)``
with P holding float32-representable coordinates (the meshes are specified in
float32).  Seeds default to the workload number so runs are reproducible, and
``mesh_digest`` gives a SHA-256 over the arrays for the bench record.

Workloads:
  grid_island     64x64 vertex grid on [-1,1]^2, z = 0.3(x^2+y^2), jitter
                  N(0, 0.002) -> 7,938 triangles (single island)
  disk_islands    islands with log-uniform triangle counts; each a Delaunay
                  triangulation of a Poisson-disk set on the unit disk lifted by
                  six Gaussian bumps (512 islands in [200, 40,000] by default)
  sliver_islands  disk islands of 5k-60k triangles with 10% of triangles
                  collapsed to aspect ratio 50-500
"""
from __future__ import annotations

import hashlib

import numpy as np


def grid_island(seed=1, nx=64, jitter=0.002):
    """64x64 grid on [-1,1]^2 lifted to z = 0.3(x^2 + y^2), xyz jitter N(0, jitter)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-1.0, 1.0, nx)
    x, y = np.meshgrid(g, g, indexing="xy")
    P = np.stack([x.ravel(), y.ravel(), 0.3 * (x.ravel() ** 2 + y.ravel() ** 2)], 1)
    P = P + rng.normal(0.0, jitter, P.shape)
    P = P.astype(np.float32).astype(np.float64)
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(nx - 1), indexing="xy")
    v00 = (j * nx + i).ravel()
    v10, v01, v11 = v00 + 1, v00 + nx, v00 + nx + 1
    T = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)]).astype(np.int64)
    return P, T


def _poisson_disk(rng, r, radius, attempts=10):
    """Maximal-ish Poisson-disk sample with spacing r inside a disk.

    Grid dart throwing on cells of size r/sqrt(2) (at most one point per
    cell), processed in 9 phases of cells 3 apart so that darts thrown in the
    same phase can never conflict; each empty cell gets ``attempts`` tries."""
    cs = r / np.sqrt(2.0)
    G = int(np.ceil(2 * radius / cs)) + 1
    W = G + 4                                            # 2-cell pad on every side
    px = np.full(W * W, np.nan)                          # cell -> point (NaN when empty)
    py = np.full(W * W, np.nan)
    ci = np.arange(G)
    lo = ci * cs - radius
    # a cell can hold a point of the disk only if its nearest point is inside
    near = np.clip(0.0, lo, lo + cs)
    inside = (near[:, None] ** 2 + near[None, :] ** 2) <= radius * radius
    gi, gj = np.nonzero(inside)
    phase = (gi % 3) * 3 + gj % 3
    flat = (gi + 2) * W + (gj + 2)
    offs = np.array([a * W + b for a in range(-2, 3) for b in range(-2, 3) if (a, b) != (0, 0)])
    r2 = r * r
    for _ in range(attempts):
        for ph in range(9):
            sel = (phase == ph) & np.isnan(px[flat])
            if not sel.any():
                continue
            cell, ii, jj = flat[sel], gi[sel], gj[sel]
            cx = (ii + rng.random(len(ii))) * cs - radius
            cy = (jj + rng.random(len(jj))) * cs - radius
            good = cx * cx + cy * cy <= radius * radius
            nb = cell[:, None] + offs[None, :]
            d2 = (px[nb] - cx[:, None]) ** 2 + (py[nb] - cy[:, None]) ** 2
            good &= ~(d2 < r2).any(1)                    # NaN neighbours compare False
            px[cell[good]] = cx[good]
            py[cell[good]] = cy[good]
    keep = ~np.isnan(px)
    return np.stack([px[keep], py[keep]], 1)


def disk_island(rng, t_target, bumps=6):
    """One disk-topology island of about ``t_target`` triangles."""
    from scipy.spatial import Delaunay

    # a maximal Poisson-disk set has about 0.64/r^2 points per unit area
    # (measured for this sampler); Delaunay gives t = 2N - h - 2
    n_target = (t_target + 2) / 2.0
    r = np.sqrt(0.64 * np.pi / max(n_target, 8.0))
    nb = max(8, int(round(2 * np.pi / r)))
    ang = (np.arange(nb) + rng.random()) * (2 * np.pi / nb)
    ring = np.stack([np.cos(ang), np.sin(ang)], 1)
    inner = _poisson_disk(rng, r, 1.0 - 0.6 * r)
    xy = np.concatenate([ring, inner])
    tri = Delaunay(xy).simplices.astype(np.int64)
    a, b, c = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
    tri = np.where((cross < 0)[:, None], tri[:, [0, 2, 1]], tri)
    z = np.zeros(len(xy))
    for _ in range(bumps):
        amp = rng.uniform(0.0, 0.5)
        s = rng.uniform(0.1, 0.4)
        rr, th = np.sqrt(rng.random()), rng.uniform(0, 2 * np.pi)
        c0 = np.array([rr * np.cos(th), rr * np.sin(th)])
        z += amp * np.exp(-((xy - c0) ** 2).sum(1) / (2 * s * s))
    P = np.concatenate([xy, z[:, None]], 1).astype(np.float32).astype(np.float64)
    return P, tri


def _disk_island_job(args):
    seq, t = args
    return disk_island(np.random.default_rng(seq), t)


def _map_islands(seed, counts, workers):
    """Generate islands with per-island child seeds (results do not depend on
    the number of worker processes)."""
    jobs = list(zip(np.random.SeedSequence(seed).spawn(len(counts)), [int(t) for t in counts]))
    if workers is None:
        import os
        workers = min(8, len(os.sched_getaffinity(0)))
    if workers <= 1 or len(jobs) < 16:
        return [_disk_island_job(j) for j in jobs]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_disk_island_job, jobs, chunksize=4)


def disk_islands(seed=2, n_islands=512, tmin=200, tmax=40000, workers=None):
    """Batched islands with log-uniform triangle counts in [tmin, tmax]."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(tmin), np.log(tmax), n_islands))).astype(np.int64)
    return _map_islands(seed, counts, workers)


def sliver_islands(seed=3, n_islands=256, tmin=5000, tmax=60000, frac=0.10, aspect=(50.0, 500.0)):
    """Disk islands (uniform triangle counts) where ``frac`` of the triangles
    are collapsed to aspect ratio U(aspect): one corner moves toward the
    midpoint of its opposite edge (in 3-D) so its height becomes
    |edge| / aspect.  Chosen triangles share no vertex, so each collapse is
    applied to an untouched triangle."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(tmin, tmax + 1, n_islands)
    out = []
    for t in counts:
        P, T = disk_island(rng, int(t))
        P = P.copy()
        used = np.zeros(len(P), bool)
        want = int(frac * len(T))
        done = 0
        for f in rng.permutation(len(T)):
            if done >= want:
                break
            if used[T[f]].any():
                continue
            k = int(rng.integers(3))
            v, a, b = T[f, k], T[f, (k + 1) % 3], T[f, (k + 2) % 3]
            m = 0.5 * (P[a] + P[b])
            e = P[b] - P[a]
            L = np.linalg.norm(e)
            h_vec = P[v] - m - e * (np.dot(P[v] - m, e) / max(L * L, 1e-30))
            h = np.linalg.norm(h_vec)
            if h <= 0:
                continue
            target = L / rng.uniform(*aspect)
            P[v] = m + h_vec * (target / h)
            used[T[f]] = True
            done += 1
        out.append((P.astype(np.float32).astype(np.float64), T))
    return out


def mesh_digest(charts):
    """SHA-256 over the float32 positions and int32 triangles of every chart."""
    h = hashlib.sha256()
    for P, T in charts:
        h.update(np.ascontiguousarray(P, np.float32).tobytes())
        h.update(np.ascontiguousarray(T, np.int32).tobytes())
    return h.hexdigest()
