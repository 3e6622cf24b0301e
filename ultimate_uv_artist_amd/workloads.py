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
  sphere_cap      one island of about 1.0M triangles: a 160-degree cap of a
                  k-frequency icosphere with radial noise, plus per-triangle
                  weights (the differentiation workload)
  heightfield_islands  a 1001x1001 fBm Perlin heightfield (2M triangles) cut by
                  a Voronoi partition into about 4,000 disk islands
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
    return _pool_map(_disk_island_job, jobs, workers)


def _pool_map(fn, jobs, workers):
    if workers is None:
        import os
        workers = min(8, len(os.sched_getaffinity(0)))
    if workers <= 1 or len(jobs) < 16:
        return [fn(j) for j in jobs]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(fn, jobs, chunksize=4)


def disk_islands(seed=2, n_islands=512, tmin=200, tmax=40000, workers=None):
    """Batched islands with log-uniform triangle counts in [tmin, tmax]."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(tmin), np.log(tmax), n_islands))).astype(np.int64)
    return _map_islands(seed, counts, workers)


def _collapse_slivers(rng, P, T, frac, aspect):
    """Collapse ``frac`` of the triangles of one island to aspect ratio
    U(aspect): one corner moves toward the midpoint of its opposite edge (in
    3-D) so its height becomes |edge| / aspect.  The collapsed triangles share
    no vertex (greedy over a random order), so each collapse applies to an
    untouched triangle."""
    want = int(frac * len(T))
    used = bytearray(len(P))
    pick = []
    for f in rng.permutation(len(T)).tolist():
        if len(pick) >= want:
            break
        a, b, c = T[f].tolist()
        if used[a] or used[b] or used[c]:
            continue
        used[a] = used[b] = used[c] = 1
        pick.append(f)
    f = np.array(pick, np.int64)
    k = rng.integers(0, 3, len(f))
    v, a, b = T[f, k], T[f, (k + 1) % 3], T[f, (k + 2) % 3]
    m = 0.5 * (P[a] + P[b])
    e = P[b] - P[a]
    L2 = np.maximum((e * e).sum(1), 1e-30)
    h_vec = P[v] - m - e * (((P[v] - m) * e).sum(1) / L2)[:, None]
    h = np.linalg.norm(h_vec, axis=1)
    target = np.sqrt(L2) / rng.uniform(aspect[0], aspect[1], len(f))
    ok = h > 0
    P = P.copy()
    P[v[ok]] = m[ok] + h_vec[ok] * (target[ok] / h[ok])[:, None]
    return P.astype(np.float32).astype(np.float64), T


def _sliver_island_job(args):
    seq, t, frac, aspect = args
    rng = np.random.default_rng(seq)
    P, T = disk_island(rng, t)
    return _collapse_slivers(rng, P, T, frac, aspect)


def sliver_islands(seed=3, n_islands=256, tmin=5000, tmax=60000, frac=0.10, aspect=(50.0, 500.0), workers=None):
    """Ill-conditioned islands: disk islands with triangle counts uniform in
    [tmin, tmax] where ``frac`` of the triangles are needles of aspect ratio
    U(aspect) (so the cotangent weights span about 1e-3 to 1e3).  Per-island
    child seeds: the result does not depend on the number of workers."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(tmin, tmax + 1, n_islands)
    jobs = [(sq, int(t), frac, tuple(aspect)) for sq, t in zip(np.random.SeedSequence(seed).spawn(n_islands), counts)]
    return _pool_map(_sliver_island_job, jobs, workers)


def _icosahedron():
    g = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0], [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return V / np.linalg.norm(V, axis=1, keepdims=True), F


def icosphere(k):
    """Unit sphere from a k-frequency subdivision of the icosahedron: k^2
    triangles per face (20 k^2 in all), outward orientation.  Points on shared
    edges and corners are keyed by their (corner, weight) pairs, so every
    point is made once and computed the same way for every face."""
    V0, F0 = _icosahedron()
    i, j = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
    m = (i + j) <= k
    i, j = i[m], j[m]
    l = k - i - j
    nloc = len(i)
    loc = -np.ones((k + 1, k + 1), np.int64)
    loc[i, j] = np.arange(nloc)
    # barycentric weights (l, i, j) on corners (a, b, c) of each face
    keys = []
    for a, b, c in F0:
        w = np.stack([l, i, j], 1)
        vid = np.broadcast_to(np.array([a, b, c]), w.shape).copy()
        vid[w == 0] = 99
        o = np.argsort(vid, axis=1, kind="stable")
        keys.append(np.concatenate([np.take_along_axis(vid, o, 1), np.take_along_axis(w, o, 1)], 1))
    keys = np.concatenate(keys)
    uk, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    W = uk[:, 3:].astype(np.float64) / k
    Vc = np.concatenate([V0, np.zeros((1, 3))])[np.minimum(uk[:, :3], 12)]
    P = (W[:, :, None] * Vc).sum(1)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    # up (i,j),(i+1,j),(i,j+1) and down (i+1,j),(i+1,j+1),(i,j+1) triangles
    ui, uj = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    mu = (ui + uj) <= k - 1
    md = (ui + uj) <= k - 2
    up = np.stack([loc[ui[mu], uj[mu]], loc[ui[mu] + 1, uj[mu]], loc[ui[mu], uj[mu] + 1]], 1)
    dn = np.stack([loc[ui[md] + 1, uj[md]], loc[ui[md] + 1, uj[md] + 1], loc[ui[md], uj[md] + 1]], 1)
    T = np.concatenate([inv[np.concatenate([up, dn]) + f * nloc] for f in range(len(F0))])
    return P, T


def _compact(P, T):
    used = np.unique(T)
    remap = -np.ones(len(P), np.int64)
    remap[used] = np.arange(len(used))
    return P[used], remap[T]


def sphere_cap(seed=4, k=348, polar_deg=80.0, noise=0.01):
    """A genus-0 surface cut to a disk: the triangles of a k-frequency
    icosphere whose corners all lie within ``polar_deg`` of the +z pole (a
    160-degree cap by default, about 1.0M triangles for k = 348), radial
    noise N(0, noise), float32; plus per-triangle weights U(0.5, 2.0) in
    float32.  Returns (P, T, w)."""
    rng = np.random.default_rng(seed)
    P, T = icosphere(k)
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]          # a random orientation of the lattice
    P = P @ R.T
    inside = P[:, 2] >= np.cos(np.radians(polar_deg))
    T = T[inside[T].all(1)]
    P, T = _compact(P, T)
    P = P * (1.0 + rng.normal(0.0, noise, (len(P), 1)))
    w = rng.uniform(0.5, 2.0, len(T)).astype(np.float32).astype(np.float64)
    return P.astype(np.float32).astype(np.float64), T, w


def _fade(t):
    return t * t * t * (t * (t * 6 - 15) + 10)


def perlin(x, y, rng, period):
    """2-D gradient (Perlin) noise with unit lattice spacing, random unit
    gradients on a ``period`` x ``period`` torus of lattice points."""
    ang = rng.uniform(0, 2 * np.pi, (period, period))
    gx, gy = np.cos(ang), np.sin(ang)
    x0, y0 = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    fx, fy = x - x0, y - y0
    out = 0.0
    u, v = _fade(fx), _fade(fy)
    for dx in (0, 1):
        for dy in (0, 1):
            ix, iy = (x0 + dx) % period, (y0 + dy) % period
            d = gx[ix, iy] * (fx - dx) + gy[ix, iy] * (fy - dy)
            out = out + d * np.where(dx, u, 1 - u) * np.where(dy, v, 1 - v)
    return out


def _edge_pairs(T, nv):
    """Pairs of triangles sharing an edge: (triangle a, triangle b, edge
    endpoint 0, edge endpoint 1)."""
    t = len(T)
    a = np.concatenate([T[:, 0], T[:, 1], T[:, 2]])
    b = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    key = np.minimum(a, b) * nv + np.maximum(a, b)
    et = np.tile(np.arange(t), 3)
    o = np.argsort(key, kind="stable")
    ks = key[o]
    d = np.nonzero(ks[1:] == ks[:-1])[0]
    return et[o[d]], et[o[d + 1]], a[o[d]], b[o[d]]


def _bowtie_split(T, nv):
    """Duplicate every vertex whose incident triangles form more than one fan
    (triangles around a vertex joined through shared edges), one copy per
    fan.  Returns the new T and, per output vertex, its source vertex."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    t = len(T)
    ta, tb, va, vb = _edge_pairs(T, nv)
    rows, cols = [], []
    for v in (va, vb):
        rows.append(3 * ta + np.argmax(T[ta] == v[:, None], 1))
        cols.append(3 * tb + np.argmax(T[tb] == v[:, None], 1))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    A = sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(3 * t, 3 * t))
    _, fan = connected_components(A, directed=False)
    pair = T.reshape(-1).astype(np.int64) * (3 * t) + fan
    up, inv = np.unique(pair, return_inverse=True)
    return inv.reshape(t, 3), up // (3 * t)


def _is_disk(T, nv):
    a = np.concatenate([T[:, 0], T[:, 1], T[:, 2]])
    b = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    e = np.unique(np.minimum(a, b) * nv + np.maximum(a, b))
    return nv - len(e) + len(T) == 1


def heightfield_islands(seed=5, n=1001, cells=4000, octaves=5, amp=0.2, tmin=50, tmax=5000):
    """A heightfield (n x n vertex grid on [0,1]^2, 2(n-1)^2 triangles, fBm of
    2-D Perlin noise with ``octaves`` octaves, amplitude ``amp``) cut into
    islands by a seeded Voronoi partition of the triangle centroids into
    ``cells`` cells.  Components smaller than ``tmin`` join the neighbouring
    island they share most edges with; vertices where an island touches
    itself only at a corner are split, so every island is a disk (checked).
    Islands keep their own copies of seam vertices.  Returns charts."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n)
    x, y = np.meshgrid(g, g, indexing="xy")
    z = np.zeros_like(x)
    for o in range(octaves):
        f = 4 * 2 ** o
        z += 0.5 ** o * perlin(x * f, y * f, rng, f)
    z *= amp / max(np.abs(z).max(), 1e-30)
    P = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32).astype(np.float64)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    v00 = (j * n + i).ravel()
    v10, v01, v11 = v00 + 1, v00 + n, v00 + n + 1
    T = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)]).astype(np.int64)
    cen = P[T, :2].mean(1)
    lab = cKDTree(rng.random((cells, 2))).query(cen)[1].astype(np.int64)
    t = len(T)
    ta, tb, _, _ = _edge_pairs(T, len(P))
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    def components(lab):
        inner = lab[ta] == lab[tb]
        A = sp.coo_matrix((np.ones(int(inner.sum())), (ta[inner], tb[inner])), shape=(t, t))
        return connected_components(A, directed=False)

    for _ in range(16):
        nc, comp = components(lab)
        small = np.bincount(comp, minlength=nc)[comp] < tmin
        if not small.any():
            break
        # a small component takes the label of the (not small) island it shares most edges with
        m1 = (lab[ta] != lab[tb]) & small[ta] & ~small[tb]
        m2 = (lab[ta] != lab[tb]) & small[tb] & ~small[ta]
        src = np.concatenate([comp[ta[m1]], comp[tb[m2]]])
        dst = np.concatenate([lab[tb[m1]], lab[ta[m2]]])
        pairs, cnt = np.unique(np.stack([src, dst], 1), axis=0, return_counts=True)
        o = np.lexsort((-cnt, pairs[:, 0]))
        pairs = pairs[o]
        first = np.r_[True, pairs[1:, 0] != pairs[:-1, 0]]
        to = -np.ones(nc, np.int64)
        to[pairs[first, 0]] = pairs[first, 1]
        lab = np.where(to[comp] >= 0, to[comp], lab)
    nc, comp = components(lab)
    order = np.argsort(comp, kind="stable")
    bounds = np.searchsorted(comp[order], np.arange(nc + 1))
    charts = []
    for c in range(nc):
        Pc, Tc = _compact(P, T[order[bounds[c]:bounds[c + 1]]])
        Tc, src = _bowtie_split(Tc, len(Pc))
        Pc = Pc[src]
        if not _is_disk(Tc, len(Pc)):
            raise RuntimeError(f"heightfield island {c} is not a disk")
        charts.append((Pc, Tc))
    return charts


def mesh_digest(charts):
    """SHA-256 over the float32 positions and int32 triangles of every chart."""
    h = hashlib.sha256()
    for P, T in charts:
        h.update(np.ascontiguousarray(P, np.float32).tobytes())
        h.update(np.ascontiguousarray(T, np.int32).tobytes())
    return h.hexdigest()
