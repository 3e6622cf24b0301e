"""CPU restatement of the reference's UV validation checks (engine/validate.py).

Test infrastructure only: the parity tests validate the layouts the GPU path
produces and the layouts of the CPU oracle with these checks and require
identical verdicts.  Nothing in the product path imports this module.

Restated (engine/validate.py):
  _orient                 validate.py:20-22
  _strict_inside          validate.py:33-45
  _proper_cross           validate.py:48-55
  _tri_overlap            validate.py:58-91
  _tri_overlap_strict     validate.py:231-254
  _overlaps, _overlaps_strict
                          validate.py:94-136, 257-299 (grid hashing only selects the
                          candidate pairs; every pair whose bounding boxes
                          meet is tested here, which flags the same pairs:
                          the exact test is false for disjoint boxes)
  validate_uvs checks     validate.py:139-210  (range, degenerate, sub-float32
                          slivers, flipped triangles against the island's
                          dominant orientation, mirrored islands, exact
                          overlaps; not the atlas gutter or the export
                          preservation checks)
  uv_islands              metrics.py:76-103
  chart_overlaps          validate.py:212-226
Pinned against the reference's own outputs on hand-built and real layouts
(tests/golden/validate_golden.npz, tests/test_golden.py) and against
hand-checked cases (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np

ULP = 2.0 ** -24          # float32 spacing just below 1.0


def _orient(ax, ay, bx, by, cx, cy):
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)


def _strict_inside(px, py, ax, ay, bx, by, cx, cy, rel):
    """Point strictly inside the triangle (either orientation), with the
    tolerance max(rel l, 4 ULP sqrt(l)), l the longest squared edge."""
    l = np.maximum((bx - ax) ** 2 + (by - ay) ** 2, np.maximum((cx - bx) ** 2 + (cy - by) ** 2, (ax - cx) ** 2 + (ay - cy) ** 2))
    eps = np.maximum(rel * l, 4.0 * ULP * np.sqrt(l)) if rel > 0 else np.zeros_like(l)
    d1 = _orient(ax, ay, bx, by, px, py)
    d2 = _orient(bx, by, cx, cy, px, py)
    d3 = _orient(cx, cy, ax, ay, px, py)
    return ((d1 > eps) & (d2 > eps) & (d3 > eps)) | ((d1 < -eps) & (d2 < -eps) & (d3 < -eps))


def _proper_cross(ax, ay, bx, by, cx, cy, dx, dy, rel):
    l1 = (bx - ax) ** 2 + (by - ay) ** 2
    l2 = (dx - cx) ** 2 + (dy - cy) ** 2
    e1 = np.maximum(rel * l1, 4.0 * ULP * np.sqrt(l1))
    e2 = np.maximum(rel * l2, 4.0 * ULP * np.sqrt(l2))
    o1 = _orient(ax, ay, bx, by, cx, cy)
    o2 = _orient(ax, ay, bx, by, dx, dy)
    o3 = _orient(cx, cy, dx, dy, ax, ay)
    o4 = _orient(cx, cy, dx, dy, bx, by)
    return (((o1 > e1) & (o2 < -e1)) | ((o1 < -e1) & (o2 > e1))) & (((o3 > e2) & (o4 < -e2)) | ((o3 < -e2) & (o4 > e2)))


def tri_overlap_pairs(U, F, s, t, rel):
    """Positive-area overlap of UV triangles s[k], t[k] (arrays of pairs).
    Corners at identical UV positions count as shared; three shared corners
    (a coincident copy) do not overlap."""
    s = np.asarray(s, np.int64)
    t = np.asarray(t, np.int64)
    a = F[s]                       # (k, 3) vertex ids
    b = F[t]
    ua = U[a]                      # (k, 3, 2)
    ub = U[b]
    same = (ua[:, :, None, 0] == ub[:, None, :, 0]) & (ua[:, :, None, 1] == ub[:, None, :, 1])   # (k, 3, 3)
    shared = same.any(2).sum(1)
    out = np.zeros(len(s), bool)
    for i in range(3):
        i1 = (i + 1) % 3
        for j in range(3):
            j1 = (j + 1) % 3
            skip = same[:, i, j] | same[:, i, j1] | same[:, i1, j] | same[:, i1, j1]
            cr = _proper_cross(ua[:, i, 0], ua[:, i, 1], ua[:, i1, 0], ua[:, i1, 1],
                               ub[:, j, 0], ub[:, j, 1], ub[:, j1, 0], ub[:, j1, 1], rel)
            out |= cr & ~skip
    for i in range(3):
        va = ~same[:, i, :].any(1)
        out |= va & _strict_inside(ua[:, i, 0], ua[:, i, 1], ub[:, 0, 0], ub[:, 0, 1], ub[:, 1, 0], ub[:, 1, 1],
                                   ub[:, 2, 0], ub[:, 2, 1], rel)
        vb = ~same[:, :, i].any(1)
        out |= vb & _strict_inside(ub[:, i, 0], ub[:, i, 1], ua[:, 0, 0], ua[:, 0, 1], ua[:, 1, 0], ua[:, 1, 1],
                                   ua[:, 2, 0], ua[:, 2, 1], rel)
    ca = ua.sum(1) / 3.0
    cb = ub.sum(1) / 3.0
    out |= _strict_inside(ca[:, 0], ca[:, 1], ub[:, 0, 0], ub[:, 0, 1], ub[:, 1, 0], ub[:, 1, 1], ub[:, 2, 0],
                          ub[:, 2, 1], rel)
    out |= _strict_inside(cb[:, 0], cb[:, 1], ua[:, 0, 0], ua[:, 0, 1], ua[:, 1, 0], ua[:, 1, 1], ua[:, 2, 0],
                          ua[:, 2, 1], rel)
    return out & (shared < 3)


def tri_cross_strict_pairs(U, F, s, t):
    """validate.py:231-254: a proper edge crossing with zero tolerance
    between triangles s[k], t[k] that are not coincident copies."""
    s = np.asarray(s, np.int64)
    t = np.asarray(t, np.int64)
    ua = U[F[s]]
    ub = U[F[t]]
    same = (ua[:, :, None, 0] == ub[:, None, :, 0]) & (ua[:, :, None, 1] == ub[:, None, :, 1])
    shared = same.any(2).sum(1)
    out = np.zeros(len(s), bool)
    for i in range(3):
        i1 = (i + 1) % 3
        for j in range(3):
            j1 = (j + 1) % 3
            skip = same[:, i, j] | same[:, i, j1] | same[:, i1, j] | same[:, i1, j1]
            p0, p1, q0, q1 = ua[:, i], ua[:, i1], ub[:, j], ub[:, j1]
            o1 = _orient(p0[:, 0], p0[:, 1], p1[:, 0], p1[:, 1], q0[:, 0], q0[:, 1])
            o2 = _orient(p0[:, 0], p0[:, 1], p1[:, 0], p1[:, 1], q1[:, 0], q1[:, 1])
            o3 = _orient(q0[:, 0], q0[:, 1], q1[:, 0], q1[:, 1], p0[:, 0], p0[:, 1])
            o4 = _orient(q0[:, 0], q0[:, 1], q1[:, 0], q1[:, 1], p1[:, 0], p1[:, 1])
            out |= (o1 * o2 < 0) & (o3 * o4 < 0) & ~skip
    return out & (shared < 3)


def candidate_pairs(U, F, skip):
    """Pairs (s < t) of non-skipped triangles whose bounding boxes meet.

    Uniform grid over the boxes (about 4 triangles per cell): each triangle is
    entered in every cell its box covers, a pair is generated in the first
    cell both boxes cover, then the exact box test keeps the meeting ones."""
    u = U[F]
    lo = u.min(1)
    hi = u.max(1)
    idx = np.nonzero(~skip)[0]
    if len(idx) < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    L = lo[idx].min(0)
    ext = float((hi[idx].max(0) - L).max())
    G = int(np.clip(np.sqrt(len(idx)) / 2, 1, 4096))
    sc = G / ext if ext > 0 else 0.0
    c0 = np.clip(np.floor((lo[idx] - L) * sc), 0, G - 1).astype(np.int64)
    c1 = np.clip(np.floor((hi[idx] - L) * sc), 0, G - 1).astype(np.int64)
    nx = c1[:, 0] - c0[:, 0] + 1
    ncell = nx * (c1[:, 1] - c0[:, 1] + 1)
    item = np.repeat(np.arange(len(idx)), ncell)
    k = np.arange(len(item)) - np.repeat(np.cumsum(ncell) - ncell, ncell)
    cell = (c0[item, 1] + k // nx[item]) * G + c0[item, 0] + k % nx[item]
    o = np.argsort(cell, kind="stable")
    cell, item = cell[o], item[o]
    _, start, cnt = np.unique(cell, return_index=True, return_counts=True)
    r = np.arange(len(item)) - np.repeat(start, cnt)
    npair = np.repeat(cnt, cnt) - 1 - r
    first = np.repeat(np.arange(len(item)), npair)
    second = first + 1 + np.arange(len(first)) - np.repeat(np.cumsum(npair) - npair, npair)
    a, b, c = item[first], item[second], cell[first]
    keep = (np.maximum(c0[a, 0], c0[b, 0]) == c % G) & (np.maximum(c0[a, 1], c0[b, 1]) == c // G)
    a, b = idx[a[keep]], idx[b[keep]]
    keep = (lo[a] <= hi[b]).all(1) & (lo[b] <= hi[a]).all(1)
    a, b = a[keep], b[keep]
    return np.minimum(a, b), np.maximum(a, b)


def overlaps(U, F, skip, rel=1e-6):
    """Per-triangle flag: overlaps another (validate.py:94-136)."""
    U = np.asarray(U, np.float64)
    F = np.asarray(F, np.int64)
    flag = np.zeros(len(F), bool)
    s, t = candidate_pairs(U, F, skip)
    if len(s):
        hit = tri_overlap_pairs(U, F, s, t, rel)
        flag[s[hit]] = True
        flag[t[hit]] = True
    return flag


def overlaps_strict(U, F, skip):
    """Per-triangle flag of zero-tolerance edge crossings (validate.py:257-299)."""
    U = np.asarray(U, np.float64)
    F = np.asarray(F, np.int64)
    flag = np.zeros(len(F), bool)
    s, t = candidate_pairs(U, F, skip)
    if len(s):
        hit = tri_cross_strict_pairs(U, F, s, t)
        flag[s[hit]] = True
        flag[t[hit]] = True
    return flag


def chart_overlaps(U, T):
    """validate.py:212-226: exact self-overlap of one chart (any UV units)."""
    U = np.asarray(U, np.float64)
    T = np.ascontiguousarray(T, np.int64)
    if len(T) < 2:
        return np.zeros(len(T), bool)
    lo = U.min(0)
    ext = float((U.max(0) - lo).max())
    if not np.isfinite(ext) or ext <= 0:
        return np.zeros(len(T), bool)
    Un = (U - lo) / ext * 0.999999
    u = Un[T]
    sa = 0.5 * ((u[:, 1, 0] - u[:, 0, 0]) * (u[:, 2, 1] - u[:, 0, 1]) - (u[:, 2, 0] - u[:, 0, 0]) * (u[:, 1, 1] - u[:, 0, 1]))
    skip = np.abs(sa) <= 1e-14
    return overlaps(Un, T, skip, 1e-6)


def uv_islands(F, n_vertices, U=None, P=None):
    """metrics.py:76-103: islands = triangles connected through shared
    vertices; with U and P given, vertices with identical float32 position and
    UV count as one.  Returns (island id per triangle, island count); the ids
    are the reference's up to a renaming of the islands, which none of the
    checks depends on."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    F = np.asarray(F, np.int64)
    if U is not None and P is not None:
        key = np.concatenate([np.asarray(P, np.float32), np.asarray(U, np.float32)], 1)
        key = np.ascontiguousarray(key).view(np.dtype((np.void, 20))).reshape(-1)
        _, inv = np.unique(key, return_inverse=True)
        F = inv.reshape(-1)[F]
        n_vertices = int(inv.max()) + 1
    m = len(F)
    if m == 0:
        return np.zeros(0, np.int32), 0
    rows = np.repeat(np.arange(m), 3)
    A = sp.csr_matrix((np.ones(3 * m), (rows, F.reshape(-1))), shape=(m, n_vertices))
    n, lab = connected_components(A @ A.T, directed=False)
    return lab.astype(np.int32), int(n)


def validate_uvs(v, f, uv):
    """validate.py:139-210 without the gutter and preservation checks."""
    U = np.asarray(uv, np.float32).astype(np.float64)
    isl, _ = uv_islands(f, len(v), U, v)
    return validate_islands(v, f, U, isl)


def validate_islands(P, T, U, island):
    """The validate_uvs checks (validate.py:139-210) of UVs in [0, 1] whose
    triangles belong to the given islands (island id per triangle): UVs are
    validated as stored, in float32.  Returns (report dict, per-triangle
    problem mask)."""
    v = np.asarray(P, np.float64)
    f = np.asarray(T, np.int64)
    U = np.asarray(U, np.float32).astype(np.float64)
    isl = np.asarray(island, np.int64)
    n = int(isl.max()) + 1 if len(isl) else 0
    rep = {}
    rep["non_finite_uv"] = int((~np.isfinite(U)).any(1).sum())
    rep["uv_outside_unit_square"] = int(((U < 0) | (U > 1)).any(1).sum())
    p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    a3 = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1)
    u = U[f]
    sa = 0.5 * ((u[:, 1, 0] - u[:, 0, 0]) * (u[:, 2, 1] - u[:, 0, 1]) - (u[:, 2, 0] - u[:, 0, 0]) * (u[:, 1, 1] - u[:, 0, 1]))
    f32_res = 2.0 ** -24
    edge = np.maximum(np.maximum(np.linalg.norm(u[:, 1] - u[:, 0], axis=1), np.linalg.norm(u[:, 2] - u[:, 1], axis=1)),
                      np.linalg.norm(u[:, 0] - u[:, 2], axis=1))
    sliver = np.abs(sa) < f32_res * np.maximum(edge, f32_res)
    has3d = a3 > 1e-14 * max(float(a3.max()), 1e-30)
    rep["degenerate_uv_triangles"] = int((has3d & (sa == 0)).sum())
    rep["sub_float32_slivers"] = int((has3d & sliver).sum())
    w = np.abs(sa)
    pos = np.bincount(isl, weights=np.where(sa > 0, w, 0), minlength=n)
    neg = np.bincount(isl, weights=np.where(sa < 0, w, 0), minlength=n)
    dom = np.where(pos >= neg, 1.0, -1.0)
    flipped = has3d & ~sliver & (np.sign(sa) != dom[isl]) & (sa != 0)
    rep["flipped_triangles"] = int(flipped.sum())
    rep["mirrored_islands"] = int((neg > pos).sum())
    rep["islands"] = n
    skip = ~has3d | sliver | ~np.isfinite(sa)
    ov = overlaps(U, f, skip, 1e-6)
    rep["overlapping_triangles_exact"] = int(ov.sum())
    rep["float32_contacts"] = int((overlaps_strict(U, f, skip) & ~ov).sum())
    if ov.any():
        rep["islands_with_overlap"] = int(len(np.unique(isl[ov])))
    bad = rep["non_finite_uv"] + rep["uv_outside_unit_square"] + rep["flipped_triangles"] + rep["overlapping_triangles_exact"]
    rep["valid"] = bool(bad == 0)
    return rep, ov | flipped


def island_verdict(P, T, U):
    """Verdict of one island's layout: the validate_uvs checks after scaling
    the layout into the unit square (as packing would, keeping its shape),
    plus the chart-level exact overlap test."""
    U = np.asarray(U, np.float64)
    lo = U.min(0)
    ext = float((U.max(0) - lo).max())
    Un = (U - lo) / ext * 0.999999 if np.isfinite(ext) and ext > 0 else U
    rep, _ = validate_islands(P, T, Un, np.zeros(len(T), np.int64))
    rep["chart_overlaps"] = int(chart_overlaps(U, T).sum())
    return rep


# ----------------------------------------------------------------- gutters
_I64_MIN = np.iinfo(np.int64).min


def _floor_int(x):
    """int(np.floor(x)) as the reference's compiled kernel converts it: NaN and
    values outside the int64 range become the int64 minimum."""
    f = np.floor(x)
    ok = np.isfinite(f) & (np.abs(f) < 2.0 ** 63)
    return np.where(ok, np.where(ok, f, 0.0).astype(np.int64), _I64_MIN)


def raster_last(U, F, res, chunk=1 << 22):
    """metrics.py:315-345 _raster_labels, as the triangle index that writes
    each texel last: per texel of a res x res image the largest triangle index
    whose conservative raster covers the texel centre, -1 where none does
    (the reference writes labels in triangle order, so the texel keeps the
    label of that triangle).  Vectorised over (triangle, texel) pairs of the
    clamped bounding boxes."""
    U = np.asarray(U, np.float64)
    F = np.asarray(F, np.int64)
    x0, y0 = U[F[:, 0], 0] * res, U[F[:, 0], 1] * res
    x1, y1 = U[F[:, 1], 0] * res, U[F[:, 1], 1] * res
    x2, y2 = U[F[:, 2], 0] * res, U[F[:, 2], 1] * res
    area = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
    sw = area < 0                                  # counter-clockwise order
    x1, x2 = np.where(sw, x2, x1), np.where(sw, x1, x2)
    y1, y2 = np.where(sw, y2, y1), np.where(sw, y1, y2)
    # Python's min/max of the reference: min(x0, min(x1, x2)) keeps x0 when it
    # is NaN and skips a NaN x1/x2 (comparisons with NaN are false)
    def pymin(a, b, c):
        bc = np.where(c < b, c, b)
        return np.where(bc < a, bc, a)

    def pymax(a, b, c):
        bc = np.where(c > b, c, b)
        return np.where(bc > a, bc, a)

    with np.errstate(invalid="ignore"):
        minx = np.maximum(_floor_int(pymin(x0, x1, x2)), 0)
        maxx = np.minimum(_floor_int(pymax(x0, x1, x2)), res - 1)
        miny = np.maximum(_floor_int(pymin(y0, y1, y2)), 0)
        maxy = np.minimum(_floor_int(pymax(y0, y1, y2)), res - 1)
    ex = (x1 - x0, x2 - x1, x0 - x2)
    ey = (y1 - y0, y2 - y1, y0 - y2)
    ox, oy = (x0, x1, x2), (y0, y1, y2)
    L = [np.sqrt(ex[k] ** 2 + ey[k] ** 2) for k in range(3)]
    wx = np.where(maxx >= minx, maxx - minx + 1, 0)   # no wrap-around for the int64 minimum
    wy = np.where(maxy >= miny, maxy - miny + 1, 0)
    cnt = wx * wy
    last = np.full(res * res, -1, np.int64)
    tri = np.repeat(np.arange(len(F)), cnt)
    start = np.cumsum(cnt) - cnt
    for a in range(0, len(tri), chunk):
        t = tri[a:a + chunk]
        k = np.arange(a, a + len(t)) - start[t]
        px = minx[t] + k % np.maximum(wx[t], 1)
        py = miny[t] + k // np.maximum(wx[t], 1)
        cx, cy = px + 0.5, py + 0.5
        inside = np.ones(len(t), bool)
        for e in range(3):
            val = ex[e][t] * (cy - oy[e][t]) - ey[e][t] * (cx - ox[e][t])
            inside &= ~(val < -0.7072 * L[e][t])
        np.maximum.at(last, (py * res + px)[inside], t[inside])
    return last.reshape(res, res)


def min_gutter(U, F, isl, res=4096):
    """metrics.py:348-373: the smallest gap in texels between two different
    islands of the conservative raster (island id per triangle in isl): NaN
    when the image is all empty or all full, 0 when two islands touch, else
    the smallest sum of the distances from two adjacent empty texels to
    their nearest filled texels of different islands, minus one.  Returns
    (gap, label image)."""
    last = raster_last(U, F, res)
    isl = np.asarray(isl, np.int64)
    img = np.where(last >= 0, isl[np.maximum(last, 0)], -1).astype(np.int32)
    return gutter_of_labels(img), img


def gutter_of_labels(img):
    """The gap measurement of metrics.py:355-373 on a label image (-1 empty)."""
    from scipy import ndimage

    empty = img < 0
    if empty.all() or not empty.any():
        return float("nan")
    d, (iy, ix) = ndimage.distance_transform_edt(empty, return_indices=True)
    near = img[iy, ix]
    touch = bool(((img[:, 1:] >= 0) & (img[:, :-1] >= 0) & (img[:, 1:] != img[:, :-1])).any() or
                 ((img[1:, :] >= 0) & (img[:-1, :] >= 0) & (img[1:, :] != img[:-1, :])).any())
    if touch:
        return 0.0
    g = np.inf
    h = near[:, 1:] != near[:, :-1]
    if h.any():
        g = min(g, float((d[:, 1:] + d[:, :-1])[h].min()))
    vv = near[1:, :] != near[:-1, :]
    if vv.any():
        g = min(g, float((d[1:, :] + d[:-1, :])[vv].min()))
    return max(g - 1.0, 0.0)


def validate_gutter(v, f, uv, atlas=4096, padding=None):
    """The gutter entries of validate.py:180-185 (min_gutter_texels, and
    gutter_ok with a padding) for UVs validated as stored (float32)."""
    U = np.asarray(uv, np.float32).astype(np.float64)
    isl, _ = uv_islands(f, len(v), U, v)
    res = min(atlas, 8192)
    g, _ = min_gutter(U, f, isl, res)
    out = {"min_gutter_texels": float(g * atlas / res)}
    if padding is not None:
        out["gutter_ok"] = bool(out["min_gutter_texels"] >= padding - 1e-6)
    return out
