"""The reference's validation of final UVs (engine/validate.py) with the exact
triangle-overlap test on the GPU (``uva_overlap_flags``).

  overlap_flags        validate.py:94-136 _overlaps (and 257-299 _overlaps_strict)
  chart_overlaps       validate.py:212-226
  validate_uvs         validate.py:139-210: range, degenerate and sub-float32
                       triangles, flips against the island's dominant
                       orientation, mirrored islands, exact overlaps, float32
                       contacts, the gutter between islands (gutter=True, the
                       default), preservation of positions / normals / colours
  uv_islands           metrics.py:76-103
  min_gutter           metrics.py:348-373: the conservative raster on the GPU
                       (uva_raster_last), scipy's distance transform on the host
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .plan import require_gpu, stream_of


def _flags_dev(U, F, skip, rel, strict, grid, dev):
    """Overlap flags (uint8 device tensor) of device tensors U (n,2) float64
    and F (t,3) int64, skip (t,) uint8 or None."""
    out = torch.empty(len(F), dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().uva_overlap_flags(U.data_ptr(), len(U), F.data_ptr(), len(F),
                                            None if skip is None else skip.data_ptr(), float(rel), int(bool(strict)),
                                            int(grid), out.data_ptr(), stream_of(dev)), "uva_overlap_flags")
    return out


def overlap_flags(U, F, skip=None, rel=1e-6, strict=False, grid=64, dev=None):
    """Per-triangle flag: the triangle overlaps another with positive area
    (``strict``: crosses another with zero tolerance).  U (n,2) in the unit
    square, F (t,3); ``grid`` cells per side of the candidate grid."""
    dev = require_gpu(dev)
    U = np.ascontiguousarray(U, np.float64).reshape(-1, 2)
    F = np.ascontiguousarray(F, np.int64).reshape(-1, 3)
    if len(F) and (F.min() < 0 or F.max() >= len(U)):
        raise ValueError("F holds vertex ids outside [0, len(U))")
    if skip is not None and len(skip) != len(F):
        raise ValueError("skip needs one entry per triangle")
    sk = None if skip is None else torch.as_tensor(np.ascontiguousarray(skip, np.uint8), device=dev)
    out = _flags_dev(torch.as_tensor(U, device=dev), torch.as_tensor(F, device=dev), sk, rel, strict, grid, dev)
    return out.cpu().numpy().astype(bool)


def _unique_rows(key):
    """Group ids of equal rows of an integer (n, c) tensor: ids follow the
    lexicographic order of the rows (stable argsort per column, last column
    first)."""
    n = key.shape[0]
    idx = torch.arange(n, device=key.device)
    for c in reversed(range(key.shape[1])):
        idx = idx[torch.argsort(key[idx, c], stable=True)]
    srt = key[idx]
    new = torch.ones(n, dtype=torch.bool, device=key.device)
    if n > 1:
        new[1:] = (srt[1:] != srt[:-1]).any(1)
    gid = torch.cumsum(new, 0) - 1
    inv = torch.empty(n, dtype=torch.int64, device=key.device)
    inv[idx] = gid
    return inv, (int(gid[-1]) + 1 if n else 0)


def _components(F, n):
    """Connected components of the vertex graph of triangles F (t,3) int64
    over n vertices: per vertex the smallest vertex id of its component.
    Rounds of hooking every root to the smallest root across an edge, then
    pointer jumping to the roots; each round at least halves the number of
    components that still have an edge between them."""
    lab = torch.arange(n, device=F.device)
    a = torch.cat([F[:, 0], F[:, 0]])
    b = torch.cat([F[:, 1], F[:, 2]])
    while True:
        la, lb = lab[a], lab[b]
        d = la != lb
        if not bool(d.any()):
            return lab
        la, lb = la[d], lb[d]
        m = torch.minimum(la, lb)
        lab.scatter_reduce_(0, la, m, "amin")      # every label is a root here: parents only decrease
        lab.scatter_reduce_(0, lb, m, "amin")
        while True:
            nxt = lab[lab]
            if torch.equal(nxt, lab):
                break
            lab = nxt


def _uv_islands_dev(F, n_vertices, U32=None, P32=None):
    """uv_islands on tensors (any device): F (t,3) int64, U32 (n,2) and P32
    (n,3) float32.  Vertices with bitwise-equal float32 position and UV (the
    reference compares the 20 bytes) are merged first."""
    if U32 is not None and P32 is not None:
        key = torch.cat([P32.contiguous().view(torch.int32), U32.contiguous().view(torch.int32)], 1)
        inv, n_vertices = _unique_rows(key)
        F = inv[F]
    if len(F) == 0:
        return torch.zeros(0, dtype=torch.int64, device=F.device), 0
    root = _components(F, int(n_vertices))[F[:, 0]]
    _, isl = torch.unique(root, return_inverse=True)
    return isl, int(isl.max()) + 1


def _triangle_terms(v, f, U):
    """Per triangle, on tensors (any device): the squared norm of the 3-D
    cross product, the signed UV area and the largest squared UV edge
    length, in numpy's operation order (np.cross; np.linalg.norm sums the
    squares left to right), so the values equal numpy's bit for bit.  The
    square roots are left to the host (``_triangle_checks``): torch's CPU
    sqrt is not correctly rounded."""
    p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e, g = p1 - p0, p2 - p0
    c0 = e[:, 1] * g[:, 2] - e[:, 2] * g[:, 1]
    c1 = e[:, 2] * g[:, 0] - e[:, 0] * g[:, 2]
    c2 = e[:, 0] * g[:, 1] - e[:, 1] * g[:, 0]
    s3 = c0 * c0 + c1 * c1 + c2 * c2
    u0, u1, u2 = U[f[:, 0]], U[f[:, 1]], U[f[:, 2]]
    sa = 0.5 * ((u1[:, 0] - u0[:, 0]) * (u2[:, 1] - u0[:, 1]) - (u2[:, 0] - u0[:, 0]) * (u1[:, 1] - u0[:, 1]))

    def sq(d):
        return d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]

    e2 = torch.maximum(torch.maximum(sq(u1 - u0), sq(u2 - u1)), sq(u0 - u2))
    return s3, sa, e2


def _triangle_checks(s3, sa, e2):
    """3-D area, sub-float32 sliver and has-3-D-area masks of
    validate.py:148-160 from ``_triangle_terms`` (numpy arrays).  The longest
    edge is sqrt of the largest squared length: a correctly rounded sqrt is
    monotone, so this equals the largest of the three norms."""
    a3 = 0.5 * np.sqrt(s3)
    r = 2.0 ** -24
    sliver = np.abs(sa) < r * np.maximum(np.sqrt(e2), r)
    has3d = a3 > 1e-14 * max(float(a3.max()) if len(a3) else 0.0, 1e-30)
    return a3, sliver, has3d


def chart_overlaps(U, T):
    """Exact positive-area self-overlap of one chart in any UV units
    (validate.py:212-226)."""
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
    G = int(np.clip(np.sqrt(len(T)) / 2, 4, 2048))
    return overlap_flags(Un, T, np.abs(sa) <= 1e-14, 1e-6, False, G)


def uv_islands(faces, n_vertices, uv=None, vertices=None, dev=None):
    """Islands = faces connected through shared vertices; with ``uv`` and
    ``vertices``, vertices with identical float32 position and UV count as one
    (metrics.py:76-103).  Returns (island id per face, island count); ids are
    the reference's up to a renaming."""
    dev = require_gpu(dev)
    F = torch.as_tensor(np.ascontiguousarray(faces, np.int64).reshape(-1, 3), device=dev)
    U32 = P32 = None
    if uv is not None and vertices is not None:
        U32 = torch.as_tensor(np.ascontiguousarray(uv, np.float32).reshape(-1, 2), device=dev)
        P32 = torch.as_tensor(np.ascontiguousarray(vertices, np.float32).reshape(-1, 3), device=dev)
    isl, n = _uv_islands_dev(F, n_vertices, U32, P32)
    return isl.to(torch.int32).cpu().numpy(), n


def _raster_last_dev(Ud, Fd, res, dev):
    """uva_raster_last on device tensors: (res, res) int32 tensor of the
    largest covering triangle index per texel, -1 where none covers it."""
    last = torch.empty(int(res) * int(res), dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().uva_raster_last(Ud.data_ptr(), len(Ud), Fd.data_ptr(), len(Fd), int(res), last.data_ptr(),
                                          stream_of(dev)), "uva_raster_last")
    return last.view(int(res), int(res))


def _gap_of_labels(img):
    """The gap measurement of metrics.py:355-373 on a label image (-1 empty):
    NaN for an image all empty or all full, 0 when texels of two islands are
    side by side, else the smallest sum of the Euclidean distances of two
    side-by-side empty texels to their nearest filled texels (scipy's exact
    distance transform, as the reference) when those belong to different
    islands, minus one."""
    from scipy import ndimage

    empty = img < 0
    if empty.all() or not empty.any():
        return float("nan")
    filled = ~empty
    if (filled[:, 1:] & filled[:, :-1] & (img[:, 1:] != img[:, :-1])).any() or \
            (filled[1:, :] & filled[:-1, :] & (img[1:, :] != img[:-1, :])).any():
        return 0.0
    dist, (iy, ix) = ndimage.distance_transform_edt(empty, return_indices=True)
    owner = img[iy, ix]
    best = np.inf
    for a, b, da, db in ((owner[:, 1:], owner[:, :-1], dist[:, 1:], dist[:, :-1]),
                         (owner[1:, :], owner[:-1, :], dist[1:, :], dist[:-1, :])):
        m = a != b
        if m.any():
            best = min(best, float((da + db)[m].min()))
    return max(best - 1.0, 0.0)


def min_gutter(uv, faces, islands, res=4096, dev=None):
    """Smallest texel gap between two different islands (metrics.py:348-373),
    with the conservative raster on the GPU (uva_raster_last) and the distance
    transform on the host.  Returns (gap, label image int32)."""
    dev = require_gpu(dev)
    Ud = torch.as_tensor(np.ascontiguousarray(uv, np.float64).reshape(-1, 2), device=dev)
    Fd = torch.as_tensor(np.ascontiguousarray(faces, np.int64).reshape(-1, 3), device=dev)
    isl = torch.as_tensor(np.ascontiguousarray(islands, np.int64).reshape(-1), device=dev)
    img = _labels_of_raster(_raster_last_dev(Ud, Fd, res, dev), isl).cpu().numpy()
    return _gap_of_labels(img), img


def _labels_of_raster(last, isl):
    """Island label per texel (int32, -1 empty) from the last-triangle image."""
    lab = isl[last.clamp_min(0).long()].to(torch.int32)
    return torch.where(last >= 0, lab, torch.full_like(lab, -1))


def validate_uvs(v, f, uv, atlas=4096, padding=None, in_v=None, in_f=None, in_normals=None, out_normals=None,
                 in_colors=None, out_colors=None, xref=None, gutter=True):
    """Checks on final UVs (validate.py:139-210), the gutter measurement
    included (``gutter``: min_gutter_texels at min(atlas, 8192) texels per
    side, and gutter_ok with a ``padding``).  Returns (report dict,
    per-triangle problem mask)."""
    dev = require_gpu(None)
    v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(f, np.int64).reshape(-1, 3)
    U32 = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    if len(f) and (f.min() < 0 or f.max() >= len(v) or f.max() >= len(U32)):
        raise ValueError("f holds vertex ids outside the vertex arrays")
    vd = torch.as_tensor(v, device=dev)
    fd = torch.as_tensor(f, device=dev)
    U32d = torch.as_tensor(U32, device=dev)
    Ud = U32d.double()
    rep = {}
    rep["non_finite_uv"] = int((~torch.isfinite(Ud)).any(1).sum())
    rep["uv_outside_unit_square"] = int(((Ud < 0) | (Ud > 1)).any(1).sum())
    s3d, sad, e2d = _triangle_terms(vd, fd, Ud)
    sa = sad.cpu().numpy()
    _, sliver, has3d = _triangle_checks(s3d.cpu().numpy(), sa, e2d.cpu().numpy())
    rep["degenerate_uv_triangles"] = int((has3d & (sa == 0)).sum())
    rep["sub_float32_slivers"] = int((has3d & sliver).sum())
    isld, n = _uv_islands_dev(fd, len(v), U32d, vd.float())
    skipd = torch.as_tensor((~has3d | sliver | ~np.isfinite(sa)).astype(np.uint8), device=dev)
    G = int(np.clip(np.sqrt(len(f)) / 2, 64, 4096))
    Uc = torch.where(torch.isfinite(Ud), Ud, torch.zeros_like(Ud))   # only skipped triangles read a zeroed corner
    ovd = _flags_dev(Uc, fd, skipd, 1e-6, False, G, dev)
    ovsd = _flags_dev(Uc, fd, skipd, 0.0, True, G, dev)
    # per-island sums in triangle order (np.bincount), as the reference adds them
    isl = isld.cpu().numpy()
    ov, ov_strict = ovd.cpu().numpy().astype(bool), ovsd.cpu().numpy().astype(bool)
    w = np.abs(sa)
    pos = np.bincount(isl, weights=np.where(sa > 0, w, 0), minlength=n)
    neg = np.bincount(isl, weights=np.where(sa < 0, w, 0), minlength=n)
    dom = np.where(pos >= neg, 1.0, -1.0)
    flipped = has3d & ~sliver & (np.sign(sa) != dom[isl]) & (sa != 0)
    rep["flipped_triangles"] = int(flipped.sum())
    rep["mirrored_islands"] = int((neg > pos).sum())
    rep["islands"] = int(n)
    rep["overlapping_triangles_exact"] = int(ov.sum())
    rep["float32_contacts"] = int((ov_strict & ~ov).sum())
    if ov.any():
        rep["islands_with_overlap"] = int(len(np.unique(isl[ov])))
    if gutter:
        res = min(int(atlas), 8192)
        img = _labels_of_raster(_raster_last_dev(Ud, fd, res, dev), isld).cpu().numpy()
        rep["min_gutter_texels"] = float(_gap_of_labels(img) * atlas / res)
        if padding is not None:
            rep["gutter_ok"] = bool(rep["min_gutter_texels"] >= padding - 1e-6)
    if in_v is not None and in_f is not None:
        in_v = np.asarray(in_v, np.float64)
        in_f = np.asarray(in_f, np.int64)
        rep["triangle_count_preserved"] = bool(len(in_f) == len(f))
        if len(in_f) == len(f):
            rep["corner_positions_exact"] = bool(np.array_equal(v[f].astype(np.float32), in_v[in_f].astype(np.float32)))
            if in_normals is not None and out_normals is not None:
                rep["corner_normals_exact"] = bool(np.array_equal(np.asarray(out_normals)[f], np.asarray(in_normals)[in_f]))
            if in_colors is not None and out_colors is not None:
                rep["corner_colors_exact"] = bool(np.array_equal(np.asarray(out_colors)[f], np.asarray(in_colors)[in_f]))
    problems = ov | flipped
    bad = (rep["non_finite_uv"] + rep["uv_outside_unit_square"] + rep["flipped_triangles"] +
           rep["overlapping_triangles_exact"])
    if gutter and padding is not None and not rep.get("gutter_ok", True):
        bad += 1
    for k in ("triangle_count_preserved", "corner_positions_exact", "corner_normals_exact", "corner_colors_exact"):
        if rep.get(k) is False:
            bad += 1
    rep["valid"] = bool(bad == 0)
    return rep, problems
