"""The reference's validation of final UVs (engine/validate.py) with the exact
triangle-overlap test on the GPU (``uva_overlap_flags``).

  overlap_flags        validate.py:94-136 _overlaps (and 257-299 _overlaps_strict)
  chart_overlaps       validate.py:212-226
  validate_uvs         validate.py:139-210 with gutter=False: range, degenerate and
                       sub-float32 triangles, flips against the island's dominant
                       orientation, mirrored islands, exact overlaps, float32
                       contacts, preservation of positions / normals / colours
  uv_islands           metrics.py:76-103

The gutter measurement (metrics.min_gutter, ``gutter=True``) is not part of
this build; the reference's pipeline validates with ``gutter=False``
(pipeline.py:1254).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .plan import require_gpu, stream_of


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
    U = torch.as_tensor(U, device=dev)
    Ft = torch.as_tensor(F, device=dev)
    sk = None if skip is None else torch.as_tensor(np.ascontiguousarray(skip, np.uint8), device=dev)
    out = torch.empty(len(Ft), dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().uva_overlap_flags(U.data_ptr(), len(U), Ft.data_ptr(), len(Ft),
                                            None if sk is None else sk.data_ptr(), float(rel), int(bool(strict)),
                                            int(grid), out.data_ptr(), stream_of(dev)), "uva_overlap_flags")
    return out.cpu().numpy().astype(bool)


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


def uv_islands(faces, n_vertices, uv=None, vertices=None):
    """Islands = faces connected through shared vertices; with ``uv`` and
    ``vertices``, vertices with identical float32 position and UV count as one
    (metrics.py:76-103).  Returns (island id per face, island count); ids are
    the reference's up to a renaming."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    F = np.asarray(faces, np.int64)
    if uv is not None and vertices is not None:
        key = np.concatenate([np.asarray(vertices, np.float32), np.asarray(uv, np.float32)], 1)
        key = np.ascontiguousarray(key).view(np.dtype((np.void, 20))).reshape(-1)
        _, inv = np.unique(key, return_inverse=True)
        F = inv.reshape(-1)[F]
        n_vertices = int(inv.max()) + 1
    if len(F) == 0:
        return np.zeros(0, np.int32), 0
    A = sp.csr_matrix((np.ones(3 * len(F)), (np.repeat(np.arange(len(F)), 3), F.reshape(-1))),
                      shape=(len(F), int(n_vertices)))
    n, lab = connected_components(A @ A.T, directed=False)
    return lab.astype(np.int32), int(n)


def validate_uvs(v, f, uv, atlas=4096, padding=None, in_v=None, in_f=None, in_normals=None, out_normals=None,
                 in_colors=None, out_colors=None, xref=None, gutter=True):
    """Checks on final UVs (validate.py:139-210).  Returns (report dict,
    per-triangle problem mask).  ``gutter=True`` is not supported."""
    if gutter:
        raise NotImplementedError("the gutter measurement is not part of this build; call with gutter=False")
    v = np.asarray(v, np.float64)
    f = np.asarray(f, np.int64)
    U = np.asarray(uv, np.float32).astype(np.float64)
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
    has3d = a3 > 1e-14 * max(float(a3.max()) if len(a3) else 0.0, 1e-30)
    rep["degenerate_uv_triangles"] = int((has3d & (sa == 0)).sum())
    rep["sub_float32_slivers"] = int((has3d & sliver).sum())
    isl, n = uv_islands(f, len(v), U, v)
    w = np.abs(sa)
    pos = np.bincount(isl, weights=np.where(sa > 0, w, 0), minlength=n)
    neg = np.bincount(isl, weights=np.where(sa < 0, w, 0), minlength=n)
    dom = np.where(pos >= neg, 1.0, -1.0)
    flipped = has3d & ~sliver & (np.sign(sa) != dom[isl]) & (sa != 0)
    rep["flipped_triangles"] = int(flipped.sum())
    rep["mirrored_islands"] = int((neg > pos).sum())
    rep["islands"] = int(n)
    skip = ~has3d | sliver | ~np.isfinite(sa)
    G = int(np.clip(np.sqrt(len(f)) / 2, 64, 4096))
    Uc = np.where(np.isfinite(U), U, 0.0)            # skipped triangles only: their cells do not matter
    ov = overlap_flags(Uc, f, skip, 1e-6, False, G)
    rep["overlapping_triangles_exact"] = int(ov.sum())
    ov_strict = overlap_flags(Uc, f, skip, 0.0, True, G)
    rep["float32_contacts"] = int((ov_strict & ~ov).sum())
    if ov.any():
        rep["islands_with_overlap"] = int(len(np.unique(isl[ov])))
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
    for k in ("triangle_count_preserved", "corner_positions_exact", "corner_normals_exact", "corner_colors_exact"):
        if rep.get(k) is False:
            bad += 1
    rep["valid"] = bool(bad == 0)
    return rep, problems
