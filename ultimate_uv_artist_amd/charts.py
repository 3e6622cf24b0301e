"""Many small ``charts.flatten`` calls as one batched GPU solve.

The reference flattens thousands of 40-90-face charts one call at a time from
several threads (grow.validate, merge_islands, chart_hier; charts.py:401-438).
``flatten_many`` runs the same sequence for a list of charts with one plan:

  LSCM of every chart                                   uva_lscm
  mirror charts with more negative than positive
  triangles (charts.py:407-410)                          uva_signed_areas + torch
  ARAP, iterations doubled and no line search for a
  folded start, energy early exit 1e-4 (charts.py:411-415)
                                                         uva_arap (iters_folded)
  sequential flip repair of charts that keep folds, and
  a second flip-safe ARAP of max(2, iters // 2)
  iterations when the repair removed them (charts.py:416-421)
                                                         uva_repair_seq, uva_arap
  boundary crossings and their local untangle,
  chart_quality (charts.py:428-437)                      host (small charts)

``FlattenQueue`` gathers flatten requests from many threads into such batches.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import Future

import numpy as np
import torch

from . import _lib, param
from .plan import IslandPlan, require_gpu


# Relative residual of the global steps.  The reference solves them exactly
# (SuperLU); small charts reach 1e-10 in a few dozen CG iterations, and a
# folded start's unconstrained iterations amplify a looser solve (1e-6 left
# 1.8e-4 on a needle chart).
FLATTEN_CG_TOL = 1e-10


def _has_boundary(T):
    return len(param.boundary_edges_oriented(np.asarray(T, np.int64))[0]) > 0


# ------------------------------------------------------------ boundary crossings
def _cells(A, B, cell, lo, nx, ny):
    """Grid cells each segment's box covers, as (segment, cell) pairs in
    segment order (charts.py:555-570)."""
    mn, mx = np.minimum(A, B), np.maximum(A, B)
    x0 = np.maximum(((mn[:, 0] - lo[0]) / cell).astype(np.int64), 0)
    x1 = np.minimum(((mx[:, 0] - lo[0]) / cell).astype(np.int64), nx - 1)
    y0 = np.maximum(((mn[:, 1] - lo[1]) / cell).astype(np.int64), 0)
    y1 = np.minimum(((mx[:, 1] - lo[1]) / cell).astype(np.int64), ny - 1)
    wx, wy = np.maximum(x1 - x0 + 1, 0), np.maximum(y1 - y0 + 1, 0)
    cnt = wx * wy
    seg = np.repeat(np.arange(len(A)), cnt)
    k = np.arange(len(seg)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    gx = x0[seg] + k // np.maximum(wy[seg], 1)
    gy = y0[seg] + k % np.maximum(wy[seg], 1)
    return seg, gy * nx + gx


def _crossing_pairs(A, B, cell, lo, nx, ny, maxk):
    """Pairs (i, j) of segments that cross properly, tested once per grid cell
    both cover (a pair in two cells is listed twice), in cell order, then
    by i, then by j; at most ``maxk`` (charts.py:553-594, overlapfix.py:88-127)."""
    seg, c = _cells(A, B, cell, lo, nx, ny)
    o = np.lexsort((seg, c))
    seg, c = seg[o], c[o]
    _, start, cnt = np.unique(c, return_index=True, return_counts=True)
    r = np.arange(len(seg)) - np.repeat(start, cnt)
    npair = np.repeat(cnt, cnt) - 1 - r
    first = np.repeat(np.arange(len(seg)), npair)
    second = first + 1 + np.arange(len(first)) - np.repeat(np.cumsum(npair) - npair, npair)
    i, j = seg[first], seg[second]
    Ai, Bi, Aj, Bj = A[i], B[i], A[j], B[j]
    touch = ((Ai == Aj).all(1) | (Ai == Bj).all(1) | (Bi == Aj).all(1) | (Bi == Bj).all(1))
    d1, d2 = Bi - Ai, Bj - Aj
    o1 = d1[:, 0] * (Aj[:, 1] - Ai[:, 1]) - d1[:, 1] * (Aj[:, 0] - Ai[:, 0])
    o2 = d1[:, 0] * (Bj[:, 1] - Ai[:, 1]) - d1[:, 1] * (Bj[:, 0] - Ai[:, 0])
    o3 = d2[:, 0] * (Ai[:, 1] - Aj[:, 1]) - d2[:, 1] * (Ai[:, 0] - Aj[:, 0])
    o4 = d2[:, 0] * (Bi[:, 1] - Aj[:, 1]) - d2[:, 1] * (Bi[:, 0] - Aj[:, 0])
    t1, t2 = 1e-9 * (d1 * d1).sum(1), 1e-9 * (d2 * d2).sum(1)
    hit = ~touch & (((o1 > t1) & (o2 < -t1)) | ((o1 < -t1) & (o2 > t1))) & \
        (((o3 > t2) & (o4 < -t2)) | ((o3 < -t2) & (o4 > t2)))
    return np.stack([i[hit], j[hit]], 1)[:maxk]


def _grid(U, bi, bj):
    A, B = U[bi], U[bj]
    lo = U.min(0)
    ext = U.max(0) - lo
    L = np.linalg.norm(B - A, axis=1)
    cell = max(float(np.median(L)) * 4, float(ext.max()) / 512, 1e-30)
    return A, B, cell, lo, int(ext[0] / cell) + 1, int(ext[1] / cell) + 1


def boundary_intersections(U, T, max_report=1000):
    """Proper crossings between non-adjacent boundary edges of the chart's
    UV layout, counted once per grid cell both cover (charts.py:597-608)."""
    bi, bj = param.boundary_edges_oriented(np.asarray(T, np.int64))
    if len(bi) < 3:
        return 0
    U = np.asarray(U, np.float64)
    return int(len(_crossing_pairs(*_grid(U, bi, bj), max_report)))


def _concat(Us, Ts):
    """Charts' layouts and faces concatenated (faces in global ids), with the
    vertex and face offsets."""
    vo = np.concatenate([[0], np.cumsum([len(U) for U in Us])]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum([len(T) for T in Ts])]).astype(np.int64)
    U = np.concatenate([np.asarray(u, np.float64).reshape(-1, 2) for u in Us]) if Us else np.zeros((0, 2))
    T = (np.concatenate([np.asarray(t, np.int64).reshape(-1, 3) + vo[i] for i, t in enumerate(Ts)])
         if Ts else np.zeros((0, 3), np.int64))
    return vo, to, U, T


def boundary_intersections_many(Us, Ts, max_report=1000):
    """boundary_intersections of every chart (Us[i], Ts[i]), equal to the
    per-chart calls: the boundary half-edges, cell grids and crossing tests
    of all charts at once, each chart's cells apart from the others'."""
    k = len(Ts)
    out = np.zeros(k, np.int64)
    if k == 0:
        return out.tolist()
    vo, to, U, T = _concat(Us, Ts)
    n = len(U)
    src = np.concatenate([T[:, 0], T[:, 1], T[:, 2]])
    dst = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    has_twin = np.isin(dst * n + src, src * n + dst)
    bi, bj = src[~has_twin], dst[~has_twin]
    vch = np.repeat(np.arange(k), np.diff(vo))
    ech = vch[bi]
    nb = np.bincount(ech, minlength=k)
    live = nb >= 3                                    # fewer than 3 boundary edges: no crossing
    # a layout with a non-finite coordinate takes the per-chart path
    finite = np.minimum.reduceat(np.isfinite(U).all(1), vo[:-1]) if n else np.ones(k, bool)
    for c in np.flatnonzero(live & ~finite):
        out[c] = boundary_intersections(Us[c], Ts[c], max_report)
    live &= finite
    keep = live[ech]
    bi, bj, ech = bi[keep], bj[keep], ech[keep]
    if len(bi) == 0:
        return out.tolist()
    A, B = U[bi], U[bj]
    # per chart: the layout's box over all its vertices, the median boundary edge
    lo = np.minimum.reduceat(U, vo[:-1], axis=0)
    ext = np.maximum.reduceat(U, vo[:-1], axis=0) - lo
    L = np.linalg.norm(B - A, axis=1)
    o = np.lexsort((L, ech))
    Ls, cs = L[o], ech[o]
    cnt = np.bincount(cs, minlength=k)
    st = np.concatenate([[0], np.cumsum(cnt)])[:-1]
    med = np.zeros(k)
    lv = np.flatnonzero(cnt > 0)
    h = cnt[lv] // 2
    odd = cnt[lv] % 2 == 1
    a_ = Ls[st[lv] + np.where(odd, h, h - 1)]
    b_ = Ls[st[lv] + h]
    med[lv] = np.where(odd, b_, (a_ + b_) / 2)
    cell = np.maximum(np.maximum(med * 4, ext.max(1) / 512), 1e-30)
    nx = (ext[:, 0] / cell).astype(np.int64) + 1
    ny = (ext[:, 1] / cell).astype(np.int64) + 1
    # cells each segment's box covers (as _cells, with its chart's grid)
    c_e, lo_e, nx_e, ny_e = cell[ech], lo[ech], nx[ech], ny[ech]
    mn, mx = np.minimum(A, B), np.maximum(A, B)
    x0 = np.maximum(((mn[:, 0] - lo_e[:, 0]) / c_e).astype(np.int64), 0)
    x1 = np.minimum(((mx[:, 0] - lo_e[:, 0]) / c_e).astype(np.int64), nx_e - 1)
    y0 = np.maximum(((mn[:, 1] - lo_e[:, 1]) / c_e).astype(np.int64), 0)
    y1 = np.minimum(((mx[:, 1] - lo_e[:, 1]) / c_e).astype(np.int64), ny_e - 1)
    wx, wy = np.maximum(x1 - x0 + 1, 0), np.maximum(y1 - y0 + 1, 0)
    ncell = wx * wy
    seg = np.repeat(np.arange(len(A)), ncell)
    q = np.arange(len(seg)) - np.repeat(np.cumsum(ncell) - ncell, ncell)
    gx = x0[seg] + q // np.maximum(wy[seg], 1)
    gy = y0[seg] + q % np.maximum(wy[seg], 1)
    cbase = np.concatenate([[0], np.cumsum(nx * ny)])[:-1]
    c = cbase[ech[seg]] + gy * nx_e[seg] + gx
    # pairs of segments in one cell, tested as _crossing_pairs does
    o = np.lexsort((seg, c))
    seg, c = seg[o], c[o]
    _, start, cc = np.unique(c, return_index=True, return_counts=True)
    r = np.arange(len(seg)) - np.repeat(start, cc)
    npair = np.repeat(cc, cc) - 1 - r
    first = np.repeat(np.arange(len(seg)), npair)
    second = first + 1 + np.arange(len(first)) - np.repeat(np.cumsum(npair) - npair, npair)
    i, j = seg[first], seg[second]
    Ai, Bi, Aj, Bj = A[i], B[i], A[j], B[j]
    touch = ((Ai == Aj).all(1) | (Ai == Bj).all(1) | (Bi == Aj).all(1) | (Bi == Bj).all(1))
    d1, d2 = Bi - Ai, Bj - Aj
    o1 = d1[:, 0] * (Aj[:, 1] - Ai[:, 1]) - d1[:, 1] * (Aj[:, 0] - Ai[:, 0])
    o2 = d1[:, 0] * (Bj[:, 1] - Ai[:, 1]) - d1[:, 1] * (Bj[:, 0] - Ai[:, 0])
    o3 = d2[:, 0] * (Ai[:, 1] - Aj[:, 1]) - d2[:, 1] * (Ai[:, 0] - Aj[:, 0])
    o4 = d2[:, 0] * (Bi[:, 1] - Aj[:, 1]) - d2[:, 1] * (Bi[:, 0] - Aj[:, 0])
    t1, t2 = 1e-9 * (d1 * d1).sum(1), 1e-9 * (d2 * d2).sum(1)
    hit = ~touch & (((o1 > t1) & (o2 < -t1)) | ((o1 < -t1) & (o2 > t1))) & \
        (((o3 > t2) & (o4 < -t2)) | ((o3 < -t2) & (o4 > t2)))
    out[live] = np.minimum(np.bincount(ech[i[hit]], minlength=k), max_report)[live]
    return out.tolist()


def chart_quality_many(charts, Us):
    """param.chart_quality of every chart (charts[i] = (P, T), layout Us[i]),
    equal to the per-chart calls: the per-triangle terms of all charts at
    once, the per-chart sums and weighted quantiles on each chart's slice."""
    k = len(charts)
    if k == 0:
        return []
    vo, to, U, T = _concat(Us, [t for _, t in charts])
    P = np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p, _ in charts])
    X, area = param.local_frames(P, T)
    sa = param.signed_areas(U, T)
    x1, x2, y2 = X[:, 1, 0], X[:, 2, 0], X[:, 2, 1]
    good = np.abs(x1 * y2) > 1e-300
    x1 = np.where(good, x1, 1)
    y2 = np.where(good, y2, 1)
    d1 = U[T[:, 1]] - U[T[:, 0]]
    d2 = U[T[:, 2]] - U[T[:, 0]]
    j00, j10 = d1[:, 0] / x1, d1[:, 1] / x1
    j01, j11 = (d2[:, 0] - j00 * x2) / y2, (d2[:, 1] - j10 * x2) / y2
    conf = np.sqrt(((j00 + j11) / 2) ** 2 + ((j10 - j01) / 2) ** 2)
    anti = np.sqrt(((j00 - j11) / 2) ** 2 + ((j10 + j01) / 2) ** 2)
    s1, s2 = conf + anti, np.abs(conf - anti)
    res = []
    for i in range(k):
        sl = slice(to[i], to[i + 1])
        ar_i, sa_i, s1_i, s2_i = area[sl], sa[sl], s1[sl], s2[sl]
        tot = ar_i.sum()
        flips = int((sa_i <= 0).sum()) if (sa_i > 0).sum() >= (sa_i < 0).sum() else int((sa_i >= 0).sum())
        sc = np.abs(sa_i).sum() / max(tot, 1e-300)
        w = ar_i / max(tot, 1e-300)
        an = s1_i / np.maximum(s2_i, 1e-12)
        ar = (s1_i * s2_i) / max(sc, 1e-300)
        lar = np.abs(np.log2(np.maximum(ar, 1e-12)))

        def wq(x, q):
            o = np.argsort(x)
            cw = np.cumsum(w[o])
            return float(x[o][min(np.searchsorted(cw, q * cw[-1]), len(o) - 1)])

        st = np.sqrt((s1_i ** 2 + s2_i ** 2) / 2 / max(sc, 1e-300))
        res.append({"flips": flips, "aniso_p95": wq(an, 0.95), "aniso_mean": float((an * w).sum()),
                    "area_p95": float(2 ** wq(lar, 0.95)), "area_logdev": float((lar * w).sum()),
                    "stretch": float(np.sqrt((st ** 2 * w).sum())), "scale": float(np.sqrt(sc)),
                    "per_tri_aniso": an, "per_tri_area": ar})
    return res


def untangle_boundary(U, T, rounds=40):
    """Remove boundary self-crossings by moving the vertices around each
    crossing toward their 1-ring centroid, never folding a triangle and never
    shrinking one below 20 % of its area (charts.py:440-544).  Returns
    (U, success)."""
    T = np.asarray(T, np.int64)
    U = np.array(U, np.float64)
    bi, bj = param.boundary_edges_oriented(T)
    if len(bi) < 3:
        return U, True
    n = len(U)
    I = np.r_[T[:, 0], T[:, 1], T[:, 2], T[:, 1], T[:, 2], T[:, 0]]
    J = np.r_[T[:, 1], T[:, 2], T[:, 0], T[:, 0], T[:, 1], T[:, 2]]
    key = np.unique(I * n + J)
    I, J = key // n, key % n
    ptr = np.searchsorted(I, np.arange(n + 1))
    deg = np.diff(ptr)
    vtri = [[] for _ in range(n)]
    for t, row in enumerate(T.tolist()):
        for v in row:
            vtri[v].append(t)
    for r in range(rounds):
        hits = _crossing_pairs(*_grid(U, bi, bj), 512)
        if len(hits) == 0:
            return U, True
        moved = False
        for sel in (hits, hits[:1]):
            seeds = np.unique(np.r_[bi[sel[:, 0]], bj[sel[:, 0]], bi[sel[:, 1]], bj[sel[:, 1]]])
            reg = np.zeros(n, bool)
            reg[seeds] = True
            cur = seeds
            for _ in range(1 + r // 8):                      # rings around the seeds
                nb = np.unique(np.concatenate([J[ptr[v]:ptr[v + 1]] for v in cur])) if len(cur) else cur
                cur = nb[~reg[nb]]
                reg[cur] = True
            rv = np.nonzero(reg)[0]
            tris = np.unique(np.concatenate([vtri[v] for v in rv]).astype(np.int64))
            Tt = T[tris]
            sa0 = param.signed_areas(U, Tt)
            was_ok = sa0 > 0
            mean = np.zeros((n, 2))
            has = deg > 0
            if has.any():
                mean[has] = np.add.reduceat(U[J], ptr[:-1][has], axis=0) / deg[has][:, None]
            for step in (0.5, 0.25, 0.1):
                Un = U.copy()
                m = reg & (deg > 0)
                Un[m] = (1 - step) * U[m] + step * mean[m]
                if ((param.signed_areas(Un, Tt) > 0.2 * sa0) | ~was_ok).all():
                    U = Un
                    moved = True
                    break
            if moved:
                break
        if not moved:
            return U, False
    return U, False


def flatten_many(charts, iters=8, dev=None, slim_iters=0):
    """charts: [(P (k,3), T (t,3))] -> [(U (k,2), quality dict)] in order, as
    ``charts.flatten(P, T, iters, slim_iters=slim_iters)`` (method "arap")
    for each chart: with ``slim_iters``, the charts without folds after the
    repair take that many SLIM iterations in one batched plan
    (charts.py:422-425; a layout with a non-finite value is not taken).
    A chart without boundary raises ValueError as param.lscm does."""
    dev = require_gpu(dev)
    charts = [(np.asarray(P, np.float64), np.asarray(T, np.int64)) for P, T in charts]
    for P, T in charts:
        if not _has_boundary(T):
            raise ValueError("chart has no boundary (closed surface): needs a cut")
    if not charts:
        return []
    plan = IslandPlan(charts, dev=dev)
    vo, to = plan.vert_off, plan.tri_off
    U = plan.device_uv(np.zeros((plan.n, 2)))
    plan.lscm(U)
    # majority orientation per chart (charts.py:407-410)
    tri_isl = torch.as_tensor(np.repeat(np.arange(len(charts)), np.diff(to)), device=dev)
    sa = plan.signed_areas(U)
    neg = torch.zeros(len(charts), dtype=torch.float64, device=dev).index_add_(0, tri_isl, (sa < 0).double())
    pos = torch.zeros(len(charts), dtype=torch.float64, device=dev).index_add_(0, tri_isl, (sa > 0).double())
    vert_isl = torch.as_tensor(np.repeat(np.arange(len(charts)), np.diff(vo)), device=dev)
    flip = (neg > pos)[vert_isl]
    U[:, 1] = torch.where(flip, -U[:, 1], U[:, 1])
    if iters:
        plan.arap(U, iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=1e-4, cg_tol=FLATTEN_CG_TOL,
                  iters_folded=2 * iters)
    Uh = U.cpu().numpy()
    out = [Uh[vo[i]:vo[i + 1]].copy() for i in range(len(charts))]
    # charts that keep folds: sequential repair, then a flip-safe ARAP if it removed them
    redo = []
    for i, (P, T) in enumerate(charts):
        if (param.signed_areas(out[i], T) <= 0).any():
            out[i] = param.repair_flips(out[i], T)
            if not (param.signed_areas(out[i], T) <= 0).any() and iters:
                redo.append(i)
    if redo:
        sub = IslandPlan([charts[i] for i in redo], dev=dev)
        U2 = sub.device_uv(np.concatenate([out[i] for i in redo]))
        sub.arap(U2, max(2, iters // 2), line_search=True, mode=_lib.UVA_MODE_REFINE, tol=1e-4,
                 cg_tol=FLATTEN_CG_TOL)
        U2 = U2.cpu().numpy()
        for k, i in enumerate(redo):
            out[i] = U2[sub.vert_off[k]:sub.vert_off[k + 1]].copy()
    if slim_iters:
        sl = [i for i, (_, T) in enumerate(charts) if not (param.signed_areas(out[i], T) <= 0).any()]
        if sl:
            sub = IslandPlan([charts[i] for i in sl], dev=dev)
            U3 = sub.device_uv(np.concatenate([out[i] for i in sl]))
            sub.slim(U3, int(slim_iters), tol=1e-5)
            U3 = U3.cpu().numpy()
            for k, i in enumerate(sl):
                Us = U3[sub.vert_off[k]:sub.vert_off[k + 1]]
                if np.isfinite(Us).all():
                    out[i] = Us.copy()
    # boundary crossings and their local untangle (charts.py:428-437), then
    # the quality, on the host for all charts at once
    hits = boundary_intersections_many(out, [T for _, T in charts])
    for i, (P, T) in enumerate(charts):
        if hits[i] and not (param.signed_areas(out[i], T) <= 0).any():
            U2, ok = untangle_boundary(out[i], T)
            if ok:
                out[i], hits[i] = U2, 0
    res = []
    for i, q in enumerate(chart_quality_many(charts, out)):
        q["boundary_hits"] = hits[i]
        q["overlap"] = 1.0 if hits[i] else 0.0
        res.append((out[i], q))
    return res


class FlattenQueue:
    """Collects ``flatten`` requests from any number of threads and runs them
    in batches of up to ``max_batch`` charts: a batch starts when it is full
    or ``max_wait_ms`` after its first request, and runs one flatten_many per
    (iters, slim_iters) setting it holds.  ``flatten(P, T, iters=None,
    slim_iters=0)`` blocks and returns what ``charts.flatten(P, T, iters,
    slim_iters=slim_iters)`` would (``iters`` None: the queue's default);
    ``submit`` returns a Future."""

    def __init__(self, iters=8, max_batch=4096, max_wait_ms=2.0, dev=None):
        self.iters = int(iters)
        self.max_batch = int(max_batch)
        self.max_wait = float(max_wait_ms) * 1e-3
        self.dev = require_gpu(dev)
        self._cv = threading.Condition()
        self._pending = []
        self._closed = False
        self._worker = threading.Thread(target=self._run, name="flatten-queue", daemon=True)
        self._worker.start()

    def submit(self, P, T, iters=None, slim_iters=0):
        fut = Future()
        key = (self.iters if iters is None else int(iters), int(slim_iters))
        with self._cv:
            if self._closed:
                raise RuntimeError("FlattenQueue is closed")
            self._pending.append((P, T, fut, key))
            self._cv.notify()
        return fut

    def flatten(self, P, T, iters=None, slim_iters=0):
        return self.submit(P, T, iters, slim_iters).result()

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify()
        self._worker.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _run(self):
        # Every request's future is resolved, with its result or with the
        # exception its call raised (as charts.flatten would raise it in the
        # caller's thread); nothing a request does ends this thread.
        try:
            if self.dev.type == "cuda":
                torch.cuda.set_device(self.dev)
        except Exception as e:
            self._fail_all(e)
            return
        while True:
            with self._cv:
                while not self._pending and not self._closed:
                    self._cv.wait()
                if not self._pending and self._closed:
                    return
                t0 = time.monotonic()
                while len(self._pending) < self.max_batch and not self._closed:
                    left = self.max_wait - (time.monotonic() - t0)
                    if left <= 0:
                        break
                    self._cv.wait(left)
                batch, self._pending = self._pending[:self.max_batch], self._pending[self.max_batch:]
            try:
                self._run_batch(batch)
            except BaseException as e:       # unexpected: every unresolved caller of the batch sees it
                for _, _, f, _ in batch:
                    if not f.done():
                        f.set_exception(e)

    def _run_batch(self, batch):
        groups = {}
        for P, T, f, key in batch:
            if not f.set_running_or_notify_cancel():
                continue
            try:
                if not _has_boundary(T):
                    raise ValueError("chart has no boundary (closed surface): needs a cut")
            except Exception as e:           # e.g. an empty or malformed triangle list
                f.set_exception(e)
                continue
            groups.setdefault(key, []).append((P, T, f))
        for (it, sl), ok in groups.items():
            try:
                res = flatten_many([(P, T) for P, T, _ in ok], it, self.dev, slim_iters=sl)
                for (_, _, f), r in zip(ok, res):
                    f.set_result(r)
            except Exception as e:        # the group failed as a whole: each of its callers sees it
                for _, _, f in ok:
                    if not f.done():
                        f.set_exception(e)

    def _fail_all(self, e):
        with self._cv:
            self._closed = True
            pending, self._pending = self._pending, []
        for _, _, f, _ in pending:
            if f.set_running_or_notify_cancel():
                f.set_exception(e)
