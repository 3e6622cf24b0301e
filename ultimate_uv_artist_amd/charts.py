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
    res = []
    for (P, T), Ui in zip(charts, out):
        # boundary crossings and their local untangle (charts.py:428-437), on the host
        hits = boundary_intersections(Ui, T)
        if hits and not (param.signed_areas(Ui, T) <= 0).any():
            U2, ok = untangle_boundary(Ui, T)
            if ok:
                Ui, hits = U2, 0
        q = param.chart_quality(P, T, Ui)
        q["boundary_hits"] = hits
        q["overlap"] = 1.0 if hits else 0.0
        res.append((Ui, q))
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
