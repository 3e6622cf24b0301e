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
  chart_quality (charts.py:432)                          host

The boundary-intersection test and its local untangle (charts.py:428-431)
are not part of this build: ``q["boundary_hits"]`` is None and callers that
need them apply the reference's functions to the returned layouts.
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


def flatten_many(charts, iters=8, dev=None):
    """charts: [(P (k,3), T (t,3))] -> [(U (k,2), quality dict)] in order, as
    ``charts.flatten(P, T, iters)`` (method "arap", no SLIM) for each chart.
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
    res = []
    for (P, T), Ui in zip(charts, out):
        q = param.chart_quality(P, T, Ui)
        q["boundary_hits"] = None
        q["overlap"] = None
        res.append((Ui, q))
    return res


class FlattenQueue:
    """Collects ``flatten`` requests from any number of threads and runs them
    in batches of up to ``max_batch`` charts: a batch starts when it is full
    or ``max_wait_ms`` after its first request.  ``flatten(P, T)`` blocks and
    returns what ``charts.flatten(P, T, iters)`` would (see flatten_many);
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

    def submit(self, P, T):
        fut = Future()
        with self._cv:
            if self._closed:
                raise RuntimeError("FlattenQueue is closed")
            self._pending.append((P, T, fut))
            self._cv.notify()
        return fut

    def flatten(self, P, T):
        return self.submit(P, T).result()

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
        torch.cuda.set_device(self.dev)
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
            good = [(P, T, f) for P, T, f in batch if f.set_running_or_notify_cancel()]
            ok = []
            for P, T, f in good:
                if _has_boundary(T):
                    ok.append((P, T, f))
                else:
                    f.set_exception(ValueError("chart has no boundary (closed surface): needs a cut"))
            try:
                res = flatten_many([(P, T) for P, T, _ in ok], self.iters, self.dev)
                for (_, _, f), r in zip(ok, res):
                    f.set_result(r)
            except Exception as e:            # the batch failed as a whole: every caller sees it
                for _, _, f in ok:
                    if not f.done():
                        f.set_exception(e)
