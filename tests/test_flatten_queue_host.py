"""FlattenQueue's request handling on the CPU (the batched solve replaced by
a stand-in): every future resolves, a malformed request fails in its own
caller only, and the worker thread survives it."""
from __future__ import annotations

import numpy as np
import torch

from ultimate_uv_artist_amd import charts, workloads


def test_queue_resolves_bad_requests_and_keeps_running(monkeypatch):
    calls = []

    def fake_flatten_many(items, iters=8, dev=None, slim_iters=0):
        calls.append((len(items), iters, slim_iters))
        return [(np.zeros((len(P), 2)), {"n": len(T)}) for P, T in items]

    monkeypatch.setattr(charts, "require_gpu", lambda dev=None: torch.device("cpu"))
    monkeypatch.setattr(charts, "flatten_many", fake_flatten_many)
    P, T = workloads.disk_island(np.random.default_rng(5), 60)
    closed = np.array([[0, 1, 2], [0, 2, 1]], np.int64)        # two faces glued: no boundary
    with charts.FlattenQueue(max_wait_ms=20.0) as q:
        good = q.submit(P, T, 6)
        empty = q.submit(np.zeros((0, 3)), np.zeros((0, 3), np.int64), 6)
        shut = q.submit(P[:3], closed, 6)
        junk = q.submit(P, "not a triangle list", 6)
        U, qd = good.result(timeout=30)
        assert U.shape == (len(P), 2) and qd["n"] == len(T)
        for f in (empty, shut, junk):
            assert f.exception(timeout=30) is not None
        assert isinstance(shut.exception(), ValueError)
        # the worker is still serving
        U2, _ = q.flatten(P, T, 4)
        assert U2.shape == (len(P), 2)
    assert all(n == 1 for n, _, _ in calls) and {it for _, it, _ in calls} == {6, 4}
