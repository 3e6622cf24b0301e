"""GPU checks of the drop-in boundary beyond the solvers: the gpu.Batch
attributes and methods the reference exposes (gpu.py:82-151, 283-295),
tensor ownership, and param.* called from eight threads at once as the
reference's pools do (pipeline.py:608-611, merge.py:114)."""
from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _items():
    from pathlib import Path

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "batch_golden.npz"))
    return [(g[f"P{i}"], g[f"T{i}"].astype(np.int64), g[f"U{i}"]) for i in range(int(g["n"]))]


def test_batch_attributes_match_the_reference_formulas():
    """area, ok, w, x1, x2, y2, X, L, dinv, wpin and jacobians() against the
    Batch oracle's restatement of gpu.py:88-151, in float64 and float32."""
    import torch

    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import gpu

    items = _items()
    r = BatchOracle(items)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-6)):
        b = gpu.Batch(items, dtype=dt)
        for name in ("area", "w", "x1", "x2", "y2", "X"):
            got = getattr(b, name)
            assert got.dtype == dt and got.device.type == "cuda", name
            ref = getattr(r, name)
            assert got.shape == ref.shape, name
            assert np.abs(got.cpu().double().numpy() - ref).max() <= tol * max(1.0, np.abs(ref).max()), name
        assert b.ok.dtype == torch.bool and np.array_equal(b.ok.cpu().numpy(), r.ok)
        L = b.L
        assert L.layout == torch.sparse_csr and L.crow_indices().dtype == torch.int32 and L.values().dtype == dt
        Ld = L.to_dense().cpu().double().numpy()
        Lr = r.L.toarray()
        assert np.abs(Ld - Lr).max() <= tol * np.abs(Lr).max()
        assert np.abs(b.dinv.cpu().double().numpy() - 1.0 / np.maximum(Lr.diagonal(), 1e-30)).max() <= \
            tol * np.abs(1.0 / Lr.diagonal()).max()
        assert b.wpin.dim() == 0 and abs(float(b.wpin) - r.wpin) <= tol * r.wpin
        U = torch.as_tensor(r.U, device=b.dev, dtype=dt)
        for got, ref in zip(b.jacobians(U), r.jacobians(r.U)):
            assert got.dtype == dt
            assert np.abs(got.cpu().double().numpy() - ref).max() <= 1e3 * tol * max(1.0, np.abs(ref).max())


def test_batch_tensors_are_the_callers_own():
    """A tensor the caller passed in or got back is never changed by later
    calls (the reference rebinds self.U to new tensors)."""
    import torch

    from ultimate_uv_artist_amd import gpu

    items = _items()
    b = gpu.Batch(items, dtype=torch.float64)
    U_in = torch.as_tensor(np.concatenate([u for _, _, u in items]), device=b.dev)
    keep_in = U_in.clone()
    b.U = U_in
    b.orient_positive()
    out1 = b.arap(2, line_search=True, cg_iters=200)
    keep1 = out1.clone()
    got = b.U
    keep_got = got.clone()
    b.repair_flips(10)
    b.arap(2, line_search=False, cg_iters=200)
    b.orient_positive()
    assert torch.equal(U_in, keep_in)
    assert torch.equal(out1, keep1)
    assert torch.equal(got, keep_got)
    assert not torch.equal(b.U, keep1)


def test_batch_shrink_and_scale():
    """shrink lowers the target scale of the given triangles and of the
    triangles within `rings` vertex rings of them (gpu.py:283-295), clamped
    at `floor`; the scale reaches the rotations (Batch oracle with the same
    scale)."""
    import torch

    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import gpu

    items = _items()
    b = gpu.Batch(items, dtype=torch.float64)
    T = b.T.cpu().numpy()
    n = b.n
    seed = np.array([5, 40, len(T) - 3])
    # expected selection by breadth-first search over the vertex graph of the
    # triangles: a triangle is selected when one of its corners lies within
    # rings - 1 steps of a seed triangle's corner
    nb = [set() for _ in range(n)]
    for a, c, d in T.tolist():
        nb[a] |= {c, d}
        nb[c] |= {a, d}
        nb[d] |= {a, c}
    dist = np.full(n, np.inf)
    front = sorted(set(T[seed].ravel().tolist()))
    dist[front] = 0
    k = 0
    while front:
        k += 1
        nxt = sorted({w for v in front for w in nb[v] if dist[w] == np.inf})
        dist[nxt] = k
        front = nxt
    for rings in (1, 3):
        want = dist[T].min(1) <= rings - 1
        b.scale = torch.ones(len(T), device=b.dev)
        b.shrink(seed, factor=0.6, floor=0.05, rings=rings)
        s = b.scale.cpu().numpy()
        assert np.array_equal(s < 1.0, want)
        assert np.allclose(s[want], 0.6)
    b.shrink(seed, factor=0.05, floor=0.05, rings=3)    # 0.6 * 0.05 clamps at the floor
    s = b.scale.cpu().numpy()
    assert np.allclose(s[want], 0.05) and np.allclose(s[~want], 1.0)
    b.shrink([], factor=0.5)
    assert np.array_equal(b.scale.cpu().numpy(), s)
    r = BatchOracle(items)
    r.scale = s.copy()
    b.arap(2, line_search=True, cg_iters=100000, cg_tol=1e-9)
    r.arap(2, line_search=True)
    for u, ur in zip(b.chart_uvs(), r.chart_uvs()):
        lo = ur.min(0)
        sc = float((ur.max(0) - lo).max())
        assert np.abs((u - lo) / sc - (ur - lo) / sc).max() <= 1e-6


def test_param_from_eight_threads():
    """param.lscm + param.refine on independent charts from eight threads at
    once, each against the oracle (unit box 1e-4 for refine; LSCM 1e-6, and
    1e-5 on the two charts of 45k-60k triangles: the FP64 CG stops at relative
    residual 1e-8, and the error that leaves grows with the system's condition
    number, i.e. with its size; the oracle solves directly).  Each thread runs on a
    stream of its own; the two large charts get teams of several workgroups,
    whose solves take turns."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    rng = np.random.default_rng(77)
    sizes = list(rng.integers(60, 4000, 14)) + [60000, 45000]
    charts = [workloads.disk_island(rng, int(t)) for t in sizes]
    teams = [IslandPlan([c], pins=[0]).info()["n_teams"] for c in charts[-2:]]
    print("teams of the two large charts:", teams, flush=True)
    assert max(teams) > 0
    refs = []
    for P, T in charts:
        U0 = oracle.lscm(P, T)
        refs.append((U0, oracle.refine(P, T, U0, iters=6, tol=1e-5)[0]))
    out = [None] * len(charts)
    errors = []
    start = threading.Barrier(8)

    def work(k):
        try:
            start.wait()
            for i in range(k, len(charts), 8):
                P, T = charts[i]
                U0 = param.lscm(P, T)
                U1, _ = param.refine(P, T, refs[i][0], iters=6, tol=1e-5)
                out[i] = (U0, U1)
        except Exception as e:       # reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for (U0, U1), (R0, R1), (_, T) in zip(out, refs, charts):
        for got, ref, tol in ((U0, R0, 1e-6 if len(T) < 20000 else 1e-5), (U1, R1, 1e-4)):
            lo = ref.min(0)
            sc = float((ref.max(0) - lo).max())
            assert np.abs((got - lo) / sc - (ref - lo) / sc).max() <= tol
