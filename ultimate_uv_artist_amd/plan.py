"""Host handle of a device plan: a batch of islands prepared for the kernels.

``IslandPlan`` owns one ``uva_plan`` (frames, sliced-ELL Laplacian, rhs
incidence, LSCM pins, task order, scratch) and runs LSCM / ARAP on device UV
tensors (n x 2 float64, torch, on the plan's GPU).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import threading

import numpy as np
import torch

from . import _lib
from ._lib import check, f64p, i64p, ptr

CG_TOL = 1e-6        # relative residual of every ARAP global step
LSCM_TOL = 1e-8      # relative residual of the LSCM solve (FP64)
# SLIM's normal equations are solved directly by the reference (param.py:432),
# so their CG runs to a tight residual
SLIM_CG_TOL = 1e-10


def require_gpu(dev=None):
    """The product path runs on a ROCm GPU only; fail loudly without one."""
    if not torch.cuda.is_available():
        raise RuntimeError("ultimate_uv_artist_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    d = torch.device(dev) if dev is not None else torch.device("cuda", torch.cuda.current_device())
    if d.type != "cuda":
        raise RuntimeError(f"ultimate_uv_artist_amd runs on a GPU device, got {d}")
    return d


_local = threading.local()
_team_solves = threading.Lock()


def own_stream(dev):
    """A stream of the calling thread's own on ``dev`` (created on first use).

    The per-chart calls (``param.lscm``, ``param.refine``) run on it: the
    reference calls them one chart per call from a pool of threads, and on one
    shared stream those calls would queue behind each other and every call's
    final synchronisation would wait for all threads' work.  Each call copies
    its result to the host before it returns, so the caller never sees the
    stream."""
    streams = getattr(_local, "streams", None)
    if streams is None:
        streams = _local.streams = {}
    s = streams.get(dev)
    if s is None:
        s = streams[dev] = torch.cuda.Stream(dev)
    return s


def solve_guard(plan):
    """Context for one call's solves on ``plan`` from a thread's own stream.

    A plan with teams of several workgroups needs each team's workgroups
    resident together; two such solves side by side could each hold CUs that
    the other's members wait for, so those take one process-wide lock until
    their results are on the host.  Plans with wavefront and one-workgroup
    tasks only never wait across workgroups and run side by side."""
    return _team_solves if plan.info()["n_teams"] > 0 else contextlib.nullcontext()


def _tuning_key(dev):
    """Cache key of a GPU model: its architecture name and compute units."""
    pr = torch.cuda.get_device_properties(dev)
    return f"{getattr(pr, 'gcnArchName', pr.name)}/{pr.multi_processor_count}cu"


def tuning_for(dev):
    """Schedule cost model to start a new plan with (microseconds per CG
    iteration: one-workgroup fixed part, team-member fixed part, per slice):
    UVART_TUNE_COST="a,b,c" when set, else the entry for this GPU model in the
    JSON file UVART_TUNE_CACHE names (written by IslandPlan.tune), else None
    (the library's defaults)."""
    import json
    import os

    env = os.environ.get("UVART_TUNE_COST")
    if env:
        vals = [float(x) for x in env.split(",")]
        if len(vals) != 3 or not all(v > 0 for v in vals):
            raise ValueError("UVART_TUNE_COST must be three positive numbers a,b,c")
        return vals
    path = os.environ.get("UVART_TUNE_CACHE")
    if not path:
        return None
    try:
        with open(path) as f:
            entry = json.load(f).get(_tuning_key(dev))
    except (OSError, ValueError):
        return None
    return list(entry["cost_us"]) if entry else None


def save_tuning(dev, cost, info=None):
    """Store a fitted cost model under this GPU model in UVART_TUNE_CACHE (no
    file named: nothing is stored).  Returns the path written or None."""
    import json
    import os
    import tempfile

    path = os.environ.get("UVART_TUNE_CACHE")
    if not path:
        return None
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[_tuning_key(dev)] = {"cost_us": [float(c) for c in cost], **(info or {})}
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d, suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
    return path


def stream_of(dev):
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def concat_charts(charts):
    """Concatenate [(P, T), ...] into offsets, P (n,3) f64 and island-local T (t,3) i64."""
    vo = np.zeros(len(charts) + 1, np.int64)
    to = np.zeros(len(charts) + 1, np.int64)
    for i, (P, T) in enumerate(charts):
        vo[i + 1] = vo[i] + len(P)
        to[i + 1] = to[i] + len(T)
    P = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p, _ in charts]))
    T = np.ascontiguousarray(np.concatenate([np.asarray(t, np.int64).reshape(-1, 3) for _, t in charts]))
    return vo, to, P, T


def half_cots(V, F):
    """Half cotangent weights per face (t,3), the weight of corner k on the
    opposite edge (k+1, k+2), in isometric face frames, clipped to
    [-5e3, 5e3]; zero on faces without area (gpu.py:88-115, hier.py:361-385)."""
    V = np.asarray(V, np.float64)
    F = np.asarray(F, np.int64)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    l1 = np.linalg.norm(e1, axis=1)
    nrm = np.cross(e1, e2)
    dbl = np.linalg.norm(nrm, axis=1)
    ok = (dbl > 1e-14 * np.maximum(l1 * l1, 1e-300)) & np.isfinite(dbl) & (l1 > 0)
    ex = e1 / np.maximum(l1, 1e-30)[:, None]
    ey = np.cross(nrm / np.maximum(dbl, 1e-30)[:, None], ex)
    X = np.zeros((len(F), 3, 2))
    X[:, 1, 0] = l1
    X[:, 2, 0] = (e2 * ex).sum(1)
    X[:, 2, 1] = (e2 * ey).sum(1)
    X[~ok] = 0.0
    w = np.empty((len(F), 3))
    for k in range(3):
        a, b = X[:, (k + 1) % 3] - X[:, k], X[:, (k + 2) % 3] - X[:, k]
        cr = np.maximum(np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]), 1e-30)
        w[:, k] = 0.5 * np.clip((a * b).sum(1) / cr, -1e4, 1e4)
    w[~ok] = 0.0
    return w


def batch_frames_dev(V, F):
    """The per-triangle quantities of gpu.Batch (gpu.py:88-121) on device
    tensors V (n,3) float64 and F (t,3) int64, in float64: frames X (t,3,2)
    (zero where the triangle is too thin for a frame), area (t,), the usable
    mask ok (t,) bool, half cotangent weights w (t,3) clipped to +-5e3 (zero
    where not usable), and the clamped frame entries x1, x2, y2 (t,) that
    Batch.jacobians divides by."""
    p0 = V[F[:, 0]]
    e1, e2 = V[F[:, 1]] - p0, V[F[:, 2]] - p0
    l1 = e1.norm(dim=1)
    nrm = torch.cross(e1, e2, dim=1)
    dbl = nrm.norm(dim=1)
    ok = (dbl > 1e-14 * (l1 * l1).clamp_min(1e-300)) & torch.isfinite(dbl) & (l1 > 0)
    ex = e1 / l1.clamp_min(1e-30)[:, None]
    ey = torch.cross(nrm / dbl.clamp_min(1e-30)[:, None], ex, dim=1)
    X = torch.zeros(len(F), 3, 2, dtype=torch.float64, device=V.device)
    X[:, 1, 0] = l1
    X[:, 2, 0] = (e2 * ex).sum(1)
    X[:, 2, 1] = (e2 * ey).sum(1)
    X[~ok] = 0.0
    w = torch.empty(len(F), 3, dtype=torch.float64, device=V.device)
    for k in range(3):
        a, b = X[:, (k + 1) % 3] - X[:, k], X[:, (k + 2) % 3] - X[:, k]
        cr = (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]).abs().clamp_min(1e-30)
        w[:, k] = 0.5 * ((a * b).sum(1) / cr).clamp(-1e4, 1e4)
    w[~ok] = 0.0
    y2 = X[:, 2, 1]
    return {"X": X, "area": 0.5 * dbl, "ok": ok, "w": w, "x1": X[:, 1, 0].clamp_min(1e-30),
            "x2": X[:, 2, 0].clone(), "y2": torch.where(y2.abs() > 1e-30, y2, torch.full_like(y2, 1e-30))}


def _half_cots_dev(V, F):
    """half_cots on device tensors V (n,3) float64, F (t,3) int64."""
    return batch_frames_dev(V, F)["w"]


def pinned_laplacian(V, F, pins, dev=None):
    """The cotangent Laplacian of a batch (F global vertex ids, rows of every
    chart together) plus a soft pin of weight max |W| (at least 1e-12) at
    each of ``pins``: the matrix of gpu.py:126-142 Batch._build_laplacian and
    hier.py:394-412.  Returns (scipy CSR, pin weight).  With ``dev`` the
    weights and the assembly run on that GPU: a stable sort of the entries by
    (row, column), then the repeated entries summed in their original order
    (deterministic)."""
    import scipy.sparse as sp

    if dev is not None:
        n = len(V)
        Vd = torch.as_tensor(np.asarray(V, np.float64), device=dev)
        Fd = torch.as_tensor(np.asarray(F, np.int64), device=dev)
        w = _half_cots_dev(Vd, Fd)
        I = torch.cat([Fd[:, 1], Fd[:, 2], Fd[:, 0]])
        J = torch.cat([Fd[:, 2], Fd[:, 0], Fd[:, 1]])
        W = torch.cat([w[:, 0], w[:, 1], w[:, 2]])
        wpin = max(float(W.abs().max()) if len(W) else 0.0, 1e-12)
        pd = torch.as_tensor(np.asarray(pins, np.int64).reshape(-1), device=dev)
        rows = torch.cat([I, J, I, J, pd])
        cols = torch.cat([J, I, I, J, pd])
        vals = torch.cat([-W, -W, W, W, torch.full((len(pd),), wpin, dtype=torch.float64, device=dev)])
        del I, J, W, w
        key, order = torch.sort(rows * n + cols, stable=True)
        vals = vals[order]
        del rows, cols, order
        uk, counts = torch.unique_consecutive(key, return_counts=True)
        data = torch.segment_reduce(vals, "sum", lengths=counts)
        r = uk // n
        indptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
        indptr[1:] = torch.cumsum(torch.bincount(r, minlength=n), 0)
        A = sp.csr_matrix((data.cpu().numpy(), (uk % n).cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
        return A, wpin
    F = np.asarray(F, np.int64)
    n = len(V)
    w = half_cots(V, F)
    I = np.concatenate([F[:, 1], F[:, 2], F[:, 0]])
    J = np.concatenate([F[:, 2], F[:, 0], F[:, 1]])
    W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
    wpin = max(float(np.abs(W).max()) if len(W) else 0.0, 1e-12)
    pins = np.asarray(pins, np.int64).reshape(-1)
    rows = np.concatenate([I, J, I, J, pins])
    cols = np.concatenate([J, I, I, J, pins])
    vals = np.concatenate([-W, -W, W, W, np.full(len(pins), wpin)])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    return A, wpin


class CsrSystem:
    """A symmetric matrix resident on the GPU (``uva_csr``) for the coupled
    Jacobi-PCG solves that span a whole batch (``uva_csr_pcg``)."""

    def __init__(self, A, dev=None):
        self.dev = require_gpu(dev)
        A = A.tocsr()
        self.n = A.shape[0]
        self.nnz = int(A.indptr[-1])
        ip = np.ascontiguousarray(A.indptr, np.int64)
        ix = np.ascontiguousarray(A.indices, np.int64)
        a = np.ascontiguousarray(A.data, np.float64)
        self._h = C.c_void_p()
        with torch.cuda.device(self.dev):
            check(_lib.lib().uva_csr_create(C.byref(self._h), self.n, i64p(ip), i64p(ix), f64p(a),
                                            stream_of(self.dev)), "uva_csr_create")

    def pcg(self, b, x, max_iters, check_every=1, tol2=0.0, strict=False):
        """x (n,) or (n,2) float64 device tensor: start in, result out, for
        rhs b of the same shape.  Stops after iteration k with k %
        check_every == 0 when every column has r.r below (``strict``) or at
        most tol2 max(b.b, 1e-30); returns the iterations run."""
        if b.shape != x.shape or b.shape[0] != self.n or b.dim() > 2 or (b.dim() == 2 and b.shape[1] > 2):
            raise ValueError("b and x must both be (n,) or (n, 1..2)")
        if x.dtype != torch.float64 or b.dtype != torch.float64 or not x.is_contiguous() or not b.is_contiguous():
            raise ValueError("b and x must be contiguous float64 tensors")
        nc = 1 if b.dim() == 1 else b.shape[1]
        its = C.c_int32(0)
        check(_lib.lib().uva_csr_pcg(self._h, ptr(b), ptr(x), int(nc), int(max_iters), int(check_every),
                                     float(tol2), int(bool(strict)), C.byref(its), stream_of(self.dev)),
              "uva_csr_pcg")
        return its.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().uva_csr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IslandPlan:
    """A batch of islands on one GPU.

    charts  list of (P (k,3), T (t,3)) with island-local vertex ids
    pins    soft-pin vertex (local id) per island, default 0
    frames  optional (X (t,3,2), area (t,), cots (t,3)) concatenated over the
            batch, used instead of recomputing them from P
    batch   gpu.Batch conventions instead of param.refine conventions
    """

    def __init__(self, charts, pins=None, frames=None, batch=False, dev=None, no_wave=False, no_teams=False,
                 keep_order=False, no_coarse=False, concat=None):
        # (concat: the charts already concatenated as concat_charts returns
        # them, in place of ``charts``)
        self.dev = require_gpu(dev)
        L = _lib.lib()
        if concat is not None:
            vo, to, P, T = concat
            self.vert_off = np.ascontiguousarray(vo, np.int64)
            self.tri_off = np.ascontiguousarray(to, np.int64)
            P = np.ascontiguousarray(P, np.float64)
            T = np.ascontiguousarray(T, np.int64)
        else:
            self.vert_off, self.tri_off, P, T = concat_charts(charts)
        self.n_islands = len(self.vert_off) - 1
        self.n = int(self.vert_off[-1])
        self.t = int(self.tri_off[-1])
        pins_a = None if pins is None else np.ascontiguousarray(np.asarray(pins, np.int64).reshape(-1))
        X = area = cots = None
        if frames is not None:
            X = np.ascontiguousarray(np.asarray(frames[0], np.float64).reshape(-1, 3, 2))
            area = np.ascontiguousarray(np.asarray(frames[1], np.float64).reshape(-1))
            cots = np.ascontiguousarray(np.asarray(frames[2], np.float64).reshape(-1, 3))
        self._h = C.c_void_p()
        with torch.cuda.device(self.dev):
            rc = L.uva_plan_create(C.byref(self._h), self.n_islands, i64p(self.vert_off), i64p(self.tri_off),
                                   f64p(P), i64p(T), i64p(pins_a), f64p(X), f64p(area), f64p(cots),
                                   (_lib.UVA_PLAN_BATCH if batch else 0) | (_lib.UVA_PLAN_NO_WAVE if no_wave else 0)
                                   | (_lib.UVA_PLAN_NO_TEAMS if no_teams else 0)
                                   | (_lib.UVA_PLAN_KEEP_ORDER if keep_order else 0)
                                   | (_lib.UVA_PLAN_NO_COARSE if no_coarse else 0),
                                   stream_of(self.dev))
        check(rc, "uva_plan_create")
        self.T_local = T
        cost = tuning_for(self.dev)
        if cost is not None and self.info()["n_group_tasks"] > 0:
            self.set_tuning(cost)

    # ------------------------------------------------------------- lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().uva_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = _lib.PlanInfo()
        check(_lib.lib().uva_plan_get_info(self._h, C.byref(out)), "uva_plan_get_info")
        return {k: getattr(out, k) for k, _ in _lib.PlanInfo._fields_}

    def stats(self):
        """Per-island statistics of the last call (synchronises the stream)."""
        arr = (_lib.IslandStats * self.n_islands)()
        check(_lib.lib().uva_plan_get_stats(self._h, arr, stream_of(self.dev)), "uva_plan_get_stats")
        names = [k for k, _ in _lib.IslandStats._fields_ if k != "pad"]
        return {k: np.array([getattr(s, k) for s in arr]) for k in names}

    def tune(self):
        """Fit the schedule's cost model to the island times of the last solve
        call and re-plan team sizes and the workgroup schedule from the CG
        iterations each island took in it (see uva_plan_tune); the fitted
        model is stored in UVART_TUNE_CACHE when that names a file."""
        check(_lib.lib().uva_plan_tune(self._h, stream_of(self.dev)), "uva_plan_tune")
        save_tuning(self.dev, self.tuning(), {"islands": self.n_islands, "triangles": self.t})

    def tuning(self):
        """The schedule's cost model (see uva_plan_get_tuning)."""
        out = np.zeros(3, np.float64)
        check(_lib.lib().uva_plan_get_tuning(self._h, f64p(out)), "uva_plan_get_tuning")
        return out

    def set_tuning(self, cost):
        """Set the schedule's cost model and re-plan (see uva_plan_set_tuning)."""
        c = np.ascontiguousarray(np.asarray(cost, np.float64).reshape(3))
        check(_lib.lib().uva_plan_set_tuning(self._h, f64p(c), stream_of(self.dev)), "uva_plan_set_tuning")

    # ------------------------------------------------------------- helpers
    def device_uv(self, U):
        """(n,2) float64 contiguous tensor on the plan's GPU (a copy of numpy input)."""
        if isinstance(U, torch.Tensor):
            U = U.to(self.dev, torch.float64)
        else:
            U = torch.as_tensor(np.asarray(U, np.float64), device=self.dev)
        U = U.reshape(self.n, 2).contiguous()
        return U

    def _check_uv(self, U):
        if not (isinstance(U, torch.Tensor) and U.dtype == torch.float64 and U.is_contiguous()
                and U.device == self.dev and tuple(U.shape) == (self.n, 2)):
            raise ValueError(f"expected a contiguous ({self.n}, 2) float64 tensor on {self.dev}")

    # ------------------------------------------------------------- solves
    def lscm(self, U, pins=None, tol=LSCM_TOL, max_iters=0):
        """LSCM of every island into U (device tensor, written in place)."""
        self._check_uv(U)
        p = None if pins is None else np.ascontiguousarray(np.asarray(pins, np.int64).reshape(-1))
        check(_lib.lib().uva_lscm(self._h, ptr(U), i64p(p), float(tol), int(max_iters), stream_of(self.dev)),
              "uva_lscm")
        return U

    def arap(self, U, iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=1e-5, cg_tol=CG_TOL,
             cg_max_iters=0, precision=_lib.UVA_PREC_MIXED, use_graph=False, iters_folded=0):
        """ARAP iterations on U (device tensor) in place; with ``iters_folded``
        > 0, islands whose start has a folded triangle run that many instead."""
        self._check_uv(U)
        o = _lib.ArapOpts(int(iters), int(bool(line_search)), int(mode), int(precision), float(tol), float(cg_tol),
                          int(cg_max_iters), int(bool(use_graph)), int(iters_folded))
        check(_lib.lib().uva_arap(self._h, ptr(U), C.byref(o), stream_of(self.dev)), "uva_arap")
        return U

    def arap_profile(self, U, iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=1e-5, cg_tol=CG_TOL,
                     cg_max_iters=0, precision=_lib.UVA_PREC_MIXED):
        """``arap`` run eagerly with HIP events around every launch: returns the
        launch durations in ms (index 0 the prologue, k the ARAP iteration k)."""
        self._check_uv(U)
        o = _lib.ArapOpts(int(iters), int(bool(line_search)), int(mode), int(precision), float(tol), float(cg_tol),
                          int(cg_max_iters), 0)
        ms = np.zeros(int(iters) + 1, np.float64)
        check(_lib.lib().uva_arap_profile(self._h, ptr(U), C.byref(o), f64p(ms), int(iters) + 1,
                                          stream_of(self.dev)), "uva_arap_profile")
        return ms

    def slim(self, U, iters, tol=1e-5, cg_tol=SLIM_CG_TOL, cg_max_iters=0):
        """SLIM (symmetric Dirichlet) iterations on U (device tensor) in place."""
        self._check_uv(U)
        o = _lib.SlimOpts(int(iters), int(cg_max_iters), float(tol), float(cg_tol))
        check(_lib.lib().uva_slim(self._h, ptr(U), C.byref(o), stream_of(self.dev)), "uva_slim")
        return U

    def local_rhs(self, U, mode=_lib.UVA_MODE_REFINE):
        self._check_uv(U)
        b = torch.empty_like(U)
        check(_lib.lib().uva_local_rhs(self._h, ptr(U), ptr(b), int(mode), stream_of(self.dev)), "uva_local_rhs")
        return b

    def cg_solve(self, b, x, tol=CG_TOL, max_iters=0, precision=_lib.UVA_PREC_MIXED):
        self._check_uv(b)
        self._check_uv(x)
        check(_lib.lib().uva_cg_solve(self._h, ptr(b), ptr(x), float(tol), int(max_iters), int(precision),
                                      stream_of(self.dev)), "uva_cg_solve")
        return x

    def max_step(self, U, D, mode=_lib.UVA_MODE_REFINE):
        self._check_uv(U)
        self._check_uv(D)
        out = torch.empty(self.n_islands, dtype=torch.float64, device=self.dev)
        check(_lib.lib().uva_max_step(self._h, ptr(U), ptr(D), ptr(out), int(mode), stream_of(self.dev)),
              "uva_max_step")
        return out

    def energy(self, U, kind=0):
        self._check_uv(U)
        out = torch.empty(self.n_islands, dtype=torch.float64, device=self.dev)
        check(_lib.lib().uva_energy(self._h, ptr(U), int(kind), ptr(out), stream_of(self.dev)), "uva_energy")
        return out

    def _check_tri(self, w):
        if not (isinstance(w, torch.Tensor) and w.dtype == torch.float64 and w.is_contiguous()
                and w.device == self.dev and tuple(w.shape) == (self.t,)):
            raise ValueError(f"expected a contiguous ({self.t},) float64 tensor on {self.dev}")

    def energy_fwd(self, U, w=None, kind=0):
        """Per-island weighted energy sum_f w_f e_f(U) (w None = 1)."""
        self._check_uv(U)
        if w is not None:
            self._check_tri(w)
        out = torch.empty(self.n_islands, dtype=torch.float64, device=self.dev)
        check(_lib.lib().uva_energy_fwd(self._h, ptr(U), None if w is None else ptr(w), int(kind), ptr(out),
                                        stream_of(self.dev)), "uva_energy_fwd")
        return out

    def energy_bwd(self, U, ge, w=None, kind=0, want_w=True):
        """(ge_i dE_i/dU, ge_i e_f or None) for the per-island upstream gradient ge."""
        self._check_uv(U)
        if w is not None:
            self._check_tri(w)
        ge = ge.to(device=self.dev, dtype=torch.float64).contiguous().reshape(self.n_islands)
        gU = torch.empty_like(U)
        gw = torch.empty(self.t, dtype=torch.float64, device=self.dev) if want_w else None
        check(_lib.lib().uva_energy_bwd(self._h, ptr(U), None if w is None else ptr(w), int(kind), ptr(ge), ptr(gU),
                                        None if gw is None else ptr(gw), stream_of(self.dev)), "uva_energy_bwd")
        return gU, gw

    def adjoint_cg(self, g, tol=1e-10, max_iters=0, precision=_lib.UVA_PREC_F64):
        """lam = L^-1 g per island (the global step's operator, zero start)."""
        self._check_uv(g)
        lam = torch.empty_like(g)
        check(_lib.lib().uva_adjoint_cg(self._h, ptr(g), ptr(lam), float(tol), int(max_iters), int(precision),
                                        stream_of(self.dev)), "uva_adjoint_cg")
        return lam

    def step_bwd(self, U, lam):
        """(db/dU)^T lam for the refine global step's right-hand side b(U)."""
        self._check_uv(U)
        self._check_uv(lam)
        gU = torch.empty_like(U)
        check(_lib.lib().uva_arap_step_bwd(self._h, ptr(U), ptr(lam), ptr(gU), stream_of(self.dev)),
              "uva_arap_step_bwd")
        return gU

    def signed_areas(self, U):
        self._check_uv(U)
        out = torch.empty(self.t, dtype=torch.float64, device=self.dev)
        check(_lib.lib().uva_signed_areas(self._h, ptr(U), ptr(out), stream_of(self.dev)), "uva_signed_areas")
        return out

    def flips_per_chart(self, U):
        self._check_uv(U)
        out = torch.empty(self.n_islands, dtype=torch.float64, device=self.dev)
        check(_lib.lib().uva_flips_per_chart(self._h, ptr(U), ptr(out), stream_of(self.dev)), "uva_flips_per_chart")
        return out

    def repair_jacobi(self, U, rounds):
        self._check_uv(U)
        check(_lib.lib().uva_repair_jacobi(self._h, ptr(U), int(rounds), stream_of(self.dev)), "uva_repair_jacobi")
        return U

    def orient_positive(self, U):
        self._check_uv(U)
        check(_lib.lib().uva_orient_positive(self._h, ptr(U), stream_of(self.dev)), "uva_orient_positive")
        return U

    def set_pins(self, pins, w=0.0):
        """Soft pins at the given vertices in place of each island's own pin
        (w = 0: the island's pin weight); see uva_plan_set_pins."""
        p = np.ascontiguousarray(np.asarray(pins, np.int64).reshape(-1))
        check(_lib.lib().uva_plan_set_pins(self._h, i64p(p), len(p), float(w), stream_of(self.dev)),
              "uva_plan_set_pins")

    def step_apply(self, U, Un, mode=_lib.UVA_MODE_BATCH, line_search=True):
        """The rest of a Batch / hier ARAP iteration after its global step
        (non-finite revert, step rule, update of U in place; uva_step_apply)."""
        self._check_uv(U)
        self._check_uv(Un)
        check(_lib.lib().uva_step_apply(self._h, ptr(U), ptr(Un), int(mode), int(bool(line_search)),
                                        stream_of(self.dev)), "uva_step_apply")
        return U

    def pcg(self, b, x, max_iters, check_every=1, tol2=0.0, strict=False, want_iters=True):
        """The reference's Jacobi-PCG over every row of the plan at once (its
        islands' Laplacians with the pins of ``set_pins``, two columns sharing
        nothing but the launches; see uva_plan_pcg).  x (n,2) float64 device
        tensor: start in, result out, for rhs b of the same shape.  Returns the
        iterations run (None with ``want_iters`` False: no wait for the stream)."""
        self._check_uv(x)
        self._check_uv(b)
        its = C.c_int32(0)
        check(_lib.lib().uva_plan_pcg(self._h, ptr(b), ptr(x), int(max_iters), int(check_every), float(tol2),
                                      int(bool(strict)), C.byref(its) if want_iters else None,
                                      stream_of(self.dev)), "uva_plan_pcg")
        return its.value if want_iters else None

    def set_scale(self, scale):
        s = scale.to(self.dev, torch.float64).contiguous()
        if s.numel() != self.t:
            raise ValueError("scale needs one value per triangle")
        check(_lib.lib().uva_plan_set_scale(self._h, ptr(s), stream_of(self.dev)), "uva_plan_set_scale")
        torch.cuda.current_stream(self.dev).synchronize()
