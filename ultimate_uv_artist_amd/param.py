"""Drop-in replacement of ``engine/param.py`` (chart parameterization).

Same names, arguments, return values and exceptions as the reference, so
``charts.flatten``, ``grow.validate``, ``pipeline._final_flatten``,
``crumbs.arap_fill`` and ``overlapfix.fix_overlaps`` call it unchanged.

Where the work goes:
  * ``lscm``, ``refine(method="arap")`` and ``SPDSolver`` run on the GPU
    through libuvart (FP64 CG for LSCM and for general SPD solves; the ARAP
    loop runs entirely on the device with FP32 CG inner iterations, FP64 dots
    and an FP64 true-residual check per global step).
  * ``repair_flips`` runs the reference's sequential update order in native
    host code inside libuvart (the order is part of its result).
  * The small geometry helpers that hand host arrays back to host callers
    (frames, cotangents, boundary walks, signed areas, chart quality and the
    private per-triangle kernels used by ``crumbs``) are host numpy.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from .plan import CG_TOL, LSCM_TOL, IslandPlan, own_stream, require_gpu, solve_guard, stream_of

DIRECT_LIMIT = 220_000        # vertices: above this the reference uses warm-started CG (param.py:15)
PARALLEL_MIN_TRIS = 20000     # param.py:328
SPD_TOL_DIRECT = 1e-12        # FP64 CG target that stands in for the reference's direct solve


# ------------------------------------------------------------- basic geometry
def local_frames(P, T):
    """Isometric 2-D coordinates of each triangle, corner 0 at the origin and
    corner 1 on +x (param.py:19-32).  Returns X (t,3,2) and area (t,)."""
    P = np.asarray(P, np.float64)
    T = np.asarray(T, np.int64)
    p0 = P[T[:, 0]]
    e1 = P[T[:, 1]] - p0
    e2 = P[T[:, 2]] - p0
    l1 = np.linalg.norm(e1, axis=1)
    nrm = np.cross(e1, e2)
    nn = np.linalg.norm(nrm, axis=1)
    ex = e1 / np.maximum(l1, 1e-300)[:, None]
    ey = np.cross(nrm / np.maximum(nn, 1e-300)[:, None], ex)
    X = np.zeros((len(T), 3, 2))
    X[:, 1, 0] = l1
    X[:, 2, 0] = np.einsum("ij,ij->i", e2, ex)
    X[:, 2, 1] = np.einsum("ij,ij->i", e2, ey)
    return X, 0.5 * nn


def cot_weights(X):
    """Cotangent of the angle at corner k, which faces edge k+1 -> k+2
    (param.py:35-44), clipped to +-1e4."""
    X = np.asarray(X, np.float64)
    c = np.empty((len(X), 3))
    for k in range(3):
        u = X[:, (k + 1) % 3] - X[:, k]
        v = X[:, (k + 2) % 3] - X[:, k]
        c[:, k] = np.einsum("ij,ij->i", u, v) / np.maximum(np.abs(u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]), 1e-300)
    return np.clip(c, -1e4, 1e4)


def _lap_triplets(T, cots):
    """COO triplets of the cotangent Laplacian (param.py:47-51)."""
    T = np.asarray(T, np.int64)
    a = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    b = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    w = 0.5 * np.concatenate([cots[:, 0], cots[:, 1], cots[:, 2]])
    return np.concatenate([a, b, a, b]), np.concatenate([b, a, a, b]), np.concatenate([-w, -w, w, w])


def cot_laplacian(n, T, cots):
    """Cotangent Laplacian as a scipy CSC matrix (param.py:54-56)."""
    I, J, W = _lap_triplets(T, cots)
    return sp.csc_matrix((W, (I, J)), shape=(n, n))


def boundary_edges_oriented(T):
    """Directed boundary half-edges i -> j: those whose reverse is absent (param.py:59-68)."""
    T = np.asarray(T, np.int64)
    src = np.concatenate([T[:, 0], T[:, 1], T[:, 2]])
    dst = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    n = int(max(src.max(), dst.max())) + 1
    has_twin = np.isin(dst * n + src, src * n + dst)
    return src[~has_twin], dst[~has_twin]


def boundary_loops(T):
    """Boundary loops in walk order (param.py:71-94)."""
    bi, bj = boundary_edges_oriented(T)
    if len(bi) == 0:
        return []
    nxt = {}
    for a, b in zip(bi.tolist(), bj.tolist()):
        nxt.setdefault(a, []).append(b)
    seen, loops = set(), []
    for s in bi.tolist():
        if s in seen:
            continue
        seen.add(s)
        loop, cur = [s], s
        while True:
            step = next((x for x in nxt.get(cur, ()) if x not in seen or x == s), None)
            if step is None or step == s:
                break
            cur = step
            loop.append(cur)
            seen.add(cur)
        loops.append(np.array(loop, np.int64))
    return loops


def signed_areas(U, T):
    """Signed UV area of every triangle (param.py:97-100)."""
    U = np.asarray(U)
    T = np.asarray(T, np.int64)
    a, b, c = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
    return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))


# ------------------------------------------------------------- solvers
class SPDSolver:
    """Reusable solver of a fixed SPD matrix (param.py:103-132), on the GPU.

    The matrix is uploaded once; ``solve`` runs the batched FP64 Jacobi-PCG
    kernel.  Where the reference factorises (n <= DIRECT_LIMIT) the solve is
    taken to relative residual 1e-12; above it the reference's CG tolerance
    and iteration cap apply."""

    def __init__(self, A):
        import ctypes as C
        import torch

        self.n = A.shape[0]
        self.A = A.tocsc()
        self.dev = require_gpu()
        csr = A.tocsr()
        indptr = np.ascontiguousarray(csr.indptr, np.int64)
        indices = np.ascontiguousarray(csr.indices, np.int64)
        data = np.ascontiguousarray(csr.data, np.float64)
        h = C.c_void_p()
        with torch.cuda.device(self.dev):
            rc = _lib.lib().uva_plan_create_matrix(C.byref(h), self.n, _lib.i64p(indptr), _lib.i64p(indices),
                                                   _lib.f64p(data), stream_of(self.dev))
        _lib.check(rc, "uva_plan_create_matrix")
        self._plan = IslandPlan.__new__(IslandPlan)
        self._plan._h = h
        self._plan.dev = self.dev
        self._plan.n = self.n
        self._plan.t = 0
        self._plan.n_islands = 1
        self.lu = None

    def solve(self, b, x0=None, tol=1e-7, maxiter=None):
        import torch

        b = np.asarray(b, np.float64)
        cols = b.reshape(self.n, -1)
        direct = self.n <= DIRECT_LIMIT
        out = np.empty_like(cols)
        for c0 in range(0, cols.shape[1], 2):
            bb = np.zeros((self.n, 2))
            k = min(2, cols.shape[1] - c0)
            bb[:, :k] = cols[:, c0:c0 + k]
            xx = np.zeros((self.n, 2))
            if x0 is not None:
                xs = np.asarray(x0, np.float64).reshape(self.n, -1)
                xx[:, :k] = xs[:, c0:c0 + k]
            bd = torch.as_tensor(bb, device=self.dev)
            xd = torch.as_tensor(xx, device=self.dev)
            self._plan.cg_solve(bd, xd, tol=SPD_TOL_DIRECT if direct else tol,
                                max_iters=0 if direct else (maxiter or 2000), precision=_lib.UVA_PREC_F64)
            out[:, c0:c0 + k] = xd.cpu().numpy()[:, :k]
        return out.reshape(b.shape)


# ------------------------------------------------------------- LSCM
def lscm(P, T, X=None, cots=None, pins=None):
    """Least-squares conformal map with two hard pins (param.py:136-169), on the GPU.

    Raises ValueError for a chart without boundary, like the reference."""
    import torch

    P = np.asarray(P, np.float64)
    T = np.asarray(T, np.int64)
    frames = None
    if X is not None and cots is not None:
        frames = (X, np.zeros(len(T)), cots)
    bi, _ = boundary_edges_oriented(T)
    if len(bi) == 0:
        raise ValueError("chart has no boundary (closed surface): needs a cut")
    dev = require_gpu()
    with torch.cuda.stream(own_stream(dev)):
        plan = IslandPlan([(P, T)], frames=frames, dev=dev)
        with solve_guard(plan):
            U = torch.zeros(len(P), 2, dtype=torch.float64, device=dev)
            plan.lscm(U, pins=None if pins is None else [int(pins[0]), int(pins[1])], tol=LSCM_TOL)
            return U.cpu().numpy()


def _far_pair(P, cand):
    """Boundary vertex farthest from the boundary centroid, then the one
    farthest from it (param.py:172-179)."""
    Q = P[cand]
    i = cand[np.argmax(((Q - Q.mean(0)) ** 2).sum(1))]
    j = cand[np.argmax(((Q - P[i]) ** 2).sum(1))]
    if i == j:
        j = cand[0] if cand[0] != i else cand[-1]
    return int(i), int(j)


def tutte(P, T):
    """Uniform Tutte embedding onto a circle (param.py:182-202); the interior
    solve runs on the GPU through ``SPDSolver``.

    The longest boundary loop goes onto the circle whose perimeter is the
    loop's length, each vertex at its arc-length angle from the loop's first
    vertex.  Every interior vertex is then the mean of its distinct
    neighbours, with one property of the reference kept because it is part
    of its result: its right-hand side takes the circle positions in loop
    order against the boundary vertices in increasing id order (a boolean
    column mask), so a boundary vertex enters its neighbours' equations with
    the position of the loop vertex whose rank in the loop equals its rank
    among the boundary ids.  The two orders agree when the loop runs in id
    order from its smallest vertex; otherwise the interior is not harmonic
    and may fold, as in the reference."""
    P = np.asarray(P, np.float64)
    T = np.asarray(T, np.int64)
    n = len(P)
    loop = max(boundary_loops(T), key=len)
    seg = np.linalg.norm(P[np.roll(loop, -1)] - P[loop], axis=1)
    length = seg.sum()
    theta = np.concatenate(([0.0], np.cumsum(seg)[:-1])) / max(length, 1e-300) * (2 * np.pi)
    circle = (length / (2 * np.pi)) * np.stack([np.cos(theta), np.sin(theta)], 1)
    on_loop = np.zeros(n, bool)
    on_loop[loop] = True
    # boundary value each boundary vertex contributes to its neighbours' rows
    rhs_value = np.zeros((n, 2))
    rhs_value[np.sort(loop)] = circle
    # distinct undirected edges; degree = number of distinct neighbours
    ends = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
    ends = np.unique(ends, axis=0)
    deg = np.bincount(ends.ravel(), minlength=n).astype(np.float64)
    free = np.flatnonzero(~on_loop)
    slot = np.full(n, -1, np.int64)
    slot[free] = np.arange(len(free))
    a, b = ends[:, 0], ends[:, 1]
    both = ~on_loop[a] & ~on_loop[b]
    rows = np.concatenate([slot[a[both]], slot[b[both]], np.arange(len(free))])
    cols = np.concatenate([slot[b[both]], slot[a[both]], np.arange(len(free))])
    vals = np.concatenate([-np.ones(2 * int(both.sum())), deg[free]])
    M = sp.csr_matrix((vals, (rows, cols)), shape=(len(free), len(free)))
    rhs = np.zeros((len(free), 2))
    for u, v in ((a, b), (b, a)):                  # u free, v on the loop
        m = ~on_loop[u] & on_loop[v]
        np.add.at(rhs, slot[u[m]], rhs_value[v[m]])
    U = np.zeros((n, 2))
    U[loop] = circle
    if len(free):
        U[free] = SPDSolver(M).solve(rhs)
    return U


# ------------------------------------------------------------- per-triangle helpers (host)
def _jac(U, T, X):
    T = np.asarray(T, np.int64)
    x1, x2, y2 = X[:, 1, 0], X[:, 2, 0], X[:, 2, 1]
    good = np.abs(x1 * y2) >= 1e-300
    x1 = np.where(good, x1, 1.0)
    y2 = np.where(good, y2, 1.0)
    d1 = U[T[:, 1]] - U[T[:, 0]]
    d2 = U[T[:, 2]] - U[T[:, 0]]
    j00, j10 = d1[:, 0] / x1, d1[:, 1] / x1
    j01 = (d2[:, 0] - j00 * x2) / y2
    j11 = (d2[:, 1] - j10 * x2) / y2
    return j00, j01, j10, j11, good


def _arap_rotations(U, T, X, cots, mode, sd_weights):
    """Per-triangle target rotation R (t,2,2) and SLIM weights Wm (t,2,2)
    (param.py:206-256): mode 0 closed-form rotation, mode 1 SVD + SLIM weights."""
    j00, j01, j10, j11, good = _jac(np.asarray(U, np.float64), T, X)
    t = len(j00)
    R = np.tile(np.eye(2), (t, 1, 1))
    Wm = np.tile(np.eye(2), (t, 1, 1))
    if mode == 0:
        a, b = j00 + j11, j10 - j01
        h = np.sqrt(a * a + b * b)
        use = good & ~(h < 1e-300)
        hs = np.where(use, h, 1.0)
        ca, cb = np.where(use, a / hs, 1.0), np.where(use, b / hs, 0.0)
        R[:, 0, 0], R[:, 0, 1], R[:, 1, 0], R[:, 1, 1] = ca, -cb, cb, ca
        return R, Wm
    idx = np.nonzero(good)[0]
    if len(idx):
        J = np.empty((len(idx), 2, 2))
        J[:, 0, 0], J[:, 0, 1], J[:, 1, 0], J[:, 1, 1] = j00[idx], j01[idx], j10[idx], j11[idx]
        Uu, S, Vt = np.linalg.svd(J)
        neg = np.linalg.det(Uu @ Vt) < 0
        Uu[neg, :, 1] = -Uu[neg, :, 1]
        S[neg, 1] = -S[neg, 1]
        R[idx] = Uu @ Vt
        s = np.maximum(np.abs(S), 1e-8)
        one = np.abs(s - 1.0) < 1e-6
        w = np.where(one, 2.0, np.sqrt(np.maximum((s - s ** -3) / np.where(one, 1.0, s - 1.0), 1e-12)))
        Wm[idx] = (Uu * w[:, None, :]) @ np.swapaxes(Uu, 1, 2)
    return R, Wm


def _max_step(U, D, T):
    """Largest t before a positively oriented triangle degenerates along
    U + t D (param.py:259-291)."""
    T = np.asarray(T, np.int64)
    ua, ub, uc = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
    da, db, dc = D[T[:, 0]], D[T[:, 1]], D[T[:, 2]]
    u1, u2, d1, d2 = ub - ua, uc - ua, db - da, dc - da
    C0 = u1[:, 0] * u2[:, 1] - u2[:, 0] * u1[:, 1]
    B = u1[:, 0] * d2[:, 1] + d1[:, 0] * u2[:, 1] - u2[:, 0] * d1[:, 1] - d2[:, 0] * u1[:, 1]
    A = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    pos = C0 > 0
    linear = pos & ((A == 0.0) | (np.abs(A) < 1e-30 * np.maximum(np.abs(B), 1e-300)))
    roots = [np.full(len(T), 1e30)]
    with np.errstate(divide="ignore", invalid="ignore"):
        roots.append(np.where(linear & (B < 0), -C0 / B, 1e30))
        disc = B * B - 4.0 * A * C0
        ok = pos & ~linear & (disc >= 0)
        sq = np.sqrt(np.where(ok, disc, 0.0))
        for r in ((-B - sq) / (2.0 * A), (-B + sq) / (2.0 * A)):
            roots.append(np.where(ok & (r > 0), r, 1e30))
    allr = np.concatenate(roots)
    allr = allr[~np.isnan(allr)]
    return float(min(allr.min(), 1e30)) if len(allr) else 1e30


def _energy(U, T, X, area, kind):
    """Summed ARAP (kind 0) or symmetric Dirichlet (kind 1) energy (param.py:294-338)."""
    j00, j01, j10, j11, good = _jac(np.asarray(U, np.float64), T, X)
    if kind == 0:
        q = np.sqrt((0.5 * (j00 + j11)) ** 2 + (0.5 * (j10 - j01)) ** 2)
        r = np.sqrt((0.5 * (j00 - j11)) ** 2 + (0.5 * (j10 + j01)) ** 2)
        e = area * ((q + r - 1) ** 2 + (q - r - 1) ** 2)
    else:
        det = j00 * j11 - j01 * j10
        fro = j00 ** 2 + j01 ** 2 + j10 ** 2 + j11 ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            e = np.where(det > 0, area * (fro + fro / (det * det) - 4.0), np.inf)
    return float(np.where(good, e, 0.0).sum())


def _arap_rhs(n, T, X, cots, R):
    """Right-hand side of the ARAP global step (param.py:457-474)."""
    T = np.asarray(T, np.int64)
    b = np.zeros((n, 2))
    for k in range(3):
        i, j = (k + 1) % 3, (k + 2) % 3
        r = np.einsum("tab,tb->ta", R, X[:, i] - X[:, j]) * (0.5 * cots[:, k])[:, None]
        np.add.at(b, T[:, i], r)
        np.add.at(b, T[:, j], -r)
    return b


# ------------------------------------------------------------- refinement
class ArapSolver:
    """What ``refine`` returns as its solver: the chart's device plan (the
    reference returns the factorised SPDSolver of the pinned Laplacian, and
    callers may pass it back in to skip the set-up)."""

    def __init__(self, plan, n, t, pin):
        self.plan, self.n, self.t, self.pin = plan, n, t, pin

    def solve(self, b, x0=None, tol=CG_TOL, maxiter=None):
        """Solve (L + wpin e_pin e_pin^T) x = b on the GPU (mixed precision)."""
        import torch

        b = np.asarray(b, np.float64).reshape(self.n, 2)
        x = torch.as_tensor(np.zeros((self.n, 2)) if x0 is None else np.asarray(x0, np.float64).reshape(self.n, 2),
                            device=self.plan.dev).contiguous()
        self.plan.cg_solve(torch.as_tensor(b, device=self.plan.dev).contiguous(), x, tol=tol,
                           max_iters=maxiter or 0)
        return x.cpu().numpy()


def refine(P, T, U, iters=10, method="arap", X=None, area=None, cots=None, solver=None, line_search=True,
           tol=1e-5, pin=None):
    """Local/global refinement (param.py:379-454).

    ``method="arap"`` runs every iteration on the GPU: rotations, rhs, CG to
    relative residual 1e-6 warm-started at U, flip-safe step (only when the
    start has no flips), energy early exit at ``tol``.  Returns (U, solver).

    ``method="slim"`` (symmetric Dirichlet, param.py:415-454) runs every
    iteration on the GPU too: SLIM weights from the closed-form 2x2 SVD, the
    normal equations by block-Jacobi PCG to relative residual 1e-10 (the
    reference solves them directly), flip-safe step halved up to 12 times
    until the energy does not increase.  Returns (U, None)."""
    P = np.asarray(P, np.float64)
    T = np.asarray(T, np.int64)
    frames = None
    if X is not None:   # caller's frames (the reference uses them instead of P's)
        frames = (X, area if area is not None else local_frames(P, T)[1], cots if cots is not None else cot_weights(X))
    if method not in ("arap", "slim"):
        raise ValueError(f"unknown refine method {method!r}")
    import torch

    dev = require_gpu()
    with torch.cuda.stream(own_stream(dev)):
        if method == "slim":
            plan = IslandPlan([(P, T)], frames=frames, dev=dev)
            with solve_guard(plan):
                Ud = plan.device_uv(U)
                plan.slim(Ud, iters, tol=tol)
                return Ud.cpu().numpy(), None
        pin = 0 if pin is None else int(pin)
        plan = None
        if (isinstance(solver, ArapSolver) and solver.n == len(P) and solver.t == len(T) and solver.pin == pin
                and solver.plan.dev == dev):
            plan = solver.plan
        if plan is None:
            plan = IslandPlan([(P, T)], pins=[pin], frames=frames, dev=dev)
            solver = ArapSolver(plan, len(P), len(T), pin)
        with solve_guard(plan):
            Ud = plan.device_uv(U)
            plan.arap(Ud, iters, line_search=line_search, mode=_lib.UVA_MODE_REFINE, tol=tol, cg_tol=CG_TOL)
            return Ud.cpu().numpy(), solver


# ------------------------------------------------------------- flip repair
def repair_flips(U, T, rounds=60):
    """Move the interior vertices of folded triangles to their 1-ring centroid
    in the reference's sequential order (param.py:478-516).  Returns a copy."""
    out = np.array(U, np.float64, order="C", copy=True)
    T = np.ascontiguousarray(T, np.int64)
    _lib.check(_lib.lib().uva_repair_seq(_lib.f64p(out), len(out), _lib.i64p(T), len(T), int(rounds)),
               "uva_repair_seq")
    return out


# ------------------------------------------------------------- chart quality
def chart_quality(P, T, U, X=None, area=None):
    """Area-weighted distortion of one chart (param.py:520-551)."""
    T = np.asarray(T, np.int64)
    if X is None:
        X, area = local_frames(P, T)
    sa = signed_areas(U, T)
    tot = area.sum()
    flips = int((sa <= 0).sum()) if (sa > 0).sum() >= (sa < 0).sum() else int((sa >= 0).sum())
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
    sc = np.abs(sa).sum() / max(tot, 1e-300)
    w = area / max(tot, 1e-300)
    an = s1 / np.maximum(s2, 1e-12)
    ar = (s1 * s2) / max(sc, 1e-300)
    lar = np.abs(np.log2(np.maximum(ar, 1e-12)))

    def wq(x, q):
        o = np.argsort(x)
        cw = np.cumsum(w[o])
        return float(x[o][min(np.searchsorted(cw, q * cw[-1]), len(o) - 1)])

    st = np.sqrt((s1 ** 2 + s2 ** 2) / 2 / max(sc, 1e-300))
    return {"flips": flips, "aniso_p95": wq(an, 0.95), "aniso_mean": float((an * w).sum()),
            "area_p95": float(2 ** wq(lar, 0.95)), "area_logdev": float((lar * w).sum()),
            "stretch": float(np.sqrt((st ** 2 * w).sum())), "scale": float(np.sqrt(sc)),
            "per_tri_aniso": an, "per_tri_area": ar}
