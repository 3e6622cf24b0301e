"""Float64 CPU restatement of ``engine/param.py`` (the reference CPU path).

TEST INFRASTRUCTURE ONLY (the oracle): used by tests, smoke() and the bench's
cpu_baseline leg as the checker, never by the product package.

The reference JIT-compiles its per-triangle loops with numba; here the same
arithmetic is written as vectorised numpy over the triangle axis.  Sparse
solves use scipy's SuperLU with the reference's options, so a global step is
solved exactly (to round-off) like the reference does for n <= DIRECT_LIMIT.

Every function cites the reference lines it follows.  Pinning of each function
(golden vectors, closed forms, library routines, brute force) is in
tests/test_oracle.py; functions only pinned by the reference's own output are
pinned by the fixtures under tests/golden/.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

DIRECT_LIMIT = 220_000        # param.py:15
PARALLEL_MIN_TRIS = 20000     # param.py:328


# ----------------------------------------------------------------- geometry
def local_frames(P, T):
    """param.py:19-32 -- isometric 2-D corner coordinates per triangle.

    Corner 0 at the origin, corner 1 on +x at distance |e1|, corner 2 at
    (e2.ex, e2.ey) with ey = n_hat x ex.  Returns X (t,3,2) and area (t,)."""
    P = np.asarray(P, np.float64)
    T = np.asarray(T, np.int64)
    e1 = P[T[:, 1]] - P[T[:, 0]]
    e2 = P[T[:, 2]] - P[T[:, 0]]
    l1 = np.sqrt((e1 * e1).sum(1))
    nrm = np.cross(e1, e2)
    nn = np.sqrt((nrm * nrm).sum(1))
    ex = e1 / np.maximum(l1, 1e-300)[:, None]
    ey = np.cross(nrm / np.maximum(nn, 1e-300)[:, None], ex)
    X = np.zeros((len(T), 3, 2))
    X[:, 1, 0] = l1
    X[:, 2, 0] = (e2 * ex).sum(1)
    X[:, 2, 1] = (e2 * ey).sum(1)
    return X, 0.5 * nn


def cot_weights(X):
    """param.py:35-44 -- cot of the angle at corner k (opposite edge k+1 -> k+2),
    dot / max(|cross|, 1e-300), clipped to [-1e4, 1e4]."""
    out = np.empty((len(X), 3))
    for k in range(3):
        a = X[:, (k + 1) % 3] - X[:, k]
        b = X[:, (k + 2) % 3] - X[:, k]
        cr = np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
        out[:, k] = (a * b).sum(1) / np.maximum(cr, 1e-300)
    return np.clip(out, -1e4, 1e4)


def lap_triplets(T, cots):
    """param.py:47-51 -- COO triplets of the cotangent Laplacian: edge k of a
    triangle joins corners k+1, k+2 with weight cots[:,k]/2."""
    T = np.asarray(T, np.int64)
    i = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    j = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    w = 0.5 * np.concatenate([cots[:, 0], cots[:, 1], cots[:, 2]])
    return (np.concatenate([i, j, i, j]), np.concatenate([j, i, i, j]),
            np.concatenate([-w, -w, w, w]))


def cot_laplacian(n, T, cots):
    """param.py:54-56 (scipy CSC, duplicates summed)."""
    I, J, W = lap_triplets(T, cots)
    return sp.csc_matrix((W, (I, J)), shape=(n, n))


def boundary_edges_oriented(T):
    """param.py:59-68 -- directed half-edges a->b whose reverse b->a is absent."""
    T = np.asarray(T, np.int64)
    a = np.concatenate([T[:, 0], T[:, 1], T[:, 2]])
    b = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    n = int(max(a.max(), b.max())) + 1
    twin = np.isin(b * n + a, np.sort(a * n + b))
    return a[~twin], b[~twin]


def boundary_loops(T):
    """param.py:71-94 -- walk the boundary half-edges into loops (first
    unused successor, starting from each unused source in edge order)."""
    bi, bj = boundary_edges_oriented(T)
    if len(bi) == 0:
        return []
    succ = {}
    for s, d in zip(bi.tolist(), bj.tolist()):
        succ.setdefault(s, []).append(d)
    loops, used = [], set()
    for start in bi.tolist():
        if start in used:
            continue
        loop = [start]
        used.add(start)
        cur = start
        while True:
            nxt = None
            for c in succ.get(cur, []):
                if c not in used or c == start:
                    nxt = c
                    break
            if nxt is None or nxt == start:
                break
            cur = nxt
            loop.append(cur)
            used.add(cur)
        loops.append(np.array(loop, np.int64))
    return loops


def signed_areas(U, T):
    """param.py:97-100."""
    T = np.asarray(T, np.int64)
    a, b, c = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
    return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))


# ----------------------------------------------------------------- solver
class SPDSolver:
    """param.py:103-132 -- SuperLU (MMD_AT_PLUS_A, symmetric mode) for
    n <= DIRECT_LIMIT, else per-column Jacobi-preconditioned CG."""

    def __init__(self, A):
        self.n = A.shape[0]
        self.A = A.tocsc()
        self.lu = None
        if self.n <= DIRECT_LIMIT:
            try:
                self.lu = spla.splu(self.A, permc_spec="MMD_AT_PLUS_A", options={"SymmetricMode": True})
            except Exception:
                self.lu = None
        if self.lu is None:
            d = self.A.diagonal()
            self.M = sp.diags(1.0 / np.where(np.abs(d) > 1e-300, d, 1.0))

    def solve(self, b, x0=None, tol=1e-7, maxiter=None):
        if self.lu is not None:
            return self.lu.solve(b)
        cols = b if b.ndim == 2 else b[:, None]
        starts = None if x0 is None else (x0 if x0.ndim == 2 else x0[:, None])
        out = np.empty_like(cols)
        for c in range(cols.shape[1]):
            out[:, c], _ = spla.cg(self.A, cols[:, c], x0=None if starts is None else starts[:, c],
                                   rtol=tol, maxiter=maxiter or 2000, M=self.M)
        return out if b.ndim == 2 else out[:, 0]


# ----------------------------------------------------------------- LSCM
def far_pair(P, cand):
    """param.py:172-179 -- boundary vertex farthest from the boundary centroid,
    then the boundary vertex farthest from it."""
    Q = P[cand]
    c = Q.mean(0)
    i = cand[np.argmax(((Q - c) ** 2).sum(1))]
    j = cand[np.argmax(((Q - P[i]) ** 2).sum(1))]
    if i == j:
        j = cand[0] if cand[0] != i else cand[-1]
    return int(i), int(j)


def lscm_system(P, T, X=None, cots=None, pins=None):
    """param.py:136-165 -- the conformal system with the two pins eliminated.

    Returns (Hff, rhs, keep, pin_val, (p0, p1)); x = pin_val; x[keep] = solve."""
    n = len(P)
    T = np.asarray(T, np.int64)
    if X is None:
        X, _ = local_frames(P, T)
    if cots is None:
        cots = cot_weights(X)
    bi, bj = boundary_edges_oriented(T)
    if len(bi) == 0:
        raise ValueError("chart has no boundary (closed surface): needs a cut")
    I, J, W = lap_triplets(T, cots)
    hb = np.full(len(bi), 0.5)
    rows = np.concatenate([I, I + n, bi, bj, bi + n, bj + n])
    cols = np.concatenate([J, J + n, bj + n, bi + n, bj, bi])
    vals = np.concatenate([W, W, -hb, hb, hb, -hb])
    if pins is None:
        pins = far_pair(P, np.unique(bi))
    p0, p1 = pins
    d = np.linalg.norm(P[p1] - P[p0])
    pinned = np.array([p0, p1, p0 + n, p1 + n])
    pin_val = np.zeros(2 * n)
    pin_val[p1] = d
    keep = np.ones(2 * n, bool)
    keep[pinned] = False
    newid = np.full(2 * n, -1)
    newid[keep] = np.arange(int(keep.sum()))
    rf, cf = newid[rows], newid[cols]
    nf = int(keep.sum())
    ff = (rf >= 0) & (cf >= 0)
    Hff = sp.csc_matrix((vals[ff], (rf[ff], cf[ff])), shape=(nf, nf))
    fp = (rf >= 0) & (cf < 0)
    rhs = -np.bincount(rf[fp], weights=vals[fp] * pin_val[cols[fp]], minlength=nf)
    return Hff, rhs, keep, pin_val, (int(p0), int(p1))


def lscm(P, T, X=None, cots=None, pins=None):
    """param.py:136-169."""
    n = len(P)
    Hff, rhs, keep, pin_val, _ = lscm_system(P, T, X, cots, pins)
    x = pin_val.copy()
    x[keep] = SPDSolver(Hff).solve(rhs)
    return np.stack([x[:n], x[n:]], 1)


def tutte(P, T):
    """param.py:182-202 -- uniform Tutte embedding of the longest boundary loop
    onto a circle of the same perimeter."""
    T = np.asarray(T, np.int64)
    loop = max(boundary_loops(T), key=len)
    n = len(P)
    seg = np.linalg.norm(P[np.roll(loop, -1)] - P[loop], axis=1)
    ang = np.r_[0, np.cumsum(seg)[:-1]] / max(seg.sum(), 1e-300) * 2 * np.pi
    Ub = np.stack([np.cos(ang), np.sin(ang)], 1) * (seg.sum() / (2 * np.pi))
    I = np.concatenate([T[:, 0], T[:, 1], T[:, 2], T[:, 1], T[:, 2], T[:, 0]])
    J = np.concatenate([T[:, 1], T[:, 2], T[:, 0], T[:, 0], T[:, 1], T[:, 2]])
    A = sp.csr_matrix((np.ones(len(I)), (I, J)), shape=(n, n))
    A.data[:] = 1.0
    Lu = sp.diags(np.asarray(A.sum(1)).ravel()) - A
    fixed = np.zeros(n, bool)
    fixed[loop] = True
    U = np.zeros((n, 2))
    U[loop] = Ub
    fr = ~fixed
    U[fr] = SPDSolver(Lu[fr][:, fr]).solve(-Lu[fr][:, fixed] @ Ub)
    return U


# ----------------------------------------------------------------- per-triangle kernels
def jacobians(U, T, X):
    """Jacobian of the map from the local frame to UV (param.py:214-224).
    Degenerate frames (|x1*y2| < 1e-300) return ok=False and zeros."""
    T = np.asarray(T, np.int64)
    x1, x2, y2 = X[:, 1, 0], X[:, 2, 0], X[:, 2, 1]
    ok = np.abs(x1 * y2) >= 1e-300
    x1s = np.where(ok, x1, 1.0)
    y2s = np.where(ok, y2, 1.0)
    du1 = U[T[:, 1]] - U[T[:, 0]]
    du2 = U[T[:, 2]] - U[T[:, 0]]
    j00 = du1[:, 0] / x1s
    j10 = du1[:, 1] / x1s
    j01 = (du2[:, 0] - j00 * x2) / y2s
    j11 = (du2[:, 1] - j10 * x2) / y2s
    return j00, j01, j10, j11, ok


def arap_rotations(U, T, X, cots=None, mode=0, sd_weights=None):
    """param.py:206-256, 331-333 -- per-triangle target rotation R (t,2,2) and
    SLIM weight matrix Wm (t,2,2).

    mode 0: R is the normalised conformal part of J (closed-form polar
    rotation, identity if degenerate or |.| < 1e-300); Wm = I.
    mode 1: R = U V^T from the SVD of J with the reflection fix; Wm = U
    diag(w) U^T with w_k = sqrt(max((s - s^-3)/(s - 1), 1e-12)), 2 near s=1."""
    t = len(T)
    j00, j01, j10, j11, ok = jacobians(U, T, X)
    R = np.zeros((t, 2, 2))
    Wm = np.zeros((t, 2, 2))
    Wm[:, 0, 0] = 1.0
    Wm[:, 1, 1] = 1.0
    if mode == 0:
        ca = j00 + j11
        cb = j10 - j01
        h = np.sqrt(ca * ca + cb * cb)
        small = h < 1e-300
        ca = np.where(small, 1.0, ca)
        cb = np.where(small, 0.0, cb)
        h = np.where(small, 1.0, h)
        ca = np.where(ok, ca / h, 1.0)
        cb = np.where(ok, cb / h, 0.0)
        R[:, 0, 0] = ca
        R[:, 0, 1] = -cb
        R[:, 1, 0] = cb
        R[:, 1, 1] = ca
        return R, Wm
    R[:, 0, 0] = 1.0
    R[:, 1, 1] = 1.0
    idx = np.nonzero(ok)[0]
    if len(idx):
        J = np.stack([np.stack([j00[idx], j01[idx]], 1), np.stack([j10[idx], j11[idx]], 1)], 1)
        Uu, S, Vt = np.linalg.svd(J)
        refl = np.linalg.det(Uu @ Vt) < 0
        Uu[refl, :, 1] *= -1.0
        S[refl, 1] *= -1.0
        R[idx] = Uu @ Vt
        s = np.maximum(np.abs(S), 1e-8)
        near = np.abs(s - 1.0) < 1e-6
        sden = np.where(near, 2.0, s - 1.0)
        w = np.where(near, 2.0, np.sqrt(np.maximum((s - s ** -3) / sden, 1e-12)))
        Wm[idx] = np.einsum("tik,tk,tjk->tij", Uu, w, Uu)
    return R, Wm


def arap_rhs(n, T, X, cots, R):
    """param.py:457-474 -- b[i] += w R (X_i - X_j), b[j] -= w R (X_i - X_j) for
    edge k = (corner k+1 -> k+2) of every triangle, w = cots[:,k]/2."""
    T = np.asarray(T, np.int64)
    b = np.zeros((n, 2))
    for k in range(3):
        ci, cj = (k + 1) % 3, (k + 2) % 3
        w = 0.5 * cots[:, k]
        d = X[:, ci] - X[:, cj]
        r = np.einsum("tab,tb->ta", R, d) * w[:, None]
        for c in range(2):
            b[:, c] += np.bincount(T[:, ci], weights=r[:, c], minlength=n)
            b[:, c] -= np.bincount(T[:, cj], weights=r[:, c], minlength=n)
    return b


def max_step(U, D, T):
    """param.py:259-291 -- largest t in (0, 1e30] before any positively oriented
    triangle's signed area reaches zero along U + t D (Smith & Schaefer)."""
    T = np.asarray(T, np.int64)
    a, b, c = T[:, 0], T[:, 1], T[:, 2]
    ux1, uy1 = U[b, 0] - U[a, 0], U[b, 1] - U[a, 1]
    ux2, uy2 = U[c, 0] - U[a, 0], U[c, 1] - U[a, 1]
    dx1, dy1 = D[b, 0] - D[a, 0], D[b, 1] - D[a, 1]
    dx2, dy2 = D[c, 0] - D[a, 0], D[c, 1] - D[a, 1]
    C = ux1 * uy2 - ux2 * uy1
    Bq = ux1 * dy2 + dx1 * uy2 - ux2 * dy1 - dx2 * uy1
    A = dx1 * dy2 - dx2 * dy1
    live = C > 0
    lin = live & ((A == 0.0) | (np.abs(A) < 1e-30 * np.maximum(np.abs(Bq), 1e-300)))
    quad = live & ~lin
    best = 1e30
    with np.errstate(divide="ignore", invalid="ignore"):
        sel = lin & (Bq < 0)
        if sel.any():
            r = -C[sel] / Bq[sel]
            r = r[r < best]
            if len(r):
                best = min(best, float(r.min()))
        disc = Bq * Bq - 4 * A * C
        sel = quad & (disc >= 0)
        if sel.any():
            sq = np.sqrt(disc[sel])
            A2 = 2 * A[sel]
            for r in ((-Bq[sel] - sq) / A2, (-Bq[sel] + sq) / A2):
                r = r[(r > 0) & (r < best)]
                if len(r):
                    best = min(best, float(r.min()))
    return best


def energy_per_tri(U, T, X, area, kind):
    """param.py:294-319 -- per-triangle energy.  kind 0 ARAP:
    area((q+r-1)^2 + (q-r-1)^2) with the signed singular values q+-r; kind 1
    symmetric Dirichlet area(|J|^2 + |J|^2/det^2 - 4), inf if det <= 0."""
    j00, j01, j10, j11, ok = jacobians(U, T, X)
    if kind == 0:
        qe, qh = 0.5 * (j00 + j11), 0.5 * (j10 - j01)
        rf, rg = 0.5 * (j00 - j11), 0.5 * (j10 + j01)
        q = np.sqrt(qe * qe + qh * qh)
        r = np.sqrt(rf * rf + rg * rg)
        e = area * ((q + r - 1) ** 2 + (q - r - 1) ** 2)
    else:
        dj = j00 * j11 - j01 * j10
        fro = j00 * j00 + j01 * j01 + j10 * j10 + j11 * j11
        with np.errstate(divide="ignore", invalid="ignore"):
            e = np.where(dj > 0, area * (fro + fro / (dj * dj) - 4.0), np.inf)
    return np.where(ok, e, 0.0)


def energy(U, T, X, area, kind):
    """param.py:294-319, 336-338 -- summed energy."""
    return float(energy_per_tri(U, T, X, area, kind).sum())


def slim_gradient_op(X):
    """param.py:341-350 -- per-triangle 2x3 gradient operator G."""
    x1, x2, y2 = X[:, 1, 0], X[:, 2, 0], X[:, 2, 1]
    G = np.zeros((len(X), 2, 3))
    G[:, 0, 0] = -1 / x1
    G[:, 0, 1] = 1 / x1
    G[:, 1, 0] = (-1 + x2 / x1) / y2
    G[:, 1, 1] = (-x2 / x1) / y2
    G[:, 1, 2] = 1 / y2
    return G


def slim_system(n, T, G, Wm, R, area):
    """param.py:353-376 -- rows sqrt(area) W (J - R) for the 4 entries (s, c),
    unknowns x = [u; v]."""
    T = np.asarray(T, np.int64)
    t = len(T)
    WR = np.einsum("tsr,trc->tsc", Wm, R)
    sq = np.sqrt(area)
    I, Jc, V, B = [], [], [], []
    blk = 0
    for s in range(2):
        for c in range(2):
            row = np.arange(t) + blk * t
            for r in range(2):
                for k in range(3):
                    I.append(row)
                    Jc.append(T[:, k] + r * n)
                    V.append(sq * Wm[:, s, r] * G[:, c, k])
            B.append(sq * WR[:, s, c])
            blk += 1
    A = sp.csr_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(Jc))), shape=(4 * t, 2 * n))
    return A, np.concatenate(B)


def refine(P, T, U, iters=10, method="arap", X=None, area=None, cots=None, solver=None, line_search=True,
           tol=1e-5, pin=None, trace=None):
    """param.py:379-454 -- local/global refinement.

    ARAP: soft pin at ``pin`` (default 0) with weight max|W|; each iteration
    rotations -> rhs -> exact solve warm-started at U -> optional flip-safe step
    min(1, 0.9 t_max) (only if the start is flip-free) -> energy -> early exit
    when |prev - e| <= tol max(prev, 1e-300).  ``trace`` (a list, optional)
    receives (energy, step) per iteration for trajectory comparisons."""
    T = np.asarray(T, np.int64)
    U = np.asarray(U, np.float64)
    n = len(P)
    if X is None:
        X, area = local_frames(P, T)
    if cots is None:
        cots = cot_weights(X)
    flips0 = int((signed_areas(U, T) <= 0).sum())
    use_ls = line_search and flips0 == 0
    pin = 0 if pin is None else pin
    if method == "arap":
        I, J, W = lap_triplets(T, cots)
        wpin = float(np.abs(W).max()) if len(W) else 1.0
        if solver is None:
            L = sp.csc_matrix((np.r_[W, wpin], (np.r_[I, pin], np.r_[J, pin])), shape=(n, n))
            solver = SPDSolver(L)
        prev = energy(U, T, X, area, 0)
        for _ in range(iters):
            R, _w = arap_rotations(U, T, X, cots, 0)
            b = arap_rhs(n, T, X, cots, R)
            b[pin] += wpin * U[pin]
            Un = solver.solve(b, x0=U)
            step = 1.0
            if use_ls:
                D = Un - U
                step = min(1.0, 0.9 * max_step(U, D, T))
                Un = U + step * D
            e = energy(Un, T, X, area, 0)
            if trace is not None:
                trace.append((e, step))
            U = Un
            if abs(prev - e) <= tol * max(prev, 1e-300):
                break
            prev = e
        return U, solver
    G = slim_gradient_op(X)
    prev = energy(U, T, X, area, 1)
    if not np.isfinite(prev):
        return U, None
    lam = 1e-6 * area.sum()
    for _ in range(iters):
        R, Wm = arap_rotations(U, T, X, cots, 1)
        A, b = slim_system(n, T, G, Wm, R, area)
        M = ((A.T @ A).tocsr() + sp.diags(np.full(2 * n, lam * 1e-6))).tocsr()
        x0 = np.r_[U[:, 0], U[:, 1]]
        Atb = A.T @ b + lam * 1e-6 * x0
        if 2 * n <= DIRECT_LIMIT:
            x = spla.spsolve(M.tocsc(), Atb)
        else:
            x, _ = spla.cg(M, Atb, x0=x0, rtol=1e-6, maxiter=500, M=sp.diags(1 / M.diagonal()))
        D = np.stack([x[:n], x[n:]], 1) - U
        step = min(1.0, 0.9 * max_step(U, D, T))
        e = energy(U + step * D, T, X, area, 1)
        k = 0
        while (not np.isfinite(e) or e > prev) and k < 12:
            step *= 0.5
            e = energy(U + step * D, T, X, area, 1)
            k += 1
        if not np.isfinite(e) or e > prev:
            break
        U = U + step * D
        if trace is not None:
            trace.append((e, step))
        if abs(prev - e) <= tol * prev:
            break
        prev = e
    return U, None


# ----------------------------------------------------------------- flip repair
def one_ring(T, n):
    """param.py:507-512 -- sorted unique neighbours of every vertex (CSR)."""
    T = np.asarray(T, np.int64)
    I = np.concatenate([T[:, 0], T[:, 1], T[:, 2], T[:, 1], T[:, 2], T[:, 0]])
    J = np.concatenate([T[:, 1], T[:, 2], T[:, 0], T[:, 0], T[:, 1], T[:, 2]])
    key = np.unique(I * n + J)
    I, J = key // n, key % n
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, I + 1, 1)
    return np.cumsum(ptr), J


def repair_flips(U, T, rounds=60):
    """param.py:478-516 -- Gauss-Seidel 1-ring relaxation of the non-boundary
    corners of every triangle with signed area <= 0, in triangle order,
    updating U in place within a round; stops after a round with none."""
    T = np.asarray(T, np.int64)
    n = len(U)
    ptr, nbr = one_ring(T, n)
    bi, bj = boundary_edges_oriented(T)
    bnd = np.zeros(n, bool)
    bnd[bi] = True
    bnd[bj] = True
    U = np.array(U, np.float64)
    Tl = T.tolist()
    for _ in range(rounds):
        bad = 0
        for a, b, c in Tl:
            ar = (U[b, 0] - U[a, 0]) * (U[c, 1] - U[a, 1]) - (U[c, 0] - U[a, 0]) * (U[b, 1] - U[a, 1])
            if ar > 0:
                continue
            bad += 1
            for v in (a, b, c):
                if bnd[v]:
                    continue
                lo, hi = ptr[v], ptr[v + 1]
                if hi > lo:
                    nb = nbr[lo:hi]
                    # sequential left-to-right sums, as the reference's loop
                    sx = 0.0
                    sy = 0.0
                    for w in nb.tolist():
                        sx += U[w, 0]
                        sy += U[w, 1]
                    U[v, 0] = sx / (hi - lo)
                    U[v, 1] = sy / (hi - lo)
        if bad == 0:
            break
    return U


# ----------------------------------------------------------------- quality
def weighted_quantile(x, w, q):
    """param.py:543-546 -- argsort, cumulative weight, searchsorted(q * total)
    (left side), clamped to the last index."""
    o = np.argsort(x)
    cw = np.cumsum(w[o])
    i = min(int(np.searchsorted(cw, q * cw[-1])), len(o) - 1)
    return float(x[o][i])


def chart_quality(P, T, U, X=None, area=None):
    """param.py:520-551 -- area-weighted distortion of one chart."""
    T = np.asarray(T, np.int64)
    if X is None:
        X, area = local_frames(P, T)
    sa = signed_areas(U, T)
    A = area.sum()
    flips = int((sa <= 0).sum()) if (sa > 0).sum() >= (sa < 0).sum() else int((sa >= 0).sum())
    x1, x2, y2 = X[:, 1, 0], X[:, 2, 0], X[:, 2, 1]
    ok = np.abs(x1 * y2) > 1e-300
    x1s = np.where(ok, x1, 1)
    y2s = np.where(ok, y2, 1)
    a0, a1, a2 = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
    j00 = (a1[:, 0] - a0[:, 0]) / x1s
    j10 = (a1[:, 1] - a0[:, 1]) / x1s
    j01 = ((a2[:, 0] - a0[:, 0]) - j00 * x2) / y2s
    j11 = ((a2[:, 1] - a0[:, 1]) - j10 * x2) / y2s
    E, F = (j00 + j11) / 2, (j00 - j11) / 2
    G, H = (j10 + j01) / 2, (j10 - j01) / 2
    Q = np.sqrt(E * E + H * H)
    Rr = np.sqrt(F * F + G * G)
    s1 = Q + Rr
    s2 = np.abs(Q - Rr)
    sc = np.abs(sa).sum() / max(A, 1e-300)
    w = area / max(A, 1e-300)
    an = s1 / np.maximum(s2, 1e-12)
    ar = (s1 * s2) / max(sc, 1e-300)
    lar = np.abs(np.log2(np.maximum(ar, 1e-12)))
    st = np.sqrt((s1 ** 2 + s2 ** 2) / 2 / max(sc, 1e-300))
    return {"flips": flips, "aniso_p95": weighted_quantile(an, w, 0.95), "aniso_mean": float((an * w).sum()),
            "area_p95": float(2 ** weighted_quantile(lar, w, 0.95)), "area_logdev": float((lar * w).sum()),
            "stretch": float(np.sqrt((st ** 2 * w).sum())), "scale": float(np.sqrt(sc)),
            "per_tri_aniso": an, "per_tri_area": ar}


def flatten_lscm_arap(P, T, iters=8, tol=1e-4):
    """charts.py:401-416 -- LSCM, majority-orientation mirror fix, then ARAP
    (iterations doubled and no line search when the start is folded)."""
    T = np.asarray(T, np.int64)
    X, area = local_frames(P, T)
    cots = cot_weights(X)
    U = lscm(P, T, X, cots)
    sa = signed_areas(U, T)
    if (sa < 0).sum() > (sa > 0).sum():
        U[:, 1] = -U[:, 1]
        sa = -sa
    folded = bool((sa <= 0).any())
    if iters:
        U, _ = refine(P, T, U, iters=iters * (2 if folded else 1), method="arap", X=X, area=area, cots=cots,
                      line_search=not folded, tol=tol)
    return U


def flatten(P, T, iters=8, slim_iters=0):
    """charts.py:401-425 -- flatten_lscm_arap, then, if folds remain, the
    sequential repair and, when it removed them, a flip-safe ARAP of
    max(2, iters // 2) iterations; with ``slim_iters`` and no fold left,
    that many SLIM iterations (kept when finite).  Method "arap"; the
    boundary untangle of charts.py:426-431 is not restated."""
    T = np.asarray(T, np.int64)
    X, area = local_frames(P, T)
    cots = cot_weights(X)
    U = flatten_lscm_arap(P, T, iters=iters, tol=1e-4)
    if (signed_areas(U, T) <= 0).any():
        U = repair_flips(U, T)
        if not (signed_areas(U, T) <= 0).any() and iters:
            U, _ = refine(P, T, U, iters=max(2, iters // 2), method="arap", X=X, area=area, cots=cots,
                          line_search=True, tol=1e-4)
    if slim_iters and not (signed_areas(U, T) <= 0).any():
        U2, _ = refine(P, T, U, iters=slim_iters, method="slim", X=X, area=area, cots=cots)
        if np.isfinite(U2).all():
            U = U2
    return U
