"""Float64 CPU restatement of the semantics of ``engine/gpu.py``'s ``Batch``.

TEST INFRASTRUCTURE ONLY (the oracle).

The reference batches charts block-diagonally on the GPU in float32 with a
CG capped at 40 iterations.  This restatement keeps every per-chart rule of
that class (frame thresholds, soft pin at each chart's first vertex with a
batch-global weight, per-chart flip-safe step, revert of non-finite charts,
Jacobi flip repair, orientation fix) but solves each global step exactly in
float64, which is what the product's CG converges to.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class BatchOracle:
    """gpu.py:66-299 restated (float64, exact solves)."""

    def __init__(self, charts):
        P, T, U, cv, ct = [], [], [], [], []
        off = 0
        for i, (p, t, u) in enumerate(charts):
            P.append(np.asarray(p, np.float64))
            U.append(np.asarray(u, np.float64))
            T.append(np.asarray(t, np.int64) + off)
            cv.append(np.full(len(p), i, np.int64))
            ct.append(np.full(len(t), i, np.int64))
            off += len(p)
        self.nc = len(charts)
        self.offsets = np.cumsum([0] + [len(p) for p, _, _ in charts])
        self.n = off
        Pn, Tn = np.concatenate(P), np.concatenate(T)
        self.T, self.U = Tn, np.concatenate(U)
        self.cv, self.ct = np.concatenate(cv), np.concatenate(ct)
        # gpu.py:88-123 frames, thin-triangle mask, half-cotangent weights
        e1 = Pn[Tn[:, 1]] - Pn[Tn[:, 0]]
        e2 = Pn[Tn[:, 2]] - Pn[Tn[:, 0]]
        l1 = np.maximum(np.linalg.norm(e1, axis=1), 1e-30)
        nrm = np.cross(e1, e2)
        dbl = np.linalg.norm(nrm, axis=1)
        ex = e1 / l1[:, None]
        ey = np.cross(nrm / np.maximum(dbl, 1e-30)[:, None], ex)
        X = np.zeros((len(Tn), 3, 2))
        X[:, 1, 0] = l1
        X[:, 2, 0] = (e2 * ex).sum(1)
        X[:, 2, 1] = (e2 * ey).sum(1)
        ok = (dbl > 1e-14 * np.maximum(l1 * l1, 1e-300)) & np.isfinite(dbl)
        X[~ok] = 0.0
        self.area = 0.5 * dbl
        self.ok = ok
        w = np.zeros((len(Tn), 3))
        for k in range(3):
            a = X[:, (k + 1) % 3] - X[:, k]
            b = X[:, (k + 2) % 3] - X[:, k]
            cr = np.maximum(np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]), 1e-30)
            w[:, k] = 0.5 * np.clip((a * b).sum(1) / cr, -1e4, 1e4)
        w[~ok] = 0.0
        self.w = w
        self.x1 = np.maximum(X[:, 1, 0], 1e-30)
        self.x2 = X[:, 2, 0]
        self.y2 = np.where(np.abs(X[:, 2, 1]) > 1e-30, X[:, 2, 1], 1e-30)
        self.X = X
        self.scale = np.ones(len(Tn))
        # gpu.py:126-142 Laplacian with a soft pin at every chart's first vertex
        n = self.n
        i = np.concatenate([Tn[:, 1], Tn[:, 2], Tn[:, 0]])
        j = np.concatenate([Tn[:, 2], Tn[:, 0], Tn[:, 1]])
        W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
        self.pins = self.offsets[:-1].astype(np.int64)
        self.wpin = max(float(np.abs(W).max()) if len(W) else 0.0, 1e-12)
        L = sp.coo_matrix((np.concatenate([-W, -W, W, W]),
                           (np.concatenate([i, j, i, j]), np.concatenate([j, i, i, j]))), shape=(n, n)).tocsc()
        L = L + sp.csc_matrix((np.full(len(self.pins), self.wpin), (self.pins, self.pins)), shape=(n, n))
        self.L = L.tocsc()
        self._lu = spla.splu(self.L, permc_spec="MMD_AT_PLUS_A", options={"SymmetricMode": True})
        key = np.unique(np.minimum(i, j) * n + np.maximum(i, j), return_counts=True)
        self._edges = (key[0] // n, key[0] % n, key[1])

    def jacobians(self, U):
        """gpu.py:145-151."""
        T = self.T
        a0, a1, a2 = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
        j00 = (a1[:, 0] - a0[:, 0]) / self.x1
        j10 = (a1[:, 1] - a0[:, 1]) / self.x1
        j01 = ((a2[:, 0] - a0[:, 0]) - j00 * self.x2) / self.y2
        j11 = ((a2[:, 1] - a0[:, 1]) - j10 * self.x2) / self.y2
        return j00, j01, j10, j11

    def signed_area(self, U):
        """gpu.py:153-156."""
        T = self.T
        a, b, c = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
        return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))

    def flips_per_chart(self, U=None):
        """gpu.py:158-163."""
        U = self.U if U is None else U
        bad = (self.signed_area(U) < 0) & self.ok
        return np.bincount(self.ct, weights=bad.astype(np.float64), minlength=self.nc)

    def rhs(self, U):
        """gpu.py:165-182 -- rotation = normalised conformal part of J (the angle
        of atan2(j10 - j01, j00 + j11)), times the per-triangle scale; zero
        angle for thin or non-finite triangles."""
        j00, j01, j10, j11 = self.jacobians(U)
        ca, cb = j00 + j11, j10 - j01
        h = np.sqrt(ca * ca + cb * cb)
        good = self.ok & np.isfinite(h)
        with np.errstate(invalid="ignore", divide="ignore"):
            c = np.where(good & (h > 0), ca / np.where(h > 0, h, 1.0), 1.0)
            s = np.where(good & (h > 0), cb / np.where(h > 0, h, 1.0), 0.0)
        c, s = c * self.scale, s * self.scale
        b = np.zeros((self.n, 2))
        X, T, w = self.X, self.T, self.w
        for k in range(3):
            i, j = (k + 1) % 3, (k + 2) % 3
            dx = X[:, i] - X[:, j]
            r = np.stack([c * dx[:, 0] - s * dx[:, 1], s * dx[:, 0] + c * dx[:, 1]], 1) * w[:, k:k + 1]
            for col in range(2):
                b[:, col] += np.bincount(T[:, i], weights=r[:, col], minlength=self.n)
                b[:, col] -= np.bincount(T[:, j], weights=r[:, col], minlength=self.n)
        b[self.pins] += self.wpin * U[self.pins]
        return b

    def max_step(self, U, D):
        """gpu.py:204-227 -- per-chart flip-safe step (1e30 when unlimited)."""
        T = self.T
        u0, u1, u2 = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
        d0, d1, d2 = D[T[:, 0]], D[T[:, 1]], D[T[:, 2]]
        ux1, uy1 = u1[:, 0] - u0[:, 0], u1[:, 1] - u0[:, 1]
        ux2, uy2 = u2[:, 0] - u0[:, 0], u2[:, 1] - u0[:, 1]
        dx1, dy1 = d1[:, 0] - d0[:, 0], d1[:, 1] - d0[:, 1]
        dx2, dy2 = d2[:, 0] - d0[:, 0], d2[:, 1] - d0[:, 1]
        C = ux1 * uy2 - ux2 * uy1
        B = ux1 * dy2 + dx1 * uy2 - ux2 * dy1 - dx2 * uy1
        A = dx1 * dy2 - dx2 * dy1
        big = np.full_like(C, 1e30)
        with np.errstate(invalid="ignore", divide="ignore"):
            lin = np.where((np.abs(A) < 1e-30) & (B < 0), -C / np.where(B != 0, B, 1.0), big)
            disc = B * B - 4 * A * C
            sq = np.sqrt(np.maximum(disc, 0))
            A2 = np.where(np.abs(A) < 1e-30, 1.0, 2 * A)
            r1, r2 = (-B - sq) / A2, (-B + sq) / A2
            quad = np.minimum(np.where(r1 > 0, r1, big), np.where(r2 > 0, r2, big))
            quad = np.where((disc >= 0) & (np.abs(A) >= 1e-30), quad, big)
        t = np.where(C > 0, np.minimum(lin, quad), big)
        out = np.full(self.nc, 1e30)
        np.minimum.at(out, self.ct, t)
        return out

    def cg(self, b, x, iters=40, tol=1e-6):
        """gpu.py:184-202 Batch._cg: one Jacobi-PCG over every chart, a step
        length per column, stopped when both columns reach r.r <= tol^2
        max(b.b, 1e-30) or after ``iters`` iterations."""
        L = self.L.tocsr()
        dinv = (1.0 / np.maximum(L.diagonal(), 1e-30))[:, None]
        r = b - L @ x
        z = r * dinv
        p = z.copy()
        rz = (r * z).sum(0)
        b2 = np.maximum((b * b).sum(0), 1e-30)
        for _ in range(iters):
            Ap = L @ p
            al = rz / np.maximum((p * Ap).sum(0), 1e-30)
            x = x + p * al
            r = r - Ap * al
            if ((r * r).sum(0) <= tol * tol * b2).all():
                break
            z = r * dinv
            rz2 = (r * z).sum(0)
            p = z + p * (rz2 / np.maximum(rz, 1e-30))
            rz = rz2
        return x

    def arap(self, iters=8, line_search=True, cg_iters=None):
        """gpu.py:230-249; global steps solved exactly, or (``cg_iters``) by
        the reference's coupled CG of that budget."""
        U = self.U
        for _ in range(iters):
            Un = self._lu.solve(self.rhs(U)) if cg_iters is None else self.cg(self.rhs(U), U.copy(), cg_iters)
            badv = ~np.isfinite(Un).all(1)
            if badv.any():
                badc = np.zeros(self.nc, bool)
                badc[self.cv[badv]] = True
                Un = np.where(badc[self.cv][:, None], U, Un)
            if line_search:
                step = np.minimum(0.9 * self.max_step(U, Un - U), 1.0)
                U = U + (Un - U) * step[self.cv][:, None]
            else:
                U = Un
        self.U = U
        return U

    def repair_flips(self, rounds=60):
        """gpu.py:251-273 -- Jacobi 1-ring averaging of the non-boundary
        vertices of folded (sa < 0, thick) triangles."""
        ea, eb, cnt = self._edges
        bnd = np.zeros(self.n, bool)
        one = cnt == 1
        bnd[ea[one]] = True
        bnd[eb[one]] = True
        A = sp.coo_matrix((np.ones(2 * len(ea)), (np.r_[ea, eb], np.r_[eb, ea])), shape=(self.n, self.n)).tocsr()
        deg = np.maximum(np.asarray(A.sum(1)).ravel(), 1)
        U = self.U
        for _ in range(rounds):
            bad = (self.signed_area(U) < 0) & self.ok
            if not bad.any():
                break
            mv = np.zeros(self.n, bool)
            mv[self.T[bad].reshape(-1)] = True
            mv &= ~bnd
            U = np.where(mv[:, None], (A @ U) / deg[:, None], U)
        self.U = U
        return U

    def orient_positive(self):
        """gpu.py:275-281."""
        sa = self.signed_area(self.U)
        s = np.bincount(self.ct, weights=np.sign(sa) * self.area, minlength=self.nc)
        flip = (s < 0)[self.cv]
        self.U = np.where(flip[:, None], self.U * np.array([1.0, -1.0]), self.U)

    def chart_uvs(self):
        """gpu.py:297-299."""
        return [self.U[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]
