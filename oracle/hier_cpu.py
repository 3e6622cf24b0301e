"""Float64 CPU restatement of ``engine/hier.py``'s ``Solver`` (relax and
bad_fraction).

TEST INFRASTRUCTURE ONLY (the oracle).

Restated (engine/hier.py):
  frames, cotangent weights, usable mask     hier.py:361-389
  relax: Laplacian with a soft pin at each   hier.py:394-417
    chart's smallest vertex id
  orientation by area-weighted sign          hier.py:423-426
  Jacobi flip repair                          hier.py:427-446
  rhs (atan2 rotation of every usable face)   hier.py:448-463
  Jacobi-PCG of exactly cg_iters iterations   hier.py:465-474
  flip-safe step over usable faces            hier.py:476-496
  iteration loop, mid and final repair        hier.py:498-510
  quality (flips, scale, per-face aniso,      hier.py:511-530
    area ratio)
  bad_fraction                                hier.py:584-590
  lscm: Dirichlet on u and v minus the area   hier.py:534-582
    term on directed boundary edges, two soft
    pins per chart (extreme start u among its
    boundary vertices, weight 10 max |W|),
    Jacobi-PCG from the start layout, at most
    ``iters`` iterations, stopping when
    r.r < 1e-12 b.b at every ``check_every``-th
    iteration (the reference: 50)
The reference runs in float32 on the GPU; this restatement keeps every rule
in float64, so it matches the reference to float32 rounding (pinned against
the reference's own CPU-torch output in tests/golden/hier_golden.npz).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def frames(V, F):
    """Isometric per-face frames X (t,3,2), half cotangent weights w (t,3),
    usable mask, 3-D area, and the Jacobian denominators x1, x2, y2."""
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
    cots = []
    for k in range(3):
        i, j = (k + 1) % 3, (k + 2) % 3
        a, b = X[:, i] - X[:, k], X[:, j] - X[:, k]
        cr = np.maximum(np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]), 1e-30)
        cots.append(np.clip((a * b).sum(1) / cr, -1e4, 1e4))
    w = 0.5 * np.stack(cots, 1)
    w[~ok] = 0.0
    x1 = np.maximum(X[:, 1, 0], 1e-30)
    x2 = X[:, 2, 0]
    y2 = np.where(np.abs(X[:, 2, 1]) > 1e-30, X[:, 2, 1], 1e-30)
    return X, w, ok, 0.5 * dbl, x1, x2, y2


def _sa(U, T):
    a, b, c = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
    return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))


def _jac(U, T, x1, x2, y2):
    a0, a1, a2 = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
    j00 = (a1[:, 0] - a0[:, 0]) / x1
    j10 = (a1[:, 1] - a0[:, 1]) / x1
    j01 = ((a2[:, 0] - a0[:, 0]) - j00 * x2) / y2
    j11 = ((a2[:, 1] - a0[:, 1]) - j10 * x2) / y2
    return j00, j01, j10, j11


class Solver:
    """hier.py:357-593 (float64, CPU)."""

    def __init__(self, V, F):
        self.X, self.w, self.ok, self.area, self.x1, self.x2, self.y2 = frames(V, F)
        self._q = None

    def relax(self, faces, CV, nv, chart, nchart, U0, iters=6, cg_iters=20, repair_rounds=24):
        fi = np.asarray(faces, np.int64)
        T = np.asarray(CV, np.int64).reshape(-1, 3)
        ct = np.asarray(chart, np.int64)
        U = np.asarray(U0, np.float64).astype(np.float32).astype(np.float64)
        X, w, ok = self.X[fi], self.w[fi], self.ok[fi]
        x1, x2, y2 = self.x1[fi], self.x2[fi], self.y2[fi]
        area = self.area[fi]
        cv = np.zeros(nv, np.int64)
        cv[T.reshape(-1)] = np.repeat(ct, 3)
        I = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
        J = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
        W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
        pins = np.full(nchart, nv, np.int64)
        np.minimum.at(pins, cv, np.arange(nv))
        pins = pins[pins < nv]
        wpin = max(float(np.abs(W).max()) if len(W) else 0.0, 1e-12)
        rows = np.concatenate([I, J, I, J, pins])
        cols = np.concatenate([J, I, I, J, pins])
        vals = np.concatenate([-W, -W, W, W, np.full(len(pins), wpin)])
        L = sp.csr_matrix((vals, (rows, cols)), shape=(nv, nv))
        diag = np.bincount(rows[rows == cols], weights=vals[rows == cols], minlength=nv)
        dinv = (1.0 / np.maximum(diag, 1e-30))[:, None]
        # orientation (area-weighted sign per chart)
        s_ = np.bincount(ct, weights=np.sign(_sa(U, T)) * area, minlength=nchart)
        U = np.where((s_ < 0)[cv][:, None], U * np.array([1.0, -1.0]), U)
        # repair adjacency: unique neighbour pairs, boundary = pairs used once
        key = np.concatenate([I, J]) * nv + np.concatenate([J, I])
        uk, cnt = np.unique(key, return_counts=True)
        bnd = np.zeros(nv, bool)
        bnd[(uk // nv)[cnt == 1]] = True
        Ar = sp.csr_matrix((np.ones(len(uk)), (uk // nv, uk % nv)), shape=(nv, nv))
        deg = np.maximum(np.asarray(Ar.sum(1)).ravel(), 1.0)

        def repair(Uv):
            for _ in range(repair_rounds):
                bad = (_sa(Uv, T) < 0) & ok
                if not bad.any():
                    break
                mv = np.zeros(nv, bool)
                mv[T[bad].reshape(-1)] = True
                mv &= ~bnd
                Uv = np.where(mv[:, None], (Ar @ Uv) / deg[:, None], Uv)
            return Uv

        def rhs(Uv):
            j00, j01, j10, j11 = _jac(Uv, T, x1, x2, y2)
            with np.errstate(invalid="ignore"):
                th = np.arctan2(j10 - j01, j00 + j11)
            th = np.where(ok & np.isfinite(th), th, 0.0)
            c, s = np.cos(th), np.sin(th)
            b = np.zeros((nv, 2))
            for k in range(3):
                i, j = (k + 1) % 3, (k + 2) % 3
                dx = X[:, i] - X[:, j]
                r = np.stack([c * dx[:, 0] - s * dx[:, 1], s * dx[:, 0] + c * dx[:, 1]], 1) * w[:, k:k + 1]
                np.add.at(b, T[:, i], r)
                np.add.at(b, T[:, j], -r)
            b[pins] += wpin * Uv[pins]
            return b

        def cg(b, x):
            r = b - L @ x
            z = r * dinv
            p = z.copy()
            rz = (r * z).sum(0)
            for _ in range(cg_iters):
                Ap = L @ p
                al = rz / np.maximum((p * Ap).sum(0), 1e-30)
                x = x + p * al
                r = r - Ap * al
                z = r * dinv
                rz2 = (r * z).sum(0)
                p = z + p * (rz2 / np.maximum(rz, 1e-30))
                rz = rz2
            return x

        def max_step(Uv, D):
            u0, u1, u2 = Uv[T[:, 0]], Uv[T[:, 1]], Uv[T[:, 2]]
            d0, d1, d2 = D[T[:, 0]], D[T[:, 1]], D[T[:, 2]]
            ux1, uy1, ux2, uy2 = u1[:, 0] - u0[:, 0], u1[:, 1] - u0[:, 1], u2[:, 0] - u0[:, 0], u2[:, 1] - u0[:, 1]
            dx1, dy1, dx2, dy2 = d1[:, 0] - d0[:, 0], d1[:, 1] - d0[:, 1], d2[:, 0] - d0[:, 0], d2[:, 1] - d0[:, 1]
            C = ux1 * uy2 - ux2 * uy1
            Bq = ux1 * dy2 + dx1 * uy2 - ux2 * dy1 - dx2 * uy1
            Aq = dx1 * dy2 - dx2 * dy1
            big = np.full_like(C, 1e30)
            with np.errstate(divide="ignore", invalid="ignore"):
                lin = np.where((np.abs(Aq) < 1e-30) & (Bq < 0), -C / np.where(Bq != 0, Bq, 1.0), big)
                disc = Bq * Bq - 4 * Aq * C
                sq = np.sqrt(np.maximum(disc, 0.0))
                A2 = np.where(np.abs(Aq) < 1e-30, 1.0, 2 * Aq)
                r1, r2 = (-Bq - sq) / A2, (-Bq + sq) / A2
            quad = np.minimum(np.where(r1 > 0, r1, big), np.where(r2 > 0, r2, big))
            quad = np.where((disc >= 0) & (np.abs(Aq) >= 1e-30), quad, big)
            t = np.where((C > 0) & ok, np.minimum(lin, quad), big)
            out = np.full(nchart, 1e30)
            np.minimum.at(out, ct, t)
            return out

        U = repair(U)
        for it in range(iters):
            Un = cg(rhs(U), U.copy())
            badv = ~np.isfinite(Un).all(1)
            if badv.any():
                badc = np.zeros(nchart, bool)
                badc[cv[badv]] = True
                Un = np.where(badc[cv][:, None], U, Un)
            fl_c = np.bincount(ct, weights=((_sa(U, T) < 0) & ok).astype(float), minlength=nchart)
            D = Un - U
            step = np.minimum(0.9 * max_step(U, D), 1.0)
            step = np.where(fl_c > 0, 1.0, step)
            U = U + D * step[cv][:, None]
            if it == iters // 2:
                U = repair(U)
        U = repair(U)
        # quality (hier.py:511-530)
        j00, j01, j10, j11 = _jac(U, T, x1, x2, y2)
        E, Fq, G, H = (j00 + j11) / 2, (j00 - j11) / 2, (j10 + j01) / 2, (j10 - j01) / 2
        Q, R = np.sqrt(E * E + H * H), np.sqrt(Fq * Fq + G * G)
        s1, s2 = Q + R, np.abs(Q - R)
        ar3 = area * ok
        sgn = _sa(U, T)
        uva = np.bincount(ct, weights=np.abs(sgn) * ok, minlength=nchart)
        a3 = np.bincount(ct, weights=ar3, minlength=nchart)
        scale = uva / np.maximum(a3, 1e-30)
        aniso = s1 / np.maximum(s2, 1e-12)
        arat = (s1 * s2) / np.maximum(scale[ct], 1e-30)
        flips = np.bincount(ct, weights=((sgn < 0) & ok).astype(float), minlength=nchart)
        self._q = (aniso, arat, ar3, ct, nchart, a3)
        return U, flips, scale

    def lscm(self, faces, CV, nv, chart, nchart, U0, ba, bb, iters=600, check_every=50):
        """hier.py:534-582.  Where several boundary vertices of a chart share
        the extreme start u, the largest vertex id is the pin (the last write
        of the reference's scatter on the CPU)."""
        fi = np.asarray(faces, np.int64)
        T = np.asarray(CV, np.int64).reshape(-1, 3)
        ct = np.asarray(chart, np.int64)
        w = self.w[fi]
        I = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
        J = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
        W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
        bi = np.asarray(ba, np.int64)
        bj = np.asarray(bb, np.int64)
        h = np.full(len(bi), 0.5)
        rows = np.concatenate([I, J, I, J, I + nv, J + nv, I + nv, J + nv, bi, bj, bi + nv, bj + nv])
        cols = np.concatenate([J, I, I, J, J + nv, I + nv, I + nv, J + nv, bj + nv, bi + nv, bj, bi])
        vals = np.concatenate([-W, -W, W, W, -W, -W, W, W, -h, h, h, -h])
        U0f = np.asarray(U0, np.float64).reshape(nv, 2).astype(np.float32).astype(np.float64)
        cv = np.zeros(nv, np.int64)
        cv[T.reshape(-1)] = np.repeat(ct, 3)
        bv = np.unique(bi)
        key = U0f[bv, 0]
        mn = np.full(nchart, 1e30)
        mx = np.full(nchart, -1e30)
        np.minimum.at(mn, cv[bv], key)
        np.maximum.at(mx, cv[bv], key)
        p0 = np.full(nchart, -1, np.int64)
        p1 = np.full(nchart, -1, np.int64)
        is_mn, is_mx = key == mn[cv[bv]], key == mx[cv[bv]]
        np.maximum.at(p0, cv[bv][is_mn], bv[is_mn])
        np.maximum.at(p1, cv[bv][is_mx], bv[is_mx])
        pins = np.concatenate([p0[p0 >= 0], p1[p1 >= 0]])
        wpin = max(float(np.abs(W).max()) if len(W) else 0.0, 1e-12) * 10
        rows = np.concatenate([rows, pins, pins + nv])
        cols = np.concatenate([cols, pins, pins + nv])
        vals = np.concatenate([vals, np.full(2 * len(pins), wpin)])
        H = sp.csr_matrix((vals, (rows, cols)), shape=(2 * nv, 2 * nv))
        diag = np.bincount(rows[rows == cols], weights=vals[rows == cols], minlength=2 * nv)
        dinv = 1.0 / np.maximum(diag, 1e-30)
        b = np.zeros(2 * nv)
        b[pins] += wpin * U0f[pins, 0]                  # no accumulation over repeated pins, as torch's b[pins] +=
        b[pins + nv] += wpin * U0f[pins, 1]
        x = np.concatenate([U0f[:, 0], U0f[:, 1]])
        r = b - H @ x
        z = r * dinv
        p = z.copy()
        rz = float(r @ z)
        b2 = max(float(b @ b), 1e-30)
        for it in range(int(iters)):
            Ap = H @ p
            al = rz / max(float(p @ Ap), 1e-30)
            x = x + al * p
            r = r - al * Ap
            if it % check_every == 0 and float(r @ r) < 1e-12 * b2:
                break
            z = r * dinv
            rz2 = float(r @ z)
            p = z + (rz2 / max(rz, 1e-30)) * p
            rz = rz2
        U = np.stack([x[:nv], x[nv:]], 1)
        return np.where(np.isfinite(U), U, U0f)

    def bad_fraction(self, max_aniso, max_area):
        aniso, arat, ar3, ct, nchart, a3 = self._q
        badA = np.bincount(ct, weights=ar3 * (aniso > max_aniso), minlength=nchart)
        badR = np.bincount(ct, weights=ar3 * ((arat > max_area) | (arat < 1.0 / max_area)), minlength=nchart)
        return badA / np.maximum(a3, 1e-30), badR / np.maximum(a3, 1e-30)
