"""hier.Solver on the GPU: ARAP relaxation of any subset of a mesh's charts
(hier.py:357-593 of the reference), with the same call signatures.

``relax`` regroups the subset into islands (each chart's vertices in
ascending local id, so the chart's soft pin -- its smallest vertex id,
hier.py:407-409 -- is the island's first vertex), builds a Batch-convention
plan and runs the reference's sequence on the GPU kernels:

  orient every chart by area-weighted sign         uva_orient_positive
  Jacobi flip repair (``repair_rounds``)            uva_repair_jacobi
  ``iters`` ARAP iterations:
    rotations and rhs of every chart                uva_local_rhs (pin weight of the
                                                      subset: uva_plan_set_pins)
    one Jacobi-PCG of exactly ``cg_iters``          uva_plan_pcg on the plan's own
    iterations over all charts together               Laplacian with the subset's
    (the reference's batched CG shares its            per-chart soft pins
    step lengths between charts), from the
    current layout
    non-finite charts keep their layout;            uva_step_apply(UVA_MODE_HIER)
    folded charts take the plain step, the
    others 0.9 of the largest fold-free step
    over usable triangles
    a repair after iteration ``iters // 2``         uva_repair_jacobi
  final repair, per-chart quality                   torch on the plan's GPU

Every vertex must belong to a face of the subset (the reference's callers
pass compacted subsets).

``lscm`` (hier.py:534-582, the free-boundary conformal start the reference
uses behind ``part_lscm``) assembles one 2n-row system over the u and v of
every chart on the host -- the cotangent Dirichlet energy minus the area term
on the given directed boundary edges, plus two soft pins per chart -- and
solves it on the GPU with ``uva_jacobi_pcg``, the reference's Jacobi-PCG
with its stopping rule (600 iterations at most, r.r < 1e-12 b.b checked
every 50).  That rule stops well short of the exact conformal map on charts
of a few thousand vertices (the reference's layout is then a few 1e-3 of the
chart's extent away from it), so the solve keeps the rule rather than
converging further.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .plan import IslandPlan, half_cots, require_gpu, stream_of


class Solver:
    """Frames come from the mesh's positions ``topo.V`` and faces ``topo.F``
    (hier.py:361-389); the GPU plans of a subset are built per ``relax`` call."""

    def __init__(self, topo, dev=None):
        self.dev = require_gpu(dev)
        self.V = np.ascontiguousarray(topo.V, np.float64)
        self.F = np.ascontiguousarray(topo.F, np.int64)
        self._q = None

    def _charts(self, faces, CV, nv, chart, nchart):
        """The subset's islands in chart order, concatenated: per chart its
        vertices in ascending local id and its faces in subset order (threaded
        host code, uva_chart_regroup).  Returns the island offsets of vertices
        and faces, positions (n,3), faces in island-local ids (t,3) (the form
        IslandPlan takes), the charts that have faces, and the vertex and face
        orders."""
        n = len(faces)
        faces = np.ascontiguousarray(faces, np.int64)
        CV = np.ascontiguousarray(CV, np.int64)
        chart = np.ascontiguousarray(chart, np.int64)
        keep = np.empty(max(nchart, 1), np.int64)
        nk = np.zeros(1, np.int64)
        vo = np.empty(max(nchart, 1) + 1, np.int64)
        to = np.empty(max(nchart, 1) + 1, np.int64)
        P = np.empty((nv, 3), np.float64)
        T = np.empty((n, 3), np.int64)
        vorder = np.empty(nv, np.int64)
        forder = np.empty(n, np.int64)
        i64, f64 = _lib.i64p, _lib.f64p
        rc = _lib.lib().uva_chart_regroup(n, i64(faces), i64(CV), nv, i64(chart), nchart, i64(self.F), len(self.F),
                                          f64(self.V), len(self.V), i64(keep), i64(nk), i64(vo), i64(to), f64(P),
                                          i64(T), i64(vorder), i64(forder))
        _lib.check(rc)
        k = int(nk[0])
        return (vo[:k + 1].copy(), to[:k + 1].copy(), P, T), keep[:k].copy(), vorder, forder

    def relax(self, faces, CV, nv, chart, nchart, U0, iters=6, cg_iters=20, repair_rounds=24):
        """faces (n,) global face ids, CV (n,3) local vertex ids, chart (n,)
        local chart ids, U0 (nv,2) start layout.  Returns the layout (nv,2)
        float64 and per-chart flips and UV/3-D area scale (hier.py:394-530)."""
        faces = np.asarray(faces, np.int64).reshape(-1)
        CV = np.asarray(CV, np.int64).reshape(-1, 3)
        chart = np.asarray(chart, np.int64).reshape(-1)
        nv, nchart = int(nv), int(nchart)
        U0 = np.asarray(U0, np.float64).reshape(nv, 2)
        cat, keep, vorder, forder = self._charts(faces, CV, nv, chart, nchart)
        flips = np.zeros(nchart)
        scale = np.zeros(nchart)
        if len(keep) == 0:
            self._q = None
            return U0.astype(np.float32).astype(np.float64), flips, scale
        # per-chart plan: orientation, repair, flips, rotations and rhs, step
        # limits, and the global step: one Jacobi-PCG over every row of the
        # plan as the reference solves the subset (shared step lengths), one
        # soft pin per chart with the subset's weight max |W| (the islands'
        # own pin weight under Batch conventions)
        plan = IslandPlan(None, batch=True, dev=self.dev, no_coarse=True, concat=cat)   # its own CG is not used
        voff = plan.vert_off
        plan.set_pins(voff[:-1], 0.0)
        # the reference holds the layout in float32 (hier.py:400); the island
        # order is gathered on the device
        vord = torch.as_tensor(vorder, device=self.dev)
        U = torch.as_tensor(U0, device=self.dev).index_select(0, vord).float().double()
        plan.orient_positive(U)
        plan.repair_jacobi(U, repair_rounds)
        Un = torch.empty_like(U)
        for it in range(int(iters)):
            if cg_iters > 0:                                   # exactly cg_iters iterations; none: the layout stays
                b = plan.local_rhs(U, _lib.UVA_MODE_BATCH)
                Un.copy_(U)
                plan.pcg(b, Un, int(cg_iters), check_every=1, tol2=0.0, strict=True, want_iters=False)
                # non-finite charts keep their layout (hier.py:500-503); folded
                # charts take the whole step, the others 0.9 of the largest
                # fold-free one over usable triangles (hier.py:504-509)
                plan.step_apply(U, Un, mode=_lib.UVA_MODE_HIER, line_search=True)
            if it == iters // 2:
                plan.repair_jacobi(U, repair_rounds)
        plan.repair_jacobi(U, repair_rounds)
        U.copy_(U.float().double())                          # the reference returns float32 values
        q = self._quality(plan, cat[2], U, len(keep))
        out = torch.empty_like(U).index_copy_(0, vord, U).cpu().numpy()
        flips[keep] = q["flips"]
        scale[keep] = q["scale"]
        self._q = (q, keep, nchart)
        return out, flips, scale

    def _quality(self, plan, Pall, U, nisl):
        """hier.py:511-530 per face and per chart, float64 on the plan's GPU."""
        d = self.dev
        P = torch.as_tensor(Pall, device=d)
        nf = torch.as_tensor(np.diff(plan.tri_off), device=d)
        T = torch.as_tensor(plan.T_local, device=d) + torch.repeat_interleave(
            torch.as_tensor(plan.vert_off[:-1], device=d), nf)[:, None]
        p0, p1, p2 = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]
        e1, e2 = p1 - p0, p2 - p0
        l1 = e1.norm(dim=1)
        nrm = torch.cross(e1, e2, dim=1)
        dbl = nrm.norm(dim=1)
        ok = (dbl > 1e-14 * (l1 * l1).clamp_min(1e-300)) & torch.isfinite(dbl)
        ex = e1 / l1.clamp_min(1e-30)[:, None]
        ey = torch.cross(nrm / dbl.clamp_min(1e-30)[:, None], ex, dim=1)
        x1 = torch.where(ok, l1, torch.zeros_like(l1)).clamp_min(1e-30)
        x2 = torch.where(ok, (e2 * ex).sum(1), torch.zeros_like(l1))
        y2 = torch.where(ok, (e2 * ey).sum(1), torch.zeros_like(l1))
        y2 = torch.where(y2.abs() > 1e-30, y2, torch.full_like(y2, 1e-30))
        a0, a1, a2 = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
        j00 = (a1[:, 0] - a0[:, 0]) / x1
        j10 = (a1[:, 1] - a0[:, 1]) / x1
        j01 = ((a2[:, 0] - a0[:, 0]) - j00 * x2) / y2
        j11 = ((a2[:, 1] - a0[:, 1]) - j10 * x2) / y2
        E, Fq, G, H = (j00 + j11) / 2, (j00 - j11) / 2, (j10 + j01) / 2, (j10 - j01) / 2
        Q, R = (E * E + H * H).sqrt(), (Fq * Fq + G * G).sqrt()
        s1, s2 = Q + R, (Q - R).abs()
        ar3 = 0.5 * dbl * ok
        sgn = 0.5 * ((a1[:, 0] - a0[:, 0]) * (a2[:, 1] - a0[:, 1]) - (a2[:, 0] - a0[:, 0]) * (a1[:, 1] - a0[:, 1]))
        ct = torch.repeat_interleave(torch.arange(nisl, device=d), nf)
        uva = torch.zeros(nisl, dtype=torch.float64, device=d).index_add_(0, ct, sgn.abs() * ok)
        a3 = torch.zeros(nisl, dtype=torch.float64, device=d).index_add_(0, ct, ar3)
        scale = uva / a3.clamp_min(1e-30)
        flips = torch.zeros(nisl, dtype=torch.float64, device=d).index_add_(0, ct, ((sgn < 0) & ok).double())
        return {"flips": flips.cpu().numpy(), "scale": scale.cpu().numpy(), "aniso": s1 / s2.clamp_min(1e-12),
                "arat": (s1 * s2) / scale[ct].clamp_min(1e-30), "ar3": ar3, "ct": ct, "a3": a3}

    def bad_fraction(self, max_aniso, max_area):
        """Per chart of the last ``relax``: area fraction of triangles with
        anisotropy above ``max_aniso`` and with area ratio outside
        [1/max_area, max_area] (hier.py:584-590)."""
        if self._q is None:
            raise ValueError("bad_fraction needs a relax call first")
        q, keep, nchart = self._q
        n = len(keep)
        d = q["ar3"].device
        badA = torch.zeros(n, dtype=torch.float64, device=d).index_add_(0, q["ct"], q["ar3"] * (q["aniso"] > max_aniso))
        badR = torch.zeros(n, dtype=torch.float64, device=d).index_add_(
            0, q["ct"], q["ar3"] * ((q["arat"] > max_area) | (q["arat"] < 1.0 / max_area)))
        a3 = q["a3"].clamp_min(1e-30)
        outA, outR = np.zeros(nchart), np.zeros(nchart)
        outA[keep] = (badA / a3).cpu().numpy()
        outR[keep] = (badR / a3).cpu().numpy()
        return outA, outR

    def lscm(self, faces, CV, nv, chart, nchart, U0, ba, bb, iters=600):
        """Free-boundary conformal map of every chart of the subset at once
        (hier.py:534-582): ba, bb the directed boundary edges (local ids), U0
        (nv,2) the start; the pins of a chart are its boundary vertices of
        smallest and largest start u (on a tie the largest vertex id).
        Returns (nv,2) float64 holding float32 values, as the reference."""
        import ctypes as C

        faces = np.asarray(faces, np.int64).reshape(-1)
        CV = np.asarray(CV, np.int64).reshape(-1, 3)
        chart = np.asarray(chart, np.int64).reshape(-1)
        nv, nchart = int(nv), int(nchart)
        U0f = np.asarray(U0, np.float64).reshape(nv, 2).astype(np.float32).astype(np.float64)
        bi = np.asarray(ba, np.int64).reshape(-1)
        bj = np.asarray(bb, np.int64).reshape(-1)
        if len(bi) != len(bj):
            raise ValueError("ba and bb need one entry per boundary edge")
        for name, a, hi in (("CV", CV, nv), ("ba", bi, nv), ("bb", bj, nv), ("chart", chart, nchart),
                            ("faces", faces, len(self.F))):
            if len(a) and (a.min() < 0 or a.max() >= hi):
                raise ValueError(f"{name} holds ids outside [0, {hi})")
        if nv == 0 or len(faces) == 0:
            return U0f
        w = half_cots(self.V, self.F[faces])
        I = np.concatenate([CV[:, 1], CV[:, 2], CV[:, 0]])
        J = np.concatenate([CV[:, 2], CV[:, 0], CV[:, 1]])
        W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
        h = np.full(len(bi), 0.5)
        # Dirichlet on u and on v, minus the area term on the boundary
        rows = [I, J, I, J, I + nv, J + nv, I + nv, J + nv, bi, bj, bi + nv, bj + nv]
        cols = [J, I, I, J, J + nv, I + nv, I + nv, J + nv, bj + nv, bi + nv, bj, bi]
        vals = [-W, -W, W, W, -W, -W, W, W, -h, h, h, -h]
        # two soft pins per chart with a boundary
        cv = np.zeros(nv, np.int64)
        cv[CV.reshape(-1)] = np.repeat(chart, 3)
        bv = np.unique(bi)
        key = U0f[bv, 0]
        mn = np.full(nchart, 1e30)
        mx = np.full(nchart, -1e30)
        np.minimum.at(mn, cv[bv], key)
        np.maximum.at(mx, cv[bv], key)
        p0 = np.full(nchart, -1, np.int64)
        p1 = np.full(nchart, -1, np.int64)
        at_mn, at_mx = key == mn[cv[bv]], key == mx[cv[bv]]
        np.maximum.at(p0, cv[bv][at_mn], bv[at_mn])
        np.maximum.at(p1, cv[bv][at_mx], bv[at_mx])
        pins = np.concatenate([p0[p0 >= 0], p1[p1 >= 0]])
        wpin = max(float(np.abs(W).max()) if len(W) else 0.0, 1e-12) * 10
        rows += [pins, pins + nv]
        cols += [pins, pins + nv]
        vals += [np.full(2 * len(pins), wpin)]
        H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(2 * nv, 2 * nv))
        H.sum_duplicates()
        b = np.zeros(2 * nv)
        b[pins] += wpin * U0f[pins, 0]                 # no accumulation over a repeated pin, as the reference
        b[pins + nv] += wpin * U0f[pins, 1]
        ip = np.ascontiguousarray(H.indptr, np.int64)
        ix = np.ascontiguousarray(H.indices, np.int64)
        data = np.ascontiguousarray(H.data, np.float64)
        bd = torch.as_tensor(b, device=self.dev)
        x = torch.as_tensor(np.concatenate([U0f[:, 0], U0f[:, 1]]), device=self.dev)
        its = C.c_int32(0)
        with torch.cuda.device(self.dev):
            _lib.check(_lib.lib().uva_jacobi_pcg(2 * nv, _lib.i64p(ip), _lib.i64p(ix), _lib.f64p(data), bd.data_ptr(),
                                                 x.data_ptr(), int(iters), 50, 1e-12, C.byref(its),
                                                 stream_of(self.dev)), "uva_jacobi_pcg")
        self.lscm_iters = its.value
        xs = x.cpu().numpy()
        U = np.stack([xs[:nv], xs[nv:]], 1)
        U = np.where(np.isfinite(U), U, U0f)
        return U.astype(np.float32).astype(np.float64)
