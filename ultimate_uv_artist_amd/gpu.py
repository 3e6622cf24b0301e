"""Drop-in replacement of ``engine/gpu.py``: batched parameterization of many
charts on one MI355X.

``Batch`` keeps the reference's constructor and methods (``orient_positive``,
``repair_flips``, ``arap``, ``chart_uvs`` ... , gpu.py:66-299).  The per-chart
work (orientation, repair, rotations and rhs, step limits) is one persistent
HIP kernel over all charts, islands pulled from a device work queue.  The
global step is the reference's one CG over the whole batch (one step length
per column, shared by every chart): grid-wide Jacobi-PCG kernels on the chart
plan's own sliced-ELL Laplacian with the batch's per-chart soft pins
(``uva_plan_pcg``), with the stop test on the device instead of a host sync
per CG iteration; the rest of each ARAP iteration (non-finite revert,
flip-safe step, update) is one more kernel (``uva_step_apply``).  The layout
is kept in FP64 on the device; ``Batch.U`` returns it in the requested dtype.

The small torch helpers (``csr32``, ``spmm``, ``device``, ``unique_edges``,
``sym_csr``) keep their reference behaviour because ``hier.py`` imports them.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .plan import IslandPlan, batch_frames_dev, pinned_laplacian, require_gpu

BATCH_CG_TOL = 1e-6   # gpu.py:184 Batch._cg(tol=1e-6)


def csr32(A):
    """CSR with int32 indices (gpu.py:16-20)."""
    A = A.to_sparse_csr() if A.layout != torch.sparse_csr else A
    return torch.sparse_csr_tensor(A.crow_indices().to(torch.int32), A.col_indices().to(torch.int32), A.values(),
                                   A.shape)


def spmm(L, X):
    """L @ X one column at a time (gpu.py:23-28)."""
    if X.dim() == 1:
        return L @ X
    return torch.stack([L @ X[:, k].contiguous() for k in range(X.shape[1])], 1)


def device():
    """The GPU this package runs on (gpu.py:30-36); the product path needs one."""
    try:
        import comfy.model_management as mm  # noqa: F401  (host application, when present)
        d = mm.get_torch_device()
        if d.type == "cuda":
            return d
    except Exception:
        pass
    return require_gpu()


def unique_edges(T, n):
    """Undirected edges (a < b), inverse map of the 3t triangle edges and use
    counts (gpu.py:39-49)."""
    a = torch.cat([T[:, 1], T[:, 2], T[:, 0]])
    b = torch.cat([T[:, 2], T[:, 0], T[:, 1]])
    key = torch.minimum(a, b) * n + torch.maximum(a, b)
    uk, inv = torch.unique(key, return_inverse=True)
    cnt = torch.bincount(inv, minlength=len(uk))
    return uk // n, uk % n, inv, cnt


def sym_csr(n, ea, eb, off, diag):
    """Symmetric CSR (int32 indices) from unique edges and a diagonal (gpu.py:52-63)."""
    dev = ea.device
    ar = torch.arange(n, device=dev)
    rows = torch.cat([ea, eb, ar])
    cols = torch.cat([eb, ea, ar])
    vals = torch.cat([off, off, diag])
    order = torch.argsort(rows * n + cols)
    rows, cols, vals = rows[order], cols[order], vals[order]
    crow = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    crow[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
    return torch.sparse_csr_tensor(crow.to(torch.int32), cols.to(torch.int32), vals, (n, n))


class Batch:
    """charts: list of (P (k,3), T (t,3), U (k,2)) (gpu.py:66-124)."""

    def __init__(self, charts, dev=None, dtype=torch.float32):
        self.dev = torch.device(dev) if dev is not None else device()
        self.dt = dtype
        charts = list(charts)
        self.nc = len(charts)
        self.offsets = np.cumsum([0] + [len(p) for p, _, _ in charts])
        self.n = int(self.offsets[-1])
        # the per-chart work, and the global step: one Jacobi-PCG over every
        # row of the plan with a soft pin at every chart's first vertex of the
        # batch-global weight max |W| (the islands' own pin weight under Batch
        # conventions; gpu.py:133-141); the per-island CG and its coarse level
        # are not used
        self._plan = IslandPlan([(p, t) for p, t, _ in charts], batch=True, dev=self.dev, no_coarse=True)
        self._P = np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p, _, _ in charts])
        self.last_cg_iters = 0
        U = np.concatenate([np.asarray(u, np.float64).reshape(-1, 2) for _, _, u in charts])
        self._U = torch.as_tensor(U, device=self.dev).contiguous()
        toff = np.cumsum([0] + [len(t) for _, t, _ in charts])
        self._Tg = self._plan.T_local + np.repeat(self.offsets[:-1], np.diff(toff))[:, None]
        self.T = torch.as_tensor(self._Tg, device=self.dev)
        self.cv = torch.as_tensor(np.repeat(np.arange(self.nc), np.diff(self.offsets)), device=self.dev)
        self.ct = torch.as_tensor(np.repeat(np.arange(self.nc), np.diff(toff)), device=self.dev)
        self.pins = torch.as_tensor(self.offsets[:-1], device=self.dev)
        self._scale = torch.ones(len(self.T), device=self.dev, dtype=torch.float64)
        self._frames = None
        self._A = None
        self._plan.set_pins(self.offsets[:-1], 0.0)

    # ---------------------------------------------------------------- layout
    # The layout lives in FP64 on the device and is updated in place by the
    # kernels; the tensors handed out and taken in are copies, so a caller's
    # tensor never changes behind its back (the reference rebinds self.U to
    # new tensors, gpu.py:230-249).
    @property
    def U(self):
        return self._U.clone() if self.dt == torch.float64 else self._U.to(self.dt)

    @U.setter
    def U(self, value):
        self._U = value.to(self.dev, torch.float64).reshape(self.n, 2).clone().contiguous()

    # ------------------------------------------------ per-triangle attributes
    # area, ok, w, x1, x2, y2, X as the reference builds them in
    # Batch.__init__ (gpu.py:88-121): float64 frames cast to the batch dtype,
    # computed on the device when first read.
    def _frame(self, key):
        if self._frames is None:
            V = torch.as_tensor(self._P, device=self.dev)
            self._frames = batch_frames_dev(V, self.T)
        v = self._frames[key]
        return v if v.dtype == torch.bool else v.to(self.dt)

    area = property(lambda self: self._frame("area"))
    ok = property(lambda self: self._frame("ok"))
    w = property(lambda self: self._frame("w"))
    x1 = property(lambda self: self._frame("x1"))
    x2 = property(lambda self: self._frame("x2"))
    y2 = property(lambda self: self._frame("y2"))
    X = property(lambda self: self._frame("X"))

    # ------------------------------------------------ the pinned Laplacian
    @property
    def L(self):
        """The batch Laplacian with every chart's soft pin (gpu.py:126-142) as
        a torch CSR tensor with int32 indices in the batch dtype."""
        self._system()
        A = self._A
        return torch.sparse_csr_tensor(torch.as_tensor(A.indptr, dtype=torch.int32, device=self.dev),
                                       torch.as_tensor(A.indices, dtype=torch.int32, device=self.dev),
                                       torch.as_tensor(A.data, device=self.dev).to(self.dt), A.shape)

    @property
    def dinv(self):
        """Jacobi weights 1 / max(diag L, 1e-30) (gpu.py:142)."""
        self._system()
        d = torch.as_tensor(self._A.diagonal(), device=self.dev)
        return (1.0 / d.clamp_min(1e-30)).to(self.dt)

    @property
    def wpin(self):
        """The batch-global soft-pin weight max |W| (gpu.py:137-139), a 0-d tensor."""
        self._system()
        return torch.tensor(self._wpin, device=self.dev, dtype=self.dt)

    @property
    def scale(self):
        return self._scale.to(self.dt)

    @scale.setter
    def scale(self, value):
        self._scale = value.to(self.dev, torch.float64).contiguous()
        self._plan.set_scale(self._scale)

    # ---------------------------------------------------------------- pieces
    def jacobians(self, U):
        """(j00, j01, j10, j11) per triangle of the map from the frames to U
        (gpu.py:145-151), in U's dtype."""
        T = self.T
        x1, x2, y2 = (self._frame(k).to(U.dtype) for k in ("x1", "x2", "y2"))
        u0 = U[T[:, 0]]
        d1, d2 = U[T[:, 1]] - u0, U[T[:, 2]] - u0
        j00, j10 = d1[:, 0] / x1, d1[:, 1] / x1
        return j00, (d2[:, 0] - j00 * x2) / y2, j10, (d2[:, 1] - j10 * x2) / y2

    def signed_area(self, U):
        """Signed UV area per triangle (gpu.py:153-156)."""
        T = self.T
        a, b, c = U[T[:, 0]], U[T[:, 1]], U[T[:, 2]]
        return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))

    def flips_per_chart(self, U=None):
        """Folded thick triangles per chart (gpu.py:158-163)."""
        U = self._U if U is None else U.to(self.dev, torch.float64).contiguous()
        return self._plan.flips_per_chart(U).to(self.dt)

    # ---------------------------------------------------------------- solvers
    def _system(self):
        """The batch Laplacian (block-diagonal over the charts) with a soft pin
        at every chart's first vertex of the batch-global weight max |W|
        (gpu.py:126-142), assembled on the GPU for the attributes L, dinv and
        wpin only: the solves run on the chart plan's own matrix."""
        if self._A is None:
            self._A, self._wpin = pinned_laplacian(self._P, self._Tg, self.offsets[:-1], dev=self.dev)
        return self._A

    def arap(self, iters=8, line_search=True, cg_iters=40, cg_tol=BATCH_CG_TOL):
        """ARAP iterations of every chart (gpu.py:230-249): rotations x scale
        and rhs, one Jacobi-PCG over all charts together from the current
        layout, stopped when both columns reach relative residual ``cg_tol``
        or after ``cg_iters`` iterations (gpu.py:184-202), revert of
        non-finite charts, per-chart flip-safe step when ``line_search``."""
        U = self._U.clone()
        Un = torch.empty_like(U)
        n = int(iters)
        for k in range(n):
            if cg_iters <= 0:                                  # no CG iteration: the layout stays
                break
            b = self._plan.local_rhs(U, _lib.UVA_MODE_BATCH)
            Un.copy_(U)
            # the CG's iteration count is read back once, after the last call
            its = self._plan.pcg(b, Un, int(cg_iters), check_every=1, tol2=float(cg_tol) ** 2, strict=False,
                                 want_iters=k == n - 1)
            if its is not None:
                self.last_cg_iters = its
            # non-finite charts keep their layout, then the per-chart
            # flip-safe step (or the whole step without line search)
            self._plan.step_apply(U, Un, mode=_lib.UVA_MODE_BATCH, line_search=line_search)
        self._U = U
        return self.U

    def repair_flips(self, rounds=60):
        """Jacobi 1-ring relaxation of the vertices of folded triangles (gpu.py:251-273)."""
        U = self._U.clone()
        self._plan.repair_jacobi(U, rounds)
        self._U = U
        return self.U

    def orient_positive(self):
        """Mirror charts whose area-weighted orientation is negative (gpu.py:275-281)."""
        U = self._U.clone()
        self._plan.orient_positive(U)
        self._U = U

    def shrink(self, tri_idx, factor=0.6, floor=0.05, rings=3):
        """Lower the target size of the given triangles and their ``rings``
        neighbourhood (gpu.py:283-295)."""
        if len(tri_idx) == 0:
            return
        T = self.T
        mark = torch.zeros(self.n, dtype=torch.bool, device=self.dev)
        sel = torch.zeros(T.shape[0], dtype=torch.bool, device=self.dev)
        sel[torch.as_tensor(tri_idx, device=self.dev)] = True
        for _ in range(rings):
            mark[T[sel].reshape(-1)] = True
            sel = mark[T].any(1)
        self.scale = torch.where(sel, (self._scale * factor).clamp_min(floor), self._scale)

    def chart_uvs(self):
        """Per-chart float64 numpy views of one host copy (gpu.py:297-299)."""
        U = self._U.cpu().numpy()
        return [U[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def stats(self):
        """Per-chart statistics of the chart plan's last call (orientation,
        repair, rhs, step limits); the coupled CG of ``arap`` reports its
        iteration count in ``last_cg_iters``."""
        return self._plan.stats()
