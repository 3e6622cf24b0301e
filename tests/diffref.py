"""Float64 torch autograd references of the weighted distortion energies and
of the refine global step (CPU, dense solve; small islands only).

Test infrastructure: the energies restate param.py:294-318 and the step
restates param.py:391-401 (soft pin) with the rhs of param.py:457-474, as
plain differentiable torch expressions, so torch's autograd supplies
reference gradients for the hand-written backward kernels."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from oracle import param_cpu as oracle


def jac(U, T, X):
    Tt = torch.as_tensor(T)
    x1, x2, y2 = (torch.as_tensor(X[:, i, j]) for i, j in ((1, 0), (2, 0), (2, 1)))
    d1 = U[Tt[:, 1]] - U[Tt[:, 0]]
    d2 = U[Tt[:, 2]] - U[Tt[:, 0]]
    j00 = d1[:, 0] / x1
    j10 = d1[:, 1] / x1
    return j00, (d2[:, 0] - j00 * x2) / y2, j10, (d2[:, 1] - j10 * x2) / y2


def energy(U, T, X, area, w, kind):
    """sum_f w_f e_f(U) (kind 0 ARAP, 1 symmetric Dirichlet)."""
    j00, j01, j10, j11 = jac(U, T, X)
    a = torch.as_tensor(area)
    if kind == 0:
        q = torch.sqrt(((j00 + j11) / 2) ** 2 + ((j10 - j01) / 2) ** 2)
        r = torch.sqrt(((j00 - j11) / 2) ** 2 + ((j10 + j01) / 2) ** 2)
        ef = a * ((q + r - 1) ** 2 + (q - r - 1) ** 2)
    else:
        det = j00 * j11 - j01 * j10
        fro = j00 ** 2 + j01 ** 2 + j10 ** 2 + j11 ** 2
        ef = a * (fro + fro / det ** 2 - 4)
    return (ef if w is None else w * ef).sum()


class Step:
    """U -> L^-1 b(U) of param.refine (soft pin at vertex 0, weight max|W|)."""

    def __init__(self, P, T):
        self.T = np.asarray(T, np.int64)
        self.X, self.area = oracle.local_frames(P, T)
        self.cots = oracle.cot_weights(self.X)
        n = len(P)
        I, J, W = oracle.lap_triplets(self.T, self.cots)
        self.wpin = float(np.abs(W).max())
        L = sp.csr_matrix((np.r_[W, self.wpin], (np.r_[I, 0], np.r_[J, 0])), shape=(n, n)).toarray()
        self.L = torch.as_tensor(L)
        self.n = n

    def __call__(self, U):
        j00, j01, j10, j11 = jac(U, self.T, self.X)
        ca, cb = j00 + j11, j10 - j01
        h = torch.sqrt(ca ** 2 + cb ** 2)
        c, s = ca / h, cb / h
        Tt = torch.as_tensor(self.T)
        Xt = torch.as_tensor(self.X)
        ct = torch.as_tensor(self.cots)
        b = torch.zeros(self.n, 2, dtype=torch.float64)
        for k in range(3):
            i, j = Tt[:, (k + 1) % 3], Tt[:, (k + 2) % 3]
            dX = Xt[:, (k + 1) % 3, :] - Xt[:, (k + 2) % 3, :]
            r = torch.stack([c * dX[:, 0] - s * dX[:, 1], s * dX[:, 0] + c * dX[:, 1]], 1) * (0.5 * ct[:, k])[:, None]
            b = b.index_add(0, i, r).index_add(0, j, -r)
        b = b.clone()
        b[0] = b[0] + self.wpin * U[0]
        return torch.linalg.solve(self.L, b)


def mesh_direction(U, T, rng):
    """Random direction scaled so that its largest vertex move equals the
    median edge length (finite differences with h = 1e-3 then perturb each
    triangle by about 0.1 % of its size)."""
    e = np.linalg.norm(U[T[:, 1]] - U[T[:, 0]], axis=1)
    D = rng.normal(size=U.shape)
    return D * (np.median(e) / np.abs(D).max())
