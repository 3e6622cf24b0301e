"""The small torch helpers of the gpu module (csr32, spmm, unique_edges,
sym_csr, device; reference gpu.py:16-63, imported by hier.py:23-25), run on
CPU torch and checked against scipy and brute force."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from ultimate_uv_artist_amd import gpu, workloads


@pytest.fixture(scope="module")
def mesh():
    P, T = workloads.disk_island(np.random.default_rng(11), 300)
    return P, T


def test_unique_edges_brute_force(mesh):
    P, T = mesh
    n = len(P)
    ea, eb, inv, cnt = gpu.unique_edges(torch.as_tensor(T), n)
    ea, eb, inv, cnt = (x.numpy() for x in (ea, eb, inv, cnt))
    # the 3t triangle edges (k+1, k+2) in the reference's order: column 1-2, 2-0, 0-1
    a = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    b = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    pairs = sorted({(min(x, y), max(x, y)) for x, y in zip(a.tolist(), b.tolist())})
    assert list(zip(ea.tolist(), eb.tolist())) == pairs          # ascending, a < b
    assert (ea[inv] == np.minimum(a, b)).all() and (eb[inv] == np.maximum(a, b)).all()
    assert (cnt == np.bincount(inv, minlength=len(pairs))).all()
    # a disk: boundary edges are used once, interior edges twice
    from ultimate_uv_artist_amd import param

    nb = len(param.boundary_edges_oriented(T)[0])
    assert int((cnt == 1).sum()) == nb and set(cnt.tolist()) == {1, 2}


def test_sym_csr_and_csr32_and_spmm(mesh):
    P, T = mesh
    n = len(P)
    Tt = torch.as_tensor(T)
    ea, eb, _, _ = gpu.unique_edges(Tt, n)
    rng = np.random.default_rng(3)
    off = torch.as_tensor(rng.normal(size=len(ea)))
    diag = torch.as_tensor(rng.uniform(1, 2, n))
    L = gpu.sym_csr(n, ea, eb, off, diag)
    assert L.layout == torch.sparse_csr and L.crow_indices().dtype == torch.int32
    assert L.col_indices().dtype == torch.int32
    ref = sp.coo_matrix((np.r_[off.numpy(), off.numpy(), diag.numpy()],
                         (np.r_[ea.numpy(), eb.numpy(), np.arange(n)], np.r_[eb.numpy(), ea.numpy(), np.arange(n)])),
                        shape=(n, n)).tocsr()
    ref.sort_indices()
    assert np.array_equal(L.crow_indices().numpy(), ref.indptr)
    assert np.array_equal(L.col_indices().numpy(), ref.indices)
    assert np.array_equal(L.values().numpy(), ref.data)
    X = torch.as_tensor(rng.normal(size=(n, 2)))
    Y = gpu.spmm(L, X)
    assert Y.shape == (n, 2)
    assert np.allclose(Y.numpy(), ref @ X.numpy(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gpu.spmm(L, X[:, 0]).numpy(), ref @ X[:, 0].numpy(), rtol=1e-12, atol=1e-12)
    # csr32: int32 indices of any sparse input, same matrix
    C64 = torch.sparse_csr_tensor(L.crow_indices().long(), L.col_indices().long(), L.values(), (n, n))
    for A in (C64, C64.to_sparse_coo()):
        B = gpu.csr32(A)
        assert B.crow_indices().dtype == torch.int32 and B.col_indices().dtype == torch.int32
        assert np.array_equal(B.to_dense().numpy(), ref.toarray())


def test_device_is_the_gpu_or_fails_loudly():
    if torch.cuda.is_available():
        assert gpu.device().type == "cuda"
    else:
        with pytest.raises(RuntimeError):
            gpu.device()
