"""uva_overlap_flags_noalloc (the overlap test without device allocation, its
boxes and cell lists in the library's per-GPU scratch array) against
uva_overlap_flags on the same inputs: flags equal bit for bit on two calls in
a row, the second smaller (the scratch keeps the first call's contents), in
the tolerant and the strict test; an input beyond the scratch is refused
with a message and leaves the flags untouched."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _atlas(rng, sizes):
    from ultimate_uv_artist_amd import workloads

    Ts, Us, off = [], [], 0
    for k, t in enumerate(sizes):
        P, T = workloads.disk_island(rng, t)
        U = (P[:, :2] - P[:, :2].min(0)) / np.ptp(P[:, :2]).max() * 0.55 + np.array([0.4 * (k % 2), 0.4 * (k // 2)])
        Ts.append(T + off)
        Us.append(U + rng.normal(0, 0.0005, U.shape))
        off += len(P)
    return np.clip(np.concatenate(Us), 0, 1), np.concatenate(Ts)


def _call(name, U, F, skip, rel, strict, grid, out):
    import torch

    from ultimate_uv_artist_amd import _lib

    L = _lib.lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return getattr(L, name)(U.data_ptr(), len(U), F.data_ptr(), len(F), None if skip is None else skip.data_ptr(),
                            float(rel), int(strict), int(grid), out.data_ptr(), st)


def test_noalloc_overlap_flags_equal_the_allocating_entry_point():
    import torch

    from ultimate_uv_artist_amd import _lib

    rng = np.random.default_rng(91)
    dev = torch.device("cuda", 0)
    for sizes in ((15000, 12000, 9000, 15000), (3000, 2500)):
        U, T = _atlas(rng, sizes)
        Ud = torch.as_tensor(U, device=dev).contiguous()
        Fd = torch.as_tensor(T, device=dev).contiguous()
        skip = torch.as_tensor((rng.random(len(T)) < 0.02).astype(np.uint8), device=dev)
        G = int(np.clip(np.sqrt(len(T)) / 2, 64, 4096))
        for rel, strict, sk in ((1e-6, 0, skip), (0.0, 1, skip), (1e-6, 0, None)):
            ref = torch.full((len(T),), 7, dtype=torch.uint8, device=dev)
            got = torch.full((len(T),), 9, dtype=torch.uint8, device=dev)
            assert _call("uva_overlap_flags", Ud, Fd, sk, rel, strict, G, ref) == 0
            rc = _call("uva_overlap_flags_noalloc", Ud, Fd, sk, rel, strict, G, got)
            assert rc == 0, _lib.lib().uva_last_error()
            torch.cuda.synchronize()
            print(f"no-alloc: {len(T)} triangles, grid {G}, strict {strict}: {int(ref.sum())} flagged", flush=True)
            assert torch.equal(got, ref)
            if strict == 0 and sk is not None:
                assert int(ref.sum()) > 50
    # beyond the scratch: refused before any work, flags untouched
    U, T = _atlas(rng, (2000,))
    Ud = torch.as_tensor(U, device=dev).contiguous()
    Fd = torch.as_tensor(T, device=dev).contiguous()
    out = torch.full((len(T),), 5, dtype=torch.uint8, device=dev)
    rc = _call("uva_overlap_flags_noalloc", Ud, Fd, None, 1e-6, 0, 4096, out)
    assert rc == _lib.UVA_EINVAL and b"scratch" in _lib.lib().uva_last_error()
    torch.cuda.synchronize()
    assert bool((out == 5).all())
