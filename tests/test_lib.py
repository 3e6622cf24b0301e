"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, and its host-only calls (argument validation, sequential flip repair)
behave like the reference without touching a GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib():
    from ultimate_uv_artist_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail("libuvart.so is not built: run __graft_entry__.build()")
    return _lib


def test_exports_every_header_symbol(lib):
    names = lib.header_symbols()
    assert len(names) >= 18
    L = lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert set(names) == set(lib.SIGNATURES), "ctypes signatures and the header disagree"
    assert L.uva_version() >= 10000


def test_invalid_arguments_are_errors_not_crashes(lib):
    L = lib.lib()
    h = C.c_void_p()
    off = np.array([0, 3], np.int64)
    toff = np.array([0, 1], np.int64)
    P = np.zeros((3, 3))
    T = np.array([[0, 1, 5]], np.int64)   # vertex outside the island
    rc = L.uva_plan_create(C.byref(h), 1, lib.i64p(off), lib.i64p(toff), lib.f64p(P), lib.i64p(T), None, None, None,
                           None, 0, None)
    assert rc == lib.UVA_EINVAL
    assert b"outside" in L.uva_last_error()
    with pytest.raises(ValueError):
        lib.check(rc)
    assert L.uva_arap(None, None, None, None) == lib.UVA_EINVAL
    assert L.uva_repair_seq(None, 0, None, 0, 1) == lib.UVA_EINVAL


def test_repair_seq_matches_oracle(lib):
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(7)
    P, T = workloads.disk_island(rng, 400)
    U = oracle.lscm(P, T)
    U = U + rng.normal(0, 0.03 * np.ptp(U), U.shape)
    assert (oracle.signed_areas(U, T) <= 0).any()
    out = np.ascontiguousarray(U.copy())
    Tc = np.ascontiguousarray(T, np.int64)
    assert lib.lib().uva_repair_seq(lib.f64p(out), len(out), lib.i64p(Tc), len(Tc), 25) == 0
    assert np.array_equal(out, oracle.repair_flips(U, T, 25))


def test_ctypes_structs_match_the_library(lib):
    """The ctypes structs have the C structs' sizes (uva_struct_size), and the
    stand-alone binding in INTEGRATION.md declares uva_arap_opts with all of
    its fields (a short struct makes the library read past the caller's)."""
    import re
    from pathlib import Path

    L = lib.lib()
    for which, st in enumerate((lib.ArapOpts, lib.SlimOpts, lib.IslandStats, lib.PlanInfo)):
        assert C.sizeof(st) == L.uva_struct_size(which), st.__name__
    assert L.uva_struct_size(99) == -1
    doc = (Path(__file__).resolve().parent.parent / "INTEGRATION.md").read_text()
    block = re.search(r"class ArapOpts\(C\.Structure\):.*?\]\n", doc, re.S).group(0)
    ns = {"C": C}
    exec(block, ns)
    assert C.sizeof(ns["ArapOpts"]) == L.uva_struct_size(0)
    assert [f for f, _ in ns["ArapOpts"]._fields_] == [f for f, _ in lib.ArapOpts._fields_]


def _regroup_reference(F, V, faces, CV, nv, chart, nchart):
    """The regrouping as in-order numpy scatters and stable sorts: per chart
    its vertices in ascending local id, its faces in subset order."""
    cv = np.full(nv, -1, np.int64)
    cv[CV.reshape(-1)] = np.repeat(chart, 3)
    corner = np.empty(nv, np.int64)
    corner[CV.reshape(-1)] = F[faces].reshape(-1)
    vorder = np.argsort(cv, kind="stable")
    vcount = np.bincount(cv, minlength=nchart)
    voff = np.concatenate([[0], np.cumsum(vcount)])
    local = np.empty(nv, np.int64)
    local[vorder] = np.arange(nv) - np.repeat(voff[:-1], vcount)
    forder = np.argsort(chart, kind="stable")
    fcount = np.bincount(chart, minlength=nchart)
    keep = np.flatnonzero(fcount > 0)
    vo = np.concatenate([[0], np.cumsum(vcount[keep])])
    to = np.concatenate([[0], np.cumsum(fcount[keep])])
    return vo, to, V[corner[vorder]], local[CV[forder]], keep, vorder, forder


def _regroup(lib, F, V, faces, CV, nv, chart, nchart):
    from ultimate_uv_artist_amd import _lib as L

    n = len(faces)
    keep = np.empty(max(nchart, 1), np.int64)
    nk = np.zeros(1, np.int64)
    vo = np.empty(max(nchart, 1) + 1, np.int64)
    to = np.empty(max(nchart, 1) + 1, np.int64)
    P = np.empty((nv, 3))
    T = np.empty((n, 3), np.int64)
    vorder = np.empty(nv, np.int64)
    forder = np.empty(n, np.int64)
    rc = lib.lib().uva_chart_regroup(n, L.i64p(faces), L.i64p(CV), nv, L.i64p(chart), nchart, L.i64p(F), len(F),
                                     L.f64p(V), len(V), L.i64p(keep), L.i64p(nk), L.i64p(vo), L.i64p(to),
                                     L.f64p(P), L.i64p(T), L.i64p(vorder), L.i64p(forder))
    k = int(nk[0])
    return rc, (vo[:k + 1], to[:k + 1], P, T, keep[:k], vorder, forder)


def test_chart_regroup_matches_the_definition(lib):
    """uva_chart_regroup (hier.Solver.relax's subset as islands) against the
    same regrouping written as numpy scatters and stable sorts: a mesh of
    several disks, a subset of their faces in shuffled order with local
    vertex ids in shuffled order, chart ids that skip some charts; bitwise."""
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(17)
    charts = [workloads.disk_island(rng, int(t)) for t in (300, 90, 1200, 40, 700)]
    vo = np.concatenate([[0], np.cumsum([len(P) for P, _ in charts])])
    V = np.concatenate([P for P, _ in charts])
    F = np.concatenate([T + vo[i] for i, (_, T) in enumerate(charts)])
    fc = np.concatenate([np.full(len(T), i) for i, (_, T) in enumerate(charts)])
    sub = np.flatnonzero(fc != 3)                        # chart 3 left out of the subset
    sub = sub[rng.permutation(len(sub))]
    used = np.unique(F[sub])
    lid = np.full(len(V), -1, np.int64)
    lid[used[rng.permutation(len(used))]] = np.arange(len(used))   # local ids in shuffled order
    CV = np.ascontiguousarray(lid[F[sub]])
    chart = np.ascontiguousarray(np.array([0, 2, 5, 9, 4])[fc[sub]], np.int64)   # 10 chart ids, 5 used
    faces = np.ascontiguousarray(sub, np.int64)
    rc, got = _regroup(lib, F, V, faces, CV, len(used), chart, 10)
    assert rc == 0
    ref = _regroup_reference(F, V, faces, CV, len(used), chart, 10)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_chart_regroup_errors(lib):
    """Out-of-range ids and a local vertex on no face are errors."""
    F = np.array([[0, 1, 2], [1, 3, 2]], np.int64)
    V = np.zeros((4, 3))
    faces = np.array([0, 1], np.int64)
    chart = np.array([0, 0], np.int64)
    CV = np.array([[0, 1, 2], [1, 3, 2]], np.int64)
    assert _regroup(lib, F, V, faces, CV, 4, chart, 1)[0] == 0
    assert _regroup(lib, F, V, faces, CV, 5, chart, 1)[0] != 0          # vertex 4 on no face
    assert _regroup(lib, F, V, faces, CV, 3, chart, 1)[0] != 0          # id 3 outside [0, 3)
    assert _regroup(lib, F, V, faces, CV, 4, np.array([0, 1], np.int64), 1)[0] != 0
    assert _regroup(lib, F, V, np.array([0, 2], np.int64), CV, 4, chart, 1)[0] != 0
