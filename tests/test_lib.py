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
