"""CPU checks of the batched host tail of charts.flatten_many: the boundary
crossing count and the chart quality of many charts at once equal the
per-chart functions (boundary_intersections, param.chart_quality) exactly."""
from __future__ import annotations

import numpy as np


def _cases():
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(23)
    charts, Us = [], []
    for k, t in enumerate(rng.integers(30, 400, 60)):
        P, T = workloads.disk_island(rng, int(t))
        U = oracle.lscm(P, T)
        if k % 3 == 1:     # crossings: the boundary pushed inward and around
            U = U + rng.normal(scale=0.3 * np.ptp(U, 0).max(), size=U.shape) * (rng.random(len(U)) < 0.1)[:, None]
        if k % 7 == 2:     # folds
            U[: len(U) // 3] *= -1.0
        charts.append((P, T))
        Us.append(U)
    return charts, Us


def test_boundary_intersections_many_matches_per_chart():
    from ultimate_uv_artist_amd import charts as ch

    charts, Us = _cases()
    Ts = [T for _, T in charts]
    got = ch.boundary_intersections_many(Us, Ts)
    ref = [ch.boundary_intersections(U, T) for U, T in zip(Us, Ts)]
    assert got == ref
    assert sum(1 for r in ref if r) >= 5              # the case has crossings to count
    assert ch.boundary_intersections_many(Us, Ts, max_report=3) == [min(r, 3) for r in ref]
    assert ch.boundary_intersections_many([], []) == []


def test_chart_quality_many_matches_per_chart():
    from ultimate_uv_artist_amd import charts as ch
    from ultimate_uv_artist_amd import param

    charts, Us = _cases()
    got = ch.chart_quality_many(charts, Us)
    for (P, T), U, g in zip(charts, Us, got):
        r = param.chart_quality(P, T, U)
        assert g.keys() == r.keys()
        for key in r:
            if isinstance(r[key], np.ndarray):
                assert np.array_equal(g[key], r[key], equal_nan=True), key
            else:
                assert g[key] == r[key] or (np.isnan(g[key]) and np.isnan(r[key])), key
