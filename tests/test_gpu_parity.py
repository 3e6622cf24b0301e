"""GPU parity: every libuvart entry point against the CPU oracle on seeded inputs.

Tolerances (written per test): UVs are compared after normalising each
island to the unit box (max |dU| <= 1e-4 for ARAP),
LSCM (FP64 CG to 1e-8) much tighter; per-island scalars relative 1e-9 where
both sides are FP64 arithmetic on identical inputs.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def unit_box_err(U, Uref):
    lo = Uref.min(0)
    s = float((Uref.max(0) - lo).max())
    return float(np.abs((U - lo) / s - (Uref - lo) / s).max())


@pytest.fixture(scope="module")
def grid():
    from ultimate_uv_artist_amd import workloads

    return workloads.grid_island(seed=1)


@pytest.fixture(scope="module")
def small_disks():
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(11)
    sizes = [200, 450, 900, 1600, 2600, 3400, 5000, 260]
    return [workloads.disk_island(rng, t) for t in sizes]


def test_lscm_matches_oracle(grid, small_disks):
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param

    for P, T in [grid] + small_disks[:4]:
        print(f"lscm: {len(T)} triangles", flush=True)
        U = param.lscm(P, T)
        Ur = oracle.lscm(P, T)
        assert unit_box_err(U, Ur) <= 1e-6


def test_lscm_given_pins_and_frames(small_disks):
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param

    P, T = small_disks[2]
    X, area = oracle.local_frames(P, T)
    cots = oracle.cot_weights(X)
    U = param.lscm(P, T, X, cots, pins=(3, 17))
    Ur = oracle.lscm(P, T, X, cots, pins=(3, 17))
    assert unit_box_err(U, Ur) <= 1e-6


def test_lscm_closed_chart_raises():
    from ultimate_uv_artist_amd import param

    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    T = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    with pytest.raises(ValueError):
        param.lscm(P, T)


def test_refine_grid_20_iters(grid):
    """The small workload: 7,938-triangle grid island, LSCM then 20 ARAP iterations."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param

    P, T = grid
    U0 = oracle.lscm(P, T)
    U, solver = param.refine(P, T, U0, iters=20, tol=0.0)
    Ur, _ = oracle.refine(P, T, U0, iters=20, tol=0.0)
    err = unit_box_err(U, Ur)
    print(f"grid 20 ARAP: unit-box err {err:.2e}", flush=True)
    assert err <= 1e-4
    q = param.chart_quality(P, T, U)
    qr = oracle.chart_quality(P, T, Ur)
    assert abs(q["aniso_p95"] / qr["aniso_p95"] - 1) <= 0.01
    assert abs(q["area_p95"] / qr["area_p95"] - 1) <= 0.01
    assert q["flips"] == qr["flips"] == 0
    # the returned solver is reusable
    U2, _ = param.refine(P, T, U0, iters=20, tol=0.0, solver=solver)
    assert np.array_equal(U, U2)


def test_refine_early_exit_and_pin(small_disks):
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib
    from ultimate_uv_artist_amd.plan import IslandPlan

    for P, T in small_disks[3:6]:
        U0 = oracle.flatten_lscm_arap(P, T, iters=0)
        trace = []
        Ur, _ = oracle.refine(P, T, U0, iters=12, tol=1e-3, pin=5, trace=trace)
        plan = IslandPlan([(P, T)], pins=[5])
        Ud = plan.device_uv(U0)
        plan.arap(Ud, 12, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=1e-3)
        st = plan.stats()
        U = Ud.cpu().numpy()
        assert st["arap_iters"][0] == len(trace)
        assert unit_box_err(U, Ur) <= 1e-4
        # the soft pin holds to the accuracy of the global solves (exactly only for a direct solve)
        assert np.abs(U[5] - U0[5]).max() <= 1e-5 * np.ptp(U0)
        assert abs(st["energy"][0] / trace[-1][0] - 1) <= 1e-6


def test_refine_folded_start_no_line_search(small_disks):
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param

    P, T = small_disks[1]
    U0 = oracle.lscm(P, T)
    rng = np.random.default_rng(5)
    U0 = U0 + rng.normal(0, 0.05 * np.ptp(U0), U0.shape)      # folds the start
    assert (oracle.signed_areas(U0, T) <= 0).any()
    U, _ = param.refine(P, T, U0, iters=6, tol=0.0)
    Ur, _ = oracle.refine(P, T, U0, iters=6, tol=0.0)
    assert unit_box_err(U, Ur) <= 1e-4


def test_building_blocks(small_disks):
    """local step + rhs, max step, energy, signed areas against the oracle (FP64 both)."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd.plan import IslandPlan

    torch = _torch()
    charts = small_disks[:6]
    plan = IslandPlan(charts)
    U0s = [oracle.lscm(P, T) for P, T in charts]
    U = plan.device_uv(np.concatenate(U0s))
    b = plan.local_rhs(U).cpu().numpy()
    rng = np.random.default_rng(2)
    D = torch.as_tensor(rng.normal(0, 0.02, (plan.n, 2)), device=plan.dev)
    tm = plan.max_step(U, D).cpu().numpy()
    e0 = plan.energy(U, 0).cpu().numpy()
    e1 = plan.energy(U, 1).cpu().numpy()
    sa = plan.signed_areas(U).cpu().numpy()
    vo, to = plan.vert_off, plan.tri_off
    Dh = D.cpu().numpy()
    for i, (P, T) in enumerate(charts):
        X, area = oracle.local_frames(P, T)
        cots = oracle.cot_weights(X)
        u = U0s[i]
        R, _ = oracle.arap_rotations(u, T, X, cots, 0)
        br = oracle.arap_rhs(len(P), T, X, cots, R)
        I, J, W = oracle.lap_triplets(T, cots)
        br[0] += np.abs(W).max() * u[0]
        assert np.abs(b[vo[i]:vo[i + 1]] - br).max() <= 1e-10 * np.abs(br).max()
        assert tm[i] == pytest.approx(oracle.max_step(u, Dh[vo[i]:vo[i + 1]], T), rel=1e-9)
        assert e0[i] == pytest.approx(oracle.energy(u, T, X, area, 0), rel=1e-9)
        assert e1[i] == pytest.approx(oracle.energy(u, T, X, area, 1), rel=1e-9)
        assert np.abs(sa[to[i]:to[i + 1]] - oracle.signed_areas(u, T)).max() <= 1e-12 * np.ptp(u) ** 2


@pytest.mark.parametrize("precision", [0, 2])
def test_cg_solve_vs_direct(small_disks, precision):
    """Batched CG (mixed and FP64) against SuperLU on the pinned Laplacians."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd.plan import IslandPlan

    torch = _torch()
    charts = small_disks
    plan = IslandPlan(charts)
    rng = np.random.default_rng(3)
    b = rng.normal(size=(plan.n, 2))
    for i in range(len(charts)):                      # rhs of an ARAP step sums to zero per island
        s = slice(plan.vert_off[i], plan.vert_off[i + 1])
        b[s] -= b[s].mean(0)
    x = torch.zeros(plan.n, 2, dtype=torch.float64, device=plan.dev)
    plan.cg_solve(torch.as_tensor(b, device=plan.dev), x, tol=1e-10 if precision == 2 else 1e-6,
                  precision=precision)
    st = plan.stats()
    xh = x.cpu().numpy()
    for i, (P, T) in enumerate(charts):
        X, _ = oracle.local_frames(P, T)
        cots = oracle.cot_weights(X)
        I, J, W = oracle.lap_triplets(T, cots)
        n = len(P)
        L = sp.csc_matrix((np.r_[W, np.abs(W).max()], (np.r_[I, 0], np.r_[J, 0])), shape=(n, n))
        s = slice(plan.vert_off[i], plan.vert_off[i + 1])
        xr = spla.splu(L).solve(b[s])
        r = np.linalg.norm(b[s] - L @ xh[s], axis=0) / np.linalg.norm(b[s], axis=0)
        assert st["converged"][i] == 1
        assert r.max() <= (1e-10 if precision == 2 else 1e-6) * 1.0001
        assert np.abs(xh[s] - xr).max() <= (1e-6 if precision == 2 else 1e-2) * np.abs(xr).max()


def test_batch_matches_batch_oracle(small_disks):
    """gpu.Batch semantics: orient, repair, free ARAP, flip-safe ARAP."""
    from oracle import param_cpu as oracle
    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import gpu

    items = [(P, T, oracle.lscm(P, T)) for P, T in small_disks]
    b = gpu.Batch(items)
    r = BatchOracle(items)
    b.orient_positive()
    r.orient_positive()
    assert np.abs(b._U.cpu().numpy() - r.U).max() == 0.0
    b.arap(3, line_search=False, cg_iters=100000, cg_tol=1e-7)
    r.arap(3, line_search=False)
    b.arap(3, line_search=True, cg_iters=100000, cg_tol=1e-7)
    r.arap(3, line_search=True)
    for u, ur in zip(b.chart_uvs(), r.chart_uvs()):
        assert unit_box_err(u, ur) <= 1e-4
    assert b.U.dtype == _torch().float32


def test_batch_repair_matches_oracle(small_disks):
    from oracle import param_cpu as oracle
    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import gpu

    rng = np.random.default_rng(9)
    items = []
    for P, T in small_disks[:5]:
        U = oracle.lscm(P, T)
        U = U + rng.normal(0, 0.02 * np.ptp(U), U.shape)
        items.append((P, T, U))
    b = gpu.Batch(items)
    r = BatchOracle(items)
    b.repair_flips(30)
    r.repair_flips(30)
    for u, ur in zip(b.chart_uvs(), r.chart_uvs()):
        assert np.abs(u - ur).max() <= 1e-12 * max(1.0, np.abs(ur).max())
    f = b.flips_per_chart().cpu().numpy()
    assert np.array_equal(f, r.flips_per_chart())


def test_repair_flips_sequential_host():
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param, workloads

    P, T = workloads.disk_island(np.random.default_rng(4), 300)
    U = oracle.lscm(P, T)
    U = U + np.random.default_rng(1).normal(0, 0.03 * np.ptp(U), U.shape)
    assert np.array_equal(param.repair_flips(U, T, 20), oracle.repair_flips(U, T, 20))


def test_spdsolver_and_tutte(small_disks):
    """param.tutte (interior solve through SPDSolver on the GPU) against the
    reference's own embeddings of three charts (tests/golden/tutte_golden.npz:
    loop in id order, loop rotated against id order, grid chart) and against
    the oracle on a disk chart."""
    from pathlib import Path

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "tutte_golden.npz"))
    for name in ("ordered", "rotated", "grid"):
        P, T = g[f"{name}_P"], g[f"{name}_T"].astype(np.int64)
        U = param.tutte(P, T)
        assert np.abs(U - g[f"{name}_U"]).max() <= 1e-8 * np.abs(g[f"{name}_U"]).max(), name
    P, T = small_disks[2]
    U = param.tutte(P, T)
    Ur = oracle.tutte(P, T)
    assert np.abs(U - Ur).max() <= 1e-8 * np.abs(Ur).max()


def test_batched_equals_single_island_plans(small_disks):
    """Islands are independent: one batched launch gives the same layout as
    solving each island alone (bitwise), on both the wavefront and the
    workgroup paths."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    charts = small_disks[:3] + [workloads.disk_island(np.random.default_rng(21), 9000)]
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in charts])
    plan = IslandPlan(charts)
    Ub = plan.device_uv(U0)
    plan.arap(Ub, 5, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0)
    Ub = Ub.cpu().numpy()
    for i, (P, T) in enumerate(charts):
        s = slice(plan.vert_off[i], plan.vert_off[i + 1])
        p1 = IslandPlan([(P, T)])
        u = p1.device_uv(U0[s])
        p1.arap(u, 5, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0)
        assert np.array_equal(u.cpu().numpy(), Ub[s])


@pytest.mark.parametrize("no_wave", [False, True])
def test_team_wave_and_workgroup_paths_agree(no_wave):
    """The same islands solved by multi-workgroup teams, single workgroups and
    single wavefronts give the same layout to CG accuracy, and match the oracle."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    rng = np.random.default_rng(31)
    charts = [workloads.disk_island(rng, t) for t in (30000, 14000, 600, 2500)]
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in charts])
    out = []
    for no_teams in (False, True):
        plan = IslandPlan(charts, no_teams=no_teams, no_wave=no_wave)
        info = plan.info()
        print(f"paths: no_teams={no_teams} no_wave={no_wave} {info['n_group_tasks']} group tasks, "
              f"{info['n_wave_tasks']} wave tasks", flush=True)
        if not no_teams:
            assert info["n_group_tasks"] > len(charts) - (0 if no_wave else 2)   # some island has a team
        U = plan.device_uv(U0)
        plan.arap(U, 3, mode=_lib.UVA_MODE_REFINE, tol=0.0)
        assert plan.stats()["converged"].all()
        out.append(U.cpu().numpy())
    vo = plan.vert_off
    for i, (P, T) in enumerate(charts):
        s = slice(vo[i], vo[i + 1])
        assert unit_box_err(out[0][s], out[1][s]) <= 1e-5
    Ur, _ = oracle.refine(*charts[0], U0[vo[0]:vo[1]], iters=3, tol=0.0)
    assert unit_box_err(out[0][vo[0]:vo[1]], Ur) <= 1e-4


def test_graph_replays_match_eager(small_disks):
    """hipGraph replays of the captured ARAP sequence, with eager work and
    synchronisations between them, reproduce the eager result every time."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib
    from ultimate_uv_artist_amd.plan import IslandPlan

    torch = _torch()
    charts = small_disks
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in charts])
    plan = IslandPlan(charts)
    ref = plan.device_uv(U0)
    plan.arap(ref, 4, mode=_lib.UVA_MODE_REFINE, tol=0.0)
    ref = ref.cpu().numpy()
    U = plan.device_uv(U0)
    for rep in range(4):
        U.copy_(torch.as_tensor(U0, device=plan.dev))
        plan.arap(U, 4, mode=_lib.UVA_MODE_REFINE, tol=0.0, use_graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(U.cpu().numpy(), ref), f"replay {rep}"
        _ = plan.energy(U)                              # eager work between replays
        torch.cuda.synchronize()


def test_nonfinite_island_reverts_in_batch(small_disks):
    """A chart with a non-finite vertex: the one CG of the batch carries the
    NaN into every chart's iterate (its dot products are shared), so every
    chart reverts and keeps its layout, as in the reference (gpu.py:184-202,
    235-240); the per-island solve of param.refine semantics leaves the other
    islands' results as they are without the bad one."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, gpu
    from ultimate_uv_artist_amd.plan import IslandPlan

    items = [(P, T, oracle.lscm(P, T)) for P, T in small_disks[:4]]
    bad_items = list(items)
    Ubad = items[1][2].copy()
    Ubad[3] = np.nan
    bad_items[1] = (items[1][0], items[1][1], Ubad)
    b = gpu.Batch(bad_items)
    b.arap(2, line_search=True, cg_iters=100000)
    c = b.chart_uvs()
    for i in (0, 2, 3):
        assert np.array_equal(c[i], items[i][2])
    assert np.isnan(c[1][3]).all()
    # per-island solves (UVA_MODE_BATCH through uva_arap): the bad island
    # reverts, the others are solved as without it
    charts = [(P, T) for P, T, _ in items]
    U0 = np.concatenate([u for _, _, u in bad_items])
    good0 = np.concatenate([u for _, _, u in items])
    plan = IslandPlan(charts, batch=True)
    Ug, Ub = plan.device_uv(good0), plan.device_uv(U0)
    plan.arap(Ug, 2, line_search=True, mode=_lib.UVA_MODE_BATCH, cg_max_iters=100000)
    plan.arap(Ub, 2, line_search=True, mode=_lib.UVA_MODE_BATCH, cg_max_iters=100000)
    Ug, Ub = Ug.cpu().numpy(), Ub.cpu().numpy()
    vo = plan.vert_off
    for i in (0, 2, 3):
        assert np.array_equal(Ug[vo[i]:vo[i + 1]], Ub[vo[i]:vo[i + 1]])
    assert np.isnan(Ub[vo[1] + 3]).all()


def _own_box(U):
    lo = U.min(0)
    return (U - lo) / float((U.max(0) - lo).max())


def test_refine_slim_matches_reference_trajectory():
    """param.refine(method="slim") on the GPU against the reference's SLIM
    trajectory (golden fixture): final energy rel 1e-9, layout 1e-8 in its own
    unit box (translation is fixed only by a 1e-12 regulariser; the GPU keeps
    the start's mean exactly), and a flipped start comes back unchanged."""
    from pathlib import Path

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import param

    d = dict(np.load(Path(__file__).parent / "golden" / "slim_golden.npz"))
    for c in (0, 1):
        P, T, U0 = d[f"P{c}"], d[f"T{c}"].astype(np.int64), d[f"U0_{c}"]
        print(f"slim case {c}: {len(T)} triangles", flush=True)
        U, solver = param.refine(P, T, U0, iters=8, method="slim", tol=1e-6)
        assert solver is None
        X, area = oracle.local_frames(P, T)
        e = oracle.energy(U, T, X, area, 1)
        assert abs(e - d[f"slim{c}_energy"][-1]) <= 1e-9 * e
        assert np.abs(_own_box(U) - _own_box(d[f"slim{c}"])).max() <= 1e-8
        assert np.abs((U - U0).mean(0)).max() <= 1e-12
    Uf = U0.copy()
    a, b, cc = (Uf[v].copy() for v in T[0])
    edge = cc - b
    Uf[T[0][0]] = 2.0 * (b + edge * ((a - b) @ edge) / (edge @ edge)) - a
    assert np.array_equal(param.refine(P, T, Uf, iters=3, method="slim")[0], Uf)


def test_tuned_schedule_matches_default():
    """uva_plan_tune re-plans team sizes and the workgroup schedule from the
    measured CG iterations; the solve gives the same layout to CG accuracy,
    repeated tuned runs are bitwise identical, and graph replays after the
    re-plan (the cached graph is rebuilt) match eager runs."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    torch = _torch()
    rng = np.random.default_rng(57)
    charts = [workloads.disk_island(rng, t) for t in (26000, 9000, 3000, 600, 300)]
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in charts])
    plan = IslandPlan(charts)
    ref = plan.device_uv(U0)
    plan.arap(ref, 3, mode=_lib.UVA_MODE_REFINE, tol=0.0, use_graph=True)
    ref = ref.cpu().numpy()
    plan.tune()
    outs = []
    for use_graph in (False, True, True):
        U = plan.device_uv(U0)
        plan.arap(U, 3, mode=_lib.UVA_MODE_REFINE, tol=0.0, use_graph=use_graph)
        torch.cuda.synchronize()
        assert plan.stats()["converged"].all()
        outs.append(U.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])
    vo = plan.vert_off
    for i in range(len(charts)):
        s = slice(vo[i], vo[i + 1])
        assert unit_box_err(outs[0][s], ref[s]) <= 1e-5


def _atlas(charts, layouts, mirror=()):
    """Islands in the cells of a k x k grid of the unit square (each scaled
    into 0.9 of its cell, keeping its shape), as one mesh for validate_uvs;
    islands in ``mirror`` are reflected."""
    k = int(np.ceil(np.sqrt(len(charts))))
    Ps, Ts, Us, off = [], [], [], 0
    for i, ((P, T), U) in enumerate(zip(charts, layouts)):
        lo = U.min(0)
        u = (U - lo) / float((U.max(0) - lo).max()) * 0.9 / k
        if i in mirror:
            u[:, 0] = 0.9 / k - u[:, 0]
        Ps.append(P + 100.0 * i)
        Ts.append(np.asarray(T, np.int64) + off)
        Us.append(u + np.array([i % k, i // k]) / k)
        off += len(P)
    return np.concatenate(Ps), np.concatenate(Ts), np.concatenate(Us)


def test_validation_verdicts_match_oracle():
    """The reference's validate.py verdicts (validate_uvs counters, the valid
    flag and chart_overlaps) of the GPU layouts equal those of the oracle's
    layouts, per island and for an atlas of all islands (one of them
    mirrored), after LSCM and after 10 ARAP iterations (CG to 1e-6), on disk
    islands and on islands with 10% needle triangles (aspect 50-500).  The
    needle islands keep one or two flipped and 5-12 overlapping triangles
    after LSCM and after ARAP in the oracle's layouts too: the verdicts must
    agree on them as well."""
    from oracle import param_cpu as oracle
    from oracle import validate_cpu as vc
    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    rng = np.random.default_rng(61)
    charts = [workloads.disk_island(rng, t) for t in (300, 800, 2000, 4500, 9000)]
    charts += workloads.sliver_islands(seed=7, n_islands=3, tmin=1500, tmax=4000)
    plan = IslandPlan(charts)
    U = plan.device_uv(np.zeros((plan.n, 2)))
    plan.lscm(U)
    Ul = U.cpu().numpy()
    plan.arap(U, 10, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0, cg_tol=1e-6)
    Ua = U.cpu().numpy()
    assert plan.stats()["converged"].all()
    vo = plan.vert_off
    gl, ga, rl, ra = [], [], [], []
    for i, (P, T) in enumerate(charts):
        s = slice(vo[i], vo[i + 1])
        url = oracle.lscm(P, T)
        ura, _ = oracle.refine(P, T, url, iters=10, tol=0.0)
        gl.append(Ul[s]), ga.append(Ua[s]), rl.append(url), ra.append(ura)
        for g, r in ((Ul[s], url), (Ua[s], ura)):
            vg, vr = vc.island_verdict(P, T, g), vc.island_verdict(P, T, r)
            print(f"island {i} ({len(T)} triangles): {vg}", flush=True)
            assert vg == vr
    for g, r in ((gl, rl), (ga, ra)):
        a = vc.validate_uvs(*_atlas(charts, g, mirror=(2,)))
        b = vc.validate_uvs(*_atlas(charts, r, mirror=(2,)))
        assert a[0] == b[0] and np.array_equal(a[1], b[1])
        assert a[0]["islands"] == len(charts) and a[0]["mirrored_islands"] == 1


@pytest.mark.parametrize("i", [0, 1, 2])
def test_hier_lscm_matches_oracle_and_reference(i):
    """hier.Solver.lscm on the GPU (uva_jacobi_pcg) against the float64
    restatement, which follows the same iteration and stopping rule (per
    chart within 1e-6 of its extent), and against the reference's float32
    run (the restatement's own agreement: 3e-4 at 60 iterations, 3e-3 at
    the stalled 600-iteration default)."""
    import types
    from pathlib import Path

    from oracle import hier_cpu
    from ultimate_uv_artist_amd import hier

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "hier_lscm_golden.npz"))
    V, F = g["V"], g["F"].astype(np.int64)
    args = (g[f"c{i}_faces"], g[f"c{i}_CV"], int(g[f"c{i}_nv"]), g[f"c{i}_chart"], int(g[f"c{i}_nchart"]),
            g[f"c{i}_U0"], g[f"c{i}_ba"], g[f"c{i}_bb"])
    iters = int(g[f"c{i}_iters"])
    print(f"hier lscm case {i}: nv {args[2]}, charts {args[4]}, iters {iters}", flush=True)
    S = hier.Solver(types.SimpleNamespace(V=V, F=F))
    U = S.lscm(*args, iters=iters)
    Uo = hier_cpu.Solver(V, F).lscm(*args, iters=iters)
    cv = np.zeros(len(U), np.int64)
    cv[args[1].reshape(-1)] = np.repeat(args[3], 3)
    for c in range(args[4]):
        m = cv == c
        assert unit_box_err(U[m], Uo[m]) <= 1e-6
        assert unit_box_err(U[m], g[f"c{i}_U"][m]) <= (3e-4 if iters <= 60 else 3e-3)
    assert 0 < S.lscm_iters <= iters
    assert U.dtype == np.float64 and np.array_equal(U, U.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("i", [0, 1])
def test_hier_solver_matches_reference(i):
    """hier.Solver.relax / bad_fraction on the GPU against the reference's own
    run (tests/golden/hier_golden.npz) and the float64 restatement: layouts
    within 1e-4 per chart (unit box), identical flip counts, scale and bad
    fractions to float32 rounding."""
    import types
    from pathlib import Path

    from oracle import hier_cpu
    from ultimate_uv_artist_amd import hier

    h = dict(np.load(Path(__file__).resolve().parent / "golden" / "hier_golden.npz"))
    V, F = h["V"], h["F"].astype(np.int64)
    kw = dict(zip(("iters", "cg_iters", "repair_rounds"), (int(x) for x in h[f"c{i}_kw"])))
    args = (h[f"c{i}_faces"], h[f"c{i}_CV"], int(h[f"c{i}_nv"]), h[f"c{i}_chart"], int(h[f"c{i}_nchart"]),
            h[f"c{i}_U0"])
    S = hier.Solver(types.SimpleNamespace(V=V, F=F))
    U, flips, scale = S.relax(*args, **kw)
    bA, bR = S.bad_fraction(2.0, 1.5)
    O = hier_cpu.Solver(V, F)
    Uo, flo, sco = O.relax(*args, **kw)
    cv = np.zeros(len(U), np.int64)
    cv[args[1].reshape(-1)] = np.repeat(args[3], 3)
    for ref in (h[f"c{i}_U"], Uo):
        for c in range(args[4]):
            m = cv == c
            assert unit_box_err(U[m], ref[m]) <= 1e-4
    assert np.array_equal(flips, h[f"c{i}_flips"]) and np.array_equal(flips, flo)
    assert np.abs(scale / h[f"c{i}_scale"] - 1).max() <= 1e-5
    assert np.abs(bA - h[f"c{i}_badA"]).max() <= 1e-4 and np.abs(bR - h[f"c{i}_badR"]).max() <= 1e-4
    assert U.dtype == np.float64 and np.array_equal(U, U.astype(np.float32).astype(np.float64))


def test_batch_default_budget_matches_reference():
    """gpu.Batch.arap at its default CG budget (40 iterations, tolerance
    1e-6, one CG shared by all charts) against the reference's own run of the
    same calls (tests/golden/batch40_golden.npz, float64 CPU torch): the
    unconverged iterates agree per chart to 1e-6 (unit box)."""
    from pathlib import Path

    from ultimate_uv_artist_amd import gpu

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "batch40_golden.npz"))
    items = [(g[f"P{i}"], g[f"T{i}"].astype(np.int64), g[f"U{i}"]) for i in range(int(g["n"]))]
    b = gpu.Batch(items, dtype=_torch().float64)
    b.arap(4, line_search=True)
    u_ls = b.U.cpu().numpy()
    b.arap(2, line_search=False)
    u_free = b.U.cpu().numpy()
    off = b.offsets
    for got, ref in ((u_ls, g["after_ls"]), (u_free, g["after_free"])):
        for i in range(len(items)):
            s = slice(off[i], off[i + 1])
            assert unit_box_err(got[s], ref[s]) <= 1e-6


def test_flatten_many_and_queue_match_reference():
    """charts.flatten for 66 small charts as one batch (charts.flatten_many)
    and through FlattenQueue from four threads, against the reference's own
    charts.flatten (tests/golden/flatten_golden.npz): layouts within 1e-4
    (unit box), identical fold counts, p95 anisotropy and area within 1 %."""
    from concurrent.futures import ThreadPoolExecutor
    from pathlib import Path

    from ultimate_uv_artist_amd import charts as gch

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "flatten_golden.npz"))
    items = [(g[f"P{i}"], g[f"T{i}"].astype(np.int64)) for i in range(int(g["n"]))]
    res = gch.flatten_many(items)
    for i, (U, q) in enumerate(res):
        Ur, qr = g[f"U{i}"], g[f"q{i}"]
        assert unit_box_err(U, Ur) <= 1e-4
        assert q["flips"] == qr[0] and q["boundary_hits"] == qr[3]
        assert abs(q["aniso_p95"] / qr[1] - 1) <= 0.01 and abs(q["area_p95"] / qr[2] - 1) <= 0.01
    with gch.FlattenQueue(max_wait_ms=20.0) as fq, ThreadPoolExecutor(4) as ex:
        out = list(ex.map(lambda it: fq.flatten(*it), items))
    for (U, _), (Uq, _) in zip(res, out):
        assert np.abs(U - Uq).max() <= 1e-9 * np.abs(U).max()
    with pytest.raises(ValueError):
        gch.flatten_many([(np.eye(4, 3), np.array([[0, 1, 2], [0, 2, 1]]))])


def _rigid_err(U, Ur):
    a, b = U - U.mean(0), Ur - Ur.mean(0)
    u, _, vt = np.linalg.svd(a.T @ b)
    if np.linalg.det(u @ vt) < 0:
        u[:, -1] *= -1
    return float(np.abs(a @ (u @ vt) - b).max() / np.ptp(Ur, 0).max())


def test_flatten_many_pipeline_settings_and_slim():
    """flatten_many at the pipeline's plan_iters 4, 6, 10, at 20 iterations,
    and the artist profile's final flatten (20 ARAP + 15 SLIM iterations)
    against the reference's charts.flatten on 24 small charts
    (tests/golden/flatten_slim_golden.npz): ARAP layouts within 1e-4 (unit
    box); SLIM layouts within 1e-4 after a rigid alignment (the SLIM system
    fixes rigid motions only through a 1e-12 x area regularisation, so
    rounding sets the rotation) with the symmetric Dirichlet energy within
    1e-6; identical fold counts.  The same requests through FlattenQueue,
    mixed settings from four threads, give the batched results."""
    from concurrent.futures import ThreadPoolExecutor
    from pathlib import Path

    from oracle import param_cpu as pc
    from ultimate_uv_artist_amd import charts as gch

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "flatten_slim_golden.npz"))
    items = [(g[f"P{i}"], g[f"T{i}"].astype(np.int64)) for i in range(int(g["n"]))]
    settings = [(int(a), int(b)) for a, b in g["settings"]]
    batched = {}
    for k, (it, sl) in enumerate(settings):
        print(f"flatten_many iters {it} slim_iters {sl}", flush=True)
        res = gch.flatten_many(items, it, slim_iters=sl)
        batched[(it, sl)] = res
        for i, (U, q) in enumerate(res):
            Ur, qr = g[f"U{i}_{k}"], g[f"q{i}_{k}"]
            if sl:
                P, T = items[i]
                X, area = pc.local_frames(P, T)
                assert _rigid_err(U, Ur) <= 1e-4
                er = pc.energy(Ur, T, X, area, 1)
                assert abs(pc.energy(U, T, X, area, 1) - er) <= 1e-6 * er
            else:
                assert unit_box_err(U, Ur) <= 1e-4
            assert q["flips"] == qr[0] and q["boundary_hits"] == qr[3]
    reqs = [(i, s) for i in range(len(items)) for s in settings]
    with gch.FlattenQueue(max_wait_ms=20.0) as fq, ThreadPoolExecutor(4) as ex:
        out = list(ex.map(lambda r: fq.flatten(*items[r[0]], iters=r[1][0], slim_iters=r[1][1]), reqs))
    for (i, s), (Uq, _) in zip(reqs, out):
        U = batched[s][i][0]
        assert np.abs(U - Uq).max() <= 1e-9 * np.abs(U).max()


def test_validation_on_gpu_matches_reference_and_oracle():
    """validate.validate_uvs (gutter=False) and chart_overlaps with the GPU
    overlap test: the reference's own reports and masks on its validation
    cases (tests/golden/validate_golden.npz), and the same flags as the CPU
    restatement on a 4-island atlas of 60k triangles with noise (both the
    tolerant test and the zero-tolerance crossings)."""
    from pathlib import Path

    from oracle import validate_cpu as vc
    from ultimate_uv_artist_amd import validate, workloads

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "validate_golden.npz"))
    keys = [str(k) for k in g["report_keys"]]
    for i in range(len(g["v_names"])):
        rep, prob = validate.validate_uvs(g[f"v{i}_P"], g[f"v{i}_F"].astype(np.int64), g[f"v{i}_U"], gutter=False)
        assert dict(zip(keys, [float(rep.get(k, -1)) for k in keys])) == dict(zip(keys, g[f"v{i}_report"]))
        assert np.array_equal(prob, g[f"v{i}_problems"])
    for i in range(len(g["c_names"])):
        assert np.array_equal(validate.chart_overlaps(g[f"c{i}_U"], g[f"c{i}_F"].astype(np.int64)), g[f"c{i}_mask"])
    rng = np.random.default_rng(72)
    Ps, Ts, Us, off = [], [], [], 0
    for k in range(4):
        P, T = workloads.disk_island(rng, 15000)
        U = (P[:, :2] - P[:, :2].min(0)) / np.ptp(P[:, :2]).max() * 0.55 + np.array([0.4 * (k % 2), 0.4 * (k // 2)])
        Ps.append(P + 5.0 * k)
        Ts.append(T + off)
        Us.append(U + rng.normal(0, 0.0005, U.shape))
        off += len(P)
    P, T, U = np.concatenate(Ps), np.concatenate(Ts), np.clip(np.concatenate(Us), 0, 1)
    skip = rng.random(len(T)) < 0.02
    G = int(np.clip(np.sqrt(len(T)) / 2, 64, 4096))
    a = validate.overlap_flags(U, T, skip, 1e-6, False, G)
    assert a.sum() > 100 and np.array_equal(a, vc.overlaps(U, T, skip, 1e-6))
    assert np.array_equal(validate.overlap_flags(U, T, skip, 0.0, True, G), vc.overlaps_strict(U, T, skip))
    r1, p1 = validate.validate_uvs(P, T, U, gutter=False)
    r2, p2 = vc.validate_uvs(P, T, U)
    assert r1 == r2 and np.array_equal(p1, p2)


def test_csr_pcg_matches_the_definition():
    """uva_csr_pcg (CsrSystem.pcg): solves in a row on one handle with other
    right-hand sides and starts, eager torch work and synchronisations
    between them, two columns at the Batch rule and one column at the lscm
    rule, each against the Jacobi-PCG definition in numpy (iterations equal,
    x to 1e-10 relative)."""
    import scipy.sparse as sp
    import torch

    from ultimate_uv_artist_amd.plan import CsrSystem

    rng = np.random.default_rng(91)
    n = 5000
    i = rng.integers(0, n, 20000)
    j = rng.integers(0, n, 20000)
    v = rng.uniform(-1, 0, 20000)
    A = sp.coo_matrix((np.r_[v, v], (np.r_[i, j], np.r_[j, i])), shape=(n, n)).tocsr()
    A = (A + sp.diags(np.asarray(abs(A).sum(1)).ravel() + rng.uniform(0.01, 0.1, n))).tocsr()
    dinv = 1.0 / A.diagonal()

    def pcg(B, X, iters, check, tol2, strict):
        B, x = B.reshape(n, -1), X.reshape(n, -1).copy()
        r = B - A @ x
        z = r * dinv[:, None]
        p = z.copy()
        rz = (r * z).sum(0)
        b2 = np.maximum((B * B).sum(0), 1e-30)
        for k in range(iters):
            Ap = A @ p
            al = rz / np.maximum((p * Ap).sum(0), 1e-30)
            x, r = x + al * p, r - al * Ap
            rr = (r * r).sum(0)
            if k % check == 0 and ((rr < tol2 * b2) if strict else (rr <= tol2 * b2)).all():
                return x.reshape(X.shape), k + 1
            z = r * dinv[:, None]
            rz2 = (r * z).sum(0)
            p = z + rz2 / np.maximum(rz, 1e-30) * p
            rz = rz2
        return x.reshape(X.shape), iters

    dev = torch.device("cuda", 0)
    S = CsrSystem(A, dev)
    cases = [((n, 2), 40, 1, 1e-12, False)] * 4 + [((n,), 600, 50, 1e-12, True)] * 2 + [((n, 2), 400, 1, 1e-24, False)]
    for k, (shape, iters, check, tol2, strict) in enumerate(cases):
        print(f"csr pcg case {k}: shape {shape}, iters {iters}", flush=True)
        B = rng.normal(size=shape) * 10.0 ** rng.integers(-2, 3)
        X0 = rng.normal(size=shape)
        b = torch.as_tensor(B, device=dev)
        x = torch.as_tensor(X0, device=dev).clone()
        its = S.pcg(b, x, iters, check, tol2, strict)
        junk = (x * 2.0).sum()                           # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        xr, kr = pcg(B, X0, iters, check, tol2, strict)
        assert its == kr, (k, its, kr)
        assert np.abs(x.cpu().numpy() - xr).max() <= 1e-10 * np.abs(xr).max(), k
    S.close()


def test_plan_pcg_matches_the_definition():
    """uva_plan_pcg (IslandPlan.pcg, the global step of gpu.Batch.arap and
    hier.Solver.relax on the chart plan's own matrix) on 40 disk charts with
    a soft pin at every chart's first vertex: solves in a row on one plan --
    the Batch rule at 40 iterations, a fixed count (hier), a run to 1e-9 --
    with eager torch work between them, each against Jacobi-PCG on the
    pinned Laplacian that scipy assembles from the reference's triplets
    (iterations equal, x to 1e-10 relative).  The run to convergence takes
    about 2,000 iterations, over which the recurrence for A p and the sum
    order move the iteration that crosses the threshold by one or two: there
    the iterations agree to 0.5 % and x meets the stop rule on the true
    residual."""
    import scipy.sparse as sp
    import torch

    from oracle import hier_cpu
    from ultimate_uv_artist_amd import workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    rng = np.random.default_rng(93)
    charts = [workloads.disk_island(rng, int(t)) for t in rng.integers(200, 6000, 40)]
    dev = torch.device("cuda", 0)
    plan = IslandPlan(charts, batch=True, dev=dev, no_coarse=True)
    voff = plan.vert_off
    plan.set_pins(voff[:-1], 0.0)
    P = np.concatenate([c[0] for c in charts])
    T = np.concatenate([c[1] + voff[i] for i, c in enumerate(charts)])
    _, w, *_ = hier_cpu.frames(P, T)
    W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
    I = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    J = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    n = len(P)
    L = sp.csr_matrix((np.concatenate([-W, -W, W, W, np.full(len(charts), max(np.abs(W).max(), 1e-12))]),
                       (np.concatenate([I, J, I, J, voff[:-1]]), np.concatenate([J, I, I, J, voff[:-1]]))),
                      shape=(n, n))
    dinv = 1.0 / np.maximum(L.diagonal(), 1e-30)[:, None]

    def pcg(B, x, iters, tol2, strict):
        x = x.copy()
        r = B - L @ x
        z = r * dinv
        p = z.copy()
        rz = (r * z).sum(0)
        b2 = np.maximum((B * B).sum(0), 1e-30)
        for k in range(iters):
            Ap = L @ p
            al = rz / np.maximum((p * Ap).sum(0), 1e-30)
            x, r = x + al * p, r - al * Ap
            rr = (r * r).sum(0)
            if ((rr < tol2 * b2) if strict else (rr <= tol2 * b2)).all():
                return x, k + 1
            z = r * dinv
            rz2 = (r * z).sum(0)
            p = z + rz2 / np.maximum(rz, 1e-30) * p
            rz = rz2
        return x, iters

    cases = [(40, 1e-12, False)] * 2 + [(20, 0.0, True), (2000, 1e-18, False)]
    for k, (iters, tol2, strict) in enumerate(cases):
        print(f"plan pcg case {k}: iters {iters}, tol2 {tol2}", flush=True)
        B = rng.normal(size=(n, 2)) * 10.0 ** rng.integers(-2, 3)
        X0 = rng.normal(size=(n, 2))
        b = torch.as_tensor(B, device=dev)
        x = torch.as_tensor(X0, device=dev).clone()
        its = plan.pcg(b, x, iters, 1, tol2, strict)
        junk = (x * 2.0).sum()                           # eager work between solves
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        xr, kr = pcg(B, X0, iters, tol2, strict)
        xg = x.cpu().numpy()
        if kr < iters and kr > 200:
            assert abs(its - kr) <= max(2, 0.005 * kr), (k, its, kr)
            rel = np.linalg.norm(B - L @ xg, axis=0) / np.linalg.norm(B, axis=0)
            assert rel.max() <= 1.01 * np.sqrt(tol2), (k, rel)
        else:
            assert its == kr, (k, its, kr)
            assert np.abs(xg - xr).max() <= 1e-10 * np.abs(xr).max(), k
    plan.close()


def test_gutter_on_gpu_matches_reference_and_oracle():
    """validate_uvs with the gutter (the default) and min_gutter: the
    reference's own entries at atlas 64 and 256 and its label images at 128
    (tests/golden/gutter_golden.npz), and the GPU raster bit for bit equal to
    the restatement's on a 60k-triangle atlas at 1024 and 4096 texels."""
    from pathlib import Path

    from oracle import validate_cpu as vc
    from ultimate_uv_artist_amd import validate, workloads

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "gutter_golden.npz"))
    keys = [str(k) for k in g["report_keys"]]
    for i, name in enumerate(g["names"]):
        P, F, U = g[f"g{i}_P"], g[f"g{i}_F"].astype(np.int64), g[f"g{i}_U"]
        U32 = np.asarray(U, np.float32).astype(np.float64)
        gap, img = validate.min_gutter(U32, F, g[f"g{i}_isl"], 128)
        assert np.array_equal(img, g[f"g{i}_img128"]), name
        ref = float(g[f"g{i}_gap128"])
        assert (np.isnan(gap) and np.isnan(ref)) or gap == ref, name
        for a in (64, 256):
            rep, _ = validate.validate_uvs(P, F, U, atlas=a, padding=2)
            want = dict(zip(keys, g[f"g{i}_a{a}"]))
            got = rep["min_gutter_texels"]
            assert got == want["min_gutter_texels"] or (np.isnan(got) and np.isnan(want["min_gutter_texels"])), name
            assert float(rep["gutter_ok"]) == want["gutter_ok"] and float(rep["valid"]) == want["valid"], name
    rng = np.random.default_rng(73)
    Ts, Us, off = [], [], 0
    for k in range(4):
        P, T = workloads.disk_island(rng, 15000)
        U = (P[:, :2] - P[:, :2].min(0)) / np.ptp(P[:, :2]).max() * 0.45 + np.array([0.5 * (k % 2), 0.5 * (k // 2)])
        Ts.append(T + off)
        Us.append(U)
        off += len(P)
    U, T = np.concatenate(Us), np.concatenate(Ts)
    isl = np.repeat(np.arange(4), [len(t) for t in Ts])
    for res in (1024, 4096):
        gap, img = validate.min_gutter(U, T, isl, res)
        gap_r, img_r = vc.min_gutter(U, T, isl, res)
        assert np.array_equal(img, img_r) and gap == gap_r, res
        print(f"gutter: 60k-triangle atlas at {res}: {gap:.3f} texels", flush=True)


def test_tuning_fit_cache_and_override(tmp_path, monkeypatch):
    """uva_plan_tune fits the schedule's cost model to measured island times
    (positive constants), stores it in UVART_TUNE_CACHE under this GPU model,
    and a new plan starts from the cached model (or UVART_TUNE_COST); the
    layouts agree with the default schedule's to CG accuracy."""
    import json

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    torch = _torch()
    rng = np.random.default_rng(58)
    charts = [workloads.disk_island(rng, t) for t in (26000, 9000, 3000, 600, 300, 14000)]
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in charts])
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("UVART_TUNE_CACHE", str(cache))
    plan = IslandPlan(charts)
    assert np.allclose(plan.tuning(), [15.0, 30.0, 0.1])
    ref = plan.device_uv(U0)
    plan.arap(ref, 3, mode=_lib.UVA_MODE_REFINE, tol=0.0)
    plan.tune()
    fitted = plan.tuning()
    print(f"tuning: fitted cost model {fitted.tolist()}", flush=True)
    assert (fitted > 0).all() and np.isfinite(fitted).all()
    data = json.loads(cache.read_text())
    assert len(data) == 1 and np.allclose(next(iter(data.values()))["cost_us"], fitted)
    plan2 = IslandPlan(charts)
    assert np.allclose(plan2.tuning(), fitted)
    U = plan2.device_uv(U0)
    plan2.arap(U, 3, mode=_lib.UVA_MODE_REFINE, tol=0.0)
    torch.cuda.synchronize()
    assert plan2.stats()["converged"].all()
    Ur, Uh = ref.cpu().numpy(), U.cpu().numpy()
    vo = plan.vert_off
    for i in range(len(charts)):
        s = slice(vo[i], vo[i + 1])
        assert unit_box_err(Uh[s], Ur[s]) <= 1e-5
    monkeypatch.setenv("UVART_TUNE_COST", "11,22,0.2")
    assert np.allclose(IslandPlan(charts[:2]).tuning(), [11, 22, 0.2])
