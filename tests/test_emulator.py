"""The island tasks' own code (run_task, cg_island, the team interface) run on
the CPU by the emulator, one thread per lane, against the oracle.

This covers the kernel logic -- recurrences, per-member slice and triangle
ranges, barrier placement, reductions -- for team shapes the GPU uses
(one wavefront, one workgroup of several wavefronts, several workgroups),
on machines without a GPU.  Tolerances are the GPU parity tests' own.
"""
from __future__ import annotations

import numpy as np
import pytest

import emu
from emu import EmuPlan

# (members, wavefronts per member): a wavefront team, a workgroup team, and
# multi-workgroup teams (3 members leave uneven slice ranges; 5 members on the
# smallest island leave some members without slices)
SHAPES = [(1, 1), (1, 4), (3, 2), (5, 1)]


def unit_box_err(U, Uref):
    lo = Uref.min(0)
    s = float((Uref.max(0) - lo).max())
    return float(np.abs((U - lo) / s - (Uref - lo) / s).max())


@pytest.fixture(scope="module")
def disks():
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(5)
    return [workloads.disk_island(rng, t) for t in (180, 700, 1500)]


def _pinned_laplacian(P, T):
    import scipy.sparse as sp

    from oracle import param_cpu as oracle

    X, _ = oracle.local_frames(P, T)
    I, J, W = oracle.lap_triplets(T, oracle.cot_weights(X))
    n = len(P)
    return sp.csc_matrix((np.r_[W, np.abs(W).max()], (np.r_[I, 0], np.r_[J, 0])), shape=(n, n))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precision", [0, 2])
def test_emu_cg_vs_direct(disks, shape, precision):
    """Mixed (FP32 inner, FP64 residual) and FP64 CG against SuperLU."""
    import scipy.sparse.linalg as spla

    e = EmuPlan(disks, *shape)
    rng = np.random.default_rng(7)
    b = rng.normal(size=(e.n, 2))
    for i in range(e.k):
        s = slice(e.vo[i], e.vo[i + 1])
        b[s] -= b[s].mean(0)
    tol = 1e-10 if precision == 2 else 1e-6
    x = e.cg(b, np.zeros_like(b), tol=tol, precision=precision)
    st = e.stats()
    for i, (P, T) in enumerate(disks):
        s = slice(e.vo[i], e.vo[i + 1])
        L = _pinned_laplacian(P, T)
        xr = spla.splu(L).solve(b[s])
        r = np.linalg.norm(b[s] - L @ x[s], axis=0) / np.linalg.norm(b[s], axis=0)
        assert st["converged"][i] == 1
        assert r.max() <= tol * 1.0001
        assert abs(st["rel_resid"][i] - r.max()) <= 1e-3 * r.max() + 1e-15
        assert np.abs(x[s] - xr).max() <= (1e-6 if precision == 2 else 1e-2) * np.abs(xr).max()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("no_coarse", [False, True])
def test_emu_lscm_vs_oracle(disks, shape, no_coarse):
    """The FP64 LSCM solve (resident where the member's rows fit, with the
    coarse level of the coupled operator unless left out) against SuperLU."""
    from oracle import param_cpu as oracle

    e = EmuPlan(disks, *shape, no_coarse=no_coarse)
    U = e.lscm(tol=1e-10, precision=2)
    assert e.stats()["converged"].all()
    for i, (P, T) in enumerate(disks):
        s = slice(e.vo[i], e.vo[i + 1])
        assert unit_box_err(U[s], oracle.lscm(P, T)) <= 1e-7


@pytest.mark.parametrize("shape", [(1, 4), (3, 2), (2, 8)])
def test_emu_lscm_resident_equals_streaming(disks, shape):
    """The resident LSCM solve (both columns coupled through the boundary
    terms, hard pins kept at zero) performs the streaming solve's arithmetic
    in the same order: the same layout bit for bit and the same iterations."""
    out, its = [], []
    for resident in (True, False):
        e = EmuPlan(disks, *shape, resident=resident, no_coarse=True)
        out.append(e.lscm(tol=1e-9, precision=2))
        st = e.stats()
        assert st["converged"].all()
        its.append(st["cg_iters"].copy())
    assert np.array_equal(its[0], its[1])
    assert np.array_equal(out[0], out[1])


def test_emu_lscm_coarse_level_is_the_inverse_of_the_restricted_operator():
    """The plan's coarse level of the LSCM operator against a construction
    from the oracle's conformal system (param.py:136-165 with the two pins
    eliminated, assembled by scipy from the reference's triplets): with Z the
    indicator of the aggregates (g consecutive 64-row slices of the plan's
    vertex order) for u and for v over the free rows, the rows of [Pc | Qc]
    are the top half of (Z^T Hff Z)^-1 = [[Pc, Qc], [-Qc, Pc]].  The coarse
    level leaves the converged layout as it is and cuts the iterations."""
    import scipy.sparse as sp

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    P, T = workloads.disk_island(np.random.default_rng(21), 3000)
    n = len(P)
    e = EmuPlan([(P, T)], members=2, waves=8)
    C_, g = e.lscm_coarse(0)
    assert C_ is not None
    nc = C_.shape[0]
    rank = np.empty(n, np.int64)
    rank[e.vperm()] = np.arange(n)       # plan position of each caller vertex
    agg = rank // (64 * g)
    assert agg.max() + 1 == nc
    Hff, _, keep, _, _ = oracle.lscm_system(P, T)
    free = np.flatnonzero(keep)           # rows of Hff: free u, then free v
    col = np.where(free < n, agg[free % n], nc + agg[free % n])
    Z = sp.csr_matrix((np.ones(len(free)), (np.arange(len(free)), col)), shape=(len(free), 2 * nc))
    S = np.linalg.inv((Z.T @ Hff @ Z).toarray())
    assert np.isfinite(S).all()
    ref = np.hstack([0.5 * (S[:nc, :nc] + S[nc:, nc:]), 0.5 * (S[:nc, nc:] - S[nc:, :nc])])
    assert np.abs(S[:nc, :nc] - S[nc:, nc:]).max() <= 1e-9 * np.abs(S).max()   # the Hermitian form
    assert np.abs(C_ - ref).max() <= 1e-6 * np.abs(ref).max()
    it, U = {}, {}
    for no_coarse in (False, True):
        e = EmuPlan([(P, T)], members=2, waves=8, no_coarse=no_coarse)
        U[no_coarse] = e.lscm(tol=1e-10, precision=2)
        st = e.stats()
        assert st["converged"].all()
        it[no_coarse] = int(st["cg_iters"][0])
    Ur = oracle.lscm(P, T)
    assert unit_box_err(U[False], Ur) <= 1e-7 and unit_box_err(U[True], Ur) <= 1e-7
    assert it[False] < it[True]


@pytest.mark.parametrize("shape", SHAPES)
def test_emu_arap_refine_vs_oracle(disks, shape):
    """param.refine semantics: soft pin, flip-safe line search, 4 iterations."""
    from oracle import param_cpu as oracle

    e = EmuPlan(disks, *shape)
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in disks])
    U = e.arap(U0, 4, tol=0.0, cg_tol=1e-8)
    assert e.stats()["converged"].all()
    for i, (P, T) in enumerate(disks):
        s = slice(e.vo[i], e.vo[i + 1])
        Ur, _ = oracle.refine(P, T, U0[s], iters=4, tol=0.0)
        assert unit_box_err(U[s], Ur) <= 1e-5


@pytest.mark.timeout(120)
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_emu_team_members_take_one_path(precision):
    """A 65-slice island on a two-member team: the members hold 32 and 33
    slices, on either side of the FP64 resident limit (4 rows per lane).
    Every member must take the same CG path (their barrier sequences differ),
    so the solve finishes and matches the oracle."""
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    P, T = workloads.disk_island(np.random.default_rng(7800), 7800)
    assert -(-len(P) // 64) == 65
    U0 = oracle.lscm(P, T)
    e = EmuPlan([(P, T)], members=2, waves=8)
    U = e.arap(U0, 2, tol=0.0, cg_tol=1e-8, precision=precision)
    if precision != 1:
        assert e.stats()["converged"].all()
        Ur, _ = oracle.refine(P, T, U0, iters=2, tol=0.0)
        assert unit_box_err(U, Ur) <= 1e-5


def _shuffled(P, T, seed):
    """The same island with its vertices in random order (a matrix bandwidth
    close to the vertex count)."""
    perm = np.random.default_rng(seed).permutation(len(P))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(P))
    return P[perm], inv[T]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("order", ["scan", "shuffled"])
def test_emu_resident_cg_equals_streaming_cg(disks, shape, precision, order):
    """The resident ARAP CG (rows in registers, u and e in LDS, halo windows
    filled from global memory) performs the streaming CG's arithmetic in the
    same order: both give the same layout bit for bit, for every team shape
    and precision.  The islands span 2 to 12 slices, so some members hold
    their rows resident and (with one wavefront per member) the largest island
    exceeds a lane's register slots and streams in both runs.  With shuffled
    vertices the largest island's bandwidth exceeds the halo window, so team
    members also gather from global memory."""
    from oracle import param_cpu as oracle

    if order == "shuffled":
        disks = [_shuffled(P, T, 40 + i) for i, (P, T) in enumerate(disks)]
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in disks])
    out = []
    for resident in (True, False):
        e = EmuPlan(disks, *shape, resident=resident, no_coarse=True)
        # all-FP32 stalls near its rounding floor, so it gets a looser tolerance
        out.append(e.arap(U0, 3, tol=0.0, cg_tol=1e-4 if precision == 1 else 1e-7, precision=precision))
        assert e.stats()["converged"].all()
    assert np.array_equal(out[0], out[1])


def test_emu_arap_batch_mode_vs_batch_oracle(disks):
    """gpu.Batch.arap semantics (free system, per-island frames)."""
    from oracle import param_cpu as oracle
    from oracle.batch_cpu import BatchOracle

    items = [(P, T, oracle.lscm(P, T)) for P, T in disks]
    r = BatchOracle(items)
    r.orient_positive()
    U0 = r.U.copy()
    e = EmuPlan(disks, 3, 2, batch=True)
    U = e.arap(U0, 3, line_search=True, batch=True, cg_tol=1e-8)
    r.arap(3, line_search=True)
    for i in range(len(disks)):
        s = slice(e.vo[i], e.vo[i + 1])
        assert unit_box_err(U[s], r.U[s]) <= 1e-4


def own_box(U):
    lo = U.min(0)
    return (U - lo) / float((U.max(0) - lo).max())


@pytest.mark.parametrize("shape", [(1, 1), (3, 2)])
def test_emu_slim_vs_reference_trajectory(shape):
    """SLIM task code against the reference's trajectory (golden fixture):
    energies rel 1e-9, layout 1e-8 in its own unit box, and the translation
    kept exactly (the mean of the layout does not move)."""
    from pathlib import Path

    d = dict(np.load(Path(__file__).parent / "golden" / "slim_golden.npz"))
    charts = [(d[f"P{c}"], d[f"T{c}"].astype(np.int64)) for c in (0, 1)]
    U0 = np.concatenate([d["U0_0"], d["U0_1"]])
    e = EmuPlan(charts, *shape)
    U = e.slim(U0, 8, tol=1e-6)
    st = e.stats()
    for c in (0, 1):
        s = slice(e.vo[c], e.vo[c + 1])
        ref_e = d[f"slim{c}_energy"]
        assert st["iters"][c] == len(ref_e) - 1
        assert abs(st["energy"][c] - ref_e[-1]) <= 1e-9 * ref_e[-1]
        assert np.abs(own_box(U[s]) - own_box(d[f"slim{c}"])).max() <= 1e-8
        assert np.abs((U[s] - U0[s]).mean(0)).max() <= 1e-12
        assert st["converged"][c] == 1 and st["rel_resid"][c] <= 1e-10


def test_emu_slim_flipped_start_and_early_exit(disks):
    """A start with a flipped triangle has infinite symmetric Dirichlet energy
    and is returned unchanged; a loose tol exits after the iteration whose
    energy change falls under it (both as in param.py:418-453)."""
    from oracle import param_cpu as oracle

    P, T = disks[1]
    U0 = oracle.refine(P, T, oracle.lscm(P, T), iters=2, tol=0.0)[0]
    Uf = U0.copy()
    a, b, c = (Uf[v].copy() for v in T[len(T) // 2])
    edge = c - b
    foot = b + edge * ((a - b) @ edge) / (edge @ edge)
    Uf[T[len(T) // 2][0]] = 2.0 * foot - a        # mirror the corner across its opposite edge
    assert (oracle.signed_areas(Uf, T) <= 0).any()
    e = EmuPlan([(P, T)])
    assert np.array_equal(e.slim(Uf, 5), Uf)
    assert e.stats()["iters"][0] == 0
    for tol in (1e-1, 1e-3):
        trace = []
        Ur, _ = oracle.refine(P, T, U0, iters=20, method="slim", tol=tol, trace=trace)
        U = e.slim(U0, 20, tol=tol)
        assert e.stats()["iters"][0] == len(trace) < 20
        assert np.abs(own_box(U) - own_box(Ur)).max() <= 1e-8


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1), (3, 2)])
def test_emu_energy_backward(disks, kind, shape):
    """Weighted energy and its hand-written backward (task code) against the
    float64 torch autograd reference (rel 1e-12) and central differences
    (h = 1e-3 along a mesh-scaled direction; their O(h^2) truncation on the
    symmetric Dirichlet energy is ~1e-5 relative, so rel 1e-4)."""
    import torch

    import diffref
    from oracle import param_cpu as oracle

    rng = np.random.default_rng(11 + kind)
    e = EmuPlan(disks[:2], *shape)
    U0 = np.concatenate([oracle.refine(P, T, oracle.lscm(P, T), iters=2, tol=0.0)[0] for P, T in disks[:2]])
    w = rng.uniform(0.5, 2.0, int(e.to[-1]))
    ge = rng.uniform(0.5, 1.5, e.k)
    E = e.energy_fwd(U0, w, kind)
    gU, gw = e.energy_bwd(U0, ge, w, kind)
    for i, (P, T) in enumerate(disks[:2]):
        sv, st = slice(e.vo[i], e.vo[i + 1]), slice(e.to[i], e.to[i + 1])
        X, area = oracle.local_frames(P, T)
        Ut = torch.tensor(U0[sv], requires_grad=True)
        wt = torch.tensor(w[st], requires_grad=True)
        Et = diffref.energy(Ut, T, X, area, wt, kind)
        (ge[i] * Et).backward()
        assert abs(E[i] - Et.item()) <= 1e-12 * abs(Et.item())
        assert np.abs(gU[sv] - Ut.grad.numpy()).max() <= 1e-12 * np.abs(Ut.grad.numpy()).max()
        assert np.abs(gw[st] - wt.grad.numpy()).max() <= 1e-12 * np.abs(wt.grad.numpy()).max()
        D = np.zeros_like(U0)
        D[sv] = diffref.mesh_direction(U0[sv], T, rng)
        h = 1e-3
        fd = (e.energy_fwd(U0 + h * D, w, kind)[i] - e.energy_fwd(U0 - h * D, w, kind)[i]) / (2 * h)
        an = (gU[sv] * D[sv]).sum() / ge[i]
        assert abs(fd - an) <= 1e-4 * abs(an)


def test_emu_global_step_backward(disks):
    """Backward of two unrolled refine global steps (adjoint CG + rhs/rotation
    backward, task code) against float64 torch autograd through dense solves
    and against central differences of the emulated steps."""
    import torch

    import diffref
    from oracle import param_cpu as oracle

    rng = np.random.default_rng(21)
    P, T = disks[1]
    U0 = oracle.refine(P, T, oracle.lscm(P, T), iters=2, tol=0.0)[0]
    e = EmuPlan([(P, T)])
    step = diffref.Step(P, T)
    gbar = rng.normal(size=U0.shape)

    def emu_steps(U):
        for _ in range(2):
            b = e.local_rhs(U)
            U = e.cg(b, U, tol=1e-13, precision=2)
        return U

    U1 = emu_steps(U0)
    Ut = torch.tensor(U0, requires_grad=True)
    Ur = step(step(Ut))
    assert np.abs(U1 - Ur.detach().numpy()).max() <= 1e-9 * np.abs(U1).max()
    (Ur * torch.as_tensor(gbar)).sum().backward()
    # backward: per step, the adjoint solve then the step backward at that step's input
    U_1 = e.cg(e.local_rhs(U0), U0, tol=1e-13, precision=2)
    g1 = e.step_bwd(U_1, e.adjoint_cg(gbar))
    g0 = e.step_bwd(U0, e.adjoint_cg(g1))
    ref = Ut.grad.numpy()
    assert np.abs(g0 - ref).max() <= 1e-9 * np.abs(ref).max()
    D = diffref.mesh_direction(U0, T, rng)
    h = 1e-3
    fd = ((emu_steps(U0 + h * D) - emu_steps(U0 - h * D)) * gbar).sum() / (2 * h)
    assert abs(fd - (g0 * D).sum()) <= 1e-5 * abs(fd)


def test_emu_hier_iterations_vs_oracle():
    """UVA_MODE_HIER (hier.Solver.relax's iterations: Jacobi-PCG of exactly
    cg_iters iterations, plain step for folded charts, flip-safe step over
    usable faces otherwise) against the restatement of hier.py with the
    repairs switched off, on one chart with a folded start."""
    from oracle import hier_cpu

    rng = np.random.default_rng(61)
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    P, T = workloads.disk_island(rng, 900)
    U0 = oracle.lscm(P, T)
    U0 = U0 + rng.normal(0, 0.01 * np.ptp(U0), U0.shape)
    assert (oracle.signed_areas(U0, T) < 0).any()
    S = hier_cpu.Solver(P, T)
    Ur, _, _ = S.relax(np.arange(len(T)), T, len(P), np.zeros(len(T), np.int64), 1, U0, iters=4, cg_iters=20,
                       repair_rounds=0)
    e = EmuPlan([(P, T)], batch=True)
    U = e.arap(U0.astype(np.float32).astype(np.float64), 4, line_search=True, tol=0.0, cg_tol=0.0,
               cg_max_iters=20, mode=2)
    assert unit_box_err(U, Ur) <= 1e-5


def test_emu_chart_pins_coupled_cg():
    """Several charts solved as one system (hier.Solver.relax / Batch): a
    one-island plan holding two charts, soft pins at each chart's first
    vertex with the island's weight max |W| (uva_plan_set_pins), and exactly
    20 Jacobi-PCG iterations shared by both charts, against plain PCG on the
    same block-diagonal matrix (hier.py:405-417, 465-474)."""
    import scipy.sparse as sp

    from oracle import hier_cpu
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(62)
    a, b = workloads.disk_island(rng, 500), workloads.disk_island(rng, 900)
    P = np.concatenate([a[0], b[0] + 5.0])
    T = np.concatenate([a[1], b[1] + len(a[0])])
    starts = np.array([0, len(a[0])])
    U0 = np.concatenate([oracle.lscm(*a), oracle.lscm(*b)])
    e = EmuPlan([(P, T)], batch=True, keep_order=True)
    e.set_pins(starts, 0.0)
    rhs = e.local_rhs(U0)
    x = e.cg(rhs, U0, tol=0.0, max_iters=20, precision=2)
    X, w, ok, area, x1, x2, y2 = hier_cpu.frames(P, T)
    W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
    I = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    J = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    wpin = np.abs(W).max()
    n = len(P)
    L = sp.csr_matrix((np.concatenate([-W, -W, W, W, [wpin, wpin]]),
                       (np.concatenate([I, J, I, J, starts]), np.concatenate([J, I, I, J, starts]))), shape=(n, n))
    dinv = 1.0 / L.diagonal()[:, None]
    xr = U0.copy()
    r = rhs - L @ xr
    z = r * dinv
    p = z.copy()
    rz = (r * z).sum(0)
    for _ in range(20):
        Ap = L @ p
        al = rz / (p * Ap).sum(0)
        xr = xr + p * al
        r = r - Ap * al
        z = r * dinv
        rz2 = (r * z).sum(0)
        p = z + p * (rz2 / rz)
        rz = rz2
    assert np.abs(x - xr).max() <= 1e-9 * np.abs(xr).max()
    # one shared CG differs from per-chart CGs after 20 iterations
    e2 = EmuPlan([a, (b[0] + 5.0, b[1])], batch=True, keep_order=True)
    x2 = e2.cg(rhs, U0, tol=0.0, max_iters=20, precision=2)
    assert np.abs(x2 - xr).max() > 1e-6 * np.abs(xr).max()


def _batch_pcg_reference(L, rhs, x0, max_iters, check_every, tol2, strict):
    """Plain Jacobi-PCG with the reference's step lengths and stop test
    (gpu.py:184-202): (x, iterations run)."""
    dinv = 1.0 / np.maximum(L.diagonal(), 1e-30)[:, None]
    x = x0.copy()
    r = rhs - L @ x
    z = r * dinv
    p = z.copy()
    rz = (r * z).sum(0)
    b2 = np.maximum((rhs * rhs).sum(0), 1e-30)
    for it in range(max_iters):
        Ap = L @ p
        al = rz / np.maximum((p * Ap).sum(0), 1e-30)
        x = x + p * al
        r = r - Ap * al
        z = r * dinv
        rr = (r * r).sum(0)
        if it % check_every == 0 and (np.all(rr < tol2 * b2) if strict else np.all(rr <= tol2 * b2)):
            return x, it + 1
        rz2 = (r * z).sum(0)
        p = z + p * (rz2 / np.maximum(rz, 1e-30))
        rz = rz2
    return x, max_iters


@pytest.mark.parametrize("cg_iters,tol,strict", [(20, 0.0, True), (40, 1e-6, False), (400, 1e-9, False)])
def test_emu_plan_pcg_matches_the_batch_cg(cg_iters, tol, strict):
    """uva_plan_pcg (the global step of gpu.Batch.arap and hier.Solver.relax
    on the chart plan's own matrix: SpMV on z, A p by recurrence) against
    plain Jacobi-PCG on the batch's pinned Laplacian assembled by scipy from
    the reference's triplets (gpu.py:126-142): three charts in one plan, a
    soft pin at every chart's first vertex, a fixed count (hier), the Batch
    stop rule at 40 iterations, and a run to convergence."""
    import scipy.sparse as sp

    from oracle import hier_cpu
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(64)
    charts = [workloads.disk_island(rng, t) for t in (300, 800, 1400)]
    U0 = np.concatenate([oracle.lscm(P, T) for P, T in charts])
    e = EmuPlan(charts, batch=True)
    starts = e.vo[:-1]
    e.set_pins(starts, 0.0)
    rhs = e.local_rhs(U0)
    P = np.concatenate([c[0] for c in charts])
    T = np.concatenate([c[1] + e.vo[i] for i, c in enumerate(charts)])
    _, w, *_ = hier_cpu.frames(P, T)
    W = np.concatenate([w[:, 0], w[:, 1], w[:, 2]])
    I = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    J = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    wpin = max(np.abs(W).max(), 1e-12)
    n = len(P)
    L = sp.csr_matrix((np.concatenate([-W, -W, W, W, np.full(len(starts), wpin)]),
                       (np.concatenate([I, J, I, J, starts]), np.concatenate([J, I, I, J, starts]))), shape=(n, n))
    xr, itr = _batch_pcg_reference(L, rhs, U0, cg_iters, 1, tol * tol, strict)
    x, its = e.plan_pcg(rhs, U0, cg_iters, check_every=1, tol2=tol * tol, strict=strict)
    assert its == itr
    assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()


def test_emu_batch_iteration_sequence_vs_batch_oracle():
    """One gpu.Batch.arap iteration as the product runs it -- rotations and
    rhs (uva_local_rhs), the coupled CG of 40 iterations on the plan's own
    matrix (uva_plan_pcg), then non-finite revert, flip-safe step and update
    (uva_step_apply) -- three times, against the Batch oracle's arap with the
    reference's coupled CG (gpu.py:184-202, 230-249)."""
    from oracle import param_cpu as oracle
    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(66)
    charts = [workloads.disk_island(rng, t) for t in (400, 900, 1600)]
    items = [(P, T, oracle.lscm(P, T)) for P, T in charts]
    r = BatchOracle(items)
    r.orient_positive()
    U = r.U.copy()
    e = EmuPlan(charts, batch=True)
    e.set_pins(e.vo[:-1], 0.0)
    for _ in range(3):
        b = e.local_rhs(U)
        Un, _ = e.plan_pcg(b, U, 40, check_every=1, tol2=1e-12, strict=False)
        U = e.step_apply(U, Un, mode=1, line_search=True)
    r.arap(3, line_search=True, cg_iters=40)
    assert np.abs(U - r.U).max() <= 1e-9 * np.abs(r.U).max()


@pytest.mark.parametrize("mode", [1, 2])
def test_emu_step_apply_reverts_and_steps_per_chart(mode):
    """uva_step_apply per chart: a chart with a non-finite row keeps its
    layout; the others move min(1, 0.9 x the flip-safe limit) of the way to
    the step's result (gpu.py:235-247); in hier mode a chart with a folded
    usable triangle takes the whole step (hier.py:500-505)."""
    from oracle import param_cpu as oracle
    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(67)
    charts = [workloads.disk_island(rng, t) for t in (300, 500, 700)]
    items = [(P, T, oracle.lscm(P, T)) for P, T in charts]
    r = BatchOracle(items)
    r.orient_positive()
    U = r.U.copy()
    vo = np.cumsum([0] + [len(P) for P, _ in charts])
    # chart 2 folded: one of its triangles turned over
    T2 = charts[2][1]
    f = T2[len(T2) // 2]
    U[vo[2] + f[0]], U[vo[2] + f[1]] = U[vo[2] + f[1]].copy(), U[vo[2] + f[0]].copy()
    Un = U + rng.normal(0, 0.05, U.shape) * np.ptp(U, 0)
    Un[vo[1] + 3] = np.nan                                  # chart 1 non-finite
    e = EmuPlan(charts, batch=True)
    got = e.step_apply(U, Un, mode=mode, line_search=True)
    D = Un - U
    D[vo[1]:vo[2]] = 0.0
    step = np.minimum(0.9 * r.max_step(U, np.nan_to_num(D)), 1.0)
    if mode == 2:
        assert r.flips_per_chart(U)[2] > 0
        step[2] = 1.0
    want = U + D * step[r.cv][:, None]
    assert np.array_equal(got[vo[1]:vo[2]], U[vo[1]:vo[2]])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_emu_folded_starts_run_more_iterations():
    """charts.flatten's ARAP (charts.py:411-415) for several charts in one
    call: a chart whose start is folded runs iters_folded iterations without
    line search, the others iters with it (per-island cap and use_ls), as the
    oracle's flatten_lscm_arap does chart by chart."""
    from pathlib import Path

    from oracle import param_cpu as oracle

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "flatten_golden.npz"))
    pick = [0, 17, 5]                          # chart 17 has a folded LSCM start
    charts = [(g[f"P{i}"], g[f"T{i}"].astype(np.int64)) for i in pick]
    U0 = []
    for P, T in charts:
        U = oracle.lscm(P, T)
        sa = oracle.signed_areas(U, T)
        if (sa < 0).sum() > (sa > 0).sum():
            U[:, 1] = -U[:, 1]
        U0.append(U)
    folded = [bool((oracle.signed_areas(u, T) <= 0).any()) for u, (_, T) in zip(U0, charts)]
    assert folded == [False, True, False]
    e = EmuPlan(charts)
    U = e.arap(np.concatenate(U0), 8, line_search=True, tol=1e-4, cg_tol=1e-10, iters_folded=16)
    st = e.stats()
    for i, (P, T) in enumerate(charts):
        s = slice(e.vo[i], e.vo[i + 1])
        tr = []
        Ur, _ = oracle.refine(P, T, U0[i], iters=16 if folded[i] else 8, line_search=not folded[i], tol=1e-4,
                              trace=tr)
        assert st["iters"][i] == len(tr)
        assert unit_box_err(U[s], Ur) <= 1e-6


def test_emu_overlap_flags_equal_restatement():
    """The GPU overlap test's pair code (validate.hip, run on the host)
    flags exactly the triangles the CPU restatement of validate.py flags, on
    the reference's own validation cases (chart_overlaps grids) and on noisy
    layouts in both modes (tolerant exact test and zero-tolerance crossings)."""
    from pathlib import Path

    from emu import overlap_flags
    from oracle import validate_cpu as vc

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "validate_golden.npz"))
    for i in range(len(g["c_names"])):
        U, T = g[f"c{i}_U"], g[f"c{i}_F"].astype(np.int64)
        lo = U.min(0)
        ext = float((U.max(0) - lo).max())
        Un = (U - lo) / ext * 0.999999
        u = Un[T]
        sa = 0.5 * ((u[:, 1, 0] - u[:, 0, 0]) * (u[:, 2, 1] - u[:, 0, 1]) - (u[:, 2, 0] - u[:, 0, 0]) * (u[:, 1, 1] - u[:, 0, 1]))
        skip = np.abs(sa) <= 1e-14
        G = int(np.clip(np.sqrt(len(T)) / 2, 4, 2048))
        assert np.array_equal(overlap_flags(Un, T, skip, 1e-6, 0, G), g[f"c{i}_mask"])
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(71)
    for k in range(4):
        P, T = workloads.disk_island(rng, 1500)
        U = (P[:, :2] - P[:, :2].min(0)) / np.ptp(P[:, :2]).max() * 0.999
        U = U + rng.normal(0, 0.002 * (k + 1), U.shape)
        U = np.clip(U, 0.0, 1.0)
        skip = rng.random(len(T)) < 0.05
        for G in (8, 64):
            assert np.array_equal(overlap_flags(U, T, skip, 1e-6, 0, G), vc.overlaps(U, T, skip, 1e-6))
            assert np.array_equal(overlap_flags(U, T, skip, 0.0, 1, G), vc.overlaps_strict(U, T, skip))


def test_emu_jacobi_pcg_iteration_and_stop_rule():
    """uva_jacobi_pcg's row code run on the host: the Jacobi-PCG iterates of
    the definition (alpha = rz / max(p.Ap, 1e-30), beta = rz' / max(rz,
    1e-30), diagonal summed over repeated entries) on a random symmetric
    positive definite matrix with repeated entries, stopping at the first
    checked iteration (every check_every-th from 0) with r.r < tol2 b.b, or
    after max_iters."""
    import scipy.sparse as sp

    rng = np.random.default_rng(77)
    n = 300
    i = rng.integers(0, n, 1500)
    j = rng.integers(0, n, 1500)
    v = rng.uniform(-1, 0, 1500)
    A = sp.coo_matrix((np.r_[v, v], (np.r_[i, j], np.r_[j, i])), shape=(n, n)).tocsr()
    d = np.asarray(abs(A).sum(1)).ravel() + rng.uniform(0.01, 0.1, n)
    # the diagonal in two entries per row (the kernel sums repeated entries)
    A = sp.csr_matrix((np.r_[A.tocoo().data, 0.5 * d, 0.5 * d],
                       (np.r_[A.tocoo().row, np.arange(n), np.arange(n)],
                        np.r_[A.tocoo().col, np.arange(n), np.arange(n)])), shape=(n, n))
    A.has_canonical_format = False
    assert A.nnz > 1500
    b = rng.normal(size=n)
    x0 = rng.normal(size=n)
    As = A.copy()
    As.sum_duplicates()
    dinv = 1.0 / As.diagonal()

    def pcg(b, x0, iters, check, tol2, strict=True):
        """the definition, per column (columns share only the stop test)"""
        B = b.reshape(n, -1)
        x = x0.reshape(n, -1).copy()
        r = B - As @ x
        z = r * dinv[:, None]
        p = z.copy()
        rz = (r * z).sum(0)
        b2 = np.maximum((B * B).sum(0), 1e-30)
        for k in range(iters):
            Ap = As @ p
            al = rz / np.maximum((p * Ap).sum(0), 1e-30)
            x, r = x + al * p, r - al * Ap
            rr = (r * r).sum(0)
            if k % check == 0 and ((rr < tol2 * b2) if strict else (rr <= tol2 * b2)).all():
                return x.reshape(x0.shape), k + 1
            z = r * dinv[:, None]
            rz2 = (r * z).sum(0)
            p = z + rz2 / np.maximum(rz, 1e-30) * p
            rz = rz2
        return x.reshape(x0.shape), iters

    for iters, check, tol2 in ((25, 50, 1e-12), (2000, 50, 1e-12), (2000, 1, 1e-12), (2000, 7, 1e-20)):
        x, k = emu.jacobi_pcg(A, b, x0, iters, check, tol2)
        xr, kr = pcg(b, x0, iters, check, tol2)
        assert k == kr and np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max(), (iters, check, k, kr)
        if iters == 2000 and check > 1 and tol2 > 1e-15:
            assert k % check == 1 or check == 1
    # two columns with their own step lengths, the Batch rule (every
    # iteration, both columns at most tol^2 b.b), and a fixed count
    b2c = np.stack([b, rng.normal(size=n) * 1e3], 1)
    x2c = np.stack([x0, rng.normal(size=n)], 1)
    for iters, tol2, strict in ((40, 1e-12, False), (2000, 1e-16, False), (20, 0.0, True)):
        x, k = emu.jacobi_pcg(A, b2c, x2c, iters, 1, tol2, strict)
        xr, kr = pcg(b2c, x2c, iters, 1, tol2, strict)
        assert k == kr and np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max(), (iters, k, kr)
    x1, _ = emu.jacobi_pcg(A, b2c[:, 1], x2c[:, 1], 20, 1, 0.0, True)
    assert np.abs(x1 - x[:, 1]).max() <= 1e-12 * np.abs(x1).max()      # columns do not mix
    x, k = emu.jacobi_pcg(A, b, x0, 0)
    assert k == 0 and np.array_equal(x, x0)


def test_emu_team_barrier_guard_stops_waiting_after_a_fault():
    """The team barrier's wait (team_wait, the loop of TeamWG::sync) with a
    member that never arrives: the first barrier gives up after the spin
    limit and sets the fault flag, and every later barrier then returns at
    once, so the launch ends in about one spin limit instead of one per
    barrier.  With every member present nothing is flagged."""
    spin, barriers = 20000, 200
    ticks, fault = emu.team_guard(4, 2, barriers, spin)
    assert fault == 1
    assert spin < ticks < spin + 20 * barriers
    ticks, fault = emu.team_guard(4, -1, barriers, spin)
    assert fault == 0 and ticks < spin


def test_emu_gutter_raster_equals_the_restatement():
    """The per-texel code of uva_raster_last (host build) equals the
    restatement of metrics.py's conservative raster on the gutter golden
    layouts and on a random atlas with huge, tiny and non-finite corners."""
    from pathlib import Path

    from oracle import validate_cpu as vc

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "gutter_golden.npz"))
    for i in range(len(g["names"])):
        U = np.asarray(g[f"g{i}_U"], np.float32).astype(np.float64)
        F = g[f"g{i}_F"].astype(np.int64)
        for res in (64, 128, 257):
            assert np.array_equal(emu.raster_last(U, F, res), vc.raster_last(U, F, res)), (g["names"][i], res)
    rng = np.random.default_rng(3)
    U = rng.random((600, 2)) * 1.2 - 0.1
    U[5] = [np.nan, 0.5]
    U[6] = [1e30, 0.2]
    U[7] = [-1e25, 0.9]
    F = rng.integers(0, 600, (400, 3))
    assert np.array_equal(emu.raster_last(U, F, 200), vc.raster_last(U, F, 200))
