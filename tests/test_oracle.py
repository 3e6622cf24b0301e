"""Pins of the CPU oracle (oracle/param_cpu.py, oracle/batch_cpu.py,
oracle/validate_cpu.py).

Each check is chosen so that a plausible restatement mistake (a dropped
factor, a wrong sign, a swapped index or corner) fails it: closed forms,
identities that hold only for the correct formula (Dirichlet energy, energy
gradients by finite differences), library routines (SVD polar factor, dense
solves) and brute force on tiny inputs.  Trajectory-level behaviour that
only the reference itself pins is covered by tests/test_golden.py.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import param_cpu as pc
from oracle.batch_cpu import BatchOracle
from ultimate_uv_artist_amd import workloads


def planar_disk(seed=0, t=400):
    P, T = workloads.disk_island(np.random.default_rng(seed), t)
    P = P.copy()
    P[:, 2] = 0.0
    return P, T


@pytest.fixture(scope="module")
def bumpy():
    return workloads.disk_island(np.random.default_rng(1), 600)


# ------------------------------------------------------------------ geometry
def test_local_frames_preserve_edges_and_area(bumpy):
    P, T = bumpy
    X, area = pc.local_frames(P, T)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        l3 = np.linalg.norm(P[T[:, b]] - P[T[:, a]], axis=1)
        l2 = np.linalg.norm(X[:, b] - X[:, a], axis=1)
        assert np.allclose(l2, l3, rtol=1e-12)
    cr = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    assert np.allclose(area, 0.5 * np.linalg.norm(cr, axis=1), rtol=1e-12)
    assert np.allclose(area, 0.5 * X[:, 1, 0] * X[:, 2, 1], rtol=1e-10)
    assert (X[:, 2, 1] > 0).all() and np.all(X[:, 0] == 0) and np.all(X[:, 1, 1] == 0)


def test_cot_weights_are_cotangents_of_opposite_angles(bumpy):
    P, T = bumpy
    X, _ = pc.local_frames(P, T)
    c = pc.cot_weights(X)
    for k in range(3):
        u = P[T[:, (k + 1) % 3]] - P[T[:, k]]
        v = P[T[:, (k + 2) % 3]] - P[T[:, k]]
        ang = np.arccos(np.clip((u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1), -1, 1))
        assert np.allclose(c[:, k], 1.0 / np.tan(ang), rtol=1e-7, atol=1e-9)


def test_laplacian_is_the_dirichlet_energy(bumpy):
    """x^T L x equals the integral of |grad x|^2 of the piecewise-linear x."""
    P, T = bumpy
    n = len(P)
    X, area = pc.local_frames(P, T)
    L = pc.cot_laplacian(n, T, pc.cot_weights(X))
    assert abs(L - L.T).max() < 1e-12
    assert np.abs(np.asarray(L.sum(1))).max() < 1e-9
    rng = np.random.default_rng(0)
    x = rng.normal(size=n)
    x1, x2, y2 = X[:, 1, 0], X[:, 2, 0], X[:, 2, 1]
    gx = (x[T[:, 1]] - x[T[:, 0]]) / x1
    gy = ((x[T[:, 2]] - x[T[:, 0]]) - gx * x2) / y2
    assert x @ (L @ x) == pytest.approx(float((area * (gx * gx + gy * gy)).sum()), rel=1e-10)


def test_laplacian_linear_precision_on_planar_mesh():
    P, T = planar_disk()
    n = len(P)
    X, _ = pc.local_frames(P, T)
    L = pc.cot_laplacian(n, T, pc.cot_weights(X))
    bi, bj = pc.boundary_edges_oriented(T)
    interior = np.ones(n, bool)
    interior[bi] = interior[bj] = False
    f = 0.3 * P[:, 0] - 1.7 * P[:, 1] + 0.5
    assert np.abs((L @ f)[interior]).max() < 1e-10


def test_boundary_edges_and_loops_on_grid():
    P, T = workloads.grid_island(nx=6)
    bi, bj = pc.boundary_edges_oriented(T)
    assert len(bi) == 4 * 5
    loops = pc.boundary_loops(T)
    assert len(loops) == 1 and len(loops[0]) == 20
    lp = loops[0]
    pairs = set(zip(bi.tolist(), bj.tolist()))
    assert all((lp[i], lp[(i + 1) % len(lp)]) in pairs for i in range(len(lp)))


def test_signed_areas_known_triangle():
    U = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 3.0]])
    assert pc.signed_areas(U, np.array([[0, 1, 2]]))[0] == 3.0
    assert pc.signed_areas(U, np.array([[0, 2, 1]]))[0] == -3.0


# ------------------------------------------------------------------ LSCM / solver
def test_lscm_reproduces_a_planar_disk():
    """A planar chart has zero conformal energy: LSCM returns it up to the
    rigid motion fixed by the pins p0 -> (0,0), p1 -> (|p1-p0|, 0)."""
    P, T = planar_disk(3, 500)
    U = pc.lscm(P, T)
    bi, _ = pc.boundary_edges_oriented(T)
    p0, p1 = pc.far_pair(P, np.unique(bi))
    d = P[p1, :2] - P[p0, :2]
    c, s = d / np.linalg.norm(d)
    R = np.array([[c, s], [-s, c]])
    Uref = (P[:, :2] - P[p0, :2]) @ R.T
    assert np.abs(U - Uref).max() < 1e-9


def test_lscm_matches_dense_solve(bumpy):
    P, T = bumpy
    H, rhs, keep, pin_val, _ = pc.lscm_system(P, T)
    x = np.linalg.solve(H.toarray(), rhs)
    U = pc.lscm(P, T)
    n = len(P)
    full = pin_val.copy()
    full[keep] = x
    assert np.abs(U - np.stack([full[:n], full[n:]], 1)).max() < 1e-9
    # H is the conformal energy: symmetric positive definite after pinning
    Hd = H.toarray()
    assert np.abs(Hd - Hd.T).max() < 1e-12
    assert np.linalg.eigvalsh(Hd).min() > 0


def test_lscm_raises_on_closed_chart():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    T = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    with pytest.raises(ValueError):
        pc.lscm(P, T)


def test_spdsolver_direct_and_cg_agree(bumpy):
    P, T = bumpy
    n = len(P)
    X, _ = pc.local_frames(P, T)
    L = pc.cot_laplacian(n, T, pc.cot_weights(X)).tolil()
    L[0, 0] += 1.0
    L = L.tocsc()
    b = np.random.default_rng(2).normal(size=(n, 2))
    s = pc.SPDSolver(L)
    x = s.solve(b)
    assert np.abs(L @ x - b).max() < 1e-9
    old = pc.DIRECT_LIMIT
    try:
        pc.DIRECT_LIMIT = 0
        xc = pc.SPDSolver(L).solve(b, tol=1e-10, maxiter=20000)
    finally:
        pc.DIRECT_LIMIT = old
    assert np.abs(xc - x).max() < 1e-6 * np.abs(x).max()
    assert np.abs(spla.spsolve(L, b) - x).max() < 1e-9 * np.abs(x).max()


# ------------------------------------------------------------------ local step pieces
def _rand_J(rng, m):
    J = rng.normal(size=(m, 2, 2))
    neg = np.linalg.det(J) < 0
    J[neg, :, 0] *= -1
    return J


def _layout_with_jacobians(J):
    """One triangle per Jacobian with frame x1=1, x2=0.3, y2=0.8 and U = J X."""
    m = len(J)
    X = np.zeros((m, 3, 2))
    X[:, 1, 0] = 1.0
    X[:, 2] = [0.3, 0.8]
    U = np.einsum("tij,tkj->tki", J, X).reshape(-1, 2)
    T = np.arange(3 * m).reshape(m, 3)
    return U, T, X


def test_rotation_is_the_polar_factor():
    rng = np.random.default_rng(4)
    J = _rand_J(rng, 200)
    U, T, X = _layout_with_jacobians(J)
    R, Wm = pc.arap_rotations(U, T, X, None, 0)
    Uu, _, Vt = np.linalg.svd(J)
    assert np.abs(R - Uu @ Vt).max() < 1e-10
    assert np.all(Wm == np.eye(2))
    th = rng.uniform(0, 2 * np.pi, 20)
    Jr = np.stack([np.stack([np.cos(th), -np.sin(th)], 1), np.stack([np.sin(th), np.cos(th)], 1)], 1)
    U, T, X = _layout_with_jacobians(Jr)
    assert np.abs(pc.arap_rotations(U, T, X, None, 0)[0] - Jr).max() < 1e-12


def test_slim_weights_from_singular_values():
    rng = np.random.default_rng(5)
    J = _rand_J(rng, 100)
    U, T, X = _layout_with_jacobians(J)
    R, Wm = pc.arap_rotations(U, T, X, None, 1)
    Uu, S, Vt = np.linalg.svd(J)
    assert np.abs(R - Uu @ Vt).max() < 1e-10
    w = np.sqrt((S - S ** -3) / (S - 1))
    # W has the left singular vectors as eigenvectors and w(s) as eigenvalues
    for f in range(len(J)):
        for k in range(2):
            assert np.allclose(Wm[f] @ Uu[f][:, k], w[f, k] * Uu[f][:, k], rtol=1e-8, atol=1e-10)


def test_rhs_is_the_negative_energy_gradient(bumpy):
    """With R fixed, E(U) = 1/2 sum w_k |u_i - u_j - R d_k|^2 has gradient L U - b."""
    P, T = bumpy
    n = len(P)
    X, _ = pc.local_frames(P, T)
    cots = pc.cot_weights(X)
    rng = np.random.default_rng(6)
    U = pc.lscm(P, T) + rng.normal(0, 0.01, (n, 2))
    R, _ = pc.arap_rotations(U, T, X, cots, 0)
    b = pc.arap_rhs(n, T, X, cots, R)
    L = pc.cot_laplacian(n, T, cots)

    def E(V):
        e = 0.0
        for k in range(3):
            i, j = (k + 1) % 3, (k + 2) % 3
            r = V[T[:, i]] - V[T[:, j]] - np.einsum("tab,tb->ta", R, X[:, i] - X[:, j])
            e += 0.5 * (0.5 * cots[:, k] * (r * r).sum(1)).sum()
        return e

    g = L @ U - b
    for v, c in [(3, 0), (17, 1), (int(T[5, 2]), 0)]:
        h = 1e-6
        Up, Um = U.copy(), U.copy()
        Up[v, c] += h
        Um[v, c] -= h
        assert (E(Up) - E(Um)) / (2 * h) == pytest.approx(g[v, c], rel=1e-5, abs=1e-8)


def test_max_step_brute_force():
    rng = np.random.default_rng(8)
    P, T = planar_disk(8, 150)
    U = P[:, :2].copy()
    for trial in range(5):
        D = rng.normal(0, 0.05, U.shape)
        tm = pc.max_step(U, D, T)
        ts = np.linspace(0, min(tm, 50.0) * 1.05, 20001)
        first = None
        for t in ts:
            if (pc.signed_areas(U + t * D, T) <= 0).any():
                first = t
                break
        if tm < 50.0:
            assert first is not None and abs(first - tm) <= (ts[1] - ts[0]) * 1.01
        else:
            assert first is None
    assert pc.max_step(U, np.ones_like(U), T) == 1e30                   # translation never flips
    c = U.mean(0)
    assert pc.max_step(U, -(U - c), T) == pytest.approx(1.0, rel=1e-6)  # double root at 1


def test_energies_at_isometry_and_flip():
    P, T = planar_disk(9, 200)
    X, area = pc.local_frames(P, T)
    U = P[:, :2]
    assert pc.energy(U, T, X, area, 0) < 1e-20
    assert pc.energy(U, T, X, area, 1) < 1e-10
    assert pc.energy(U * 2.0, T, X, area, 0) == pytest.approx(2 * area.sum(), rel=1e-12)   # s1=s2=2
    Uf = U.copy()
    Uf[:, 1] *= -1
    assert pc.energy(Uf, T, X, area, 1) == np.inf


# ------------------------------------------------------------------ refine / repair / quality
def test_refine_isometry_is_a_fixed_point_and_pin_holds():
    P, T = planar_disk(10, 300)
    U0 = P[:, :2] + 0.5
    U, _ = pc.refine(P, T, U0, iters=3, tol=0.0, pin=7)
    assert np.abs(U - U0).max() < 1e-10


def test_refine_energy_decreases_and_stays_flip_free(bumpy):
    P, T = bumpy
    U0 = pc.lscm(P, T)
    trace = []
    U, _ = pc.refine(P, T, U0, iters=15, tol=0.0, trace=trace)
    e = [t[0] for t in trace]
    assert all(e[i + 1] <= e[i] * (1 + 1e-12) for i in range(len(e) - 1))
    assert (pc.signed_areas(U, T) > 0).all()
    trace2 = []
    pc.refine(P, T, U0, iters=50, tol=1e-3, trace=trace2)
    assert len(trace2) < 50
    assert abs(trace2[-1][0] - trace2[-2][0]) <= 1e-3 * trace2[-2][0]


def test_repair_moves_interior_vertex_to_ring_mean():
    # a fan around vertex 0 with vertex 0 pushed outside: one pass fixes it
    ang = np.linspace(0, 2 * np.pi, 7)[:-1]
    P = np.concatenate([[[0.0, 0.0, 0.0]], np.stack([np.cos(ang), np.sin(ang), 0 * ang], 1)])
    T = np.array([[0, 1 + k, 1 + (k + 1) % 6] for k in range(6)])
    U = P[:, :2].copy()
    U[0] = [1.5, 0.2]
    assert (pc.signed_areas(U, T) <= 0).any()
    out = pc.repair_flips(U, T, rounds=5)
    assert np.allclose(out[0], U[1:].mean(0))
    assert (pc.signed_areas(out, T) > 0).all()
    assert np.array_equal(out[1:], U[1:])


def test_weighted_quantile_and_quality_at_isometry():
    x = np.array([5.0, 1.0, 3.0, 2.0])
    w = np.array([0.1, 0.4, 0.3, 0.2])
    assert pc.weighted_quantile(x, w, 0.5) == 2.0
    assert pc.weighted_quantile(x, w, 0.95) == 5.0
    assert pc.weighted_quantile(x, w, 0.0) == 1.0
    P, T = planar_disk(11, 200)
    q = pc.chart_quality(P, T, P[:, :2] * 3.0)
    assert q["flips"] == 0
    assert q["aniso_p95"] == pytest.approx(1.0, abs=1e-9)
    assert q["area_p95"] == pytest.approx(1.0, abs=1e-9)
    assert q["scale"] == pytest.approx(3.0, rel=1e-12)


# ------------------------------------------------------------------ Batch semantics
def test_batch_oracle_equals_refine_semantics_on_one_chart(bumpy):
    """With exact solves, a single-chart Batch.arap with line search is
    param.refine without early exit (the batch-global pin weight changes the
    matrix, not the solution: the ARAP rhs sums to zero)."""
    P, T = bumpy
    U0 = pc.lscm(P, T)
    b = BatchOracle([(P, T, U0)])
    b.arap(5, line_search=True)
    U, _ = pc.refine(P, T, U0, iters=5, tol=-1.0)
    assert np.abs(b.U - U).max() < 1e-9 * np.ptp(U)


def test_batch_orient_and_flip_count():
    P, T = planar_disk(12, 200)
    U = P[:, :2].copy()
    U[:, 1] *= -1
    b = BatchOracle([(P, T, U)])
    assert b.flips_per_chart()[0] == len(T)
    b.orient_positive()
    assert b.flips_per_chart()[0] == 0
    assert np.allclose(b.U, P[:, :2])


# ------------------------------------------------------------ validation checks
def _clip_area(a, b):
    """Area of the intersection of triangles a, b (3x2 arrays) by
    Sutherland-Hodgman clipping of a against b's half-planes."""
    if (b[1, 0] - b[0, 0]) * (b[2, 1] - b[0, 1]) - (b[2, 0] - b[0, 0]) * (b[1, 1] - b[0, 1]) < 0:
        b = b[::-1]
    poly = [tuple(p) for p in a]
    for k in range(3):
        p, q = b[k], b[(k + 1) % 3]
        side = lambda x: (q[0] - p[0]) * (x[1] - p[1]) - (q[1] - p[1]) * (x[0] - p[0])
        out = []
        for i in range(len(poly)):
            c, d = np.array(poly[i]), np.array(poly[(i + 1) % len(poly)])
            sc, sd = side(c), side(d)
            if sc >= 0:
                out.append(tuple(c))
            if (sc >= 0) != (sd >= 0):
                out.append(tuple(c + (d - c) * (sc / (sc - sd))))
        poly = out
        if not poly:
            return 0.0
    x = np.array([p[0] for p in poly])
    y = np.array([p[1] for p in poly])
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_triangle_overlap_equals_positive_intersection_area():
    """On integer corners (orientation tolerances far below 1) and no shared
    corner positions: a reported overlap always has positive intersection
    area (computed independently by clipping), and in general position (no
    corner on the line of an edge of the other triangle) the test is exactly
    'positive intersection area'.  With corners on edges the reference's test
    can miss an overlap, e.g. (6,3),(2,3),(6,6) against (4,3),(6,4),(6,2),
    which share the area-1 triangle (4,3),(6,3),(6,4): no crossing is proper,
    no corner or centroid is strictly inside."""
    from oracle import validate_cpu as vc

    rng = np.random.default_rng(41)
    U = rng.integers(0, 7, (6 * 6000, 2)).astype(np.float64)
    F = np.arange(len(U)).reshape(-1, 3)
    u = U[F]
    sa = (u[:, 1, 0] - u[:, 0, 0]) * (u[:, 2, 1] - u[:, 0, 1]) - (u[:, 2, 0] - u[:, 0, 0]) * (u[:, 1, 1] - u[:, 0, 1])
    F = F[sa != 0]
    s, t = 2 * np.arange(len(F) // 2), 2 * np.arange(len(F) // 2) + 1
    Fp = F[: 2 * len(s)]
    same = (U[Fp[s]][:, :, None, :] == U[Fp[t]][:, None, :, :]).all(-1).any((1, 2))
    s, t = s[~same], t[~same]
    got = vc.tri_overlap_pairs(U, Fp, s, t, 1e-6)
    want = np.array([_clip_area(U[Fp[i]], U[Fp[j]]) > 1e-9 for i, j in zip(s, t)])
    assert not (got & ~want).any()
    ua, ub = U[Fp[s]], U[Fp[t]]

    def on_line(x, y):             # any corner of y on the line of an edge of x
        hit = np.zeros(len(x), bool)
        for i in range(3):
            a, b = x[:, i], x[:, (i + 1) % 3]
            for j in range(3):
                hit |= vc._orient(a[:, 0], a[:, 1], b[:, 0], b[:, 1], y[:, j, 0], y[:, j, 1]) == 0
        return hit

    gen = ~on_line(ua, ub) & ~on_line(ub, ua)
    assert got[gen].sum() > 200 and (~got[gen]).sum() > 200
    assert np.array_equal(got[gen], want[gen])
    A = np.array([[6.0, 3], [2, 3], [6, 6], [4, 3], [6, 4], [6, 2]])
    assert _clip_area(A[:3], A[3:]) == 1.0
    assert not vc.tri_overlap_pairs(A, np.array([[0, 1, 2], [3, 4, 5]]), [0], [1], 1e-6)[0]
    # symmetric in the pair and in the corner order of either triangle
    assert np.array_equal(vc.tri_overlap_pairs(U, Fp, t, s, 1e-6), got)
    Fr = Fp[:, [1, 2, 0]]
    assert np.array_equal(vc.tri_overlap_pairs(U, Fr, s, t, 1e-6), got)


def test_candidate_pairs_are_all_meeting_bounding_boxes():
    from oracle import validate_cpu as vc

    rng = np.random.default_rng(42)
    U = rng.random((300, 2))
    F = rng.integers(0, 300, (250, 3))
    skip = rng.random(250) < 0.1
    s, t = vc.candidate_pairs(U, F, skip)
    lo, hi = U[F].min(1), U[F].max(1)
    want = {(i, j) for i in range(250) for j in range(i + 1, 250)
            if not skip[i] and not skip[j] and (lo[i] <= hi[j]).all() and (lo[j] <= hi[i]).all()}
    got = set(zip(s.tolist(), t.tolist()))
    assert got == want and len(got) == len(s)


def test_chart_overlaps_invariant_under_similarity_and_reorder():
    from oracle import validate_cpu as vc

    rng = np.random.default_rng(43)
    P, T = planar_disk(44, 300)
    U = P[:, :2] + rng.normal(0, 0.03 * np.ptp(P[:, :2]), (len(P), 2))
    m = vc.chart_overlaps(U, T)
    assert m.any() and not m.all()
    perm = rng.permutation(len(T))
    assert np.array_equal(vc.chart_overlaps(U * 3.7 + np.array([5.0, -2.0]), T), m)
    assert np.array_equal(vc.chart_overlaps(U, T[perm]), m[perm])
    assert not vc.chart_overlaps(P[:, :2], T).any()          # a planar embedding


def test_validate_orientation_and_island_counts():
    from oracle import validate_cpu as vc

    P, T = planar_disk(45, 200)
    U = (P[:, :2] - P[:, :2].min(0)) / np.ptp(P[:, :2]).max() * 0.4 + 0.05
    rep, _ = vc.validate_uvs(P, T, U)
    assert rep["valid"] and rep["islands"] == 1 and rep["flipped_triangles"] == 0
    Um = U * np.array([-1.0, 1.0]) + np.array([1.0, 0.0])
    rep, _ = vc.validate_uvs(P, T, Um)
    assert rep["mirrored_islands"] == 1 and rep["flipped_triangles"] == 0 and rep["valid"]
    # a second copy of the chart beside the first: two islands; one triangle reversed
    P2 = np.r_[P, P + 10.0]
    T2 = np.r_[T, T[:, [0, 2, 1]] + len(P)]
    T2[len(T)] = T2[len(T)][[0, 2, 1]]
    U2 = np.r_[U, U + np.array([0.5, 0.0])]
    rep, prob = vc.validate_uvs(P2, T2, U2)
    assert rep["islands"] == 2 and rep["mirrored_islands"] == 1
    # the second copy has reversed winding everywhere except the one triangle
    assert rep["flipped_triangles"] == 1 and prob[len(T)] and not rep["valid"]
    isl, n = vc.uv_islands(T2, len(P2), U2, P2)
    assert n == 2 and len(set(isl[: len(T)])) == 1 and len(set(isl[len(T):])) == 1


def test_hier_lscm_keeps_a_conformal_start_and_moves_a_mirrored_one():
    """hier.Solver.lscm restatement: a planar chart whose start is a
    similarity of its own positions is a minimum of the Dirichlet-minus-area
    energy (zero energy, pins satisfied), so the converged solve returns it;
    its mirror image is anti-conformal and is moved.  A flipped sign of the
    area term, or an area term on the wrong boundary orientation, swaps the
    two outcomes."""
    from oracle import hier_cpu

    rng = np.random.default_rng(41)
    P, T = workloads.disk_island(rng, 600)
    P = P.copy()
    P[:, 2] = 0.0
    th, s = 0.7, 1.9
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) * s
    a = T.reshape(-1)
    b = T[:, [1, 2, 0]].reshape(-1)
    key = a * len(P) + b
    m = ~np.isin(b * len(P) + a, key)
    S = hier_cpu.Solver(P, T)
    args = (np.arange(len(T)), T, len(P), np.zeros(len(T), np.int64), 1)
    for mirror, moves in ((False, False), (True, True)):
        U0 = P[:, :2] @ R.T + np.array([0.3, -0.2])
        if mirror:
            U0[:, 0] = -U0[:, 0]
        U0 = U0.astype(np.float32).astype(np.float64)
        U = S.lscm(*args, U0, a[m], b[m], iters=20000, check_every=1)
        err = np.abs(U - U0).max() / np.ptp(U0, 0).max()
        assert (err > 1e-2) if moves else (err < 1e-7), (mirror, err)


def test_hier_lscm_pins_are_extreme_boundary_vertices():
    """The pins of each chart are the boundary vertices of smallest and
    largest start u: with a start far from conformal, the converged layout
    keeps those two vertices near their start (weight 10 max |W|) and no
    other vertex pinned."""
    from oracle import hier_cpu

    rng = np.random.default_rng(42)
    P, T = workloads.disk_island(rng, 400)
    a = T.reshape(-1)
    b = T[:, [1, 2, 0]].reshape(-1)
    m = ~np.isin(b * len(P) + a, a * len(P) + b)
    U0 = (P[:, :2] * np.array([3.0, 0.5])).astype(np.float32).astype(np.float64)
    bv = np.unique(a[m])
    lo, hi = bv[np.argmin(U0[bv, 0])], bv[np.argmax(U0[bv, 0])]
    S = hier_cpu.Solver(P, T)
    U = S.lscm(np.arange(len(T)), T, len(P), np.zeros(len(T), np.int64), 1, U0, a[m], b[m], iters=20000,
               check_every=1)
    d = np.linalg.norm(U - U0, axis=1) / np.ptp(U0, 0).max()
    assert d[lo] < 0.02 and d[hi] < 0.02 and np.median(d) > 5 * max(d[lo], d[hi])


def _neighbour_sets(T, n):
    nb = [set() for _ in range(n)]
    for a, b, c in np.asarray(T, np.int64).tolist():
        nb[a] |= {b, c}
        nb[b] |= {a, c}
        nb[c] |= {a, b}
    return nb


def test_tutte_is_harmonic_with_the_boundary_on_a_circle():
    """A chart whose boundary loop runs in vertex-id order (so the reference's
    pairing of loop-order positions with id-order columns is consistent):
    every interior vertex is the mean of its distinct neighbours, and the
    boundary vertices sit on the circle of the perimeter's radius at their
    arc-length angles, the first at angle 0."""
    from pathlib import Path

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "tutte_golden.npz"))
    P, T = g["ordered_P"], g["ordered_T"].astype(np.int64)
    U = pc.tutte(P, T)
    loop = max(pc.boundary_loops(T), key=len)
    assert np.array_equal(loop, np.sort(loop))
    seg = np.linalg.norm(P[np.roll(loop, -1)] - P[loop], axis=1)
    rad = seg.sum() / (2 * np.pi)
    assert np.allclose(np.hypot(U[loop, 0], U[loop, 1]), rad, rtol=1e-12)
    ang = np.mod(np.arctan2(U[loop, 1], U[loop, 0]), 2 * np.pi)
    want = np.r_[0.0, np.cumsum(seg)[:-1]] / rad
    assert np.allclose(ang[1:], want[1:], atol=1e-10) and abs(ang[0]) < 1e-12
    nb = _neighbour_sets(T, len(P))
    interior = np.setdiff1d(np.arange(len(P)), loop)
    mean = np.array([U[sorted(nb[v])].mean(0) for v in interior])
    assert np.abs(U[interior] - mean).max() <= 1e-10 * rad
    # a disk chart's harmonic map onto a convex boundary has no fold (Tutte)
    assert (pc.signed_areas(U, T) > 0).all() or (pc.signed_areas(U, T) < 0).all()


def test_tutte_rotated_loop_is_not_harmonic():
    """The same chart with its loop starting elsewhere: the boundary still lies
    on the circle, but the interior solve uses the id-order pairing, so the
    interior is no longer the mean of its neighbours' actual positions (the
    behaviour the reference's own output pins)."""
    from pathlib import Path

    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "tutte_golden.npz"))
    P, T = g["rotated_P"], g["rotated_T"].astype(np.int64)
    U = pc.tutte(P, T)
    loop = max(pc.boundary_loops(T), key=len)
    seg = np.linalg.norm(P[np.roll(loop, -1)] - P[loop], axis=1)
    assert np.allclose(np.hypot(U[loop, 0], U[loop, 1]), seg.sum() / (2 * np.pi), rtol=1e-12)
    nb = _neighbour_sets(T, len(P))
    interior = np.setdiff1d(np.arange(len(P)), loop)
    mean = np.array([U[sorted(nb[v])].mean(0) for v in interior])
    assert np.abs(U[interior] - mean).max() > 1e-3 * seg.sum()


def test_gutter_of_two_aligned_squares():
    """Two axis-aligned squares side by side on texel boundaries: the raster
    visits the texels of the floored bounding box and keeps those whose
    centre lies within 0.7072 texel of the triangle, so a square spanning
    [10, 30] texels fills columns 10-30 (centre 30.5 is half a texel out).
    Squares at [10, 30] and [40, 60] leave 9 empty columns between them; the
    measure (the smallest sum of distances to different islands over
    adjacent empty texels, minus one) is 8.  One square alone has no gap
    (inf); touching squares give 0."""
    from oracle import validate_cpu as vc

    sq = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    F = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]])
    res = 128
    for x2, want in ((40.0, 8.0), (31.0, 0.0)):
        U = np.r_[np.array([10.0, 40.0]) + 20.0 * sq, np.array([x2, 40.0]) + 20.0 * sq] / res
        gap, img = vc.min_gutter(U, F, np.array([0, 0, 1, 1]), res)
        assert gap == want
    cols = np.nonzero((img >= 0).any(0))[0]
    assert cols.min() == 10 and cols.max() == 51 and len(cols) == 42
    gap, _ = vc.min_gutter(U[:4], F[:2], np.array([0, 0]), res)
    assert gap == np.inf
    # last writer: where the two squares overlap the later triangles' island wins
    U = np.r_[np.array([10.0, 40.0]) + 20.0 * sq, np.array([20.0, 40.0]) + 20.0 * sq] / res
    _, img = vc.min_gutter(U, F, np.array([0, 0, 1, 1]), res)
    assert (img[45, 20:41] == 1).all() and (img[45, 10:20] == 0).all()
