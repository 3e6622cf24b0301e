"""GPU differentiation: energy backward and adjoint global-step backward
through the torch.autograd.Functions of ultimate_uv_artist_amd.diff, against
float64 torch autograd references (dense solves) and central differences
(h = 1e-3 along mesh-scaled directions, FP64)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def charts():
    from ultimate_uv_artist_amd import workloads

    rng = np.random.default_rng(17)
    return [workloads.disk_island(rng, t) for t in (300, 900, 2600)]


def _start(charts):
    from oracle import param_cpu as oracle

    return np.concatenate([oracle.refine(P, T, oracle.lscm(P, T), iters=2, tol=0.0)[0] for P, T in charts])


@pytest.mark.parametrize("kind", [0, 1])
def test_energy_autograd_matches_f64_reference(charts, kind):
    import torch

    import diffref
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import diff
    from ultimate_uv_artist_amd.plan import IslandPlan

    plan = IslandPlan(charts)
    rng = np.random.default_rng(5 + kind)
    U0 = _start(charts)
    w0 = rng.uniform(0.5, 2.0, plan.t)
    U = torch.tensor(U0, device=plan.dev, requires_grad=True)
    w = torch.tensor(w0, device=plan.dev, requires_grad=True)
    ge = torch.as_tensor(rng.uniform(0.5, 1.5, plan.n_islands), device=plan.dev)
    E = diff.island_energy(plan, U, w, kind)
    (E * ge).sum().backward()
    gU, gw = U.grad.cpu().numpy(), w.grad.cpu().numpy()
    for i, (P, T) in enumerate(charts):
        sv = slice(plan.vert_off[i], plan.vert_off[i + 1])
        st = slice(plan.tri_off[i], plan.tri_off[i + 1])
        X, area = oracle.local_frames(P, T)
        Ut = torch.tensor(U0[sv], requires_grad=True)
        wt = torch.tensor(w0[st], requires_grad=True)
        Et = diffref.energy(Ut, T, X, area, wt, kind)
        (float(ge[i]) * Et).backward()
        assert abs(float(E[i].detach()) - Et.item()) <= 1e-12 * abs(Et.item())
        assert np.abs(gU[sv] - Ut.grad.numpy()).max() <= 1e-11 * np.abs(Ut.grad.numpy()).max()
        assert np.abs(gw[st] - wt.grad.numpy()).max() <= 1e-11 * np.abs(wt.grad.numpy()).max()
        D = np.zeros_like(U0)
        D[sv] = diffref.mesh_direction(U0[sv], T, rng)
        Dt = torch.as_tensor(D, device=plan.dev)
        with torch.no_grad():
            ep = plan.energy_fwd((U + 1e-3 * Dt).contiguous(), w, kind)[i]
            em = plan.energy_fwd((U - 1e-3 * Dt).contiguous(), w, kind)[i]
        fd = float(ep - em) / 2e-3
        an = float((gU[sv] * D[sv]).sum()) / float(ge[i])
        assert abs(fd - an) <= 1e-4 * abs(an)


def test_unrolled_steps_adjoint_matches_f64_reference(charts):
    """Loss = weighted ARAP energy after 3 global steps; gradient wrt the
    start layout and the weights through the adjoint solves."""
    import torch

    import diffref
    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, diff
    from ultimate_uv_artist_amd.plan import IslandPlan

    plan = IslandPlan(charts)
    rng = np.random.default_rng(9)
    U0 = _start(charts)
    w0 = rng.uniform(0.5, 2.0, plan.t)
    U = torch.tensor(U0, device=plan.dev, requires_grad=True)
    w = torch.tensor(w0, device=plan.dev, requires_grad=True)
    U3 = diff.arap_unrolled(plan, U, 3, tol=1e-12, precision=_lib.UVA_PREC_F64, adjoint_tol=1e-12)
    loss = diff.island_energy(plan, U3, w, 0).sum()
    loss.backward()
    gU, gw = U.grad.cpu().numpy(), w.grad.cpu().numpy()
    for i, (P, T) in enumerate(charts[:2]):          # dense references on the two smaller islands
        sv = slice(plan.vert_off[i], plan.vert_off[i + 1])
        st = slice(plan.tri_off[i], plan.tri_off[i + 1])
        step = diffref.Step(P, T)
        Ut = torch.tensor(U0[sv], requires_grad=True)
        wt = torch.tensor(w0[st], requires_grad=True)
        U3t = step(step(step(Ut)))
        assert np.abs(U3.detach().cpu().numpy()[sv] - U3t.detach().numpy()).max() <= 1e-9
        diffref.energy(U3t, T, step.X, step.area, wt, 0).backward()
        assert np.abs(gU[sv] - Ut.grad.numpy()).max() <= 1e-7 * np.abs(Ut.grad.numpy()).max()
        assert np.abs(gw[st] - wt.grad.numpy()).max() <= 1e-9 * np.abs(wt.grad.numpy()).max()
    # central differences on the largest island
    i = 2
    sv = slice(plan.vert_off[i], plan.vert_off[i + 1])
    D = np.zeros_like(U0)
    D[sv] = diffref.mesh_direction(U0[sv], charts[i][1], rng)
    Dt = torch.as_tensor(D, device=plan.dev)
    with torch.no_grad():
        vals = []
        for sgn in (1.0, -1.0):
            Us = diff.arap_unrolled(plan, (U + sgn * 1e-3 * Dt).contiguous(), 3, tol=1e-12,
                                    precision=_lib.UVA_PREC_F64)
            vals.append(float(plan.energy_fwd(Us, w, 0)[i]))
    fd = (vals[0] - vals[1]) / 2e-3
    an = float((gU[sv] * D[sv]).sum())
    assert abs(fd - an) <= 1e-4 * abs(an)


def test_stretch_refine_lowers_weighted_stretch(charts):
    """200 gradient steps on the weighted symmetric Dirichlet energy: the
    energy never increases, ends lower, and no triangle folds."""
    import torch

    from ultimate_uv_artist_amd import diff
    from ultimate_uv_artist_amd.plan import IslandPlan

    plan = IslandPlan(charts)
    rng = np.random.default_rng(3)
    U = torch.tensor(_start(charts), device=plan.dev)
    w = torch.as_tensor(rng.uniform(0.5, 2.0, plan.t), device=plan.dev)
    U, hist = diff.stretch_refine(plan, U, w, steps=200)
    h = hist.cpu().numpy()
    print("stretch energies start", h[0], "end", h[-1], flush=True)
    assert np.all(np.diff(h, axis=0) <= 1e-12 * np.abs(h[:-1]))
    assert np.all(h[-1] < 0.97 * h[0])
    assert float(plan.signed_areas(U).min()) > 0.0
