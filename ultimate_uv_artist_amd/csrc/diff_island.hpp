// Backward passes of the distortion energies and of the ARAP global step
// (included by kernels.hip).
//
// Energy.  E_i(U, w) = sum over island i's triangles of w_f e_f(U), with e_f
// the ARAP or symmetric Dirichlet energy of param.py:294-318 (frames with
// |x1 y2| < 1e-300 add nothing).  e_f depends on U through the Jacobian
// J = Q G^T (Q = corner UVs, G the triangle's 2x3 gradient operator), so
// dE/dQ = w_f (de/dJ) G and dE/dw_f = e_f.  With q = |conformal part|,
//   ARAP: e = area ((q+r-1)^2 + (q-r-1)^2) = area |J - R|_F^2 with R the
//         closest rotation, so de/dJ = 2 area (J - R) (R = I at q = 0, where
//         every rotation is closest);
//   SD:   e = area (|J|^2 + |J|^2/det^2 - 4), de/dJ = area (2J + 2J/det^2
//         - 2 |J|^2/det^3 cof(J)); a folded triangle (det <= 0) has e = inf
//         and contributes no U gradient.
//
// Global step.  U' = L^-1 b(U) with b = sum over corners of R_f(U) g_{f,c}
// plus the soft-pin term wpin U[pin] (param.py:391-401, 457-474).  Given
// lambda = L^-1 dLoss/dU' (the adjoint solve, same SPD operator), the step's
// backward is dLoss/dU = (db/dU)^T lambda: per triangle,
//   A = sum_c lambda_c . g_c,  B = sum_c lambda_c . (-g_c.y, g_c.x)
// are dLoss/dcos and dLoss/dsin of R; through (cos, sin) = (ca, cb)/h with
// ca = j00 + j11, cb = j10 - j01, h = |(ca, cb)|,
//   dca = (A - cos (cos A + sin B))/h,  dcb = (B - sin (cos A + sin B))/h,
// and dLoss/dJ = [[dca, -dcb], [dcb, dca]].  Triangles whose rotation is
// the constant identity (degenerate frame, h < 1e-300) pass nothing.
//
// Both backward passes write per-triangle corner gradients into P.sy and
// gather them per vertex in incidence order (no atomics: deterministic).
#pragma once

// e_f and de/dJ of one triangle (kind 0 ARAP, 1 symmetric Dirichlet).
UVA_HD __forceinline__ double energy_and_grad(const Jac& J, double area, int kind, Jac& g) {
    if (kind == 0) {
        const double E = 0.5 * (J.j00 + J.j11), H = 0.5 * (J.j10 - J.j01);
        const double q = sqrt(E * E + H * H);
        double c = 1.0, s = 0.0;
        if (q > 0.0) {
            c = E / q;
            s = H / q;
        }
        g.j00 = 2.0 * area * (J.j00 - c);
        g.j01 = 2.0 * area * (J.j01 + s);
        g.j10 = 2.0 * area * (J.j10 - s);
        g.j11 = 2.0 * area * (J.j11 - c);
        return arap_energy_tri(J, area);
    }
    const double det = J.j00 * J.j11 - J.j01 * J.j10;
    if (!(det > 0.0)) {
        g.j00 = g.j01 = g.j10 = g.j11 = 0.0;
        return __builtin_huge_val();
    }
    const double fro = J.j00 * J.j00 + J.j01 * J.j01 + J.j10 * J.j10 + J.j11 * J.j11;
    const double id2 = 1.0 / (det * det);
    const double a = 2.0 * area * (1.0 + id2), b = -2.0 * area * fro * id2 / det;
    g.j00 = a * J.j00 + b * J.j11;
    g.j01 = a * J.j01 - b * J.j10;
    g.j10 = a * J.j10 - b * J.j01;
    g.j11 = a * J.j11 + b * J.j00;
    return area * (fro + fro * id2 - 4.0);
}

// Corner gradients (dQ = dJ G) of triangle f into P.sy.
UVA_HD __forceinline__ void store_corner_grads(const DevPlan& P, int f, const Jac& dJ) {
    double g[2][3];
    slim_G(P, f, g);
    for (int k = 0; k < 3; ++k)
        P.sy[3 * f + k] = make_double2(dJ.j00 * g[0][k] + dJ.j01 * g[1][k], dJ.j10 * g[0][k] + dJ.j11 * g[1][k]);
}

// out[v] = sum of the corner gradients of v's incident triangles (+ extra).
template <class Team>
UVA_HD void gather_corner_grads(Team& tm, const DevPlan& P, const Isl& d, double2* __restrict__ out,
                                const double2* __restrict__ lam) {
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        double gx = 0.0, gy = 0.0;
        for (int q = P.inc_ptr[row]; q < P.inc_ptr[row + 1]; ++q) {
            const int f = P.inc_tri[q];
            const double2 y = P.sy[3 * f + corner_of(P, f, row)];
            gx += y.x;
            gy += y.y;
        }
        if (lam) {   // soft-pin term of b: wpin U[pin]
            const double2 l = lam[row];
            const double pw = pin_weight(P, d, row);
            if (pw != 0.0) {
                gx += pw * l.x;
                gy += pw * l.y;
            }
        }
        out[row] = make_double2(gx, gy);
    }
}

// Weighted energy of the island (forward).
template <class Team>
UVA_HD double energy_fwd_island(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U,
                                const double* __restrict__ w, int kind) {
    double en = 0.0;
    FOR_TRIS(tm, d, f) {
        if (!P.fok[f]) continue;
        const Jac J = jacobian(P, f, U);
        const double e = kind == 0 ? arap_energy_tri(J, P.farea[f]) : sd_energy_tri(J, P.farea[f]);
        en += w ? w[f] * e : e;
    }
    return tm.sum4(make_double4(en, 0.0, 0.0, 0.0)).x;
}

// dE/dU into gU and dE/dw into gw (when given), scaled by the island's
// upstream gradient ge.
template <class Team>
UVA_HD void energy_bwd_island(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U,
                              const double* __restrict__ w, int kind, double ge, double2* __restrict__ gU,
                              double* __restrict__ gw) {
    FOR_TRIS(tm, d, f) {
        Jac dJ;
        double e = 0.0;
        if (P.fok[f]) {
            e = energy_and_grad(jacobian(P, f, U), P.farea[f], kind, dJ);
            const double s = ge * (w ? w[f] : 1.0);
            dJ.j00 *= s; dJ.j01 *= s; dJ.j10 *= s; dJ.j11 *= s;
        } else {
            dJ.j00 = dJ.j01 = dJ.j10 = dJ.j11 = 0.0;
        }
        if (gw) gw[f] = ge * e;
        store_corner_grads(P, f, dJ);
    }
    tm.sync();
    gather_corner_grads(tm, P, d, gU, nullptr);
    tm.sync();
}

// dLoss/dU of one global step U' = L^-1 b(U), given lambda = L^-1 dLoss/dU'.
template <class Team>
UVA_HD void step_bwd_island(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U,
                            const double2* __restrict__ lam, double2* __restrict__ gU) {
    FOR_TRIS(tm, d, f) {
        Jac dJ;
        dJ.j00 = dJ.j01 = dJ.j10 = dJ.j11 = 0.0;
        if (P.fok[f]) {
            const Jac J = jacobian(P, f, U);
            const double ca = J.j00 + J.j11, cb = J.j10 - J.j01;
            const double h = sqrt(ca * ca + cb * cb);
            if (!(h < 1e-300)) {
                const double c = ca / h, s = cb / h;
                const int vs[3] = {P.tv0[f], P.tv1[f], P.tv2[f]};
                double A = 0.0, B = 0.0;
                for (int k = 0; k < 3; ++k) {
                    const double2 g = P.tri_g[3 * f + k], l = lam[vs[k]];
                    A += l.x * g.x + l.y * g.y;
                    B += -l.x * g.y + l.y * g.x;
                }
                const double pr = c * A + s * B;
                const double dca = (A - c * pr) / h, dcb = (B - s * pr) / h;
                dJ.j00 = dca;
                dJ.j11 = dca;
                dJ.j10 = dcb;
                dJ.j01 = -dcb;
            }
        }
        store_corner_grads(P, f, dJ);
    }
    tm.sync();
    gather_corner_grads(tm, P, d, gU, lam);
    tm.sync();
}
