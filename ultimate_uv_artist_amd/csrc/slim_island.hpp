// SLIM (symmetric Dirichlet) on one island (included by kernels.hip).
//
// param.py:415-454 minimises  sum_f area_f |W_f (J_f - R_f)|_F^2
// + 1e-12 (island area) |x - x0|^2  over the UVs x through its normal
// equations.  With the per-triangle gradient operator G (2x3,
// param.py:341-350) and J = Q G^T (Q = the triangle's corner UVs as a 2x3
// matrix), the normal-equation operator and right-hand side are, per
// triangle,
//     y = area (W^T W) J G,      rhs = area (W^T W) R G,
// column k of which belongs to corner k.  Both are applied matrix-free: a
// triangle pass writes y, a vertex pass sums the incident corners in a fixed
// order (no atomics), so results are deterministic.  Only W^T W enters, and
// the local step stores its three distinct entries per triangle.
#pragma once

// G of triangle f: rows (-1/x1, 1/x1, 0) and ((-1 + x2/x1)/y2, -(x2/x1)/y2, 1/y2).
UVA_HD __forceinline__ void slim_G(const DevPlan& P, int f, double g[2][3]) {
    const double x1 = P.fx1[f], x2 = P.fx2[f], y2 = P.fy2[f];
    g[0][0] = -1.0 / x1;
    g[0][1] = 1.0 / x1;
    g[0][2] = 0.0;
    g[1][0] = (-1.0 + x2 / x1) / y2;
    g[1][1] = (-x2 / x1) / y2;
    g[1][2] = 1.0 / y2;
}

// SLIM weight of one signed singular value (param.py:246-252).
UVA_HD __forceinline__ double slim_w(double sv) {
    double s = fabs(sv);
    s = s > 1e-8 ? s : 1e-8;
    if (fabs(s - 1.0) < 1e-6) return 2.0;
    const double q = (s - 1.0 / (s * s * s)) / (s - 1.0);
    return sqrt(q > 1e-12 ? q : 1e-12);
}

// SLIM local step of one triangle (param.py:234-256): R = U V^T of the SVD of
// J with the reflection fix and W = U diag(w) U^T.  Closed form of the 2x2
// SVD: with E = (j00+j11)/2, H = (j10-j01)/2, F = (j00-j11)/2,
// K = (j10+j01)/2, q = |(E,H)|, r = |(F,K)|, the signed singular values are
// q + r and q - r (negative for a reflection), U V^T is the rotation by
// atan2(H, E) and U the rotation by (atan2(H,E) + atan2(K,F))/2, so
//     W^T W = W^2 = (w1^2 + w2^2)/2 I + (w1^2 - w2^2)/2 [[c, s], [s, -c]]
// with (c, s) = (cos, sin)(atan2(H,E) + atan2(K,F)) = ((E F - H K), (E K + H F)) / (q r).
// Degenerate frames keep R = W = I (param.py:216-219).
UVA_HD __forceinline__ void slim_local(const DevPlan& P, int f, const double2* __restrict__ U, double& cr,
                                       double& sr, double4& mw) {
    cr = 1.0;
    sr = 0.0;
    mw = make_double4(1.0, 0.0, 1.0, 0.0);
    if (!P.fok[f]) return;
    const Jac J = jacobian(P, f, U);
    const double E = 0.5 * (J.j00 + J.j11), H = 0.5 * (J.j10 - J.j01);
    const double F = 0.5 * (J.j00 - J.j11), K = 0.5 * (J.j10 + J.j01);
    const double q = sqrt(E * E + H * H), r = sqrt(F * F + K * K);
    if (q >= 1e-300) {
        cr = E / q;
        sr = H / q;
    }
    const double w1 = slim_w(q + r), w2 = slim_w(q - r);
    const double a = 0.5 * (w1 * w1 + w2 * w2), b = 0.5 * (w1 * w1 - w2 * w2);
    double c = 0.0, s = 0.0;
    if (b != 0.0 && q * r > 0.0) {
        c = (E * F - H * K) / (q * r);
        s = (E * K + H * F) / (q * r);
    }
    mw = make_double4(a + b * c, b * s, a - b * c, 0.0);
}

// y = area M J G for the corner values v (M = W^T W as xx, xy, yy).
UVA_HD __forceinline__ void slim_tri_apply(const DevPlan& P, int f, const double4& mw, const double2 v[3],
                                           double2 y[3]) {
    double g[2][3];
    slim_G(P, f, g);
    double j[2][2];
    for (int c = 0; c < 2; ++c) {
        j[0][c] = g[c][0] * v[0].x + g[c][1] * v[1].x + g[c][2] * v[2].x;
        j[1][c] = g[c][0] * v[0].y + g[c][1] * v[1].y + g[c][2] * v[2].y;
    }
    const double ar = P.farea[f];
    double k[2][2];
    for (int c = 0; c < 2; ++c) {
        k[0][c] = ar * (mw.x * j[0][c] + mw.y * j[1][c]);
        k[1][c] = ar * (mw.y * j[0][c] + mw.z * j[1][c]);
    }
    for (int q = 0; q < 3; ++q)
        y[q] = make_double2(k[0][0] * g[0][q] + k[0][1] * g[1][q], k[1][0] * g[0][q] + k[1][1] * g[1][q]);
}

// Corner of triangle f that is vertex v.
UVA_HD __forceinline__ int corner_of(const DevPlan& P, int f, int v) {
    return P.tv0[f] == v ? 0 : (P.tv1[f] == v ? 1 : 2);
}

// out = M v (+ reg v): triangle pass into P.sy, barrier, vertex gather.
// Returns the partial dot of v with the result over this thread's rows.
template <class Team>
UVA_HD double slim_apply(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ v,
                         double2* __restrict__ out, double reg) {
    FOR_TRIS(tm, d, f) {
        const double2 c3[3] = {v[P.tv0[f]], v[P.tv1[f]], v[P.tv2[f]]};
        double2 y[3];
        slim_tri_apply(P, f, P.sw[f], c3, y);
        P.sy[3 * f] = y[0];
        P.sy[3 * f + 1] = y[1];
        P.sy[3 * f + 2] = y[2];
    }
    tm.sync();
    double dot = 0.0;
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        const double2 vr = v[row];
        double ox = reg * vr.x, oy = reg * vr.y;
        for (int q = P.inc_ptr[row]; q < P.inc_ptr[row + 1]; ++q) {
            const int f = P.inc_tri[q];
            const double2 y = P.sy[3 * f + corner_of(P, f, row)];
            ox += y.x;
            oy += y.y;
        }
        out[row] = make_double2(ox, oy);
        dot += vr.x * ox + vr.y * oy;
    }
    return dot;
}

// One SLIM global step: local step of U, normal equations by block-Jacobi
// PCG (FP64) warm-started at U into X64.  Returns the CG iterations and sets
// conv/rel.
template <class Team>
UVA_HD int slim_solve(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U, double reg,
                      double tol, int maxit, int& conv, double& rel) {
    double2* __restrict__ r = (double2*)P.vr;
    double2* __restrict__ p = (double2*)P.vp;
    double2* __restrict__ Ap = (double2*)P.vAp;
    double2* __restrict__ z = (double2*)P.vu;
    double2* __restrict__ x = P.X64;
    double2* __restrict__ b = P.b64;
    // local step: W^T W per triangle, right-hand side contributions into sy
    FOR_TRIS(tm, d, f) {
        double cr, sr;
        double4 mw;
        slim_local(P, f, U, cr, sr, mw);
        P.sw[f] = mw;
        // rhs contribution area M R G, R = [[cr, -sr], [sr, cr]]
        double g[2][3];
        slim_G(P, f, g);
        const double ar = P.farea[f];
        const double k00 = ar * (mw.x * cr + mw.y * sr), k01 = ar * (-mw.x * sr + mw.y * cr);
        const double k10 = ar * (mw.y * cr + mw.z * sr), k11 = ar * (-mw.y * sr + mw.z * cr);
        for (int q = 0; q < 3; ++q)
            P.sy[3 * f + q] = make_double2(k00 * g[0][q] + k01 * g[1][q], k10 * g[0][q] + k11 * g[1][q]);
    }
    tm.sync();
    // b = gathered rhs + reg x0; inverse 2x2 diagonal blocks; x = U
    double bb = 0.0;
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        const double2 u = U[row];
        double bx = reg * u.x, by = reg * u.y;
        double dxx = reg, dxy = 0.0, dyy = reg;
        for (int q = P.inc_ptr[row]; q < P.inc_ptr[row + 1]; ++q) {
            const int f = P.inc_tri[q];
            const int k = corner_of(P, f, row);
            const double2 y = P.sy[3 * f + k];
            bx += y.x;
            by += y.y;
            double g[2][3];
            slim_G(P, f, g);
            const double gg = P.farea[f] * (g[0][k] * g[0][k] + g[1][k] * g[1][k]);
            const double4 mw = P.sw[f];
            dxx += gg * mw.x;
            dxy += gg * mw.y;
            dyy += gg * mw.z;
        }
        const double det = dxx * dyy - dxy * dxy;
        const double id = det != 0.0 ? 1.0 / det : 0.0;
        P.sdi[row] = make_double4(dyy * id, -dxy * id, dxx * id, 0.0);
        b[row] = make_double2(bx, by);
        x[row] = u;
        bb += bx * bx + by * by;
    }
    tm.sync();
    // r = b - M x, z = D^-1 r, p = z
    (void)slim_apply(tm, P, d, x, Ap, reg);
    double4 a1 = make_double4(bb, 0.0, 0.0, 0.0);
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        const double2 bv = b[row], av = Ap[row];
        const double2 rv = make_double2(bv.x - av.x, bv.y - av.y);
        const double4 di = P.sdi[row];
        const double2 zv = make_double2(di.x * rv.x + di.y * rv.y, di.y * rv.x + di.z * rv.y);
        r[row] = rv;
        z[row] = zv;
        p[row] = zv;
        a1.y += rv.x * rv.x + rv.y * rv.y;
        a1.z += rv.x * zv.x + rv.y * zv.y;
    }
    const double4 s1 = tm.sum4(a1);
    const double b2 = s1.x > 1e-300 ? s1.x : 1e-300;
    double rr = s1.y, rz = s1.z;
    const double t2 = tol * tol;
    int it = 0;
    conv = 0;
    while (true) {
        if (rr <= t2 * b2) { conv = 1; break; }
        if (it >= maxit || !__builtin_isfinite(rr) || !__builtin_isfinite(rz)) break;
        const double pap = tm.sum4(make_double4(slim_apply(tm, P, d, p, Ap, reg), 0.0, 0.0, 0.0)).x;
        const double alpha = pap > 0.0 ? rz / pap : 0.0;
        double4 a2 = make_double4(0.0, 0.0, 0.0, 0.0);
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 pv = p[row], av = Ap[row];
            double2 xv = x[row], rv = r[row];
            xv.x += alpha * pv.x;
            xv.y += alpha * pv.y;
            rv.x -= alpha * av.x;
            rv.y -= alpha * av.y;
            const double4 di = P.sdi[row];
            const double2 zv = make_double2(di.x * rv.x + di.y * rv.y, di.y * rv.x + di.z * rv.y);
            x[row] = xv;
            r[row] = rv;
            z[row] = zv;
            a2.x += rv.x * rv.x + rv.y * rv.y;
            a2.y += rv.x * zv.x + rv.y * zv.y;
        }
        const double4 s2 = tm.sum4(a2);
        ++it;
        rr = s2.x;
        const double beta = rz > 0.0 ? s2.y / rz : 0.0;
        rz = s2.y;
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 zv = z[row], pv = p[row];
            p[row] = make_double2(zv.x + beta * pv.x, zv.y + beta * pv.y);
        }
        tm.sync();
    }
    rel = sqrt(rr / b2);
    // Translations are in the null space of the least-squares operator, so
    // the exact regularised solution keeps the mean of x0 = U; CG leaves an
    // arbitrary translation there (weighted by 1e-12 only), so project.
    double4 a3 = make_double4(0.0, 0.0, 0.0, 0.0);
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        const double2 xv = x[row], u = U[row];
        a3.x += xv.x - u.x;
        a3.y += xv.y - u.y;
    }
    const double4 s3 = tm.sum4(a3);
    const double mx = s3.x / d.nv, my = s3.y / d.nv;
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        const double2 xv = x[row];
        x[row] = make_double2(xv.x - mx, xv.y - my);
    }
    tm.sync();
    return it;
}

// Symmetric Dirichlet energy of U + step (X64 - U) (param.py:294-318; frames
// with |x1 y2| < 1e-300 add nothing).
template <class Team>
UVA_HD double slim_energy_at(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U,
                             double step) {
    double en = 0.0;
    FOR_TRIS(tm, d, f) {
        if (!P.fok[f]) continue;
        const int ia = P.tv0[f], ib = P.tv1[f], ic = P.tv2[f];
        double2 c[3];
        const int idx[3] = {ia, ib, ic};
        for (int q = 0; q < 3; ++q) {
            const double2 u = U[idx[q]], xn = P.X64[idx[q]];
            const double dx = xn.x - u.x, dy = xn.y - u.y;
            c[q] = make_double2(u.x + step * dx, u.y + step * dy);
        }
        const double x1 = P.fx1[f], x2 = P.fx2[f], y2 = P.fy2[f];
        Jac J;
        J.j00 = (c[1].x - c[0].x) / x1;
        J.j10 = (c[1].y - c[0].y) / x1;
        J.j01 = ((c[2].x - c[0].x) - J.j00 * x2) / y2;
        J.j11 = ((c[2].y - c[0].y) - J.j10 * x2) / y2;
        en += sd_energy_tri(J, P.farea[f]);
    }
    return tm.sum4(make_double4(en, 0.0, 0.0, 0.0)).x;
}
