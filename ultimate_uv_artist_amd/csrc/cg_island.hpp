// Island-level preconditioned conjugate gradients (included by kernels.hip).
//
// Chronopoulos-Gear form of Jacobi-PCG: each iteration is one pass of vector
// updates over the island's rows, one barrier, one SpMV pass that also forms
// the dot products, and ONE team reduction (r.r, r.z, w.z for both columns),
// instead of three passes and two reductions.  With u = D^-1 r, w = A u:
//   p = u + beta p,  s = w + beta s (= A p),  e += alpha p,  r -= alpha s,
//   u = D^-1 r,  w = A u,  gamma' = r.u,  delta' = w.u,
//   beta = gamma'/gamma,  alpha = gamma' / (delta' - beta gamma'/alpha).
// (The pipelined Ghysels-Vanroose form, one barrier per iteration, amplifies
// FP32 rounding in its extra recurrences: on cotangent Laplacians of a few
// thousand rows it needs 10-100x the iterations or diverges, so it is not
// used with FP32 vectors.)
// Mixed precision: the inner iterations work in FP32 on the correction e
// with FP64 dots; every outer pass recomputes the true residual b - A X in
// FP64 (hi+lo matrix values, FP64 iterate) and folds e into X (iterative
// refinement).  Each lane handles two rows per step (slices s and s + W) so
// two rows' loads are in flight together.
#pragma once

// Residual reduction of one inner FP32 pass of the mixed-precision LSCM
// solve: an FP32 recurrence cannot reach the FP64 target (1e-8) in one pass,
// so each pass stops here and the next outer pass restarts from the FP64
// true residual.
constexpr double kInnerRel = 1e-4;

struct CGOut {
    int iters;
    int conv;
    double rel;
};

// Rows owned by this lane, two at a time: slice s and slice s + nwaves of the
// lane's wavefront.  okA/okB say whether each row exists; sbv says whether
// slice B exists (uniform across the wavefront).
template <class Team, class F>
UVA_HD __forceinline__ void for_row_pairs(const Team& tm, const Isl& d, F&& f) {
    const int nw = tm.nwaves(), lane = tm.lane();
    const int hi = tm.sl_hi(d.ns);
    for (int s = tm.sl_lo(d.ns) + tm.wave(); s < hi; s += 2 * nw) {
        const int sb = s + nw;
        const bool sbv = sb < hi;
        const bool oka = s * kSlice + lane < d.nv;
        const bool okb = sbv && sb * kSlice + lane < d.nv;
        f(d.v0 + s * kSlice + lane, d.s0 + s, oka, d.v0 + sb * kSlice + lane, d.s0 + sb, sbv, okb);
    }
}

// y = A x for the two rows of a pair (see spmv_row); absent rows give 0.
template <typename V, int OP>
UVA_HD __forceinline__ void spmv_row2(const DevPlan& P, int ra, int sa, bool oka, int rb, int sb, bool sbv,
                                          bool okb, int lane, const typename Vec2<V>::type* __restrict__ x, V dga,
                                          V dgb, typename Vec2<V>::type& ya, typename Vec2<V>::type& yb) {
    using V2 = typename Vec2<V>::type;
    const int baseA = P.sl_ptr[sa];
    const int wA = (P.sl_ptr[sa + 1] - baseA) >> 6;
    int baseB = 0, wB = 0;
    if (sbv) {
        baseB = P.sl_ptr[sb];
        wB = (P.sl_ptr[sb + 1] - baseB) >> 6;
    }
    V2 xa, xb;
    xa.x = 0; xa.y = 0; xb.x = 0; xb.y = 0;
    if (oka) xa = x[ra];
    if (okb) xb = x[rb];
    V yax = dga * xa.x, yay = dga * xa.y, ybx = dgb * xb.x, yby = dgb * xb.y;
    const int wmax = wA > wB ? wA : wB;
    for (int k0 = 0; k0 < wmax; k0 += kGather) {
        int ca[kGather], cb[kGather];
        V va[kGather], vb[kGather], ba[kGather], bb[kGather];
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            const bool ia = oka && k0 + q < wA;
            const bool ib = okb && k0 + q < wB;
            const int oa = baseA + (k0 + q) * kSlice + lane;
            const int ob = baseB + (k0 + q) * kSlice + lane;
            ca[q] = ia ? P.m_col[oa] : 0;
            va[q] = ia ? mval<V>(P, oa) : (V)0;
            cb[q] = ib ? P.m_col[ob] : 0;
            vb[q] = ib ? mval<V>(P, ob) : (V)0;
            ba[q] = (OP == OP_LSCM && ia) ? (V)P.m_b[oa] : (V)0;
            bb[q] = (OP == OP_LSCM && ib) ? (V)P.m_b[ob] : (V)0;
        }
        // entries past a row's width gather row 0 and are left out by select
        // (a multiply by zero would let a non-finite value of row 0 through)
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            const bool ia = oka && k0 + q < wA;
            const bool ib = okb && k0 + q < wB;
            const V2 ga = x[ca[q]];
            const V2 gb = x[cb[q]];
            if (OP == OP_ARAP) {
                yax = ia ? fma(va[q], ga.x, yax) : yax;
                yay = ia ? fma(va[q], ga.y, yay) : yay;
                ybx = ib ? fma(vb[q], gb.x, ybx) : ybx;
                yby = ib ? fma(vb[q], gb.y, yby) : yby;
            } else {
                yax = ia ? fma(va[q], ga.x, fma(ba[q], ga.y, yax)) : yax;
                yay = ia ? fma(va[q], ga.y, -fma(ba[q], ga.x, -yay)) : yay;
                ybx = ib ? fma(vb[q], gb.x, fma(bb[q], gb.y, ybx)) : ybx;
                yby = ib ? fma(vb[q], gb.y, -fma(bb[q], gb.x, -yby)) : yby;
            }
        }
    }
    ya.x = yax; ya.y = yay;
    yb.x = ybx; yb.y = yby;
}

template <int PREC, int OP, class Team>
UVA_HD __forceinline__ CGOut cg_island(Team& tm, const DevPlan& P, const Isl& d, double2 b2, double tol, int maxit, int lp0,
                           int lp1) {
    using PT = Prec<PREC>;
    using V = typename PT::V;
    using V2 = typename Vec2<V>::type;
    V2* __restrict__ r = (V2*)P.vr;
    V2* __restrict__ p = (V2*)P.vp;
    V2* __restrict__ s = (V2*)P.vAp;
    V2* __restrict__ e = (V2*)P.ve;
    V2* __restrict__ u = (V2*)P.vu;
    V2* __restrict__ w = (V2*)P.vw;
    const bool coupled = (OP == OP_LSCM);
    constexpr bool kRelInner = OP == OP_LSCM && sizeof(V) == 4 && PT::kResid64;
    const double bx2 = coupled ? fmax(b2.x + b2.y, 1e-300) : fmax(b2.x, 1e-300);
    const double by2 = coupled ? bx2 : fmax(b2.y, 1e-300);
    const double t2 = tol * tol;
    const int lane = tm.lane();
    auto is_pin = [&](int row) { return OP == OP_LSCM && (row == lp0 || row == lp1); };
    int it = 0;
    CGOut out{0, 0, 0.0};
    double relx = 0.0, rely = 0.0;
    for (int outer = 0;; ++outer) {
        // true residual r = b - A X, u = D^-1 r, e = 0
        double4 a1 = make_double4(0.0, 0.0, 0.0, 0.0);
        FOR_ROWS(tm, d, row, sg) {
            double rx, ry;
            if (PT::kResid64) {
                double2 y = spmv_row<double, OP>(P, row, sg, lane, P.X64, diag_of<OP>(P, d, row));
                double2 bb = P.b64[row];
                rx = bb.x - y.x;
                ry = bb.y - y.y;
            } else {
                // all-FP32: residual from the FP32 matrix and the FP32-rounded iterate
                const int base = P.sl_ptr[sg];
                const int width = (P.sl_ptr[sg + 1] - base) >> 6;
                double2 xr = P.X64[row];
                float dg = (float)diag_of<OP>(P, d, row);
                float yx = dg * (float)xr.x, yy = dg * (float)xr.y;
                int o = base + lane;
                for (int k = 0; k < width; ++k, o += kSlice) {
                    double2 xc = P.X64[P.m_col[o]];
                    yx = fmaf(P.m_hi[o], (float)xc.x, yx);
                    yy = fmaf(P.m_hi[o], (float)xc.y, yy);
                }
                double2 bb = P.b64[row];
                rx = (double)((float)bb.x - yx);
                ry = (double)((float)bb.y - yy);
            }
            if (is_pin(row)) { rx = 0.0; ry = 0.0; }
            V2 rv;
            rv.x = (V)rx;
            rv.y = (V)ry;
            const V di = (V)dinv_of<OP>(P, d, row);
            V2 uv;
            uv.x = di * rv.x;
            uv.y = di * rv.y;
            r[row] = rv;
            u[row] = uv;
            V2 zero;
            zero.x = 0;
            zero.y = 0;
            e[row] = zero;
            a1.x += PT::kResid64 ? rx * rx : (double)((float)rx * (float)rx);
            a1.y += PT::kResid64 ? ry * ry : (double)((float)ry * (float)ry);
            a1.z += acc_mul<PREC>(rv.x, uv.x);
            a1.w += acc_mul<PREC>(rv.y, uv.y);
        }
        tm.sync();
        double4 a2 = make_double4(0.0, 0.0, 0.0, 0.0);
        for_row_pairs(tm, d, [&](int ra, int sa, bool oka, int rb, int sb, bool sbv, bool okb) {
            V2 wa, wb;
            spmv_row2<V, OP>(P, ra, sa, oka, rb, sb, sbv, okb, lane, u, oka ? (V)diag_of<OP>(P, d, ra) : (V)0,
                             okb ? (V)diag_of<OP>(P, d, rb) : (V)0, wa, wb);
            if (oka) {
                if (is_pin(ra)) { wa.x = 0; wa.y = 0; }
                w[ra] = wa;
                const V2 ua = u[ra];
                a2.x += acc_mul<PREC>(wa.x, ua.x);
                a2.y += acc_mul<PREC>(wa.y, ua.y);
            }
            if (okb) {
                if (is_pin(rb)) { wb.x = 0; wb.y = 0; }
                w[rb] = wb;
                const V2 ub = u[rb];
                a2.x += acc_mul<PREC>(wb.x, ub.x);
                a2.y += acc_mul<PREC>(wb.y, ub.y);
            }
        });
        double4 S1, S2;
        tm.sum8(a1, a2, S1, S2);
        if (coupled) {
            relx = rely = (S1.x + S1.y) / bx2;
        } else {
            relx = S1.x / bx2;
            rely = S1.y / by2;
        }
        if (relx <= t2 && rely <= t2) { out.conv = 1; break; }
        if (outer >= PT::kMaxOuter || it >= maxit || !__builtin_isfinite(relx) || !__builtin_isfinite(rely)) break;
        const double rr0 = S1.x + S1.y;   // squared true residual at the start of this pass
        double gx = S1.z, gy = S1.w, dx = S2.x, dy = S2.y;
        if (coupled) { gx = gy = S1.z + S1.w; dx = dy = S2.x + S2.y; }
        double alx = cg_alpha<PREC>(gx, dx), aly = cg_alpha<PREC>(gy, dy);
        double bex = 0.0, bey = 0.0;
        bool first = true;
        for (;;) {
            const V ax = (V)alx, ay = (V)aly, bx = (V)bex, by = (V)bey;
            double4 g1 = make_double4(0.0, 0.0, 0.0, 0.0);
            for_row_pairs(tm, d, [&](int ra, int sa, bool oka, int rb, int sb, bool sbv, bool okb) {
                (void)sa; (void)sb; (void)sbv;
                V2 uA, wA, rA, eA, pA, sA, uB, wB, rB, eB, pB, sB;
                V dA = 0, dB = 0;
                if (oka) {
                    uA = u[ra]; wA = w[ra]; rA = r[ra]; eA = e[ra]; dA = (V)dinv_of<OP>(P, d, ra);
                    if (!first) { pA = p[ra]; sA = s[ra]; }
                }
                if (okb) {
                    uB = u[rb]; wB = w[rb]; rB = r[rb]; eB = e[rb]; dB = (V)dinv_of<OP>(P, d, rb);
                    if (!first) { pB = p[rb]; sB = s[rb]; }
                }
                if (oka) {
                    if (first) { pA = uA; sA = wA; }
                    else {
                        pA.x = fma(bx, pA.x, uA.x); pA.y = fma(by, pA.y, uA.y);
                        sA.x = fma(bx, sA.x, wA.x); sA.y = fma(by, sA.y, wA.y);
                    }
                    eA.x = fma(ax, pA.x, eA.x); eA.y = fma(ay, pA.y, eA.y);
                    rA.x = fma(-ax, sA.x, rA.x); rA.y = fma(-ay, sA.y, rA.y);
                    uA.x = dA * rA.x; uA.y = dA * rA.y;
                    p[ra] = pA; s[ra] = sA; e[ra] = eA; r[ra] = rA; u[ra] = uA;
                    g1.x += acc_mul<PREC>(rA.x, rA.x);
                    g1.y += acc_mul<PREC>(rA.y, rA.y);
                    g1.z += acc_mul<PREC>(rA.x, uA.x);
                    g1.w += acc_mul<PREC>(rA.y, uA.y);
                }
                if (okb) {
                    if (first) { pB = uB; sB = wB; }
                    else {
                        pB.x = fma(bx, pB.x, uB.x); pB.y = fma(by, pB.y, uB.y);
                        sB.x = fma(bx, sB.x, wB.x); sB.y = fma(by, sB.y, wB.y);
                    }
                    eB.x = fma(ax, pB.x, eB.x); eB.y = fma(ay, pB.y, eB.y);
                    rB.x = fma(-ax, sB.x, rB.x); rB.y = fma(-ay, sB.y, rB.y);
                    uB.x = dB * rB.x; uB.y = dB * rB.y;
                    p[rb] = pB; s[rb] = sB; e[rb] = eB; r[rb] = rB; u[rb] = uB;
                    g1.x += acc_mul<PREC>(rB.x, rB.x);
                    g1.y += acc_mul<PREC>(rB.y, rB.y);
                    g1.z += acc_mul<PREC>(rB.x, uB.x);
                    g1.w += acc_mul<PREC>(rB.y, uB.y);
                }
            });
            tm.sync();
            double4 g2 = make_double4(0.0, 0.0, 0.0, 0.0);
            for_row_pairs(tm, d, [&](int ra, int sa, bool oka, int rb, int sb, bool sbv, bool okb) {
                V2 wa, wb;
                spmv_row2<V, OP>(P, ra, sa, oka, rb, sb, sbv, okb, lane, u, oka ? (V)diag_of<OP>(P, d, ra) : (V)0,
                                 okb ? (V)diag_of<OP>(P, d, rb) : (V)0, wa, wb);
                if (oka) {
                    if (is_pin(ra)) { wa.x = 0; wa.y = 0; }
                    w[ra] = wa;
                    const V2 ua = u[ra];
                    g2.x += acc_mul<PREC>(wa.x, ua.x);
                    g2.y += acc_mul<PREC>(wa.y, ua.y);
                }
                if (okb) {
                    if (is_pin(rb)) { wb.x = 0; wb.y = 0; }
                    w[rb] = wb;
                    const V2 ub = u[rb];
                    g2.x += acc_mul<PREC>(wb.x, ub.x);
                    g2.y += acc_mul<PREC>(wb.y, ub.y);
                }
            });
            double4 T1, T2;
            tm.sum8(g1, g2, T1, T2);
            ++it;
            first = false;
            double rrx = T1.x, rry = T1.y;
            if (coupled) rrx = rry = T1.x + T1.y;
            if ((rrx <= PT::kInnerFrac2 * t2 * bx2 && rry <= PT::kInnerFrac2 * t2 * by2) || it >= maxit) break;
            if (kRelInner && rrx <= kInnerRel * kInnerRel * rr0) break;
            if (!__builtin_isfinite(rrx) || !__builtin_isfinite(rry)) break;   // a non-finite island stops (it is not converged)
            double ngx = T1.z, ngy = T1.w, ndx = T2.x, ndy = T2.y;
            if (coupled) { ngx = ngy = T1.z + T1.w; ndx = ndy = T2.x + T2.y; }
            bex = gx > 0.0 ? ngx / gx : 0.0;
            bey = gy > 0.0 ? ngy / gy : 0.0;
            const double denx = ndx - (alx != 0.0 ? bex * ngx / alx : 0.0);
            const double deny = ndy - (aly != 0.0 ? bey * ngy / aly : 0.0);
            alx = cg_alpha<PREC>(ngx, denx, ndx);
            aly = cg_alpha<PREC>(ngy, deny, ndy);
            gx = ngx;
            gy = ngy;
        }
        // fold the correction into the FP64 iterate
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            double2 x = P.X64[row];
            V2 ee = e[row];
            x.x += (double)ee.x;
            x.y += (double)ee.y;
            if (PT::kRoundX) { x.x = (double)(float)x.x; x.y = (double)(float)x.y; }
            P.X64[row] = x;
        }
        tm.sync();
    }
    out.iters = it;
    out.rel = sqrt(fmax(relx, rely));
    return out;
}
