// Island-level preconditioned conjugate gradients (included by kernels.hip).
//
// Pipelined Jacobi-PCG (Ghysels & Vanroose): every iteration is ONE pass over
// the island's rows and ONE team barrier.  The pass computes the SpMV
// n = A m of the vector published by the previous barrier and, row by row,
// all recurrences with the scalars of that barrier's reduction:
//   z = n + beta z,  s = w + beta s,  p = u + beta p,
//   e += alpha p,  r -= alpha s,  w -= alpha z,  m' = D^-1 w,
//   gamma' = r.u', delta' = w.u', r.r   (u = D^-1 r computed on the fly)
//   beta = gamma'/gamma,  alpha = gamma' / (delta' - beta gamma'/alpha)
// The barrier both reduces (gamma', delta', r.r) for both columns and
// publishes m' (double-buffered, so the pass never overwrites the vector
// other rows are still gathering).
// Mixed precision: the inner iterations work in FP32 on the correction e
// with FP64 dots; every outer pass recomputes the true residual b - A X in
// FP64 (hi+lo matrix values, FP64 iterate), restarts the recurrences and
// folds e into X (iterative refinement, which also bounds the pipelined
// recurrences' drift).  Each lane handles two rows per step (slices s and
// s + W) so two rows' loads are in flight together.
#pragma once

struct CGOut {
    int iters;
    int conv;
    double rel;
};

// Rows owned by this lane, two at a time: slice s and slice s + nwaves of the
// lane's wavefront.  okA/okB say whether each row exists; sbv says whether
// slice B exists (uniform across the wavefront).
template <class Team, class F>
__device__ __forceinline__ void for_row_pairs(const Team& tm, const Isl& d, F&& f) {
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
__device__ __forceinline__ void spmv_row2(const DevPlan& P, int ra, int sa, bool oka, int rb, int sb, bool sbv,
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
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            const V2 ga = x[ca[q]];
            const V2 gb = x[cb[q]];
            if (OP == OP_ARAP) {
                yax = fma(va[q], ga.x, yax);
                yay = fma(va[q], ga.y, yay);
                ybx = fma(vb[q], gb.x, ybx);
                yby = fma(vb[q], gb.y, yby);
            } else {
                yax = fma(va[q], ga.x, fma(ba[q], ga.y, yax));
                yay = fma(va[q], ga.y, -fma(ba[q], ga.x, -yay));
                ybx = fma(vb[q], gb.x, fma(bb[q], gb.y, ybx));
                yby = fma(vb[q], gb.y, -fma(bb[q], gb.x, -yby));
            }
        }
    }
    ya.x = yax; ya.y = yay;
    yb.x = ybx; yb.y = yby;
}

template <int PREC, int OP, class Team>
__device__ CGOut cg_island(Team& tm, const DevPlan& P, const Isl& d, double2 b2, double tol, int maxit, int lp0,
                           int lp1) {
    using PT = Prec<PREC>;
    using V = typename PT::V;
    using V2 = typename Vec2<V>::type;
    V2* __restrict__ r = (V2*)P.vr;
    V2* __restrict__ p = (V2*)P.vp;
    V2* __restrict__ s = (V2*)P.vAp;
    V2* __restrict__ e = (V2*)P.ve;
    V2* __restrict__ z = (V2*)P.vu;
    V2* __restrict__ w = (V2*)P.vw;
    V2* __restrict__ mbuf[2] = {(V2*)P.vm0, (V2*)P.vm1};
    const bool coupled = (OP == OP_LSCM);
    const double bx2 = coupled ? fmax(b2.x + b2.y, 1e-300) : fmax(b2.x, 1e-300);
    const double by2 = coupled ? bx2 : fmax(b2.y, 1e-300);
    const double t2 = tol * tol;
    const int lane = tm.lane();
    auto is_pin = [&](int row) { return OP == OP_LSCM && (row == lp0 || row == lp1); };
    V2 zero;
    zero.x = 0;
    zero.y = 0;
    int it = 0;
    CGOut out{0, 0, 0.0};
    double relx = 0.0, rely = 0.0;
    for (int outer = 0;; ++outer) {
        // --- restart: r = b - A X (true residual), e = 0, u = D^-1 r (kept in z)
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
            z[row] = uv;
            e[row] = zero;
            a1.x += PT::kResid64 ? rx * rx : (double)((float)rx * (float)rx);
            a1.y += PT::kResid64 ? ry * ry : (double)((float)ry * (float)ry);
            a1.z += acc_mul<PREC>(rv.x, uv.x);
            a1.w += acc_mul<PREC>(rv.y, uv.y);
        }
        tm.sync();
        // --- w = A u, m = D^-1 w, delta = w.u
        double4 a2 = make_double4(0.0, 0.0, 0.0, 0.0);
        FOR_ROWS(tm, d, row, sg) {
            V2 wv = spmv_row<V, OP>(P, row, sg, lane, z, (V)diag_of<OP>(P, d, row));
            if (is_pin(row)) wv = zero;
            const V di = (V)dinv_of<OP>(P, d, row);
            V2 mv;
            mv.x = di * wv.x;
            mv.y = di * wv.y;
            w[row] = wv;
            mbuf[0][row] = mv;
            const V2 uv = z[row];
            a2.x += acc_mul<PREC>(wv.x, uv.x);
            a2.y += acc_mul<PREC>(wv.y, uv.y);
        }
        double4 S1, S2;
        tm.sum8(a1, a2, S1, S2);
        if (coupled) {
            relx = rely = (S1.x + S1.y) / bx2;
        } else {
            relx = S1.x / bx2;
            rely = S1.y / by2;
        }
        if (relx <= t2 && rely <= t2) { out.conv = 1; break; }
        if (outer >= PT::kMaxOuter || it >= maxit || !isfinite(relx) || !isfinite(rely)) break;
        double gx = S1.z, gy = S1.w, dx = S2.x, dy = S2.y;
        if (coupled) { gx = gy = S1.z + S1.w; dx = dy = S2.x + S2.y; }
        double alx = dx > 0.0 ? gx / dx : 0.0, aly = dy > 0.0 ? gy / dy : 0.0;
        double bex = 0.0, bey = 0.0;
        bool first = true;
        int cur = 0;
        // --- pipelined iterations: one pass and one barrier each
        for (;;) {
            const V ax = (V)alx, ay = (V)aly, bx = (V)bex, by = (V)bey;
            const V2* __restrict__ mc = mbuf[cur];
            V2* __restrict__ mn = mbuf[cur ^ 1];
            double4 g1 = make_double4(0.0, 0.0, 0.0, 0.0), g2 = make_double4(0.0, 0.0, 0.0, 0.0);
            auto step_row = [&](int row, const V2 nv) {
                const V di = (V)dinv_of<OP>(P, d, row);
                V2 rv = r[row], wv = w[row], ev = e[row], zv, sv, pv;
                const V2 uv = V2{di * rv.x, di * rv.y};
                if (first) {
                    zv = nv;
                    sv = wv;
                    pv = uv;
                } else {
                    zv = z[row];
                    sv = s[row];
                    pv = p[row];
                    zv.x = fma(bx, zv.x, nv.x); zv.y = fma(by, zv.y, nv.y);
                    sv.x = fma(bx, sv.x, wv.x); sv.y = fma(by, sv.y, wv.y);
                    pv.x = fma(bx, pv.x, uv.x); pv.y = fma(by, pv.y, uv.y);
                }
                ev.x = fma(ax, pv.x, ev.x); ev.y = fma(ay, pv.y, ev.y);
                rv.x = fma(-ax, sv.x, rv.x); rv.y = fma(-ay, sv.y, rv.y);
                wv.x = fma(-ax, zv.x, wv.x); wv.y = fma(-ay, zv.y, wv.y);
                const V2 un = V2{di * rv.x, di * rv.y};
                z[row] = zv; s[row] = sv; p[row] = pv; e[row] = ev; r[row] = rv; w[row] = wv;
                mn[row] = V2{di * wv.x, di * wv.y};
                g1.x += acc_mul<PREC>(rv.x, rv.x);
                g1.y += acc_mul<PREC>(rv.y, rv.y);
                g1.z += acc_mul<PREC>(rv.x, un.x);
                g1.w += acc_mul<PREC>(rv.y, un.y);
                g2.x += acc_mul<PREC>(wv.x, un.x);
                g2.y += acc_mul<PREC>(wv.y, un.y);
            };
            for_row_pairs(tm, d, [&](int ra, int sa, bool oka, int rb, int sb, bool sbv, bool okb) {
                V2 na, nb;
                spmv_row2<V, OP>(P, ra, sa, oka, rb, sb, sbv, okb, lane, mc, oka ? (V)diag_of<OP>(P, d, ra) : (V)0,
                                 okb ? (V)diag_of<OP>(P, d, rb) : (V)0, na, nb);
                if (oka) step_row(ra, is_pin(ra) ? zero : na);
                if (okb) step_row(rb, is_pin(rb) ? zero : nb);
            });
            double4 T1, T2;
            tm.sum8(g1, g2, T1, T2);
            ++it;
            first = false;
            cur ^= 1;
            double rrx = T1.x, rry = T1.y;
            if (coupled) rrx = rry = T1.x + T1.y;
            if ((rrx <= PT::kInnerFrac2 * t2 * bx2 && rry <= PT::kInnerFrac2 * t2 * by2) || it >= maxit) break;
            if (!isfinite(rrx) || !isfinite(rry)) break;   // a non-finite island stops (it is not converged)
            double ngx = T1.z, ngy = T1.w, ndx = T2.x, ndy = T2.y;
            if (coupled) { ngx = ngy = T1.z + T1.w; ndx = ndy = T2.x + T2.y; }
            bex = gx > 0.0 ? ngx / gx : 0.0;
            bey = gy > 0.0 ? ngy / gy : 0.0;
            const double denx = ndx - (alx != 0.0 ? bex * ngx / alx : 0.0);
            const double deny = ndy - (aly != 0.0 ? bey * ngy / aly : 0.0);
            alx = denx > 0.0 ? ngx / denx : 0.0;
            aly = deny > 0.0 ? ngy / deny : 0.0;
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
