// Island-resident preconditioned CG (included by kernels.hip).
//
// The same Chronopoulos-Gear Jacobi-PCG as cg_island (same recurrences, same
// per-lane accumulation order), for islands whose rows fit on the team's
// compute units.  Each lane keeps r, p, s = A p, w = A u, the diagonal and
// its inverse for its rows in registers (at most 8 rows in FP32, 4 in FP64 at
// 512 threads): slot j holds row lane of the local slice j*W + wave of the
// member's slice range.  The preconditioned residual
// u = D^-1 r, which the SpMV gathers, and the correction e live in LDS, so a
// CG iteration reads only the matrix from memory.  Members of a
// multi-workgroup team also publish u to global memory; their SpMV gathers
// rows owned by other members from there (the halo) and their own rows from
// LDS.
#pragma once

template <int PREC, int RPT, class Team>
struct ResCG {
    using PT = Prec<PREC>;
    using V = typename PT::V;
    using V2 = typename Vec2<V>::type;
    static constexpr int kG = 4;   // matrix entries whose loads are issued together

    Team& tm;
    const DevPlan& P;
    const Isl& d;
    int W, wv, lane, slo, shi, rlo, rhi;
    bool multi;
    V2* lu;     // LDS: u of the member's rows
    V2* le;     // LDS: e of the member's rows
    V2* gu;     // global: u of every row (teams only)
    V2 r[RPT], p[RPT], s[RPT], w[RPT];
    V dv[RPT], dg[RPT];

    UVA_HD __forceinline__ ResCG(Team& t, const DevPlan& P_, const Isl& d_) : tm(t), P(P_), d(d_) {
        W = tm.nwaves();
        wv = tm.wave();
        lane = tm.lane();
        slo = tm.sl_lo(d.ns);
        shi = tm.sl_hi(d.ns);
        rlo = d.v0 + slo * kSlice;
        rhi = min(d.v0 + shi * kSlice, d.v0 + d.nv);
        multi = tm.multi();
        lu = (V2*)tm.lvec();
        le = lu + tm.lvec_bytes() / (2 * (int)sizeof(V2));
        gu = (V2*)P.vu;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int l = (j * W + wv) * kSlice + lane;
            const int row = rlo + l;
            const bool ok = slot(j) && row < rhi;
            dv[j] = ok ? (V)dinv_of<OP_ARAP>(P, d, row) : (V)0;
            dg[j] = ok ? (V)diag_of<OP_ARAP>(P, d, row) : (V)0;
            r[j].x = 0; r[j].y = 0;
            p[j] = r[j]; s[j] = r[j]; w[j] = r[j];
        }
    }
    // slot j exists for this wavefront (uniform)
    UVA_HD __forceinline__ bool slot(int j) const { return slo + j * W + wv < shi; }
    UVA_HD __forceinline__ int lrow(int j) const { return (j * W + wv) * kSlice + lane; }
    // Row addresses are cheap to recompute but costly to keep: each pass
    // starts from an opaque copy of the lane id, so the compiler cannot hoist
    // per-slot addresses out of the CG loop and hold them in registers.
    UVA_HD __forceinline__ void refresh() {
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" : "+v"(lane));
#endif
    }

    UVA_HD __forceinline__ V2 gather(int c) const {
        if (multi && (c < rlo || c >= rhi)) return gu[c];
        return lu[c - rlo];
    }

    // w = A u over the member's rows (u = D^-1 r from registers for the own
    // row, LDS / global for the columns) and the dot w.u.
    UVA_HD __forceinline__ double2 spmv() {
        refresh();
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            if (!slot(j)) continue;
            const int l = lrow(j);
            const int row = rlo + l;
            const bool ok = row < rhi;
            const int sg = d.s0 + slo + j * W + wv;
            const int base = P.sl_ptr[sg];
            const int width = (P.sl_ptr[sg + 1] - base) >> 6;
            V2 uo;
            uo.x = dv[j] * r[j].x;
            uo.y = dv[j] * r[j].y;
            V yx = dg[j] * uo.x, yy = dg[j] * uo.y;
            for (int k0 = 0; k0 < width; k0 += kG) {
                int c[kG];
                V v[kG];
#pragma unroll
                for (int q = 0; q < kG; ++q) {
                    const bool in = ok && k0 + q < width;
                    const int o = base + (k0 + q) * kSlice + lane;
                    c[q] = in ? P.m_col[o] : rlo;
                    v[q] = in ? mval<V>(P, o) : (V)0;
                }
#pragma unroll
                for (int q = 0; q < kG; ++q) {
                    const bool in = ok && k0 + q < width;
                    const V2 x = gather(c[q]);
                    yx = in ? fma(v[q], x.x, yx) : yx;
                    yy = in ? fma(v[q], x.y, yy) : yy;
                }
            }
            if (!ok) { yx = 0; yy = 0; }
            w[j].x = yx;
            w[j].y = yy;
            if (ok) {
                acc.x += acc_mul<PREC>(yx, uo.x);
                acc.y += acc_mul<PREC>(yy, uo.y);
            }
        }
        return acc;
    }

    UVA_HD __forceinline__ CGOut solve(double2 b2, double tol, int maxit) {
        const double bx2 = fmax(b2.x, 1e-300), by2 = fmax(b2.y, 1e-300);
        const double t2 = tol * tol;
        int it = 0;
        CGOut out{0, 0, 0.0};
        double relx = 0.0, rely = 0.0;
        for (int outer = 0;; ++outer) {
            // true residual r = b - A X (FP64 from the hi+lo matrix, or the
            // all-FP32 form), u = D^-1 r, e = 0
            double4 a1 = make_double4(0.0, 0.0, 0.0, 0.0);
            refresh();
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                if (!slot(j)) continue;
                const int l = lrow(j);
                const int row = rlo + l;
                if (row >= rhi) continue;
                const int sg = d.s0 + slo + j * W + wv;
                double rx, ry;
                if (PT::kResid64) {
                    const double2 y = spmv_row<double, OP_ARAP>(P, row, sg, lane, P.X64, diag_of<OP_ARAP>(P, d, row));
                    const double2 bb = P.b64[row];
                    rx = bb.x - y.x;
                    ry = bb.y - y.y;
                } else {
                    const int base = P.sl_ptr[sg];
                    const int width = (P.sl_ptr[sg + 1] - base) >> 6;
                    const double2 xr = P.X64[row];
                    const float dgf = (float)diag_of<OP_ARAP>(P, d, row);
                    float yx = dgf * (float)xr.x, yy = dgf * (float)xr.y;
                    int o = base + lane;
                    for (int k = 0; k < width; ++k, o += kSlice) {
                        const double2 xc = P.X64[P.m_col[o]];
                        yx = fmaf(P.m_hi[o], (float)xc.x, yx);
                        yy = fmaf(P.m_hi[o], (float)xc.y, yy);
                    }
                    const double2 bb = P.b64[row];
                    rx = (double)((float)bb.x - yx);
                    ry = (double)((float)bb.y - yy);
                }
                r[j].x = (V)rx;
                r[j].y = (V)ry;
                V2 uv;
                uv.x = dv[j] * r[j].x;
                uv.y = dv[j] * r[j].y;
                lu[l] = uv;
                if (multi) gu[row] = uv;
                V2 zero;
                zero.x = 0;
                zero.y = 0;
                le[l] = zero;
                a1.x += PT::kResid64 ? rx * rx : (double)((float)rx * (float)rx);
                a1.y += PT::kResid64 ? ry * ry : (double)((float)ry * (float)ry);
                a1.z += acc_mul<PREC>(r[j].x, uv.x);
                a1.w += acc_mul<PREC>(r[j].y, uv.y);
            }
            tm.sync();
            const double2 wu = spmv();
            double4 S1, S2;
            tm.sum8(a1, make_double4(wu.x, wu.y, 0.0, 0.0), S1, S2);
            relx = S1.x / bx2;
            rely = S1.y / by2;
            if (relx <= t2 && rely <= t2) { out.conv = 1; break; }
            if (outer >= PT::kMaxOuter || it >= maxit || !__builtin_isfinite(relx) || !__builtin_isfinite(rely)) break;
            double gx = S1.z, gy = S1.w, dx = S2.x, dy = S2.y;
            double alx = dx > 0.0 ? gx / dx : 0.0, aly = dy > 0.0 ? gy / dy : 0.0;
            double bex = 0.0, bey = 0.0;
            bool first = true;
            for (;;) {
                const V ax = (V)alx, ay = (V)aly, bx = (V)bex, by = (V)bey;
                double4 g1 = make_double4(0.0, 0.0, 0.0, 0.0);
                refresh();
#pragma unroll
                for (int j = 0; j < RPT; ++j) {
                    if (!slot(j)) continue;
                    const int l = lrow(j);
                    const int row = rlo + l;
                    if (row >= rhi) continue;
                    V2 uA;
                    uA.x = dv[j] * r[j].x;
                    uA.y = dv[j] * r[j].y;
                    if (first) {
                        p[j] = uA;
                        s[j] = w[j];
                    } else {
                        p[j].x = fma(bx, p[j].x, uA.x); p[j].y = fma(by, p[j].y, uA.y);
                        s[j].x = fma(bx, s[j].x, w[j].x); s[j].y = fma(by, s[j].y, w[j].y);
                    }
                    V2 eA = le[l];
                    eA.x = fma(ax, p[j].x, eA.x);
                    eA.y = fma(ay, p[j].y, eA.y);
                    le[l] = eA;
                    r[j].x = fma(-ax, s[j].x, r[j].x);
                    r[j].y = fma(-ay, s[j].y, r[j].y);
                    uA.x = dv[j] * r[j].x;
                    uA.y = dv[j] * r[j].y;
                    lu[l] = uA;
                    if (multi) gu[row] = uA;
                    g1.x += acc_mul<PREC>(r[j].x, r[j].x);
                    g1.y += acc_mul<PREC>(r[j].y, r[j].y);
                    g1.z += acc_mul<PREC>(r[j].x, uA.x);
                    g1.w += acc_mul<PREC>(r[j].y, uA.y);
                }
                tm.sync();
                const double2 wu2 = spmv();
                double4 T1, T2;
                tm.sum8(g1, make_double4(wu2.x, wu2.y, 0.0, 0.0), T1, T2);
                ++it;
                first = false;
                const double rrx = T1.x, rry = T1.y;
                if ((rrx <= PT::kInnerFrac2 * t2 * bx2 && rry <= PT::kInnerFrac2 * t2 * by2) || it >= maxit) break;
                if (!__builtin_isfinite(rrx) || !__builtin_isfinite(rry)) break;
                const double ngx = T1.z, ngy = T1.w, ndx = T2.x, ndy = T2.y;
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
            refresh();
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                if (!slot(j)) continue;
                const int l = lrow(j);
                const int row = rlo + l;
                if (row >= rhi) continue;
                double2 x = P.X64[row];
                const V2 ee = le[l];
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
};

// ARAP global-step CG of one island: the resident form when the member's
// rows fit in the team's registers and LDS, the streaming form otherwise.
template <int PREC, class Team>
UVA_HD __forceinline__ CGOut cg_arap(Team& tm, const DevPlan& P, const Isl& d, double2 b2, double tol, int maxit) {
    using V = typename Prec<PREC>::V;
    using V2 = typename Vec2<V>::type;
    // rows per lane the registers hold without spilling at 512 threads
    constexpr int kMaxRpt = sizeof(V) == 4 ? 8 : 4;
    const int nsl = tm.sl_hi(d.ns) - tm.sl_lo(d.ns);
    const int W = tm.nwaves();
    const int rpt = (nsl + W - 1) / W;
    const int cap = tm.lvec_bytes() / (2 * (int)sizeof(V2));
    if (tm.lvec() != nullptr && nsl * kSlice <= cap && rpt <= kMaxRpt) {
        if (rpt <= 2) return ResCG<PREC, 2, Team>(tm, P, d).solve(b2, tol, maxit);
        if (kMaxRpt == 4 || rpt <= 4) return ResCG<PREC, 4, Team>(tm, P, d).solve(b2, tol, maxit);
        return ResCG<PREC, kMaxRpt, Team>(tm, P, d).solve(b2, tol, maxit);
    }
    return cg_island<PREC, OP_ARAP>(tm, P, d, b2, tol, maxit, -1, -1);
}
