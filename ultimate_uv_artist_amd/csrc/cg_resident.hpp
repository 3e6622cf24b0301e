// Island-resident preconditioned CG (included by kernels.hip).
//
// The same Chronopoulos-Gear PCG as cg_island (same recurrences, same
// per-lane accumulation order), for islands whose rows fit on the team's
// compute units.  Each lane keeps r, p, s = A p and w = A u for its rows in
// registers (at most 6 rows in FP32, 4 in FP64 at 512 threads): slot j holds
// row lane of the local slice j*W + wave of the member's slice range.  The
// preconditioned residual u = M^-1 r, which the SpMV gathers, the correction
// e and the diagonal D and D^-1 live in LDS, so a CG iteration reads only the
// matrix from memory.
//
// Preconditioner: Jacobi, plus (when the plan has one for the island) an
// additive coarse correction
//     M^-1 r = D^-1 r + Z Ac^-1 Z^T r,   Ac = Z^T A Z,
// over aggregates of g consecutive 64-row slices.  The plan orders vertices
// by Morton code, so slices are compact surface patches and the coarse level
// removes the smooth error that Jacobi-PCG needs O(sqrt(n)) iterations for.
// Z^T r: a slice is one wavefront's slot (lane = row), so its sum is a
// butterfly reduction across the wavefront's lanes of the r each lane holds
// in registers (FP64).  The slices' sums go to LDS (one workgroup) or through
// global memory to the whole team; then kPart threads per aggregate of the
// member each apply a strided share of the aggregate's row of Ac^-1 (kept in
// LDS for the whole solve) and one thread adds the shares in order.
//
// The same code solves the LSCM system (OP_LSCM): the operator couples the
// two columns through the boundary terms, [[K, B], [-B, K]], so both columns
// share one step length and one stop test; the two hard-pinned rows keep
// r = u = w = 0; and the coarse operator is [[Kc, Bc], [-Bc, Kc]] over the
// same aggregates (pin rows left out), whose inverse [[Pc, Qc], [-Qc, Pc]]
// is applied from the rows of [Pc | Qc].  In mixed precision each inner
// FP32 pass of the LSCM solve stops once it has reduced its residual by
// kInnerRel (an FP32 recurrence cannot reach the FP64 target of 1e-8 in one
// pass); the FP64 true residual of the next outer pass decides convergence.
//
// Members of a multi-workgroup team keep u for a window of rows that extends
// H = min(bandwidth, halo_rows) rows beyond each end of their own range.  They
// publish the rows other members' windows cover to global memory, and after
// each team barrier copy their window's foreign rows from there into LDS in
// one coalesced pass, so every gather of the SpMV reads LDS.  (An island whose
// bandwidth exceeds the halo publishes all rows and gathers the columns outside
// the window from global memory.)
#pragma once

// LDS bytes of a member holding `rows` rows (its slice range): the u window
// (rows + 2 * halo), e, D^-1 and D (the coarse level's share:
// res_coarse_bytes).
template <class V2, class V>
UVA_HD __forceinline__ int res_vec_bytes(int rows, int halo) {
    return res_vec_bytes_of(rows, halo, (int)sizeof(V2), (int)sizeof(V));
}

template <int PREC, int RPT, class Team, int OP = OP_ARAP>
struct ResCG {
    using PT = Prec<PREC>;
    using V = typename PT::V;
    using V2 = typename Vec2<V>::type;
    static constexpr bool kCoupled = OP == OP_LSCM;
    // matrix entries per slot whose loads are issued together (as many as the
    // registers hold beside the slots' vectors)
    static constexpr int kG = RPT * (int)sizeof(V) * (kCoupled ? 2 : 1) <= 16 ? 2 : 1;
    // halo rows of a team member's u window
    static constexpr int kH = halo_rows((int)sizeof(V2));

    Team& tm;
    const DevPlan& P;
    const Isl& d;
    int W, wv, lane, slo, shi, rlo, rhi;
    int H, wlo, whi;   // u window [wlo, whi) in LDS covers [rlo, rhi) and H rows each side
    bool multi;        // several members (halo exchange through global memory)
    bool gfull;        // the bandwidth exceeds the window: publish every row
    bool coarse;       // additive coarse correction
    int g, nc, ncp, alo, na;   // g slices per aggregate, nc aggregates (row stride ncp); the member's are [alo, alo + na)
    int crow;          // LDS stride of a row of the coarse inverse (ncp, or 2 ncp for [Pc | Qc])
    int lp0, lp1;      // hard-pinned rows of the LSCM system (-1 for ARAP)
    // LDS pointers are typed as such (UVA_LDS): the accesses are then LDS
    // instructions with 32-bit addresses, not generic (flat) ones, whose
    // waits also drain every global load in flight
    V2 UVA_LDS* lu;     // LDS: u of the window, row - wlo
    V2 UVA_LDS* le;     // LDS: e of the member's rows
    V UVA_LDS* ldv;     // LDS: D^-1
    V UVA_LDS* ldg;     // LDS: D (with the pin weight)
    float UVA_LDS* lci;     // LDS: rows alo .. alo + na - 1 of Ac^-1
    double2 UVA_LDS* lrc;   // LDS: Z^T r of every aggregate
    double2 UVA_LDS* ly;    // LDS: the member's aggregates' coarse correction
    double2 UVA_LDS* lpt;   // LDS: kPart shares of each of the member's mat-vec rows
    double2 UVA_LDS* lcs;   // LDS: Z^T r of the member's slices (one workgroup)
    V2* gu;     // global: u of the published rows (teams only)
    V2 r[RPT], p[RPT], s[RPT], w[RPT];

    UVA_HD __forceinline__ ResCG(Team& t, const DevPlan& P_, const Isl& d_, bool use_coarse, int p0 = -1, int p1 = -1)
        : tm(t), P(P_), d(d_), lp0(p0), lp1(p1) {
        W = tm.nwaves();
        wv = tm.wave();
        lane = tm.lane();
        slo = tm.sl_lo(d.ns);
        shi = tm.sl_hi(d.ns);
        rlo = d.v0 + slo * kSlice;
        rhi = min(d.v0 + shi * kSlice, d.v0 + d.nv);
        multi = tm.multi();
        H = multi ? min(d.bw, kH) : 0;
        gfull = multi && d.bw > kH;
        wlo = max(rlo - H, d.v0);
        whi = min(rhi + H, d.v0 + d.nv);
        coarse = use_coarse;
        g = kCoupled ? d.lcg : d.cg;
        nc = (d.ns + g - 1) / g;
        ncp = coarse_stride(nc);
        crow = kCoupled ? 2 * ncp : ncp;
        alo = slo / g;
        na = shi > slo ? (shi - 1) / g - alo + 1 : 0;
        const int rows = (shi - slo) * kSlice;
        char UVA_LDS* base = (char UVA_LDS*)tm.lvec();
        lu = (V2 UVA_LDS*)base;
        le = lu + rows + 2 * (multi ? kH : 0);
        ldv = (V UVA_LDS*)(le + rows);
        ldg = ldv + rows;
        lci = (float UVA_LDS*)(base + res_vec_bytes<V2, V>(rows, multi ? kH : 0));
        lrc = (double2 UVA_LDS*)((char UVA_LDS*)lci + align16(na * crow * 4));
        ly = lrc + nc;
        lpt = ly + na;
        lcs = lpt + na * kPart;
        gu = (V2*)P.vu;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int l = (j * W + wv) * kSlice + lane;
            const int row = rlo + l;
            if (slot(j) && row < rhi) {
                ldv[l] = (V)dinv_of<OP>(P, d, row);
                ldg[l] = (V)diag_of<OP>(P, d, row);
            }
            r[j].x = 0; r[j].y = 0;
            p[j] = r[j]; s[j] = r[j]; w[j] = r[j];
        }
        if (coarse) {
            // global rows: nc values (ARAP), or nc of Pc then nc of Qc (LSCM)
            const int gl = kCoupled ? 2 * nc : nc;
            const float* gci = kCoupled ? P.lcinv + d.lcofs + (long long)alo * gl : P.cinv + d.cofs + (long long)alo * nc;
            for (int i = tm.tid(); i < na * gl; i += tm.tstride()) {
                const int a = i / gl, b = i - a * gl;
                lci[a * crow + (b < nc ? b : ncp + (b - nc))] = gci[i];
            }
        }
        // each pass reads D^-1 and D of its own rows only, and the coarse
        // rows after a team barrier
    }
    UVA_HD __forceinline__ bool is_pin(int row) const { return kCoupled && (row == lp0 || row == lp1); }
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
    // u of an own row into the window (and to global memory when another
    // member's window or gathers may need it)
    UVA_HD __forceinline__ void publish(int row, V2 u) {
        lu[row - wlo] = u;
        if (multi && (gfull || row < rlo + H || row >= rhi - H)) gu[row] = u;
    }
    // After the team barrier: the window's foreign rows from global memory.
    UVA_HD __forceinline__ void fill_halo() {
        if (!multi) return;
        const int nlo = rlo - wlo, n = nlo + (whi - rhi);
        // kFill loads of each thread in flight together
        constexpr int kFill = 4;
        const int st = tm.tstride();
        for (int i0 = tm.tid(); i0 < n; i0 += kFill * st) {
            V2 t[kFill];
#pragma unroll
            for (int q = 0; q < kFill; ++q) {
                const int i = i0 + q * st;
                if (i < n) t[q] = gu[i < nlo ? wlo + i : rhi + (i - nlo)];
            }
#pragma unroll
            for (int q = 0; q < kFill; ++q) {
                const int i = i0 + q * st;
                if (i < n) lu[(i < nlo ? wlo + i : rhi + (i - nlo)) - wlo] = t[q];
            }
        }
        tm.bsync();
    }
    // Z^T r of slot j's slice: the sum over the wavefront's lanes of the r
    // each holds (0 for rows past the island).  Called by every lane of a
    // wavefront whose slot j exists; lane 0 stores the sum, to LDS for one
    // workgroup or to global memory for a team.
    UVA_HD __forceinline__ void restrict_slot(int j, V2 rv) {
        const double2 sm = tm.lane_sum2(make_double2((double)rv.x, (double)rv.y));
        if (lane == 0) {
            const int t = j * W + wv;
            if (multi) P.crc[d.s0 + slo + t] = sm;
            else lcs[t] = sm;
        }
    }
    // After every slice's Z^T r is stored: Z^T r per aggregate, then
    // y = Ac^-1 Z^T r for the member's aggregates, into LDS.
    UVA_HD __forceinline__ void coarse_solve() {
        if (multi) tm.sync();
        else tm.bsync();
        for (int b = tm.tid(); b < nc; b += tm.tstride()) {
            double2 acc = make_double2(0.0, 0.0);
            const int s1 = min((b + 1) * g, d.ns);
            for (int sl = b * g; sl < s1; ++sl) {
                const double2 v = multi ? P.crc[d.s0 + sl] : lcs[sl - slo];
                acc.x += v.x;
                acc.y += v.y;
            }
            lrc[b] = acc;
        }
        tm.bsync();
        // share q of row a: columns q, q + kPart, ...
        for (int i = tm.tid(); i < na * kPart; i += tm.tstride()) {
            const int a = i / kPart, q = i - a * kPart;
            const float UVA_LDS* ci = lci + a * crow;
            double yx = 0.0, yy = 0.0;
            for (int b = q; b < nc; b += kPart) {
                const double c = (double)ci[b];
                const double2 v = lrc[b];
                if (kCoupled) {
                    // (Pc r_u + Qc r_v, -Qc r_u + Pc r_v)
                    const double e = (double)ci[ncp + b];
                    yx = fma(c, v.x, fma(e, v.y, yx));
                    yy = fma(c, v.y, fma(-e, v.x, yy));
                } else {
                    yx = fma(c, v.x, yx);
                    yy = fma(c, v.y, yy);
                }
            }
            lpt[i] = make_double2(yx, yy);
        }
        tm.bsync();
        for (int a = tm.tid(); a < na; a += tm.tstride()) {
            double yx = 0.0, yy = 0.0;
            for (int q = 0; q < kPart; ++q) {
                const double2 v = lpt[a * kPart + q];
                yx += v.x;
                yy += v.y;
            }
            ly[a] = make_double2(yx, yy);
        }
        tm.bsync();
    }
    // u = M^-1 r of the member's rows into the window; returns the dots r.u
    UVA_HD __forceinline__ double2 precondition() {
        if (coarse) coarse_solve();
        refresh();
        double2 ru = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            if (!slot(j)) continue;
            const int l = lrow(j);
            const int row = rlo + l;
            if (row >= rhi) continue;
            const V dvj = ldv[l];
            V2 u;
            u.x = dvj * r[j].x;
            u.y = dvj * r[j].y;
            if (coarse) {
                const double2 y = ly[(slo + j * W + wv) / g - alo];
                u.x += (V)y.x;
                u.y += (V)y.y;
            }
            if (is_pin(row)) { u.x = 0; u.y = 0; }
            publish(row, u);
            ru.x += acc_mul<PREC>(r[j].x, u.x);
            ru.y += acc_mul<PREC>(r[j].y, u.y);
        }
        return ru;
    }

    // w = A u over the member's rows (own u from the window, columns from the
    // window or, beyond it, global memory) and the dot w.u.  The matrix loads
    // of every slot are issued together, kG entries per slot at a time, so a
    // wavefront waits for one memory round trip per kG entries of its widest
    // slice rather than one per slot.  Entries are summed in row order, as in
    // spmv_row2.
    UVA_HD __forceinline__ double2 spmv() {
        refresh();
        int base[RPT], wid[RPT];    // uniform per wavefront
        bool ok[RPT];               // the lane's row of slot j exists
        int wmax = 0;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            base[j] = 0;
            wid[j] = 0;
            if (slot(j)) {
                const int sg = d.s0 + slo + j * W + wv;
                base[j] = uniform_i(P.sl_ptr[sg]);
                wid[j] = (uniform_i(P.sl_ptr[sg + 1]) - base[j]) >> 6;
                wmax = wid[j] > wmax ? wid[j] : wmax;
            }
            const int l = lrow(j);
            ok[j] = slot(j) && rlo + l < rhi;
            V2 uo;
            uo.x = 0; uo.y = 0;
            if (ok[j]) uo = lu[rlo + l - wlo];
            const V dgj = ok[j] ? ldg[l] : (V)0;
            w[j].x = dgj * uo.x;
            w[j].y = dgj * uo.y;
        }
        const unsigned ulane = (unsigned)lane;
        const unsigned zero = (unsigned)P.m_zero + ulane;
        for (int k0 = 0; k0 < wmax; k0 += kG) {
            // every lane loads, so the loads of all slots form one straight
            // run the compiler issues back to back; offsets are unsigned.  A
            // step past a slot's width reads the zero entries after the last
            // slice and weighs the window's first row by 0: the FMAs run
            // unconditionally (no branch around each gather, no select on a
            // loaded value that would make the scheduler wait for it early).
            // A non-finite u there would make that 0 weight non-finite, but u
            // is the island's own preconditioned residual, and a non-finite
            // value in it already makes the next step length, and so every
            // row of the island, non-finite.
            int c[RPT][kG];
            V v[RPT][kG];
            V bv[RPT][kG];
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
#pragma unroll
                for (int q = 0; q < kG; ++q) {
                    const bool in = ok[j] && k0 + q < wid[j];
                    const unsigned o = in ? (unsigned)(base[j] + (k0 + q) * kSlice) + ulane : zero;
                    c[j][q] = P.m_col[o];
                    v[j][q] = mval<V>(P, o);
                    if (kCoupled) bv[j][q] = (V)P.m_b[o];
                    c[j][q] = in ? c[j][q] : wlo;
                }
            }
            // columns beyond the window (teams whose bandwidth exceeds the
            // halo): their global loads are issued first, all together, and
            // only by the lanes that need them; every other gather is an LDS
            // read
            V2 xf[RPT][kG];
            bool far[RPT][kG];
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
#pragma unroll
                for (int q = 0; q < kG; ++q) {
                    far[j][q] = gfull && (c[j][q] < wlo || c[j][q] >= whi);
                    xf[j][q].x = 0;
                    xf[j][q].y = 0;
                    if (far[j][q]) xf[j][q] = gu[c[j][q]];
                }
            }
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
#pragma unroll
                for (int q = 0; q < kG; ++q) {
                    const V2 xn = lu[far[j][q] ? 0 : c[j][q] - wlo];
                    const V2 x = far[j][q] ? xf[j][q] : xn;
                    if (kCoupled) {
                        w[j].x = fma(v[j][q], x.x, fma(bv[j][q], x.y, w[j].x));
                        w[j].y = fma(v[j][q], x.y, -fma(bv[j][q], x.x, -w[j].y));
                    } else {
                        w[j].x = fma(v[j][q], x.x, w[j].x);
                        w[j].y = fma(v[j][q], x.y, w[j].y);
                    }
                }
            }
        }
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            if (!ok[j]) continue;
            if (is_pin(rlo + lrow(j))) { w[j].x = 0; w[j].y = 0; }
            const V2 uo = lu[rlo + lrow(j) - wlo];
            acc.x += acc_mul<PREC>(w[j].x, uo.x);
            acc.y += acc_mul<PREC>(w[j].y, uo.y);
        }
        return acc;
    }

    UVA_HD __forceinline__ CGOut solve(double2 b2, double tol, int maxit) {
        const double bx2 = kCoupled ? fmax(b2.x + b2.y, 1e-300) : fmax(b2.x, 1e-300);
        const double by2 = kCoupled ? bx2 : fmax(b2.y, 1e-300);
        // an inner FP32 pass of the coupled solve also stops at kInnerRel of its start
        constexpr bool kRelInner = kCoupled && sizeof(V) == 4 && PT::kResid64;
        const double t2 = tol * tol;
        int it = 0;
        CGOut out{0, 0, 0.0};
        double relx = 0.0, rely = 0.0;
        for (int outer = 0;; ++outer) {
            // true residual r = b - A X (FP64 from the hi+lo matrix, or the
            // all-FP32 form), e = 0, u = M^-1 r
            double4 a1 = make_double4(0.0, 0.0, 0.0, 0.0);
            refresh();
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                if (!slot(j)) continue;
                const int l = lrow(j);
                const int row = rlo + l;
                const bool ok = row < rhi;
                if (ok) {
                    const int sg = d.s0 + slo + j * W + wv;
                    double rx, ry;
                    if (PT::kResid64) {
                        const double2 y = spmv_row<double, OP>(P, row, sg, lane, P.X64, diag_of<OP>(P, d, row));
                        const double2 bb = P.b64[row];
                        rx = bb.x - y.x;
                        ry = bb.y - y.y;
                    } else {
                        const int base = P.sl_ptr[sg];
                        const int width = (P.sl_ptr[sg + 1] - base) >> 6;
                        const double2 xr = P.X64[row];
                        const float dgf = (float)diag_of<OP>(P, d, row);
                        float yx = dgf * (float)xr.x, yy = dgf * (float)xr.y;
                        int o = base + lane;
                        for (int k = 0; k < width; ++k, o += kSlice) {
                            const double2 xc = P.X64[P.m_col[o]];
                            yx = fmaf(P.m_hi[o], (float)xc.x, yx);
                            yy = fmaf(P.m_hi[o], (float)xc.y, yy);
                            if (kCoupled) {
                                yx = fmaf(P.m_b[o], (float)xc.y, yx);
                                yy = fmaf(-P.m_b[o], (float)xc.x, yy);
                            }
                        }
                        const double2 bb = P.b64[row];
                        rx = (double)((float)bb.x - yx);
                        ry = (double)((float)bb.y - yy);
                    }
                    if (is_pin(row)) { rx = 0.0; ry = 0.0; }
                    r[j].x = (V)rx;
                    r[j].y = (V)ry;
                    V2 zero;
                    zero.x = 0;
                    zero.y = 0;
                    le[l] = zero;
                    a1.x += PT::kResid64 ? rx * rx : (double)((float)rx * (float)rx);
                    a1.y += PT::kResid64 ? ry * ry : (double)((float)ry * (float)ry);
                }
                if (coarse) restrict_slot(j, r[j]);
            }
            const double2 ru = precondition();
            a1.z += ru.x;
            a1.w += ru.y;
            tm.sync();
            fill_halo();
            const double2 wu = spmv();
            double4 S1, S2;
            tm.sum8(a1, make_double4(wu.x, wu.y, 0.0, 0.0), S1, S2);
            if (kCoupled) {
                relx = rely = (S1.x + S1.y) / bx2;
            } else {
                relx = S1.x / bx2;
                rely = S1.y / by2;
            }
            if (relx <= t2 && rely <= t2) { out.conv = 1; break; }
            if (outer >= PT::kMaxOuter || it >= maxit || !__builtin_isfinite(relx) || !__builtin_isfinite(rely)) break;
            const double rr0 = S1.x + S1.y;   // squared true residual at the start of this pass
            double gx = S1.z, gy = S1.w, dx = S2.x, dy = S2.y;
            if (kCoupled) { gx = gy = S1.z + S1.w; dx = dy = S2.x + S2.y; }
            double alx = cg_alpha<PREC>(gx, dx), aly = cg_alpha<PREC>(gy, dy);
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
                    const bool ok = rlo + l < rhi;
                    if (ok) {
                        // u of the last precondition step (each lane reads
                        // and rewrites only its own rows' slots here)
                        const V2 uA = lu[rlo + l - wlo];
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
                        g1.x += acc_mul<PREC>(r[j].x, r[j].x);
                        g1.y += acc_mul<PREC>(r[j].y, r[j].y);
                    }
                    if (coarse) restrict_slot(j, r[j]);
                }
                const double2 ru2 = precondition();
                g1.z += ru2.x;
                g1.w += ru2.y;
                tm.sync();
                fill_halo();
                const double2 wu2 = spmv();
                double4 T1, T2;
                tm.sum8(g1, make_double4(wu2.x, wu2.y, 0.0, 0.0), T1, T2);
                ++it;
                first = false;
                double rrx = T1.x, rry = T1.y;
                if (kCoupled) rrx = rry = T1.x + T1.y;
                if ((rrx <= PT::kInnerFrac2 * t2 * bx2 && rry <= PT::kInnerFrac2 * t2 * by2) || it >= maxit) break;
                if (kRelInner && rrx <= kInnerRel * kInnerRel * rr0) break;
                if (!__builtin_isfinite(rrx) || !__builtin_isfinite(rry)) break;
                double ngx = T1.z, ngy = T1.w, ndx = T2.x, ndy = T2.y;
                if (kCoupled) { ngx = ngy = T1.z + T1.w; ndx = ndy = T2.x + T2.y; }
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

// CG of one island (the ARAP global step, or the LSCM system with its hard
// pins p0, p1): the resident form when the member's rows fit in the team's
// registers and LDS, the streaming form otherwise.  The coarse level is used
// when the plan has one for the island and operator, `coarse` is set and its
// rows fit in LDS beside the vectors.
template <int PREC, int OP, class Team>
UVA_HD __forceinline__ CGOut cg_solve(Team& tm, const DevPlan& P, const Isl& d, double2 b2, double tol, int maxit,
                                      bool coarse, int p0, int p1) {
    using V = typename Prec<PREC>::V;
    using V2 = typename Vec2<V>::type;
    // rows per lane the registers hold without spilling at 512 threads
    constexpr int kMaxRpt = sizeof(V) == 4 ? 6 : 4;
    // Every member of a team must take the same path (resident or streaming,
    // with or without the coarse level): the paths pass different sequences
    // of team barriers.  The choice is made for the largest member share
    // (members' shares differ by one slice), which every member computes alike.
    const int m = tm.members();
    const int g = OP == OP_LSCM ? d.lcg : d.cg;     // slices per coarse aggregate
    int nsl = 0, na = 0;
    for (int k = 0; k < m; ++k) {
        const int lo = (int)((long long)d.ns * k / m), hi = (int)((long long)d.ns * (k + 1) / m);
        nsl = hi - lo > nsl ? hi - lo : nsl;
        if (g > 0 && hi > lo) {
            const int a = (hi - 1) / g - lo / g + 1;
            na = a > na ? a : na;
        }
    }
    const int W = tm.nwaves();
    const int rpt = (nsl + W - 1) / W;
    const int vec = res_vec_bytes<V2, V>(nsl * kSlice, tm.multi() ? halo_rows((int)sizeof(V2)) : 0);
    if (tm.lvec() != nullptr && vec <= tm.lvec_bytes() && rpt <= kMaxRpt) {
        constexpr bool cp = OP == OP_LSCM;
        const int nc = g > 0 ? (d.ns + g - 1) / g : 0;
        const bool co = coarse && (cp ? d.lcofs : d.cofs) >= 0 &&
                        vec + res_coarse_bytes(nc, na, nsl, cp) <= tm.lvec_bytes();
        if (rpt <= 2) return ResCG<PREC, 2, Team, OP>(tm, P, d, co, p0, p1).solve(b2, tol, maxit);
        if (rpt <= 4) return ResCG<PREC, 4, Team, OP>(tm, P, d, co, p0, p1).solve(b2, tol, maxit);
        return ResCG<PREC, kMaxRpt, Team, OP>(tm, P, d, co, p0, p1).solve(b2, tol, maxit);
    }
    return cg_island<PREC, OP>(tm, P, d, b2, tol, maxit, p0, p1);
}

template <int PREC, class Team>
UVA_HD __forceinline__ CGOut cg_arap(Team& tm, const DevPlan& P, const Isl& d, double2 b2, double tol, int maxit,
                                     bool coarse) {
    return cg_solve<PREC, OP_ARAP>(tm, P, d, b2, tol, maxit, coarse, -1, -1);
}
