// Jacobi-PCG over every row of a plan (uva_plan_pcg): the pieces shared by
// the device kernels (pcg.hip) and the CPU emulation (emu.inc).
#pragma once
#include <hip/hip_runtime.h>

#include "plan.hpp"

namespace uva {

// ---------------------------------------------------------------------------
// The same Jacobi-PCG over every row of a plan (uva_plan_pcg): the plan's own
// sliced-ELL Laplacian (values hi + lo, the islands' blocks side by side) with
// the per-row soft pins of uva_plan_set_pins, two columns.  This is the global
// step of gpu.Batch.arap and hier.Solver.relax without a second copy of the
// matrix.  Two launches per iteration instead of five: the SpMV runs on z and
// forms A p = A z + beta A p by recurrence next to p = z + beta p, and every
// workgroup sums the previous launch's block partials itself (a fixed-order
// sum of one value set per workgroup, identical in every workgroup), so no
// launch is spent on a one-workgroup reduction:
//   K1(it): r.r, r.z of the last update -> stop test (of iteration it - 1),
//           beta; q = A z, p = z + beta p, A p = q + beta A p; block p.Ap
//   K2(it): p.Ap -> alpha = r.z / max(p.Ap, 1e-30); x += alpha p,
//           r -= alpha A p, z = r / max(diag, 1e-30); block r.r, r.z
// The step lengths are the reference's (gpu.py:184-202): alpha = rz /
// max(p.Ap, 1e-30), beta = rz' / max(rz, 1e-30); the stop test after
// iteration it (it % check_every == 0) is r.r <= tol2 max(b.b, 1e-30) per
// column (< with strict).
namespace ppcg {

constexpr int kT = 512;         // threads per workgroup: eight 64-row slices at a time
constexpr int kSlicesPerWG = kT / 64;

struct State {
    double b2[2];
    double rz[2][2];            // r.z of iteration it in slot it & 1
    int done, its;
};

__host__ __device__ inline double pinned_diag(const DevPlan& P, int row) { return P.pdiag[row]; }

// y = (A + pins) v of `row` (lane `lane` of slice s): the diagonal, then the
// row's entries in column order; padding entries past the slice's width of
// the row are left out by select
__host__ __device__ inline double2 row_mv(const DevPlan& P, int s, int lane, int row, const double2* v) {
    const int base = P.sl_ptr[s];
    const int w = (P.sl_ptr[s + 1] - base) >> 6;
    const double dg = pinned_diag(P, row);
    const double2 vr = v[row];
    double yx = dg * vr.x, yy = dg * vr.y;
    for (int k0 = 0; k0 < w; k0 += 8) {
        int c[8];
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool in = k0 + u < w;
            const int o = base + (k0 + u) * 64 + lane;
            c[u] = in ? P.m_col[o] : row;
            a[u] = in ? (double)P.m_hi[o] + (double)P.m_lo[o] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool in = k0 + u < w;
            const double2 xv = v[c[u]];
            yx = in ? fma(a[u], xv.x, yx) : yx;
            yy = in ? fma(a[u], xv.y, yy) : yy;
        }
    }
    return make_double2(yx, yy);
}

__host__ __device__ inline bool stop_now(const double* rr, const double* b2, double tol2, int strict) {
    bool all = true;
    for (int c = 0; c < 2; ++c) all = all && (strict ? rr[c] < tol2 * b2[c] : rr[c] <= tol2 * b2[c]);
    return all;
}

}  // namespace ppcg

}  // namespace uva
