// Jacobi-preconditioned conjugate gradients on one symmetric matrix in CSR
// form, for one or two right-hand-side columns with their own step lengths:
// the coupled solves of the reference that span a whole batch of charts,
//   engine/gpu.py:184-202 Batch._cg     two columns (u, v), both checked
//                                       every iteration, stop when
//                                       r.r <= tol^2 max(b.b, 1e-30) for both
//   engine/hier.py:565-577 Solver.lscm  one 2n-row column (u and v coupled by
//                                       the boundary area term), checked every
//                                       50 iterations, stop when r.r < 1e-12 b.b
//   engine/hier.py:465-474 Solver.relax two columns, a fixed iteration count
// From the start x: alpha = rz / max(p.Ap, 1e-30), beta = rz' / max(rz,
// 1e-30), Jacobi weights 1 / max(diag, 1e-30) with the diagonal summed over
// repeated entries.
//
// These systems span every chart of a batch (millions of rows, one shared
// step length per column), so they run as ordinary grid-wide kernels over all
// CUs, five launches per iteration: SpMV + partial p.Ap; a one-workgroup
// reduction to alpha; the x, r, z update + partial r.r and r.z; a reduction
// to beta and the stop test; the direction update.  The stop test is decided
// on the device (a flag every later launch reads first), so a solve queues
// all its launches without a host round trip.  Dot products are block
// partial sums in a fixed tree order followed by a fixed-order reduction, so
// repeated runs are bitwise identical.  (Folding the reductions into the
// producing kernels with a last-workgroup-done counter needs a device-scope
// fence per workgroup, which on this multi-XCD part writes back L2 and made
// the solve three times slower.)
//
// The matrix is held in sliced ELL: per slice of 64 rows (one wavefront) the
// entries of its rows column-major, width = the slice's longest row, 32-bit
// column ids (-1 pads), so each load of a wavefront is one contiguous 256 or
// 512-byte run.  Entries keep the CSR order of their row, so a row's sum is
// the CSR sum.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/uvart.h"
#include "plan_pcg.hpp"

namespace uva {

namespace pcg {

constexpr int kB = 256;        // rows per workgroup
constexpr int kR = 1024;       // threads of the final reductions

struct Scalars {
    double rz[2], b2[2], al[2], beta[2], rr[2];
    int done, its;
};

__host__ __device__ inline double clamp30(double v) { return v > 1e-30 ? v : 1e-30; }

__host__ __device__ inline double row_diag_inv(const int64_t* ip, const int64_t* ix, const double* a, int64_t i) {
    double g = 0.0;
    for (int64_t q = ip[i]; q < ip[i + 1]; ++q)
        if (ix[q] == i) g += a[q];
    return 1.0 / clamp30(g);
}

constexpr int kSl = 64;       // rows per slice of the sliced-ELL matrix

// y = A v of row i from the sliced-ELL arrays (slice offsets sp in entries)
template <int NC>
__host__ __device__ inline void row_mv_sell(const int64_t* sp, const int* col, const double* val, const double* v,
                                            int64_t i, double* out) {
    double s[NC];
    for (int c = 0; c < NC; ++c) s[c] = 0.0;
    const int64_t sl = i / kSl, lane = i % kSl;
    const int64_t w = (sp[sl + 1] - sp[sl]) / kSl;
    for (int64_t k = 0; k < w; ++k) {
        const int64_t q = sp[sl] + k * kSl + lane;
        const int j = col[q];
        if (j >= 0) {
            const double a = val[q];
            for (int c = 0; c < NC; ++c) s[c] += a * v[(int64_t)j * NC + c];
        }
    }
    for (int c = 0; c < NC; ++c) out[c] = s[c];
}

// sliced ELL of a CSR matrix (n rows): slice offsets (ns + 1), columns, values
inline void build_sell(int64_t n, const int64_t* ip, const int64_t* ix, const double* a, std::vector<int64_t>& sp,
                       std::vector<int>& col, std::vector<double>& val) {
    const int64_t ns = (n + kSl - 1) / kSl;
    sp.assign(ns + 1, 0);
    for (int64_t sl = 0; sl < ns; ++sl) {
        int64_t w = 0;
        for (int64_t i = sl * kSl; i < n && i < (sl + 1) * kSl; ++i) w = ip[i + 1] - ip[i] > w ? ip[i + 1] - ip[i] : w;
        sp[sl + 1] = sp[sl] + w * kSl;
    }
    col.assign(sp[ns], -1);
    val.assign(sp[ns], 0.0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t sl = i / kSl, lane = i % kSl;
        for (int64_t q = ip[i]; q < ip[i + 1]; ++q) {
            const int64_t o = sp[sl] + (q - ip[i]) * kSl + lane;
            col[o] = (int)ix[q];
            val[o] = a[q];
        }
    }
}

// stop test of one check: every column below its threshold
__host__ __device__ inline bool converged(const double* rr, const double* b2, int nc, double tol2, int strict) {
    bool all = true;
    for (int c = 0; c < nc; ++c) all = all && (strict ? rr[c] < tol2 * b2[c] : rr[c] <= tol2 * b2[c]);
    return all;
}

// host checks of a CSR matrix: "" when valid
inline const char* check_csr(int64_t n, const int64_t* indptr, const int64_t* indices, const double* data) {
    if (n < 1) return "matrix must have at least one row";
    if (!indptr) return "null indptr";
    if (indptr[0] != 0) return "indptr must start at 0";
    for (int64_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) return "indptr must be non-decreasing";
    if (indptr[n] > 0 && (!indices || !data)) return "null indices or data";
    for (int64_t q = 0; q < indptr[n]; ++q)
        if (indices[q] < 0 || indices[q] >= n) return "column index out of range";
    if (n > 0x7fffffffLL) return "matrix too large for 32-bit column ids";
    return "";
}

}  // namespace pcg
}  // namespace uva

#if !defined(UVA_EMULATOR)
struct uva_csr {
    int64_t n = 0, nnz = 0, nb = 0, nsell = 0;
    int64_t* sp = nullptr;           // slice offsets (entries)
    int* col = nullptr;              // column ids, -1 pads
    double *val = nullptr, *dinv = nullptr, *work = nullptr, *part = nullptr;
    uva::pcg::Scalars* sc = nullptr;
};

namespace uva {
namespace pcg {

// block sums of K values per thread into part[K * block + k]
template <int K>
__device__ inline void block_sum(const double (&v)[K], double* part) {
    __shared__ double s[K][kB];
    for (int k = 0; k < K; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int h = kB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < K; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) part[K * blockIdx.x + k] = s[k][0];
}

// fixed-order sums of the nb block partials (K per block) for NC columns:
// stage 0 (after init, K = 2 NC):   rz, b2; done = 0; its = max_iters
// stage 1 (after spmv, K = NC):     al
// stage 2 (after update, K = 2 NC): rr, beta, rz; stop test at check iterations
template <int NC, int K>
__global__ void reduce_kernel(const double* part, int64_t nb, int stage, int it, int max_iters, int check_every,
                              double tol2, int strict, Scalars* s) {
    if (stage != 0 && s->done) return;
    __shared__ double sh[K][kR];
    double u[K];
    for (int k = 0; k < K; ++k) u[k] = 0.0;
    for (int64_t q = threadIdx.x; q < nb; q += kR)
        for (int k = 0; k < K; ++k) u[k] += part[K * q + k];
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = u[k];
    __syncthreads();
    for (int h = kR / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        for (int c = 0; c < NC; ++c) {
            s->rz[c] = sh[c][0];
            s->b2[c] = clamp30(sh[NC + c][0]);
        }
        s->done = 0;
        s->its = max_iters;
    } else if (stage == 1) {
        for (int c = 0; c < NC; ++c) s->al[c] = s->rz[c] / clamp30(sh[c][0]);
    } else {
        for (int c = 0; c < NC; ++c) {
            s->rr[c] = sh[c][0];
            s->beta[c] = sh[NC + c][0] / clamp30(s->rz[c]);
            s->rz[c] = sh[NC + c][0];
        }
        if (it % check_every == 0 && converged(s->rr, s->b2, NC, tol2, strict)) {
            s->done = 1;
            s->its = it + 1;
        }
    }
}

template <int NC>
__global__ void init_kernel(int64_t n, const int64_t* sp, const int* col, const double* val, const double* dinv,
                            const double* b, const double* x, double* r, double* z, double* p, double* part) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double v[2 * NC];
    for (int k = 0; k < 2 * NC; ++k) v[k] = 0.0;
    if (i < n) {
        double ax[NC];
        row_mv_sell<NC>(sp, col, val, x, i, ax);
        const double di = dinv[i];
        for (int c = 0; c < NC; ++c) {
            const double bi = b[i * NC + c];
            const double ri = bi - ax[c];
            r[i * NC + c] = ri;
            z[i * NC + c] = ri * di;
            p[i * NC + c] = ri * di;
            v[c] = ri * (ri * di);
            v[NC + c] = bi * bi;
        }
    }
    block_sum<2 * NC>(v, part);
}

// Ap = A p; partial p.Ap per column
template <int NC>
__global__ void spmv_kernel(int64_t n, const int64_t* sp, const int* col, const double* val, const double* p,
                            double* Ap, double* part, const Scalars* s) {
    if (s->done) return;
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double v[NC];
    for (int c = 0; c < NC; ++c) v[c] = 0.0;
    if (i < n) {
        double ap[NC];
        row_mv_sell<NC>(sp, col, val, p, i, ap);
        for (int c = 0; c < NC; ++c) {
            Ap[i * NC + c] = ap[c];
            v[c] = p[i * NC + c] * ap[c];
        }
    }
    block_sum<NC>(v, part);
}

// x += al p, r -= al Ap, z = D r; partial r.r and r.z per column
template <int NC>
__global__ void update_kernel(int64_t n, const double* dinv, const double* p, const double* Ap, double* x, double* r,
                              double* z, double* part, const Scalars* s) {
    if (s->done) return;
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double v[2 * NC];
    for (int k = 0; k < 2 * NC; ++k) v[k] = 0.0;
    if (i < n) {
        const double di = dinv[i];
        for (int c = 0; c < NC; ++c) {
            const double al = s->al[c];
            x[i * NC + c] += al * p[i * NC + c];
            const double ri = r[i * NC + c] - al * Ap[i * NC + c];
            const double zi = ri * di;
            r[i * NC + c] = ri;
            z[i * NC + c] = zi;
            v[c] = ri * ri;
            v[NC + c] = ri * zi;
        }
    }
    block_sum<2 * NC>(v, part);
}

// p = z + beta p
template <int NC>
__global__ void direction_kernel(int64_t n, const double* z, double* p, const Scalars* s) {
    if (s->done) return;
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n)
        for (int c = 0; c < NC; ++c) p[i * NC + c] = z[i * NC + c] + s->beta[c] * p[i * NC + c];
}

template <int NC>
int solve(uva_csr* m, const double* b, double* x, int max_iters, int check_every, double tol2, int strict,
          hipStream_t st) {
    const int64_t n = m->n, nb = m->nb;
    double *r = m->work, *z = r + NC * n, *p = z + NC * n, *Ap = p + NC * n;
    const dim3 grid((unsigned)nb), blk(kB), one(1), red(kR);
    hipLaunchKernelGGL(init_kernel<NC>, grid, blk, 0, st, n, m->sp, m->col, m->val, m->dinv, b, x, r, z, p, m->part);
    hipLaunchKernelGGL((reduce_kernel<NC, 2 * NC>), one, red, 0, st, m->part, nb, 0, 0, max_iters, check_every,
                       tol2, strict, m->sc);
    for (int it = 0; it < max_iters; ++it) {
        hipLaunchKernelGGL(spmv_kernel<NC>, grid, blk, 0, st, n, m->sp, m->col, m->val, p, Ap, m->part, m->sc);
        hipLaunchKernelGGL((reduce_kernel<NC, NC>), one, red, 0, st, m->part, nb, 1, it, max_iters, check_every,
                           tol2, strict, m->sc);
        hipLaunchKernelGGL(update_kernel<NC>, grid, blk, 0, st, n, m->dinv, p, Ap, x, r, z, m->part, m->sc);
        hipLaunchKernelGGL((reduce_kernel<NC, 2 * NC>), one, red, 0, st, m->part, nb, 2, it, max_iters, check_every,
                           tol2, strict, m->sc);
        hipLaunchKernelGGL(direction_kernel<NC>, grid, blk, 0, st, n, z, p, m->sc);
    }
    return hipGetLastError() == hipSuccess ? UVA_OK : UVA_EHIP;
}

}  // namespace pcg

namespace ppcg {

// fixed-order sum of the G block partials (K values each) into every thread
template <int K>
__device__ inline void sum_parts(const double* part, int G, double* sh, double (&out)[K]) {
    double u[K];
    for (int k = 0; k < K; ++k) u[k] = 0.0;
    for (int q = threadIdx.x; q < G; q += kT)
        for (int k = 0; k < K; ++k) u[k] += part[K * q + k];
    for (int k = 0; k < K; ++k) sh[k * kT + threadIdx.x] = u[k];
    __syncthreads();
    for (int h = kT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < K; ++k) sh[k * kT + threadIdx.x] += sh[k * kT + threadIdx.x + h];
        __syncthreads();
    }
    for (int k = 0; k < K; ++k) out[k] = sh[k * kT];
    __syncthreads();
}

// block sum of K values per thread into part[K * blockIdx.x + k]
template <int K>
__device__ inline void block_part(const double (&v)[K], double* sh, double* part) {
    for (int k = 0; k < K; ++k) sh[k * kT + threadIdx.x] = v[k];
    __syncthreads();
    for (int h = kT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < K; ++k) sh[k * kT + threadIdx.x] += sh[k * kT + threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) part[K * blockIdx.x + k] = sh[k * kT];
}

// Rows of this workgroup: slices blockIdx.x * 8 + wave, then every 8 gridDim.x.
#define PPCG_ROWS(P, s, lane, row)                                                                     \
    for (int s = blockIdx.x * kSlicesPerWG + (int)(threadIdx.x >> 6); s < (P).n_slices;                \
         s += gridDim.x * kSlicesPerWG)                                                                \
        for (int lane = (int)(threadIdx.x & 63), row = (P).sl_row[s] + lane, _once = 1;                \
             _once && row < (P).sl_row[s + 1]; _once = 0)

// r = b - (A + pins) x, z = r / max(diag, 1e-30); block r.z and b.b
__global__ __launch_bounds__(kT) void init_kernel(DevPlan P, const double2* __restrict__ b,
                                                   const double2* __restrict__ x, double2* __restrict__ r,
                                                   double2* __restrict__ z, double* partB, State* st, int max_iters) {
    __shared__ double sh[4 * kT];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    PPCG_ROWS(P, s, lane, row) {
        const double2 ax = row_mv(P, s, lane, row, x);
        const double2 bi = b[row];
        const double di = 1.0 / fmax(pinned_diag(P, row), 1e-30);
        const double rx = bi.x - ax.x, ry = bi.y - ax.y;
        const double zx = rx * di, zy = ry * di;
        r[row] = make_double2(rx, ry);
        z[row] = make_double2(zx, zy);
        v[0] += rx * zx;
        v[1] += ry * zy;
        v[2] += bi.x * bi.x;
        v[3] += bi.y * bi.y;
    }
    block_part<4>(v, sh, partB);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->done = 0;
        st->its = max_iters;
    }
}

// K1(it): see the namespace comment
__global__ __launch_bounds__(kT) void spmv_kernel(DevPlan P, const double2* __restrict__ z, double2* __restrict__ p,
                                                   double2* __restrict__ Ap, const double* partB, double* partA,
                                                   State* st, int it, int check_every, double tol2, int strict) {
    __shared__ double sh[4 * kT];
    if (st->done) return;
    double red[4];
    sum_parts<4>(partB, gridDim.x, sh, red);
    double rz[2], beta[2] = {0.0, 0.0};
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (it == 0) {
        rz[0] = red[0];
        rz[1] = red[1];
        if (lead) {
            st->b2[0] = red[2] > 1e-30 ? red[2] : 1e-30;
            st->b2[1] = red[3] > 1e-30 ? red[3] : 1e-30;
        }
    } else {
        const double rr[2] = {red[0], red[1]};
        const double b2[2] = {st->b2[0], st->b2[1]};
        if ((it - 1) % check_every == 0 && stop_now(rr, b2, tol2, strict)) {
            if (lead) {
                st->done = 1;
                st->its = it;
            }
            return;
        }
        rz[0] = red[2];
        rz[1] = red[3];
        for (int c = 0; c < 2; ++c) {
            const double prev = st->rz[(it - 1) & 1][c];
            beta[c] = rz[c] / (prev > 1e-30 ? prev : 1e-30);
        }
    }
    if (lead) {
        st->rz[it & 1][0] = rz[0];
        st->rz[it & 1][1] = rz[1];
    }
    double v[2] = {0.0, 0.0};
    PPCG_ROWS(P, s, lane, row) {
        const double2 q = row_mv(P, s, lane, row, z);
        const double2 zi = z[row];
        double2 pi, api;
        if (it == 0) {
            pi = zi;
            api = q;
        } else {
            const double2 po = p[row], ao = Ap[row];
            pi = make_double2(zi.x + beta[0] * po.x, zi.y + beta[1] * po.y);
            api = make_double2(q.x + beta[0] * ao.x, q.y + beta[1] * ao.y);
        }
        p[row] = pi;
        Ap[row] = api;
        v[0] += pi.x * api.x;
        v[1] += pi.y * api.y;
    }
    block_part<2>(v, sh, partA);
}

// K2(it): see the namespace comment
__global__ __launch_bounds__(kT) void update_kernel(DevPlan P, const double2* __restrict__ p,
                                                     const double2* __restrict__ Ap, double2* __restrict__ x,
                                                     double2* __restrict__ r, double2* __restrict__ z,
                                                     const double* partA, double* partB, const State* st, int it) {
    __shared__ double sh[4 * kT];
    if (st->done) return;
    double red[2];
    sum_parts<2>(partA, gridDim.x, sh, red);
    double al[2];
    for (int c = 0; c < 2; ++c) al[c] = st->rz[it & 1][c] / (red[c] > 1e-30 ? red[c] : 1e-30);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    PPCG_ROWS(P, s, lane, row) {
        const double2 pi = p[row], api = Ap[row];
        double2 xi = x[row];
        xi.x += al[0] * pi.x;
        xi.y += al[1] * pi.y;
        x[row] = xi;
        double2 ri = r[row];
        ri.x -= al[0] * api.x;
        ri.y -= al[1] * api.y;
        r[row] = ri;
        const double di = 1.0 / fmax(pinned_diag(P, row), 1e-30);
        const double zx = ri.x * di, zy = ri.y * di;
        z[row] = make_double2(zx, zy);
        v[0] += ri.x * ri.x;
        v[1] += ri.y * ri.y;
        v[2] += ri.x * zx;
        v[3] += ri.y * zy;
    }
    block_part<4>(v, sh, partB);
}
#undef PPCG_ROWS

}  // namespace ppcg

// Workgroups of the plan PCG on a device of `cus` compute units: four per CU
// at most (32 wavefronts per CU hide the gathers' latency; every workgroup
// reads all block partials, so more would cost L2 traffic), one per eight
// slices at least.
int plan_pcg_grid(const DevPlan& P, int cus) {
    const int need = (P.n_slices + ppcg::kSlicesPerWG - 1) / ppcg::kSlicesPerWG;
    const int cap = 4 * (cus > 0 ? cus : 1);
    return need < cap ? (need > 0 ? need : 1) : cap;
}

size_t plan_pcg_scratch_doubles(int grid) { return (size_t)grid * 6 + sizeof(ppcg::State) / sizeof(double) + 1; }

// The launches of one plan PCG (x, b, the four work vectors in plan order;
// scratch: plan_pcg_scratch_doubles(grid) doubles).  Returns the launch error.
hipError_t launch_plan_pcg(const DevPlan& P, const double2* b, double2* x, double2* r, double2* z, double2* p,
                           double2* Ap, double* scratch, int grid, int max_iters, int check_every, double tol2,
                           int strict, hipStream_t s) {
    using namespace ppcg;
    double* partA = scratch;
    double* partB = scratch + 2 * (size_t)grid;
    State* st = (State*)(scratch + 6 * (size_t)grid);
    hipLaunchKernelGGL(init_kernel, dim3(grid), dim3(kT), 0, s, P, b, x, r, z, partB, st, max_iters);
    for (int it = 0; it < max_iters; ++it) {
        hipLaunchKernelGGL(spmv_kernel, dim3(grid), dim3(kT), 0, s, P, z, p, Ap, partB, partA, st, it, check_every,
                           tol2, strict);
        hipLaunchKernelGGL(update_kernel, dim3(grid), dim3(kT), 0, s, P, p, Ap, x, r, z, partA, partB, st, it);
    }
    return hipGetLastError();
}

// iterations the last plan PCG ran (the State at the end of the scratch)
int* plan_pcg_iters_ptr(double* scratch, int grid) { return &((ppcg::State*)(scratch + 6 * (size_t)grid))->its; }

}  // namespace uva

static void csr_free(uva_csr* m) {
    (void)hipFree(m->sp);
    (void)hipFree(m->col);
    (void)hipFree(m->val);
    (void)hipFree(m->dinv);
    (void)hipFree(m->work);
    (void)hipFree(m->part);
    (void)hipFree(m->sc);
    delete m;
}

extern "C" int uva_csr_create(uva_csr** out, int64_t n, const int64_t* indptr, const int64_t* indices,
                              const double* data, void* stream) {
    using namespace uva::pcg;
    if (!out) return UVA_EINVAL;
    *out = nullptr;
    if (*check_csr(n, indptr, indices, data)) return UVA_EINVAL;
    const int64_t nb = (n + kB - 1) / kB;
    if (nb > (int64_t)0x7fffffff) return UVA_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int64_t> sp;
    std::vector<int> col;
    std::vector<double> val, dinv((size_t)n);
    build_sell(n, indptr, indices, data, sp, col, val);
    for (int64_t i = 0; i < n; ++i) dinv[i] = row_diag_inv(indptr, indices, data, i);
    uva_csr* m = new uva_csr();
    m->n = n;
    m->nnz = indptr[n];
    m->nb = nb;
    m->nsell = (int64_t)col.size();
    const size_t ne = col.size() + 1;
    // work: r, z, p, Ap for up to two columns
    if (hipMalloc(&m->sp, sp.size() * 8) != hipSuccess || hipMalloc(&m->col, ne * 4) != hipSuccess ||
        hipMalloc(&m->val, ne * 8) != hipSuccess || hipMalloc(&m->dinv, (size_t)n * 8) != hipSuccess ||
        hipMalloc(&m->work, (size_t)n * 8 * 8) != hipSuccess || hipMalloc(&m->part, (size_t)nb * 4 * 8) != hipSuccess ||
        hipMalloc(&m->sc, sizeof(Scalars)) != hipSuccess) {
        csr_free(m);
        return UVA_ENOMEM;
    }
    if (hipMemcpyAsync(m->sp, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        (!col.empty() && hipMemcpyAsync(m->col, col.data(), col.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (!val.empty() && hipMemcpyAsync(m->val, val.data(), val.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemcpyAsync(m->dinv, dinv.data(), (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        csr_free(m);
        return UVA_EHIP;
    }
    *out = m;
    return UVA_OK;
}

extern "C" int uva_csr_destroy(uva_csr* m) {
    if (!m) return UVA_OK;
    (void)hipDeviceSynchronize();
    csr_free(m);
    return UVA_OK;
}

extern "C" int uva_csr_pcg(uva_csr* m, const double* b, double* x, int32_t ncols, int32_t max_iters,
                           int32_t check_every, double tol2, int32_t strict, int32_t* iters_out, void* stream) {
    using namespace uva::pcg;
    if (!m || !b || !x || (ncols != 1 && ncols != 2) || max_iters < 0 || check_every < 1 || !(tol2 >= 0.0))
        return UVA_EINVAL;
    if (iters_out) *iters_out = 0;
    if (max_iters == 0) return UVA_OK;
    hipStream_t st = (hipStream_t)stream;
    const int rc = ncols == 1 ? solve<1>(m, b, x, max_iters, check_every, tol2, strict, st)
                              : solve<2>(m, b, x, max_iters, check_every, tol2, strict, st);
    if (rc == UVA_OK && iters_out) {
        int its = 0;
        if (hipMemcpyAsync(&its, &m->sc->its, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return UVA_EHIP;
        *iters_out = its;
    }
    return rc;
}

extern "C" int uva_jacobi_pcg(int64_t n, const int64_t* indptr, const int64_t* indices, const double* data,
                              const double* b, double* x, int32_t max_iters, int32_t check_every, double tol2,
                              int32_t* iters_out, void* stream) {
    if (n < 0 || max_iters < 0 || check_every < 1 || !(tol2 >= 0.0)) return UVA_EINVAL;
    if (iters_out) *iters_out = 0;
    if (n == 0) return UVA_OK;
    if (!b || !x) return UVA_EINVAL;
    uva_csr* m = nullptr;
    int rc = uva_csr_create(&m, n, indptr, indices, data, stream);
    if (rc) return rc;
    rc = uva_csr_pcg(m, b, x, 1, max_iters, check_every, tol2, 1, iters_out, stream);
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess && rc == UVA_OK) rc = UVA_EHIP;
    uva_csr_destroy(m);
    return rc;
}
#else
// The same iteration on the host, on the same sliced-ELL arrays and row code,
// rows in order and sequential sums (test-only emulation; equal to the device
// result up to the order of the dot-product sums).  b and x are host arrays, n x ncols.
extern "C" int uva_emu_csr_pcg(int64_t n, const int64_t* indptr, const int64_t* indices, const double* data,
                               const double* b, double* x, int32_t ncols, int32_t max_iters, int32_t check_every,
                               double tol2, int32_t strict, int32_t* iters_out) {
    using namespace uva::pcg;
    if ((ncols != 1 && ncols != 2) || max_iters < 0 || check_every < 1 || !(tol2 >= 0.0)) return UVA_EINVAL;
    if (*check_csr(n, indptr, indices, data) || !b || !x) return UVA_EINVAL;
    if (iters_out) *iters_out = 0;
    if (max_iters == 0) return UVA_OK;
    const int nc = ncols;
    std::vector<int64_t> sp;
    std::vector<int> col;
    std::vector<double> val;
    build_sell(n, indptr, indices, data, sp, col, val);
    std::vector<double> dinv(n), r(n * nc), z(n * nc), p(n * nc), Ap(n * nc);
    double rz[2] = {0, 0}, b2[2] = {0, 0}, t[2];
    for (int64_t i = 0; i < n; ++i) {
        dinv[i] = row_diag_inv(indptr, indices, data, i);
        if (nc == 1) row_mv_sell<1>(sp.data(), col.data(), val.data(), x, i, t);
        else row_mv_sell<2>(sp.data(), col.data(), val.data(), x, i, t);
        for (int c = 0; c < nc; ++c) {
            r[i * nc + c] = b[i * nc + c] - t[c];
            z[i * nc + c] = p[i * nc + c] = r[i * nc + c] * dinv[i];
            rz[c] += r[i * nc + c] * z[i * nc + c];
            b2[c] += b[i * nc + c] * b[i * nc + c];
        }
    }
    for (int c = 0; c < nc; ++c) b2[c] = clamp30(b2[c]);
    int it_run = max_iters;
    for (int it = 0; it < max_iters; ++it) {
        double pap[2] = {0, 0}, rr[2] = {0, 0}, rz2[2] = {0, 0}, al[2];
        for (int64_t i = 0; i < n; ++i) {
            if (nc == 1) row_mv_sell<1>(sp.data(), col.data(), val.data(), p.data(), i, &Ap[i]);
            else row_mv_sell<2>(sp.data(), col.data(), val.data(), p.data(), i, &Ap[2 * i]);
            for (int c = 0; c < nc; ++c) pap[c] += p[i * nc + c] * Ap[i * nc + c];
        }
        for (int c = 0; c < nc; ++c) al[c] = rz[c] / clamp30(pap[c]);
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < nc; ++c) {
                x[i * nc + c] += al[c] * p[i * nc + c];
                r[i * nc + c] -= al[c] * Ap[i * nc + c];
                z[i * nc + c] = r[i * nc + c] * dinv[i];
                rr[c] += r[i * nc + c] * r[i * nc + c];
                rz2[c] += r[i * nc + c] * z[i * nc + c];
            }
        if (it % check_every == 0 && converged(rr, b2, nc, tol2, strict)) {
            it_run = it + 1;
            break;
        }
        for (int c = 0; c < nc; ++c) {
            const double beta = rz2[c] / clamp30(rz[c]);
            rz[c] = rz2[c];
            for (int64_t i = 0; i < n; ++i) p[i * nc + c] = z[i * nc + c] + beta * p[i * nc + c];
        }
    }
    if (iters_out) *iters_out = it_run;
    return UVA_OK;
}
#endif
