// Jacobi-preconditioned conjugate gradients on one symmetric matrix in CSR
// form, with the iteration and stopping rule of the reference's free-boundary
// LSCM solve (engine/hier.py:565-577): from the start x, at most max_iters
// iterations; after iteration k (0-based) with k % check_every == 0 the solve
// stops when r.r < tol2 * max(b.b, 1e-30).  alpha = rz / max(p.Ap, 1e-30),
// beta = rz' / max(rz, 1e-30), Jacobi weights 1 / max(diag, 1e-30) with the
// diagonal summed over repeated entries.
//
// The operator of that solve couples the u and v columns of every chart (the
// area term on the boundary), so it is one 2n-row system, not the island
// kernel's two-column Laplacian.  The matrix comes from the host (checked
// there, then copied); b and x are device arrays.  Rows are independent per thread (one thread
// per row, CSR gather); dot products are block partial sums in a fixed tree
// order followed by a one-workgroup reduction in a fixed order, so repeated
// runs are bitwise identical.  The scalars stay on the device; the host reads
// the residual only at check iterations.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/uvart.h"

namespace uva {
namespace pcg {

constexpr int kB = 256;        // rows per workgroup
constexpr int kR = 1024;       // threads of the final reduction

struct Scalars {
    double rz, b2, al, rr, beta;
};

__host__ __device__ inline double row_dot(const int64_t* ip, const int64_t* ix, const double* a, const double* v,
                                          int64_t i) {
    double s = 0.0;
    for (int64_t q = ip[i]; q < ip[i + 1]; ++q) s += a[q] * v[ix[q]];
    return s;
}

__host__ __device__ inline double row_diag(const int64_t* ip, const int64_t* ix, const double* a, int64_t i) {
    double g = 0.0;
    for (int64_t q = ip[i]; q < ip[i + 1]; ++q)
        if (ix[q] == i) g += a[q];
    return 1.0 / (g > 1e-30 ? g : 1e-30);
}

__host__ __device__ inline double clamp30(double v) { return v > 1e-30 ? v : 1e-30; }

#if !defined(UVA_EMULATOR)
// block sum of two values per thread into part[2*block], part[2*block+1]
__device__ inline void block_sum2(double u, double v, double* part) {
    __shared__ double su[kB], sv[kB];
    su[threadIdx.x] = u;
    sv[threadIdx.x] = v;
    __syncthreads();
    for (int h = kB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            su[threadIdx.x] += su[threadIdx.x + h];
            sv[threadIdx.x] += sv[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = su[0];
        part[2 * blockIdx.x + 1] = sv[0];
    }
}

// r = b - A x, z = D r, p = z; partial r.z and b.b
__global__ void init_kernel(int64_t n, const int64_t* ip, const int64_t* ix, const double* a, const double* b,
                            const double* x, double* dinv, double* r, double* z, double* p, double* part) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double rz = 0.0, bb = 0.0;
    if (i < n) {
        const double di = row_diag(ip, ix, a, i);
        const double ri = b[i] - row_dot(ip, ix, a, x, i);
        dinv[i] = di;
        r[i] = ri;
        z[i] = ri * di;
        p[i] = ri * di;
        rz = ri * (ri * di);
        bb = b[i] * b[i];
    }
    block_sum2(rz, bb, part);
}

// Ap = A p; partial p.Ap
__global__ void spmv_kernel(int64_t n, const int64_t* ip, const int64_t* ix, const double* a, const double* p,
                            double* Ap, double* part) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double pap = 0.0;
    if (i < n) {
        const double v = row_dot(ip, ix, a, p, i);
        Ap[i] = v;
        pap = p[i] * v;
    }
    block_sum2(pap, 0.0, part);
}

// x += al p, r -= al Ap, z = D r; partial r.r and r.z
__global__ void update_kernel(int64_t n, const Scalars* s, const double* dinv, const double* p, const double* Ap,
                              double* x, double* r, double* z, double* part) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < n) {
        const double al = s->al;
        x[i] += al * p[i];
        const double ri = r[i] - al * Ap[i];
        r[i] = ri;
        const double zi = ri * dinv[i];
        z[i] = zi;
        rr = ri * ri;
        rz = ri * zi;
    }
    block_sum2(rr, rz, part);
}

// p = z + beta p
__global__ void direction_kernel(int64_t n, const Scalars* s, const double* z, double* p) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = z[i] + s->beta * p[i];
}

// sums of the nb block pairs in a fixed order, then the step scalars:
// stage 0 (after init): rz, b2; stage 1 (after spmv): al; stage 2 (after
// update): rr, beta, rz
__global__ void reduce_kernel(const double* part, int64_t nb, int stage, Scalars* s) {
    __shared__ double su[kR], sv[kR];
    double u = 0.0, v = 0.0;
    for (int64_t k = threadIdx.x; k < nb; k += kR) {
        u += part[2 * k];
        v += part[2 * k + 1];
    }
    su[threadIdx.x] = u;
    sv[threadIdx.x] = v;
    __syncthreads();
    for (int h = kR / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            su[threadIdx.x] += su[threadIdx.x + h];
            sv[threadIdx.x] += sv[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        s->rz = su[0];
        s->b2 = clamp30(sv[0]);
    } else if (stage == 1) {
        s->al = s->rz / clamp30(su[0]);
    } else {
        s->rr = su[0];
        s->beta = sv[0] / clamp30(s->rz);
        s->rz = sv[0];
    }
}
#endif

}  // namespace pcg
}  // namespace uva

#if !defined(UVA_EMULATOR)
extern "C" int uva_jacobi_pcg(int64_t n, const int64_t* indptr, const int64_t* indices, const double* data,
                              const double* b, double* x, int32_t max_iters, int32_t check_every, double tol2,
                              int32_t* iters_out, void* stream) {
    using namespace uva::pcg;
    if (n < 0 || max_iters < 0 || check_every < 1 || !(tol2 >= 0.0)) return UVA_EINVAL;
    if (n > 0 && (!indptr || !b || !x)) return UVA_EINVAL;
    if (iters_out) *iters_out = 0;
    if (n == 0) return UVA_OK;
    if (indptr[0] != 0) return UVA_EINVAL;
    for (int64_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) return UVA_EINVAL;
    const int64_t nnz = indptr[n];
    if (nnz > 0 && (!indices || !data)) return UVA_EINVAL;
    for (int64_t q = 0; q < nnz; ++q)
        if (indices[q] < 0 || indices[q] >= n) return UVA_EINVAL;
    if (max_iters == 0) return UVA_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nb = (n + kB - 1) / kB;
    if (nb > (int64_t)0x7fffffff) return UVA_EINVAL;
    double* buf = nullptr;
    int64_t* idx = nullptr;
    Scalars* sc = nullptr;
    const size_t nvec = (size_t)n, nz = (size_t)nnz;
    if (hipMalloc(&buf, (5 * nvec + 2 * (size_t)nb + nz + 1) * sizeof(double)) != hipSuccess) return UVA_ENOMEM;
    if (hipMalloc(&idx, (nvec + 1 + nz + 1) * sizeof(int64_t)) != hipSuccess) {
        (void)hipFree(buf);
        return UVA_ENOMEM;
    }
    if (hipMalloc(&sc, sizeof(Scalars)) != hipSuccess) {
        (void)hipFree(buf);
        (void)hipFree(idx);
        return UVA_ENOMEM;
    }
    double *dinv = buf, *r = buf + nvec, *z = buf + 2 * nvec, *p = buf + 3 * nvec, *Ap = buf + 4 * nvec;
    double* part = buf + 5 * nvec;
    double* a = part + 2 * nb;
    int64_t *ip = idx, *ix = idx + nvec + 1;
    const dim3 grid((unsigned)nb), blk(kB);
    int rc = UVA_OK;
    int it_run = max_iters;
    if (hipMemcpyAsync(ip, indptr, (nvec + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess ||
        (nz && hipMemcpyAsync(ix, indices, nz * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess) ||
        (nz && hipMemcpyAsync(a, data, nz * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess))
        rc = UVA_EHIP;
    if (rc == UVA_OK)
        hipLaunchKernelGGL(init_kernel, grid, blk, 0, st, n, ip, ix, a, b, x, dinv, r, z, p, part);
    if (rc == UVA_OK) hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kR), 0, st, part, nb, 0, sc);
    for (int it = 0; it < max_iters && rc == UVA_OK; ++it) {
        hipLaunchKernelGGL(spmv_kernel, grid, blk, 0, st, n, ip, ix, a, p, Ap, part);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kR), 0, st, part, nb, 1, sc);
        hipLaunchKernelGGL(update_kernel, grid, blk, 0, st, n, sc, dinv, p, Ap, x, r, z, part);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kR), 0, st, part, nb, 2, sc);
        if (hipGetLastError() != hipSuccess) rc = UVA_EHIP;
        if (rc == UVA_OK && it % check_every == 0) {
            Scalars h;
            if (hipMemcpyAsync(&h, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = UVA_EHIP;
            else if (h.rr < tol2 * h.b2) {
                it_run = it + 1;
                break;
            }
        }
        hipLaunchKernelGGL(direction_kernel, grid, blk, 0, st, n, sc, z, p);
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == UVA_OK) rc = UVA_EHIP;
    (void)hipFree(buf);
    (void)hipFree(idx);
    (void)hipFree(sc);
    if (rc == UVA_OK && iters_out) *iters_out = it_run;
    return rc;
}

#else
// The same iteration on the host, rows in order and sequential sums
// (test-only emulation; equal to the device result up to the order of the
// dot-product sums).
extern "C" int uva_emu_jacobi_pcg(int64_t n, const int64_t* indptr, const int64_t* indices, const double* data,
                                  const double* b, double* x, int32_t max_iters, int32_t check_every, double tol2,
                                  int32_t* iters_out) {
    using namespace uva::pcg;
    if (n < 0 || max_iters < 0 || check_every < 1 || !(tol2 >= 0.0)) return UVA_EINVAL;
    if (iters_out) *iters_out = 0;
    if (n == 0 || max_iters == 0) return UVA_OK;
    if (!indptr || indptr[0] != 0 || !b || !x) return UVA_EINVAL;
    for (int64_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) return UVA_EINVAL;
    for (int64_t q = 0; q < indptr[n]; ++q)
        if (indices[q] < 0 || indices[q] >= n) return UVA_EINVAL;
    std::vector<double> dinv(n), r(n), z(n), p(n), Ap(n);
    double rz = 0.0, b2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        dinv[i] = row_diag(indptr, indices, data, i);
        r[i] = b[i] - row_dot(indptr, indices, data, x, i);
        z[i] = p[i] = r[i] * dinv[i];
        rz += r[i] * z[i];
        b2 += b[i] * b[i];
    }
    b2 = clamp30(b2);
    int it_run = max_iters;
    for (int it = 0; it < max_iters; ++it) {
        double pap = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            Ap[i] = row_dot(indptr, indices, data, p.data(), i);
            pap += p[i] * Ap[i];
        }
        const double al = rz / clamp30(pap);
        double rr = 0.0, rz2 = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += al * p[i];
            r[i] -= al * Ap[i];
            z[i] = r[i] * dinv[i];
            rr += r[i] * r[i];
            rz2 += r[i] * z[i];
        }
        if (it % check_every == 0 && rr < tol2 * b2) {
            it_run = it + 1;
            break;
        }
        const double beta = rz2 / clamp30(rz);
        rz = rz2;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    if (iters_out) *iters_out = it_run;
    return UVA_OK;
}
#endif
