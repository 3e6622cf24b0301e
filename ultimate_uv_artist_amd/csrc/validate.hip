// Exact UV triangle overlap test of the reference's validation
// (engine/validate.py:58-136) on the GPU.
//
// A pair of triangles overlaps with positive area when an edge of one
// properly crosses an edge of the other, or a corner or the centroid of one
// lies strictly inside the other, with orientation tolerances relative to the
// squared edge lengths and never below float32 storage resolution; corners at
// identical UV positions count as shared, and three shared corners (a stacked
// copy) do not overlap.  Candidate pairs come from a uniform g x g grid over
// the unit square (cells clamped to the grid): each triangle is entered in the
// cells its box covers and a pair is tested in the first cell both cover, so
// every pair whose boxes meet is tested exactly once.
//
// Kernels: cell ranges and counts per cell (atomics), one-workgroup exclusive
// scan, fill (atomics; the order inside a cell does not change the result),
// pair tests with one thread per cell entry.  Arithmetic is plain IEEE double
// without contraction, in the reference's expression order, so flags match
// the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/uvart.h"

#pragma clang fp contract(off)

namespace uva {
namespace ov {

constexpr double kUlp = 1.0 / 16777216.0;   // 2^-24: float32 spacing just below 1.0

__host__ __device__ inline double orient(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}

// Python's max(a, b): b only when b > a
__host__ __device__ inline double dmax(double a, double b) { return b > a ? b : a; }

__host__ __device__ inline bool strict_inside(double px, double py, double ax, double ay, double bx, double by,
                                              double cx, double cy, double rel) {
    const double l = dmax((bx - ax) * (bx - ax) + (by - ay) * (by - ay),
                          dmax((cx - bx) * (cx - bx) + (cy - by) * (cy - by), (ax - cx) * (ax - cx) + (ay - cy) * (ay - cy)));
    const double eps = rel > 0.0 ? dmax(rel * l, 4.0 * kUlp * sqrt(l)) : 0.0;
    const double d1 = orient(ax, ay, bx, by, px, py);
    const double d2 = orient(bx, by, cx, cy, px, py);
    const double d3 = orient(cx, cy, ax, ay, px, py);
    return (d1 > eps && d2 > eps && d3 > eps) || (d1 < -eps && d2 < -eps && d3 < -eps);
}

__host__ __device__ inline bool proper_cross(double ax, double ay, double bx, double by, double cx, double cy,
                                             double dx, double dy, double rel) {
    const double l1 = (bx - ax) * (bx - ax) + (by - ay) * (by - ay);
    const double l2 = (dx - cx) * (dx - cx) + (dy - cy) * (dy - cy);
    const double e1 = dmax(rel * l1, 4.0 * kUlp * sqrt(l1));
    const double e2 = dmax(rel * l2, 4.0 * kUlp * sqrt(l2));
    const double o1 = orient(ax, ay, bx, by, cx, cy), o2 = orient(ax, ay, bx, by, dx, dy);
    const double o3 = orient(cx, cy, dx, dy, ax, ay), o4 = orient(cx, cy, dx, dy, bx, by);
    return ((o1 > e1 && o2 < -e1) || (o1 < -e1 && o2 > e1)) && ((o3 > e2 && o4 < -e2) || (o3 < -e2 && o4 > e2));
}

// validate.py:58-91 for triangles s, t
__host__ __device__ inline bool tri_overlap(const double2* U, const int64_t* F, int64_t s, int64_t t, double rel) {
    double2 a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = U[F[3 * s + k]];
        b[k] = U[F[3 * t + k]];
    }
    bool same[3][3];
    int shared = 0;
    for (int i = 0; i < 3; ++i) {
        bool any = false;
        for (int j = 0; j < 3; ++j) {
            same[i][j] = a[i].x == b[j].x && a[i].y == b[j].y;
            any = any || same[i][j];
        }
        shared += any ? 1 : 0;
    }
    if (shared == 3) return false;
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3;
            if (same[i][j] || same[i][j1] || same[i1][j] || same[i1][j1]) continue;
            if (proper_cross(a[i].x, a[i].y, a[i1].x, a[i1].y, b[j].x, b[j].y, b[j1].x, b[j1].y, rel)) return true;
        }
    }
    for (int i = 0; i < 3; ++i) {
        if (!(same[i][0] || same[i][1] || same[i][2]) &&
            strict_inside(a[i].x, a[i].y, b[0].x, b[0].y, b[1].x, b[1].y, b[2].x, b[2].y, rel))
            return true;
        if (!(same[0][i] || same[1][i] || same[2][i]) &&
            strict_inside(b[i].x, b[i].y, a[0].x, a[0].y, a[1].x, a[1].y, a[2].x, a[2].y, rel))
            return true;
    }
    const double cax = (a[0].x + a[1].x + a[2].x) / 3.0, cay = (a[0].y + a[1].y + a[2].y) / 3.0;
    if (strict_inside(cax, cay, b[0].x, b[0].y, b[1].x, b[1].y, b[2].x, b[2].y, rel)) return true;
    const double cbx = (b[0].x + b[1].x + b[2].x) / 3.0, cby = (b[0].y + b[1].y + b[2].y) / 3.0;
    return strict_inside(cbx, cby, a[0].x, a[0].y, a[1].x, a[1].y, a[2].x, a[2].y, rel);
}

// validate.py:229-254: a proper crossing with zero tolerance (float32-level
// contacts), for triangles that are not stacked copies
__host__ __device__ inline bool tri_cross_strict(const double2* U, const int64_t* F, int64_t s, int64_t t) {
    double2 a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = U[F[3 * s + k]];
        b[k] = U[F[3 * t + k]];
    }
    bool same[3][3];
    int shared = 0;
    for (int i = 0; i < 3; ++i) {
        bool any = false;
        for (int j = 0; j < 3; ++j) {
            same[i][j] = a[i].x == b[j].x && a[i].y == b[j].y;
            any = any || same[i][j];
        }
        shared += any ? 1 : 0;
    }
    if (shared == 3) return false;
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3;
            if (same[i][j] || same[i][j1] || same[i1][j] || same[i1][j1]) continue;
            const double o1 = orient(a[i].x, a[i].y, a[i1].x, a[i1].y, b[j].x, b[j].y);
            const double o2 = orient(a[i].x, a[i].y, a[i1].x, a[i1].y, b[j1].x, b[j1].y);
            const double o3 = orient(b[j].x, b[j].y, b[j1].x, b[j1].y, a[i].x, a[i].y);
            const double o4 = orient(b[j].x, b[j].y, b[j1].x, b[j1].y, a[i1].x, a[i1].y);
            if (o1 * o2 < 0.0 && o3 * o4 < 0.0) return true;
        }
    }
    return false;
}

__host__ __device__ inline bool pair_hits(const double2* U, const int64_t* F, int64_t s, int64_t t, double rel,
                                          int strict) {
    return strict ? tri_cross_strict(U, F, s, t) : tri_overlap(U, F, s, t, rel);
}

// grid cell of a coordinate: int(x * g) truncated, clamped to [0, g-1]
__host__ __device__ inline int cell_of(double x, int g) {
    double v = x * (double)g;
    if (!(v > 0.0)) return 0;                       // negatives and NaN
    if (v >= (double)(g - 1)) return g - 1;
    return (int)v;
}

__host__ __device__ inline void tri_box(const double2* U, const int64_t* F, int64_t f, int g, int& x0, int& x1,
                                        int& y0, int& y1) {
    const double2 a = U[F[3 * f]], b = U[F[3 * f + 1]], c = U[F[3 * f + 2]];
    x0 = cell_of(fmin(a.x, fmin(b.x, c.x)), g);
    x1 = cell_of(fmax(a.x, fmax(b.x, c.x)), g);
    y0 = cell_of(fmin(a.y, fmin(b.y, c.y)), g);
    y1 = cell_of(fmax(a.y, fmax(b.y, c.y)), g);
}

// ---------------------------------------------------------------- gutter raster
// Conservative raster of metrics.py:315-345 (_raster_labels): a texel (px,
// py) of the res x res image belongs to a triangle when it lies in the
// floored bounding box of the triangle's corners (in texel units, clamped to
// the image) and its centre is within 0.7072 texel of every edge line (on the
// inside or up to that far outside).  The reference writes island labels in
// triangle order, so a texel keeps the label of the largest triangle index
// covering it: the kernels take the maximum index per texel with atomics.
struct RasterTri {
    double ox[3], oy[3], ex[3], ey[3], L[3];
    long long minx, maxx, miny, maxy;
};

// int(np.floor(x)) of the reference's compiled kernel: NaN and values outside
// the int64 range convert to the int64 minimum.
__host__ __device__ inline long long floor_i64(double x) {
    const double f = floor(x);
    if (!(fabs(f) < 9223372036854775808.0)) return (long long)0x8000000000000000ULL;
    return (long long)f;
}

// Python's min(a, min(b, c)) / max(...): a later argument replaces an
// earlier one only when strictly smaller / larger.
__host__ __device__ inline double pymin3(double a, double b, double c) {
    const double bc = c < b ? c : b;
    return bc < a ? bc : a;
}
__host__ __device__ inline double pymax3(double a, double b, double c) {
    const double bc = c > b ? c : b;
    return bc > a ? bc : a;
}

__host__ __device__ inline void raster_setup(const double2* U, const int64_t* F, int64_t f, int res, RasterTri& r) {
    const double2 a = U[F[3 * f]], b = U[F[3 * f + 1]], c = U[F[3 * f + 2]];
    double x0 = a.x * res, y0 = a.y * res, x1 = b.x * res, y1 = b.y * res, x2 = c.x * res, y2 = c.y * res;
    const double area = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    if (area < 0) {
        double t = x1; x1 = x2; x2 = t;
        t = y1; y1 = y2; y2 = t;
    }
    long long v;
    v = floor_i64(pymin3(x0, x1, x2)); r.minx = v > 0 ? v : 0;
    v = floor_i64(pymax3(x0, x1, x2)); r.maxx = v < res - 1 ? v : res - 1;
    v = floor_i64(pymin3(y0, y1, y2)); r.miny = v > 0 ? v : 0;
    v = floor_i64(pymax3(y0, y1, y2)); r.maxy = v < res - 1 ? v : res - 1;
    r.ex[0] = x1 - x0; r.ex[1] = x2 - x1; r.ex[2] = x0 - x2;
    r.ey[0] = y1 - y0; r.ey[1] = y2 - y1; r.ey[2] = y0 - y2;
    r.ox[0] = x0; r.ox[1] = x1; r.ox[2] = x2;
    r.oy[0] = y0; r.oy[1] = y1; r.oy[2] = y2;
    for (int k = 0; k < 3; ++k) r.L[k] = sqrt(r.ex[k] * r.ex[k] + r.ey[k] * r.ey[k]);
}

__host__ __device__ inline bool raster_covers(const RasterTri& r, long long px, long long py) {
    const double cx = px + 0.5, cy = py + 0.5;
    for (int k = 0; k < 3; ++k) {
        const double e = r.ex[k] * (cy - r.oy[k]) - r.ey[k] * (cx - r.ox[k]);
        if (e < -0.7072 * r.L[k]) return false;
    }
    return true;
}

constexpr long long kRasterSmall = 256;   // texels of a box one thread covers alone

#if !defined(UVA_EMULATOR)
__global__ void boxes_kernel(const double2* U, const int64_t* F, int64_t t, const uint8_t* skip, int g, int4* box,
                             unsigned long long* cnt) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= t) return;
    int x0, x1, y0, y1;
    tri_box(U, F, f, g, x0, x1, y0, y1);
    box[f] = make_int4(x0, x1, y0, y1);
    if (skip && skip[f]) return;
    for (int gy = y0; gy <= y1; ++gy)
        for (int gx = x0; gx <= x1; ++gx) atomicAdd(cnt + (size_t)gy * g + gx, 1ull);
}

// exclusive scan of n counts into start[0..n] with one workgroup of 1024 threads
__global__ void scan_kernel(const unsigned long long* cnt, unsigned long long* start, int64_t n) {
    __shared__ unsigned long long part[1024];
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t lo = (int64_t)threadIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    unsigned long long s = 0;
    for (int64_t i = lo; i < hi; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int k = 0; k < 1024; ++k) {
            const unsigned long long v = part[k];
            part[k] = run;
            run += v;
        }
        start[n] = run;
    }
    __syncthreads();
    unsigned long long run = part[threadIdx.x];
    for (int64_t i = lo; i < hi; ++i) {
        start[i] = run;
        run += cnt[i];
    }
}

__global__ void fill_kernel(int64_t t, const uint8_t* skip, int g, const int4* box, unsigned long long* fill,
                            int64_t* items, int* item_cell) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= t || (skip && skip[f])) return;
    const int4 bx = box[f];
    for (int gy = bx.z; gy <= bx.w; ++gy)
        for (int gx = bx.x; gx <= bx.y; ++gx) {
            const int c = gy * g + gx;
            const unsigned long long p = atomicAdd(fill + c, 1ull);
            items[p] = f;
            item_cell[p] = c;
        }
}

__global__ void fill_i32_kernel(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// One thread per triangle: boxes of at most kRasterSmall texels are covered
// here, larger ones are listed for raster_big_kernel.
__global__ void raster_kernel(const double2* U, const int64_t* F, int64_t t, int res, int* last, int* big,
                              int* nbig) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= t) return;
    RasterTri r;
    raster_setup(U, F, f, res, r);
    if (r.maxx < r.minx || r.maxy < r.miny) return;
    const long long w = r.maxx - r.minx + 1, h = r.maxy - r.miny + 1;
    if (w * h > kRasterSmall) {
        big[atomicAdd(nbig, 1)] = (int)f;
        return;
    }
    for (long long py = r.miny; py <= r.maxy; ++py)
        for (long long px = r.minx; px <= r.maxx; ++px)
            if (raster_covers(r, px, py)) atomicMax(last + py * res + px, (int)f);
}

// One workgroup per listed triangle at a time, its threads over the box.
__global__ void raster_big_kernel(const double2* U, const int64_t* F, int res, int* last, const int* big,
                                  const int* nbig) {
    const int n = *nbig;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int f = big[i];
        RasterTri r;
        raster_setup(U, F, f, res, r);
        const long long w = r.maxx - r.minx + 1, h = r.maxy - r.miny + 1;
        for (long long q = threadIdx.x; q < w * h; q += blockDim.x) {
            const long long py = r.miny + q / w, px = r.minx + q % w;
            if (raster_covers(r, px, py)) atomicMax(last + py * res + px, f);
        }
    }
}

__global__ void pairs_kernel(const double2* U, const int64_t* F, int g, const int4* box,
                             const unsigned long long* start, const int64_t* items, const int* item_cell,
                             int64_t n_items, double rel, int strict, uint8_t* flag) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_items) return;
    const int c = item_cell[p];
    const int cy = c / g, cx = c % g;
    const int64_t s = items[p];
    const int4 bs = box[s];
    const int64_t end = (int64_t)start[c + 1];
    for (int64_t q = p + 1; q < end; ++q) {
        const int64_t t = items[q];
        const int4 bt = box[t];
        // once per pair: the first cell both boxes cover
        if ((bs.x > bt.x ? bs.x : bt.x) != cx || (bs.z > bt.z ? bs.z : bt.z) != cy) continue;
        if (bs.x > bt.y || bt.x > bs.y || bs.z > bt.w || bt.z > bs.w) continue;
        if (pair_hits(U, F, s, t, rel, strict)) {
            flag[s] = 1;
            flag[t] = 1;
        }
    }
}
#endif

}  // namespace ov
}  // namespace uva

#if !defined(UVA_EMULATOR)
namespace uva {
// the calling thread's message for uva_last_error (capi.cpp); returns code
int set_error(int code, const char* msg);

namespace ov {

__global__ void zero_bytes_kernel(uint8_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

inline unsigned blocks_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Device memory of one overlap test: per triangle its cell box, per cell the
// count and the start of its entries; `items` and `item_cell` (one per cell
// entry) are sized after the scan, by get_items(total).
struct Work {
    int4* box;
    unsigned long long* cnt;
    unsigned long long* start;
};

// The kernels of one overlap test on the stream; flags and the counters are
// cleared by fill kernels.  Synchronises the stream once the entry count is
// known and at the end.
template <class GetItems>
int run_overlap(const double* U, const int64_t* F, int64_t n_tri, const uint8_t* skip, double rel, int strict,
                int grid, uint8_t* flags, const Work& w, GetItems&& get_items, hipStream_t s) {
    const int64_t ncell = (int64_t)grid * grid;
    const int B = 256;
    const unsigned nb = blocks_of(n_tri, B);
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(nb), dim3(B), 0, s, flags, n_tri);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(blocks_of(ncell, B)), dim3(B), 0, s, w.cnt, ncell);
    hipLaunchKernelGGL(boxes_kernel, dim3(nb), dim3(B), 0, s, (const double2*)U, F, n_tri, skip, grid, w.box, w.cnt);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, w.cnt, w.start, ncell);
    if (hipGetLastError() != hipSuccess) return UVA_EHIP;
    unsigned long long total = 0;
    if (hipMemcpyAsync(&total, w.start + ncell, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return UVA_EHIP;
    if (total == 0) return UVA_OK;
    int64_t* items = nullptr;
    int* item_cell = nullptr;
    int rc = get_items(total, items, item_cell);
    if (rc) return rc;
    if (hipMemcpyAsync(w.cnt, w.start, (size_t)ncell * 8, hipMemcpyDeviceToDevice, s) != hipSuccess) return UVA_EHIP;
    hipLaunchKernelGGL(fill_kernel, dim3(nb), dim3(B), 0, s, n_tri, skip, grid, w.box, w.cnt, items, item_cell);
    hipLaunchKernelGGL(pairs_kernel, dim3(blocks_of((int64_t)total, B)), dim3(B), 0, s, (const double2*)U, F, grid,
                       w.box, w.start, items, item_cell, (int64_t)total, rel, strict, flags);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return UVA_EHIP;
    return UVA_OK;
}

// Scratch of uva_overlap_flags_noalloc: one array per GPU for the life of
// the process.  Host code never uses the symbol itself as an address (on the
// host it only names the array); each device's own address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(UVA_EHIP, "hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return set_error(UVA_EINVAL, "device id beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(UVA_EHIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return UVA_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace ov
}  // namespace uva

extern "C" int uva_overlap_flags(const double* U, int64_t n_vert, const int64_t* F, int64_t n_tri,
                                 const uint8_t* skip, double rel, int32_t strict, int32_t grid, uint8_t* flags,
                                 void* stream) {
    using namespace uva::ov;
    if (!U || !F || !flags || n_vert < 0 || n_tri < 0) return UVA_EINVAL;
    if (grid < 1 || grid > 8192 || !(rel >= 0.0)) return UVA_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    if (n_tri == 0) return UVA_OK;
    const int64_t ncell = (int64_t)grid * grid;
    Work w{nullptr, nullptr, nullptr};
    int64_t* items = nullptr;
    int* item_cell = nullptr;
    int rc = UVA_OK;
    auto mem = [](hipError_t e) { return e == hipErrorOutOfMemory ? UVA_ENOMEM : UVA_EHIP; };
    hipError_t e;
    if ((e = hipMalloc(&w.box, (size_t)n_tri * sizeof(int4))) != hipSuccess ||
        (e = hipMalloc(&w.cnt, (size_t)ncell * 8)) != hipSuccess ||
        (e = hipMalloc(&w.start, (size_t)(ncell + 1) * 8)) != hipSuccess) {
        rc = mem(e);
    } else {
        rc = run_overlap(U, F, n_tri, skip, rel, strict, grid, flags, w,
                         [&](unsigned long long total, int64_t*& it, int*& ic) {
                             hipError_t e2;
                             if ((e2 = hipMalloc(&items, (size_t)total * 8)) != hipSuccess ||
                                 (e2 = hipMalloc(&item_cell, (size_t)total * 4)) != hipSuccess)
                                 return mem(e2);
                             it = items;
                             ic = item_cell;
                             return (int)UVA_OK;
                         },
                         s);
        if (rc) (void)hipStreamSynchronize(s);
    }
    if (items) (void)hipFree(items);
    if (item_cell) (void)hipFree(item_cell);
    if (w.box) (void)hipFree(w.box);
    if (w.cnt) (void)hipFree(w.cnt);
    if (w.start) (void)hipFree(w.start);
    return rc;
}

extern "C" int uva_overlap_flags_noalloc(const double* U, int64_t n_vert, const int64_t* F, int64_t n_tri,
                                         const uint8_t* skip, double rel, int32_t strict, int32_t grid,
                                         uint8_t* flags, void* stream) {
    using namespace uva::ov;
    if (!U || !F || !flags || n_vert < 0 || n_tri < 0) return uva::set_error(UVA_EINVAL, "bad argument");
    if (grid < 1 || grid > 8192 || !(rel >= 0.0)) return uva::set_error(UVA_EINVAL, "grid or rel out of range");
    if (n_tri == 0) return UVA_OK;
    const int64_t ncell = (int64_t)grid * grid;
    const size_t fixed = align256((size_t)n_tri * sizeof(int4)) + align256((size_t)ncell * 8) +
                         align256((size_t)(ncell + 1) * 8);
    if (fixed > kScratchBytes)
        return uva::set_error(UVA_EINVAL, "triangles and grid cells exceed the 64 MiB scratch of "
                                          "uva_overlap_flags_noalloc");
    unsigned char* base = nullptr;
    int rc = scratch_of_current_device(&base);
    if (rc) return rc;
    Work w;
    w.box = (int4*)base;
    w.cnt = (unsigned long long*)(base + align256((size_t)n_tri * sizeof(int4)));
    w.start = (unsigned long long*)((unsigned char*)w.cnt + align256((size_t)ncell * 8));
    unsigned char* rest = base + fixed;
    const size_t left = kScratchBytes - fixed;
    rc = run_overlap(U, F, n_tri, skip, rel, strict, grid, flags, w,
                     [&](unsigned long long total, int64_t*& it, int*& ic) {
                         if (align256((size_t)total * 8) + (size_t)total * 4 > left)
                             return uva::set_error(UVA_EINVAL, "cell entries exceed the 64 MiB scratch of "
                                                               "uva_overlap_flags_noalloc");
                         it = (int64_t*)rest;
                         ic = (int*)(rest + align256((size_t)total * 8));
                         return (int)UVA_OK;
                     },
                     (hipStream_t)stream);
    if (rc == UVA_EHIP) uva::set_error(UVA_EHIP, "HIP error in uva_overlap_flags_noalloc");
    return rc;
}
extern "C" int uva_raster_last(const double* U, int64_t n_vert, const int64_t* F, int64_t n_tri, int32_t res,
                               int32_t* last, void* stream) {
    using namespace uva::ov;
    if (!U || !F || !last || n_vert < 0 || n_tri < 0 || n_tri > 0x7fffffffLL)
        return uva::set_error(UVA_EINVAL, "bad argument");
    if (res < 1 || res > 8192) return uva::set_error(UVA_EINVAL, "res must be in [1, 8192]");
    hipStream_t s = (hipStream_t)stream;
    const int64_t npix = (int64_t)res * res;
    const int B = 256;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_of(npix, B)), dim3(B), 0, s, last, npix, -1);
    if (n_tri == 0) return hipGetLastError() == hipSuccess ? UVA_OK : uva::set_error(UVA_EHIP, "raster fill");
    int* big = nullptr;
    hipError_t e = hipMalloc(&big, ((size_t)n_tri + 1) * sizeof(int));
    if (e != hipSuccess) return uva::set_error(e == hipErrorOutOfMemory ? UVA_ENOMEM : UVA_EHIP, "raster list");
    int* nbig = big + n_tri;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(1), 0, s, nbig, (int64_t)1, 0);
    hipLaunchKernelGGL(raster_kernel, dim3(blocks_of(n_tri, B)), dim3(B), 0, s, (const double2*)U, F, n_tri, (int)res,
                       last, big, nbig);
    hipLaunchKernelGGL(raster_big_kernel, dim3(1024), dim3(B), 0, s, (const double2*)U, F, (int)res, last, big, nbig);
    int rc = UVA_OK;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        rc = uva::set_error(UVA_EHIP, "raster kernels");
    (void)hipFree(big);
    return rc;
}
#else
// The same algorithm on the host, one pair at a time (test-only emulation).
extern "C" int uva_emu_overlap_flags(const double* U, int64_t n_vert, const int64_t* F, int64_t n_tri,
                                     const uint8_t* skip, double rel, int32_t strict, int32_t grid, uint8_t* flags) {
    using namespace uva::ov;
    (void)n_vert;
    const int g = grid;
    std::vector<int> x0(n_tri), x1(n_tri), y0(n_tri), y1(n_tri);
    std::vector<std::vector<int64_t>> cells((size_t)g * g);
    for (int64_t f = 0; f < n_tri; ++f) {
        flags[f] = 0;
        tri_box((const double2*)U, F, f, g, x0[f], x1[f], y0[f], y1[f]);
        if (skip && skip[f]) continue;
        for (int gy = y0[f]; gy <= y1[f]; ++gy)
            for (int gx = x0[f]; gx <= x1[f]; ++gx) cells[(size_t)gy * g + gx].push_back(f);
    }
    for (int c = 0; c < g * g; ++c) {
        const auto& it = cells[c];
        for (size_t p = 0; p < it.size(); ++p)
            for (size_t q = p + 1; q < it.size(); ++q) {
                const int64_t s = it[p], t = it[q];
                if (std::max(x0[s], x0[t]) != c % g || std::max(y0[s], y0[t]) != c / g) continue;
                if (x0[s] > x1[t] || x0[t] > x1[s] || y0[s] > y1[t] || y0[t] > y1[s]) continue;
                if (pair_hits((const double2*)U, F, s, t, rel, strict)) flags[s] = flags[t] = 1;
            }
    }
    return 0;
}
// The raster of uva_raster_last on the host, triangles in order (test-only
// emulation of the same per-texel code): last (res*res) gets the largest
// covering triangle index, -1 elsewhere.
extern "C" int uva_emu_raster_last(const double* U, const int64_t* F, int64_t n_tri, int32_t res, int32_t* last) {
    using namespace uva::ov;
    for (int64_t i = 0; i < (int64_t)res * res; ++i) last[i] = -1;
    for (int64_t f = 0; f < n_tri; ++f) {
        RasterTri r;
        raster_setup((const double2*)U, F, f, res, r);
        for (long long py = r.miny; py <= r.maxy; ++py)
            for (long long px = r.minx; px <= r.maxx; ++px)
                if (raster_covers(r, px, py)) last[py * res + px] = (int32_t)f;
    }
    return 0;
}
#endif
