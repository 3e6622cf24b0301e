// Device-side layout of a batch of islands (shared by host setup and kernels).
//
// Vertices, triangles and matrix rows of all islands are concatenated; each
// island owns contiguous ranges.  Vectors are indexed by global vertex id.
//
// Matrix: the off-diagonal part of each island's cotangent Laplacian in
// sliced ELL with 64-row slices (one slice = one wavefront, lane = row), so
// every matrix load of a wavefront is one contiguous 256-byte segment.  Slice
// s covers rows [slice_row[s], slice_row[s] + 64) of one island; its entries
// live at sl_ptr[s] + k*64 + lane for k < (sl_ptr[s+1] - sl_ptr[s]) / 64.
// Values are stored as an FP32 pair hi + lo (hi = fp32(v), lo = fp32(v - hi))
// so the FP32 inner iterations read 4 bytes per entry and the FP64 residual
// reads 8.  Padding entries point at the row itself with value 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace uva {

constexpr int kSlice = 64;          // rows per slice (= wavefront width)
constexpr int kWaveMaxRows = 384;   // islands up to this many vertices run on one wavefront
constexpr int kResSlices = 48;      // slices one workgroup keeps resident in the FP32 ARAP CG (8 waves x 6)
constexpr int kResSlicesF64 = 32;   // the same for FP64 vectors (8 waves x 4): the LSCM solve
constexpr int kMaxTeam = 16;        // most workgroups on one island
constexpr int kCoarseN = 128;       // most aggregates of a general matrix's coarse level (64 KB inverse)
constexpr int kTeamMinRows = 2048;  // fewest rows per team member
constexpr int kMaxXcd = 8;          // XCDs (own L2 each) a device may have
// Task counters of one launch: one per XCD list of workgroup tasks, then one
// for the wavefront tasks.
constexpr int kQueueStride = kMaxXcd + 1;

// LDS of the resident CG per workgroup: u with its halo, e, D^-1 and D (80 KB
// for 3072 FP32 rows) and the coarse level (rows of the inverse coarse
// operator for the member's aggregates, the restricted residual of every
// aggregate, the member's correction, mat-vec shares and slice sums)
constexpr int kLdsVec = 151552;
// rows of u a team member keeps beyond each end of its range, for FP32 and
// for FP64 u (the FP64 rows of the LSCM solve leave less LDS for a halo)
constexpr int kHalo = 512;
constexpr int kHaloF64 = 512;
__host__ __device__ constexpr int halo_rows(int v2) { return v2 <= 8 ? kHalo : kHaloF64; }
constexpr int kPart = 8;            // threads sharing one row of the coarse mat-vec

__host__ __device__ constexpr int align16(int b) { return (b + 15) & ~15; }
// rows of Ac^-1 are stored with an odd stride, so the threads of different
// aggregates read different LDS banks
__host__ __device__ constexpr int coarse_stride(int nc) { return nc | 1; }
// vectors of the resident CG: u (rows + both halos) and e of v2 bytes per
// row, D^-1 and D of v bytes
__host__ __device__ constexpr int res_vec_bytes_of(int rows, int halo, int v2, int v) {
    return align16((rows + 2 * halo) * v2 + rows * (v2 + 2 * v));
}
// the coarse level of a member with na of the nc aggregates and nsl slices
// (rows of Ac^-1 twice as long for the coupled LSCM operator)
__host__ __device__ constexpr int res_coarse_bytes(int nc, int na, int nsl, bool coupled = false) {
    return nc > 0 ? align16(na * coarse_stride(nc) * 4 * (coupled ? 2 : 1)) + (nc + na + na * kPart + nsl) * 16 : 0;
}

struct IslandState {
    double e_prev;       // ARAP energy of the current layout
    double rel_resid;    // last true relative residual
    double step;         // last line-search step
    double tmax;         // last max step (before the 0.9 factor)
    int done;            // energy early exit reached
    int use_ls;
    int cg_iters;        // summed over the call
    int cg_iters_max;
    int arap_iters;
    int converged;
    int flips0;
    int status;
    long long ticks;     // device wall-clock ticks spent on this island in the call
    double reg;          // SLIM regularisation weight 1e-12 * island area
    int iters_cap;       // ARAP iterations this island runs in the call (0 = all launched)
    int pad_;
};

struct DevPlan {
    int n_isl, n_vert, n_tri;
    int mode_batch;              // UVA_PLAN_BATCH conventions
    // islands
    const int* isl_v0;
    const int* isl_nv;
    const int* isl_t0;
    const int* isl_nt;
    const int* isl_s0;           // first slice
    const int* isl_ns;           // number of slices
    const int* isl_pin;          // global id of the soft pin
    const int* isl_bw;           // matrix bandwidth: largest |column - row| of the island
    const double* isl_wpin;
    const int* isl_lpin0;        // LSCM pins (global ids), -1 when the island has no boundary
    const int* isl_lpin1;
    const double* isl_ldist;     // |P[p1] - P[p0]|
    // triangles (global vertex ids)
    const int* tv0;
    const int* tv1;
    const int* tv2;
    const double* fx1;           // Jacobian frame: x1, x2, y2 (clamped in batch mode)
    const double* fx2;
    const double* fy2;
    const double* farea;
    const unsigned char* fok;    // frame usable (non-degenerate / thick)
    double* fscale;              // per-triangle target scale (batch mode)
    // rhs gather: corners incident to each vertex
    const int* inc_ptr;          // n_vert + 1
    const int* inc_tri;
    const double2* inc_g;        // b[v] += R(tri) g
    // matrix
    const int* sl_ptr;           // n_slices + 1
    const int* sl_row;           // first row of each slice; sl_row[s + 1] - sl_row[s] rows (n_slices + 1)
    int n_slices;
    const int* m_col;
    const float* m_hi;
    const float* m_lo;
    const float* m_b;            // LSCM boundary coupling B_ij (exact in fp32)
    int m_zero;                  // offset of kSlice entries of value 0 after the last slice
    const double* diag;          // Laplacian diagonal without pin
    const double* rpin;          // per-row soft-pin weights replacing the islands' pins (NULL = unused)
    const double* pdiag;         // diag + rpin (set with rpin)
    const float* dinv32;         // 1 / (diag + pin) for the ARAP operator
    // adjacency (unique neighbours) for Jacobi flip repair; boundary flags
    const int* adj_ptr;
    const int* adj_idx;
    const unsigned char* vbnd;
    // scratch (16 bytes per vertex each so FP32 and FP64 vectors fit)
    double2* R;                  // per-triangle rotation (cos, sin) x scale
    double2* b64;
    double2* X64;                // CG iterate (warm start = U)
    void* vr;
    void* vp;
    void* vAp;                   // s = A p in the CG recurrence
    void* ve;
    void* vu;                    // u = D^-1 r
    void* vw;                    // w = A u
    const double2* tri_g;        // rhs corner vectors per triangle (3 per triangle)
    // SLIM scratch (allocated on the first SLIM call): per triangle the
    // weight matrix W^T W (xx, xy, yy) and the 2x3 operator output, per
    // vertex the inverse of the 2x2 diagonal block (xx, xy, yy)
    double4* sw;
    double2* sy;
    double4* sdi;
    double* red;                 // per-island scalars for multi-launch paths
    IslandState* st;
    // vertex order: plan vertex v holds the caller's vertex vperm[v]
    const int* vperm;
    // coarse level of the ARAP CG (aggregates of isl_cg[i] consecutive 64-row
    // slices): isl_cofs[i] is the offset of island i's inverse coarse operator
    // (nc x nc FP32, row-major) in cinv, or -1; crc holds the restricted
    // residual (FP64, two columns) of every slice during a solve
    const int64_t* isl_cofs;
    const int* isl_cg;           // slices per aggregate
    const int* isl_lcg;          // slices per aggregate of the LSCM coarse level
    const float* cinv;
    // the LSCM operator's coarse level over the same aggregates: island i's
    // inverse [[Pc, Qc], [-Qc, Pc]] as nc rows of [Pc | Qc] at isl_lcofs[i]
    // in lcinv (-1: none)
    const int64_t* isl_lcofs;
    const float* lcinv;
    double2* crc;
    int* queue;                  // task counters
    const int* order_group;      // group tasks: island ids, longest first
    const int* grp_mem;          // member index of the task in its team
    const int* grp_m;            // team size (1 = one workgroup, no cross-workgroup barrier)
    const int* grp_team;         // team slot (teams of two or more)
    const int* order_wave;
    int n_group, n_wave;
    int n_teams;
    int* team_cnt;               // [launch slot][team] barrier counters, zeroed per call sequence
    double* team_part;           // [team][2][kMaxTeam][8] partial sums
    // Schedule of the workgroup tasks, slots_per_xcd workgroups per XCD.  A
    // workgroup takes tasks of the XCD it runs on (hardware id -> list x
    // through xcc_map), so all members of a team share one L2 and their
    // barrier needs no L2 write-back or invalidation (xcd_local).  Without a
    // probed XCD map there is one list and teams use agent-scope
    // release/acquire.
    //   claim_tasks = 0: static slot lists; slot q = x * slots_per_xcd + k runs
    //     tasks [slot_ptr[q], slot_ptr[q+1]) of order_group in order, and a
    //     workgroup claims the next slot of its XCD.
    //   claim_tasks = 1: one list per XCD, tasks [slot_ptr[x], slot_ptr[x+1]),
    //     longest first, a team's members consecutive; a workgroup claims the
    //     next task, so a team is formed by the next workgroups of the XCD that
    //     come free, whatever the earlier tasks took.
    int n_xcd;
    int xcd_local;
    int max_m;                   // largest team
    int slots_per_xcd;
    int claim_tasks;
    const int* slot_ptr;
    int xcc_map[16];
    long long spin_ticks;        // hang guard of a team barrier (device clock ticks)
    int* fault;                  // set when a team barrier gave up waiting
};

}  // namespace uva
