// Kernel launch interface (host side of kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/uvart.h"
#include "plan.hpp"

namespace uva {

constexpr int kBlock = 1024;    // workgroup size of the island kernel (16 wavefronts)
constexpr int kBlockF64 = 512;  // FP64 instantiations (more registers per lane)
constexpr int kBlockRes = 512;  // kinds with the resident CG (up to 8 rows per lane in registers)

enum TaskKind {
    TASK_PROLOGUE = 0,   // flips of the start, line-search decision, rotations + energy
    TASK_ARAP_ITER = 1,  // rhs -> CG -> (revert) -> line search -> update -> rotations + energy
    TASK_LSCM = 2,
    TASK_LOCAL_RHS = 3,
    TASK_CG = 4,
    TASK_MAX_STEP = 5,
    TASK_ENERGY = 6,
    TASK_REPAIR = 7,
    TASK_ORIENT = 8,
    TASK_FLIPS = 9,
    TASK_SLIM_PROLOGUE = 10,  // symmetric Dirichlet energy of the start, regularisation weight
    TASK_SLIM_ITER = 11,      // SLIM local step -> normal equations by PCG -> step with backtracking
    TASK_ENERGY_BWD = 12,     // dE/dU and dE/dw of the weighted energy
    TASK_STEP_BWD = 13,       // backward of one ARAP global step given the adjoint solution
    TASK_STEP_APPLY = 14,     // Batch / hier: non-finite revert, step rule and update from a global step's result
};

// Kinds whose tasks solve the ARAP or LSCM system (resident CG when the
// island fits).
constexpr bool solves_resident(int kind) { return kind == TASK_ARAP_ITER || kind == TASK_CG || kind == TASK_LSCM; }
constexpr int lds_vec_bytes(int kind) { return solves_resident(kind) ? kLdsVec : 0; }
constexpr int block_size(int kind, int prec) {
    return solves_resident(kind) ? kBlockRes : (prec == UVA_PREC_F64 ? kBlockF64 : kBlock);
}

struct TaskArgs {
    double2* U;              // layout (in/out)
    const double2* in;       // rhs (CG) or direction (max step)
    double2* out;            // rhs output (local_rhs)
    double* out_isl;         // per-island scalar output
    double cg_tol;
    double tol;              // energy early exit
    int cg_max;
    int line_search;
    int batch;               // Batch conventions (gpu.py): batch-global pin weight, non-finite revert
    int hier;                // hier.Solver.relax step rule (hier.py:498-509)
    int kind;                // energy kind
    int rounds;              // repair rounds
    const double* w;         // per-triangle energy weights (NULL = 1)
    const double* gin;       // per-island upstream gradient (energy backward)
    double* gw;              // per-triangle weight gradient (NULL = not wanted)
    int zero_x;              // CG starts from x = 0 instead of U (adjoint solve)
    int iters_plain;         // ARAP iterations of islands whose start has no fold (0 = all launched)
    int iters_folded;        // ... and of islands whose start has one (charts.py:411-415)
};

hipError_t launch_island_tasks(int kind, int prec, const DevPlan& P, const TaskArgs& a, int slot, int cus,
                               hipStream_t s);
hipError_t launch_fill_int(int* p, int n, int v, hipStream_t s);
// out[b] = hardware XCD id of block b (blocks of 64 threads)
hipError_t launch_xcc_probe(int* out, int blocks, hipStream_t s);
hipError_t launch_signed_areas(const DevPlan& P, const double2* U, double* sa, hipStream_t s);
// per-row pin weights of m pinned rows (ascending, distinct) -> rpin, pdiag,
// dinv32 over n rows
hipError_t launch_set_pins(const double* diag, const int* rows, const double* w, int m, double* rpin, double* pdiag,
                           float* dinv32, int n, hipStream_t s);

// Jacobi-PCG over every row of a plan (pcg.hip): workgroups for a device of
// `cus` CUs, scratch doubles they need, the launches of one solve (vectors in
// plan order; r, z, p, Ap work vectors), and where the iteration count lands.
int plan_pcg_grid(const DevPlan& P, int cus);
size_t plan_pcg_scratch_doubles(int grid);
hipError_t launch_plan_pcg(const DevPlan& P, const double2* b, double2* x, double2* r, double2* z, double2* p,
                           double2* Ap, double* scratch, int grid, int max_iters, int check_every, double tol2,
                           int strict, hipStream_t s);
int* plan_pcg_iters_ptr(double* scratch, int grid);
// caller-order vertex array -> plan order (gather), or back (scatter)
hipError_t launch_permute(const DevPlan& P, const double2* src, double2* dst, bool to_plan, hipStream_t s);

}  // namespace uva
