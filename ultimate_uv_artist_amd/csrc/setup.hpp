// Host-side construction of an island batch (frames, Laplacian in sliced ELL,
// rhs incidence, LSCM pins, task order).  Runs once per plan.
#pragma once
#include <stdint.h>

#include <sys/mman.h>

#include <cstdlib>
#include <new>
#include <string>
#include <memory>
#include <utility>
#include <vector>

#include "plan.hpp"

namespace uva {

// Vectors whose resize leaves new elements uninitialised (default-init, no
// zero fill): the plan build writes every element of its large arrays in
// threaded passes, so the pages are first touched there, in parallel, and
// not by a serial fill.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    // arrays of 4 MiB or more on 2 MiB boundaries, marked for huge pages: the
    // first touch then faults once per 2 MiB instead of once per 4 KiB
    T* allocate(std::size_t n) {
        const std::size_t bytes = n * sizeof(T);
        if (bytes < (std::size_t(4) << 20)) return std::allocator<T>::allocate(n);
        const std::size_t huge = std::size_t(2) << 20;
        void* p = std::aligned_alloc(huge, (bytes + huge - 1) / huge * huge);
        if (!p) throw std::bad_alloc();
        (void)madvise(p, bytes, MADV_HUGEPAGE);
        return static_cast<T*>(p);
    }
    void deallocate(T* p, std::size_t n) {
        if (n * sizeof(T) < (std::size_t(4) << 20)) std::allocator<T>::deallocate(p, n);
        else std::free(p);
    }
    template <class U>
    void construct(U* p) noexcept { ::new ((void*)p) U; }
    template <class U, class... A>
    void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
};
template <class T>
using hvec = std::vector<T, NoInitAlloc<T>>;

struct HostPlan {
    int n_isl = 0, n_vert = 0, n_tri = 0;
    bool batch = false;
    std::vector<int> isl_v0, isl_nv, isl_t0, isl_nt, isl_s0, isl_ns, isl_pin, isl_lpin0, isl_lpin1;
    std::vector<int> isl_bw;     // matrix bandwidth: largest |column - row| of the island
    std::vector<double> isl_wpin, isl_ldist;
    hvec<int> tv0, tv1, tv2;
    hvec<double> fx1, fx2, fy2, farea;
    hvec<unsigned char> fok;
    hvec<int> inc_ptr, inc_tri;
    hvec<double> inc_g;          // 2 per incidence
    hvec<double> tri_g;          // the same corner vectors triangle-major: 6 per triangle (not in Batch plans)
    std::vector<int> sl_ptr;
    hvec<int> m_col;
    std::vector<int> sl_row;            // first row of each slice (n_slices + 1, the last = n_vert)
    hvec<float> m_hi, m_lo, m_b;        // m_b: LSCM coupling (empty in Batch plans)
    hvec<double> diag;
    hvec<float> dinv32;
    hvec<int> adj_ptr, adj_idx;
    hvec<unsigned char> vbnd;
    std::vector<int> order_group, order_wave;
    // group tasks: island order_group[k] is solved by member grp_mem[k] of a
    // team of grp_m[k] workgroups; teams with more than one member have a
    // slot grp_team[k] for their barrier counter and partial sums
    std::vector<int> grp_mem, grp_m, grp_team;
    int n_teams = 0;
    int max_m = 1;
    // schedule: n_xcd x slots_per_xcd workgroup slots; slot q runs the tasks
    // [slot_ptr[q], slot_ptr[q+1]) of order_group in order, or (claim_tasks)
    // XCD x's workgroups claim the tasks [slot_ptr[x], slot_ptr[x+1]) one at a
    // time
    int n_xcd = 1;
    int slots_per_xcd = 1;
    bool claim_tasks = false;
    std::vector<int> slot_ptr{0, 0};
    // inputs of the schedule, kept to re-plan it from measured iterations
    std::vector<int> group_isl;
    int sched_slots = 1;
    bool sched_teams = false;
    // cost model of the schedule, microseconds per CG iteration: fixed part
    // of a one-workgroup island, fixed part of a team member, and per 64-row
    // slice of a member (the defaults until uva_plan_tune fits them to
    // measured island times, or a caller sets them)
    double cost[3] = {15.0, 30.0, 0.1};
    int64_t nnz_true = 0;
    // plan vertex -> caller's global vertex id (islands keep their offsets;
    // inside an island the plan orders vertices by Morton code)
    hvec<int> vperm;
    // coarse level: isl_cofs[i] is the offset of island i's inverse coarse
    // operator (nc x nc, FP32, nc = ceil(ns / isl_cg[i]) aggregates of
    // isl_cg[i] slices) in cinv, or -1 without a coarse level
    std::vector<int64_t> isl_cofs;
    std::vector<int> isl_cg;      // slices per aggregate
    std::vector<int> isl_lcg;     // slices per aggregate of the LSCM coarse level
    std::vector<float> cinv;
    // coarse level of the LSCM operator over aggregates of isl_lcg[i] slices
    // (for the plan's LSCM pins): isl_lcofs[i] is the offset of island i's nc
    // rows of [Pc | Qc] (2 nc FP32 values each) in lcinv, or -1
    std::vector<int64_t> isl_lcofs;
    std::vector<float> lcinv;
};

// Returns an empty string on success, else an error message.
std::string build_host_plan(HostPlan& hp, int64_t n_islands, const int64_t* vert_off, const int64_t* tri_off,
                            const double* P, const int64_t* T, const int64_t* pins, const double* X,
                            const double* area, const double* cots, bool batch, const int64_t* lscm_pins,
                            int wave_max_rows = kWaveMaxRows, int slots = 1, int n_xcd = 1, bool teams = false,
                            bool reorder = true, bool coarse = true);

// Static gang schedule of the workgroup tasks.  `slots` workgroups run at
// once (one per CU for the CG kernels), slots / n_xcd per XCD.  Team sizes
// (teams = true): enough members that each member's rows stay resident in the
// ARAP CG, more for islands whose estimated time exceeds the batch's ideal
// makespan.  The estimate is est_iters[i] CG iterations per global step (or a
// size-based guess when NULL) times a per-iteration cost.  Islands go to the
// least-loaded XCD; within an XCD, longest first, an island's m members take
// the m slots that free up first and start together.  Every slot lists its
// tasks in one global order, so no two teams wait on each other.
// res_slices: slices one member keeps resident (kResSlices for the FP32 ARAP
// CG, kResSlicesF64 for the FP64 LSCM solve).
void plan_teams(HostPlan& hp, const std::vector<int>& group_islands, int slots, int n_xcd, bool teams,
                const double* est_iters, int res_slices = kResSlices);

// One island holding a general symmetric matrix (no triangles, no pin).
std::string build_matrix_plan(HostPlan& hp, int64_t n, const int64_t* indptr, const int64_t* indices,
                              const double* data);

// hier.Solver.relax's regrouping of a subset's charts into islands
// (uva_chart_regroup).  Returns an empty string on success, else an error
// message; *nk receives the number of charts with faces.
std::string chart_regroup(int64_t n, const int64_t* faces, const int64_t* CV, int64_t nv, const int64_t* chart,
                          int64_t nchart, const int64_t* mesh_F, int64_t n_mesh_f, const double* mesh_V,
                          int64_t n_mesh_v, int64_t* keep, int64_t* nk, int64_t* vert_off, int64_t* tri_off, double* P,
                          int64_t* T, int64_t* vorder, int64_t* forder);

// param.repair_flips (sequential Gauss-Seidel order) on one chart.
void repair_sequential(double* U, int64_t n, const int64_t* T, int64_t t, int rounds);

}  // namespace uva
