// Host-side construction of an island batch (frames, Laplacian in sliced ELL,
// rhs incidence, LSCM pins, task order).  Runs once per plan.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "plan.hpp"

namespace uva {

struct HostPlan {
    int n_isl = 0, n_vert = 0, n_tri = 0;
    bool batch = false;
    std::vector<int> isl_v0, isl_nv, isl_t0, isl_nt, isl_s0, isl_ns, isl_pin, isl_lpin0, isl_lpin1;
    std::vector<int> isl_bw;     // matrix bandwidth: largest |column - row| of the island
    std::vector<double> isl_wpin, isl_ldist;
    std::vector<int> tv0, tv1, tv2;
    std::vector<double> fx1, fx2, fy2, farea;
    std::vector<unsigned char> fok;
    std::vector<int> inc_ptr, inc_tri;
    std::vector<double> inc_g;          // 2 per incidence
    std::vector<double> tri_g;          // the same corner vectors triangle-major: 6 per triangle
    std::vector<int> sl_ptr, m_col;
    std::vector<float> m_hi, m_lo, m_b;
    std::vector<double> diag;
    std::vector<float> dinv32;
    std::vector<int> adj_ptr, adj_idx;
    std::vector<unsigned char> vbnd;
    std::vector<int> order_group, order_wave;
    // group tasks: island order_group[k] is solved by member grp_mem[k] of a
    // team of grp_m[k] workgroups; teams with more than one member have a
    // slot grp_team[k] for their barrier counter and partial sums
    std::vector<int> grp_mem, grp_m, grp_team;
    int n_teams = 0;
    int max_m = 1;
    // order_group is split into n_xcd lists: list x is [grp_xoff[x], grp_xoff[x+1])
    int n_xcd = 1;
    std::vector<int> grp_xoff{0, 0};
    int64_t nnz_true = 0;
    // plan vertex -> caller's global vertex id (islands keep their offsets;
    // inside an island the plan orders vertices by Morton code)
    std::vector<int> vperm;
    // coarse level: isl_cofs[i] is the offset of island i's inverse coarse
    // operator (nc x nc, FP32, nc = ceil(ns / isl_cg[i]) aggregates of
    // isl_cg[i] slices) in cinv, or -1 without a coarse level
    std::vector<int64_t> isl_cofs;
    std::vector<int> isl_cg;      // slices per aggregate
    std::vector<float> cinv;
};

// Returns an empty string on success, else an error message.
std::string build_host_plan(HostPlan& hp, int64_t n_islands, const int64_t* vert_off, const int64_t* tri_off,
                            const double* P, const int64_t* T, const int64_t* pins, const double* X,
                            const double* area, const double* cots, bool batch, const int64_t* lscm_pins,
                            int wave_max_rows = kWaveMaxRows, int team_slots = 0, int n_xcd = 1,
                            bool reorder = true, bool coarse = true);

// Group-task order and team sizes: islands whose cost exceeds their share of
// `team_slots` concurrent workgroups get several workgroups (at most kMaxTeam,
// at least kTeamMinRows rows each).  team_slots = 0 gives every island one.
// Islands are dealt to n_xcd lists (one per XCD), largest first onto the
// least-loaded list; within a list, longest time per member first.
void plan_teams(HostPlan& hp, const std::vector<int>& group_islands, int team_slots, int n_xcd = 1);

// One island holding a general symmetric matrix (no triangles, no pin).
std::string build_matrix_plan(HostPlan& hp, int64_t n, const int64_t* indptr, const int64_t* indices,
                              const double* data);

// param.repair_flips (sequential Gauss-Seidel order) on one chart.
void repair_sequential(double* U, int64_t n, const int64_t* T, int64_t t, int rounds);

}  // namespace uva
