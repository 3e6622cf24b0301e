/*
 * uvart -- MI355X (gfx950) UV parameterization hot path: LSCM initialisation
 * and ARAP local/global iterations whose global step is a batched
 * Jacobi-preconditioned conjugate-gradient solve on per-island cotangent
 * Laplacians.
 *
 * Plain C ABI over raw pointers.  Host arrays are read during plan creation
 * only; the UV arrays passed to the solve calls are DEVICE pointers owned by
 * the caller (n_vertices x 2 float64, row-major).  Streams are hipStream_t
 * passed as void* (NULL = the default stream).
 *
 * Every call returns UVA_OK (0) or a negative status; uva_last_error() gives a
 * message for the calling thread.  A plan is not safe to use from two threads
 * at once; distinct plans are independent.
 *
 * Reference interfaces these entry points replace (Python, engine/):
 *   uva_plan_create     gpu.py:69-142 Batch.__init__ / _build_laplacian,
 *                       param.py:19-56 local_frames, cot_weights, cot_laplacian
 *   uva_lscm            param.py:136-179 lscm (+ _far_pair)
 *   uva_arap            param.py:379-414 refine(method="arap"),
 *                       gpu.py:230-249 Batch.arap and the iterations of
 *                       hier.py:394-509 Solver.relax (UVA_MODE_HIER)
 *   uva_slim            param.py:415-454 refine(method="slim") with
 *                       param.py:234-256 _arap_rotations (mode 1) and
 *                       param.py:341-376 _assemble_weighted / _slim_system
 *   uva_local_rhs       param.py:206-233 _arap_rotations (mode 0) +
 *                       param.py:457-474 _arap_rhs; gpu.py:165-182 Batch._rhs
 *   uva_cg_solve        param.py:103-132 SPDSolver.solve; gpu.py:184-202 Batch._cg
 *   uva_max_step        param.py:259-291 _max_step; gpu.py:204-227 Batch._max_step
 *   uva_step_apply      gpu.py:235-247 (the rest of a Batch.arap iteration after
 *                       its CG); hier.py:498-509 (the same in Solver.relax)
 *   uva_energy          param.py:294-338 _energy
 *   uva_energy_fwd      param.py:294-338 _energy with per-triangle weights
 *   uva_energy_bwd,     no counterpart in the reference (it has no
 *   uva_adjoint_cg,     gradients): the backward of _energy and of one
 *   uva_arap_step_bwd   refine global step (param.py:391-401, 457-474)
 *   uva_signed_areas    param.py:97-100 signed_areas; gpu.py:153-156
 *   uva_flips_per_chart gpu.py:158-163 Batch.flips_per_chart
 *   uva_repair_jacobi   gpu.py:251-273 Batch.repair_flips
 *   uva_orient_positive gpu.py:275-281 Batch.orient_positive
 *   uva_repair_seq      param.py:478-516 repair_flips (host, exact order)
 *   uva_chart_regroup   hier.py:398-408 (the subset's faces, local vertex ids
 *                       and chart ids that Solver.relax starts from; host)
 *   uva_overlap_flags   validate.py:94-136 _overlaps, 257-299 _overlaps_strict
 *   uva_jacobi_pcg      hier.py:565-577 (the CG of Solver.lscm)
 *   uva_csr_pcg         gpu.py:184-202 Batch._cg; hier.py:465-474, 565-577
 *   uva_plan_pcg        gpu.py:184-202 Batch._cg; hier.py:465-474 (the
 *                       CG of Solver.relax), on a plan's own matrix
 */
#ifndef UVART_H
#define UVART_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UVA_OK 0
#define UVA_EINVAL (-1)     /* bad argument or shape */
#define UVA_EHIP (-2)       /* HIP runtime error */
#define UVA_ENOMEM (-3)     /* device or host allocation failed */
#define UVA_ENOBOUNDARY (-5) /* LSCM on an island without boundary (closed) */

/* plan flags */
#define UVA_PLAN_BATCH 1u   /* gpu.Batch conventions: thin-triangle mask,
                               1e-30 clamps, batch-global pin weight */
#define UVA_PLAN_NO_WAVE 2u /* solve every island with a whole workgroup
                               (no single-wavefront tasks) */
#define UVA_PLAN_NO_TEAMS 4u /* one workgroup per island at most (no
                               multi-workgroup teams for large islands) */
#define UVA_PLAN_KEEP_ORDER 8u /* keep the caller's vertex order inside the
                               plan (default: Morton order per island) */
#define UVA_PLAN_NO_COARSE 16u /* Jacobi-preconditioned CG only (no coarse
                               level over 64-vertex aggregates) */

/* ARAP semantics */
#define UVA_MODE_REFINE 0   /* param.refine: per-island pin weight, energy early
                               exit, line search only if the start has no
                               flips, exact-root step rule */
#define UVA_MODE_BATCH 1    /* Batch.arap: no early exit, line search when
                               asked, non-finite islands keep their layout,
                               Jacobi-preconditioned CG */
#define UVA_MODE_HIER 2     /* hier.Solver.relax (hier.py:498-509): Batch
                               conventions, and each iteration a chart with a
                               folded usable triangle takes the plain step;
                               only usable triangles limit the others' step */

/* precision of the CG global step */
#define UVA_PREC_MIXED 0    /* FP32 SpMV/vectors, FP64 dots, FP64 true residual */
#define UVA_PREC_F32 1      /* FP32 everywhere (precision study) */
#define UVA_PREC_F64 2      /* FP64 everywhere */

typedef struct uva_plan uva_plan;

typedef struct {
    int64_t n_islands;
    int64_t n_vertices;
    int64_t n_triangles;
    int64_t nnz;             /* stored off-diagonal entries incl. padding */
    int64_t nnz_true;        /* off-diagonal nonzeros without padding */
    int64_t device_bytes;    /* device memory held by the plan */
    int32_t n_group_tasks;   /* islands solved by a whole workgroup */
    int32_t n_wave_tasks;    /* islands solved by a single wavefront */
    int32_t n_teams;         /* teams of two or more workgroups (their members
                                wait on each other: such a plan's solves need
                                its workgroups resident together) */
    int32_t reserved;
} uva_plan_info;

typedef struct {
    int32_t cg_iters;        /* inner CG iterations, summed over the call */
    int32_t cg_iters_max;    /* largest single global-step iteration count */
    int32_t arap_iters;      /* ARAP iterations run (early exit stops early) */
    int32_t converged;       /* last global step met the tolerance */
    int32_t flips0;          /* triangles with signed area <= 0 at the start */
    int32_t use_ls;          /* line search was active */
    int32_t status;          /* 0, or UVA_ENOBOUNDARY for LSCM */
    int32_t members;         /* workgroups of the island's team in the
                                schedule (0: one wavefront) */
    double rel_resid;        /* last true relative residual (max over u, v) */
    double energy;           /* ARAP energy of the returned layout (REFINE) */
    double step;             /* last line-search step */
    double busy_ms;          /* time a team spent on this island in the call */
} uva_island_stats;

typedef struct {
    int32_t iters;           /* ARAP iterations */
    int32_t line_search;     /* 0/1 */
    int32_t mode;            /* UVA_MODE_* */
    int32_t precision;       /* UVA_PREC_* */
    double tol;              /* energy early-exit tolerance (REFINE) */
    double cg_tol;           /* relative residual per global step; 0 runs
                                exactly cg_max_iters CG iterations (the
                                fixed budgets of Batch and hier.Solver) */
    int32_t cg_max_iters;    /* cap per global step (Batch's cg_iters) */
    int32_t use_graph;       /* capture the ARAP iterations in a hipGraph */
    int32_t iters_folded;    /* 0: every island runs iters iterations; > 0:
                                islands whose start has a folded triangle
                                run iters_folded, the others iters
                                (charts.py:411-415 doubles the iterations of
                                folded starts) */
} uva_arap_opts;

int uva_version(void);
const char *uva_last_error(void);
/* sizeof of the ABI structs, for bindings that declare them by hand:
 * which = 0 uva_arap_opts, 1 uva_slim_opts, 2 uva_island_stats,
 * 3 uva_plan_info; -1 for another value.  No counterpart in the reference. */
int uva_struct_size(int32_t which);

/* Build the device plan of a batch of islands.
 *   vert_off, tri_off  n_islands+1 prefix offsets (host, int64)
 *   P                  n_vertices x 3 float64 (host)
 *   T                  n_triangles x 3 int64, island-local vertex ids (host)
 *   pins               per-island soft-pin vertex (local id) or NULL (= 0)
 *   X, area, cots      optional precomputed frames (t x 3 x 2), areas (t) and
 *                      cotangents (t x 3) in float64 (host), all NULL or all set
 *   flags              UVA_PLAN_* */
int uva_plan_create(uva_plan **out, int64_t n_islands, const int64_t *vert_off, const int64_t *tri_off,
                    const double *P, const int64_t *T, const int64_t *pins, const double *X,
                    const double *area, const double *cots, uint32_t flags, void *stream);
/* Plan of one symmetric positive definite matrix (CSR, host, int64 indices,
 * float64 values, both triangles stored) for uva_cg_solve: the GPU form of
 * param.py:103-132 SPDSolver(A). */
int uva_plan_create_matrix(uva_plan **out, int64_t n, const int64_t *indptr, const int64_t *indices,
                           const double *data, void *stream);
/* Release a plan: its device memory (taken from the device's stream-ordered
 * pool) returns to the pool in the order of the stream of the plan's latest
 * call, without waiting for the device; that stream must still exist. */
int uva_plan_destroy(uva_plan *plan);
int uva_plan_get_info(const uva_plan *plan, uva_plan_info *info);
/* Per-island statistics of the last solve call (host array of n_islands). */
int uva_plan_get_stats(uva_plan *plan, uva_island_stats *stats, void *stream);
/* Autotune the schedule: fit the schedule's cost model to the island times
 * of the last solve call and re-plan team sizes and the workgroup schedule
 * from the CG iterations per global step each island took in it (the
 * iteration counts of a repeated solve are reproducible).
 * Later calls run the new schedule; results are unchanged up to the order of
 * FP64 reductions across team members.  No counterpart in the reference. */
int uva_plan_tune(uva_plan *plan, void *stream);
/* The schedule's cost model, microseconds per CG iteration: cost[0] fixed
 * part of an island on one workgroup, cost[1] fixed part of a team member,
 * cost[2] per 64-row slice of a member.  uva_plan_tune fits it to the
 * measured island times of the last solve; uva_plan_set_tuning sets it (a
 * value measured on this GPU model earlier, say) and re-plans with the
 * iterations uva_plan_tune last measured, or the size-based guess.  No
 * counterpart in the reference. */
int uva_plan_get_tuning(const uva_plan *plan, double *cost);
int uva_plan_set_tuning(uva_plan *plan, const double *cost, void *stream);
/* Soft pins of weight w (w = 0: the island's own pin weight, max |W| of the
 * island) at the given vertices (caller ids) in place of each island's own
 * pin, for a plan with Batch conventions whose single island holds several
 * charts as one block-diagonal system: hier.py:407-412 and gpu.py:133-141 pin
 * every chart of a batch that one CG solves together.  Turns the coarse level
 * off (Jacobi-preconditioned CG). */
int uva_plan_set_pins(uva_plan *plan, const int64_t *pins, int64_t n_pins, double w, void *stream);
/* The Jacobi-PCG of uva_csr_pcg (same step lengths, stop test and
 * iteration count; two columns) over every row of the plan at once: its
 * islands' Laplacians side by side with the per-row soft pins of
 * uva_plan_set_pins (required), so the islands share the step lengths as the
 * reference's batched CG over a whole Batch or hier subset does (gpu.py:
 * 184-202 Batch._cg, hier.py:465-474).  b and x: device arrays n x 2 float64
 * in the caller's vertex order; x is the start and the result.  *iters_out
 * (may be NULL, then the call does not wait for the stream) gets the
 * iterations run.  Two launches per iteration. */
int uva_plan_pcg(uva_plan *plan, const double *b, double *x, int32_t max_iters, int32_t check_every, double tol2,
                 int32_t strict, int32_t *iters_out, void *stream);
/* Set the per-triangle target scale (gpu.py:123 Batch.scale); device array. */
int uva_plan_set_scale(uva_plan *plan, const double *scale_dev, void *stream);

/* LSCM of every island into U (device, n x 2 float64).  pins: 2 local ids per
 * island (host) or NULL for the reference's farthest boundary pair.  FP64 CG
 * to relative residual tol.  Plans built with refine conventions only
 * (UVA_PLAN_BATCH plans hold no LSCM boundary coupling). */
int uva_lscm(uva_plan *plan, double *U, const int64_t *pins, double tol, int32_t max_iters, void *stream);

/* ARAP iterations on U in place (device, n x 2 float64). */
int uva_arap(uva_plan *plan, double *U, const uva_arap_opts *opts, void *stream);

/* uva_arap run eagerly with a HIP event pair around every launch on the
 * stream: launch_ms[0] is the prologue, launch_ms[k] ARAP iteration k
 * (n_launch = iters + 1).  Measurement only; no counterpart in the reference. */
int uva_arap_profile(uva_plan *plan, double *U, const uva_arap_opts *opts, double *launch_ms, int32_t n_launch,
                     void *stream);

/* Exact UV overlap flag per triangle (validate.py:58-136 _overlaps; with
 * strict = 1 the zero-tolerance crossings of validate.py:229-299
 * _overlaps_strict).  U (n_vert x 2 float64), F (n_tri x 3 int64), skip
 * (n_tri bytes, 1 = left out; may be NULL) and flags (n_tri bytes, out) are
 * device arrays; grid is the g of the g x g candidate grid over [0,1]^2 and
 * rel the orientation tolerance relative to squared edge lengths.  Pairs in
 * one crowded cell are tested one by one, so a layout that stacks most
 * triangles in one cell costs time quadratic in their number, as in the
 * reference.  Synchronises the stream. */
int uva_overlap_flags(const double *U, int64_t n_vert, const int64_t *F, int64_t n_tri, const uint8_t *skip,
                      double rel, int32_t strict, int32_t grid, uint8_t *flags, void *stream);

/* Conservative raster of the gutter measurement (metrics.py:315-345
 * _raster_labels, used by validate.py:180-185 min_gutter): for each texel of
 * a res x res image over [0,1]^2 (row py, column px), last[py * res + px]
 * gets the largest triangle index whose raster covers it, -1 where none does
 * (the reference writes island labels in triangle order, so the texel's
 * label is that triangle's island).  A triangle covers the texels of its
 * floored, clamped bounding box (texel units) whose centre lies within
 * 0.7072 texel of every edge line; NaN or out-of-range corners cover
 * nothing.  U (n_vert x 2 float64), F (n_tri x 3 int64) and last (res * res
 * int32) are device arrays; 1 <= res <= 8192.  Synchronises the stream. */
int uva_raster_last(const double *U, int64_t n_vert, const int64_t *F, int64_t n_tri, int32_t res, int32_t *last,
                    void *stream);

/* uva_overlap_flags without any device allocation, for callers that cannot
 * allocate (code that must not call hipMalloc, such as tight serving loops or
 * harnesses that forbid it).  No counterpart in the reference.  The boxes,
 * cell counters and cell entries live in one 64 MiB array per GPU, held from
 * the library's load until the process ends: 16 B per triangle, 16 B per grid
 * cell and 12 B per cell entry must fit in it (about 1.5 million triangles
 * at the grid validate.py picks); larger inputs return UVA_EINVAL with a
 * message and touch nothing.  Calls on one GPU share that array, so calls
 * that overlap on one GPU (other threads or streams) race; serialise them.
 * Same arguments and result as uva_overlap_flags; synchronises the stream. */
int uva_overlap_flags_noalloc(const double *U, int64_t n_vert, const int64_t *F, int64_t n_tri,
                              const uint8_t *skip, double rel, int32_t strict, int32_t grid, uint8_t *flags,
                              void *stream);

/* Jacobi-preconditioned CG on one symmetric n x n matrix, the solve of
 * hier.py:534-582 Solver.lscm (one system over the u and v of every chart,
 * coupled by the boundary area term): the CSR arrays (int64 indptr and
 * indices, float64 data, repeated entries summed) are host memory, b and x
 * (n float64, x the start on entry and the result on exit) device memory.
 * At most max_iters iterations; after iteration k (from 0) with
 * k % check_every == 0 it stops when r.r < tol2 * max(b.b, 1e-30) (the
 * reference: check_every 50, tol2 1e-12).  *iters_out (may be NULL) gets the
 * iterations run.  Synchronises the stream. */
int uva_jacobi_pcg(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data, const double *b,
                   double *x, int32_t max_iters, int32_t check_every, double tol2, int32_t *iters_out, void *stream);

/* A symmetric matrix resident on the device for repeated coupled solves
 * (CSR from host memory as above; the handle holds the matrix, its Jacobi
 * weights and the work vectors of a two-column solve).  uva_csr_pcg: the
 * same iteration for ncols = 1 or 2 columns with their own step lengths
 * (b and x device arrays, n x ncols row-major); after iteration k with
 * k % check_every == 0 it stops when every column has r.r < tol2 max(b.b,
 * 1e-30) (strict = 1) or <= (strict = 0) -- gpu.py:184-202 Batch._cg is
 * check_every 1, tol2 = tol^2, strict 0; hier.py:465-474 (a fixed count) is
 * tol2 0, strict 1.  The stop test runs on the device; *iters_out (may be
 * NULL, then the call does not wait for the stream) gets the iterations
 * run.  One handle serves one solve at a time. */
typedef struct uva_csr uva_csr;
int uva_csr_create(uva_csr **out, int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
                   void *stream);
int uva_csr_destroy(uva_csr *m);
int uva_csr_pcg(uva_csr *m, const double *b, double *x, int32_t ncols, int32_t max_iters, int32_t check_every,
                double tol2, int32_t strict, int32_t *iters_out, void *stream);

/* SLIM (symmetric Dirichlet) options: param.refine(method="slim"). */
typedef struct uva_slim_opts {
    int32_t iters;           /* SLIM iterations (refine's iters) */
    int32_t cg_max_iters;    /* PCG cap per global step (0 = 20000) */
    double tol;              /* energy early exit |prev - e| <= tol prev */
    double cg_tol;           /* relative residual of the normal equations;
                                the reference solves them directly, so keep
                                it tight (1e-10) */
} uva_slim_opts;

/* SLIM iterations on U in place (device, n x 2 float64).  Plans built
 * without UVA_PLAN_BATCH only (refine conventions). */
int uva_slim(uva_plan *plan, double *U, const uva_slim_opts *opts, void *stream);

/* Building blocks (one launch each). */
/* local step + rhs: rotations of U and b (device, n x 2 float64); mode picks
 * the rotation rule (REFINE: identity when degenerate; BATCH: times scale) */
int uva_local_rhs(uva_plan *plan, const double *U, double *b, int32_t mode, void *stream);
/* batched CG: x (device, n x 2) warm start in, solution out, for rhs b. */
int uva_cg_solve(uva_plan *plan, const double *b, double *x, double tol, int32_t max_iters, int32_t precision,
                 void *stream);
/* flip-preventing step of U along D per island -> tmax (device, n_islands) */
int uva_max_step(uva_plan *plan, const double *U, const double *D, double *tmax, int32_t mode, void *stream);
/* The rest of one Batch (mode UVA_MODE_BATCH) or hier (UVA_MODE_HIER) ARAP
 * iteration after its global step, per island: Un (device, n x 2, the step's
 * result) is replaced by U on an island with a non-finite row; with
 * line_search (hier: unless a usable triangle of U is folded) the step is
 * min(1, 0.9 x the flip-safe limit along Un - U), else 1; U += step (Un - U)
 * in place.  No host synchronisation. */
int uva_step_apply(uva_plan *plan, double *U, const double *Un, int32_t mode, int32_t line_search, void *stream);
/* per-island energy (kind 0 ARAP, 1 symmetric Dirichlet) -> e (device) */
int uva_energy(uva_plan *plan, const double *U, int32_t kind, double *e, void *stream);
/* Differentiation (refine plans).  Per-island weighted energy
 * E_i = sum_f w_f e_f(U) (w: device, n_triangles, or NULL = 1) -> e. */
int uva_energy_fwd(uva_plan *plan, const double *U, const double *w, int32_t kind, double *e, void *stream);
/* gU = ge_i dE_i/dU (device, n x 2) and, if gw is not NULL, gw_f = ge_i e_f
 * (device, n_triangles), for the per-island upstream gradient ge (device). */
int uva_energy_bwd(uva_plan *plan, const double *U, const double *w, int32_t kind, const double *ge, double *gU,
                   double *gw, void *stream);
/* Adjoint of the global step: lam = L^-1 g per island (same SPD operator as
 * the ARAP solve, started from zero). */
int uva_adjoint_cg(uva_plan *plan, const double *g, double *lam, double tol, int32_t max_iters, int32_t precision,
                   void *stream);
/* Backward of one global step U' = L^-1 b(U): gU = (db/dU)^T lam, with lam
 * from uva_adjoint_cg on dLoss/dU'. */
int uva_arap_step_bwd(uva_plan *plan, const double *U, const double *lam, double *gU, void *stream);
/* per-triangle signed UV area -> sa (device, n_triangles) */
int uva_signed_areas(uva_plan *plan, const double *U, double *sa, void *stream);
/* Batch.flips_per_chart -> counts (device, n_islands float64) */
int uva_flips_per_chart(uva_plan *plan, const double *U, double *counts, void *stream);
/* Batch.repair_flips: Jacobi 1-ring relaxation, at most `rounds` rounds */
int uva_repair_jacobi(uva_plan *plan, double *U, int32_t rounds, void *stream);
/* Batch.orient_positive: mirror islands with negative area-weighted sign sum */
int uva_orient_positive(uva_plan *plan, double *U, void *stream);

/* param.repair_flips: sequential Gauss-Seidel flip repair of one chart on the
 * host (the reference's exact update order).  U (n x 2 float64) is updated
 * in place; T is t x 3 int64. */
int uva_repair_seq(double *U, int64_t n, const int64_t *T, int64_t t, int32_t rounds);

/* hier.Solver.relax's subset as islands, on the host with threads: n faces
 * (faces: their ids in the mesh mesh_F, n_mesh_f x 3 global vertex ids into
 * mesh_V, n_mesh_v x 3), their local vertex ids CV (n x 3, in [0, nv)) and
 * chart ids (n, in [0, nchart)).  Island k is the k-th chart with faces
 * (keep[k], *nk of them): its vertices in ascending local id, its faces in
 * subset order.  Outputs (caller-allocated): keep (nchart), vert_off and
 * tri_off (nk + 1), P (nv x 3, each vertex's mesh position), T (n x 3,
 * island-local ids), vorder (nv: local id of each island-order vertex), forder
 * (n).  Every local vertex must lie on one of the faces: UVA_EINVAL for ids out
 * of range or an uncovered vertex. */
int uva_chart_regroup(int64_t n, const int64_t *faces, const int64_t *CV, int64_t nv, const int64_t *chart,
                      int64_t nchart, const int64_t *mesh_F, int64_t n_mesh_f, const double *mesh_V, int64_t n_mesh_v,
                      int64_t *keep, int64_t *nk, int64_t *vert_off, int64_t *tri_off, double *P, int64_t *T,
                      int64_t *vorder, int64_t *forder);

#ifdef __cplusplus
}
#endif
#endif
