// Host memory / undefined-behaviour check of the library's host code and of
// the island tasks' logic: this driver is linked with the CPU emulation of the
// kernels (one thread per lane, -DUVA_EMULATOR) and built with AddressSanitizer
// and UBSan by tests/test_sanitize_host.py.  It runs the plan builder in both
// conventions and schedules, the emulated LSCM / ARAP / SLIM / CG / plan-PCG
// tasks for three team shapes on charts of 1 to 384 triangles, the
// sequential flip repair and the chart regrouping (valid and invalid input).
// Exit status 0 and "host_check ok" mean every call returned its expected
// status with finite output and neither sanitizer reported anything.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../ultimate_uv_artist_amd/csrc/setup.hpp"
#include "uvart.h"

extern "C" {
void* uva_emu_create(int64_t n_islands, const int64_t* vert_off, const int64_t* tri_off, const double* P,
                     const int64_t* T, const int64_t* pins, uint32_t flags, int32_t team_members, int32_t team_waves);
void uva_emu_destroy(void* p);
int uva_emu_lscm(void* p, double* U, double tol, int32_t max_iters, int32_t precision);
int uva_emu_arap(void* p, double* U, const uva_arap_opts* o);
int uva_emu_slim(void* p, double* U, const uva_slim_opts* o);
int uva_emu_local_rhs(void* p, double* U, double* b);
int uva_emu_step_apply(void* p, double* U, const double* Un, int32_t mode, int32_t line_search);
int uva_emu_set_pins(void* p, const int64_t* pins, int64_t n_pins, double w);
int uva_emu_cg(void* p, const double* b, double* x, double tol, int32_t max_iters, int32_t precision);
int uva_emu_plan_pcg(void* p, const double* b, double* x, int32_t max_iters, int32_t check_every, double tol2,
                     int32_t strict);
}

namespace {

int failures = 0;

void stage(const char* what) {   // progress line with the seconds since start
    static const auto t0 = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("[%7.2f s] %s\n", s, what);
    std::fflush(stdout);
}

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

uint64_t rng_state = 0x9e3779b97f4a7c15ull;
double urand() {   // xorshift64*, uniform in [0, 1)
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 0x2545f4914f6cdd1dull) >> 11) * (1.0 / 9007199254740992.0);
}

// A batch of charts, island-local triangle ids, concatenated.
struct Batch {
    std::vector<int64_t> vo{0}, to{0};
    std::vector<double> P;
    std::vector<int64_t> T;
    int64_t n() const { return vo.back(); }

    void add(const std::vector<double>& p, const std::vector<int64_t>& t) {
        P.insert(P.end(), p.begin(), p.end());
        T.insert(T.end(), t.begin(), t.end());
        vo.push_back(vo.back() + (int64_t)p.size() / 3);
        to.push_back(to.back() + (int64_t)t.size() / 3);
    }
    // an nx x ny grid of jittered quads on a bumped sheet (2 nx ny triangles);
    // nx = 0 gives a single triangle
    void grid(int nx, int ny) {
        std::vector<double> p;
        std::vector<int64_t> t;
        if (nx == 0) {
            p = {0, 0, 0, 1, 0.1, 0, 0.3, 0.9, 0.2};
            t = {0, 1, 2};
        } else {
            for (int j = 0; j <= ny; ++j)
                for (int i = 0; i <= nx; ++i) {
                    const double x = i + 0.2 * (urand() - 0.5), y = j + 0.2 * (urand() - 0.5);
                    p.insert(p.end(), {x, y, 0.3 * std::sin(0.4 * x) * std::cos(0.3 * y)});
                }
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int64_t a = j * (nx + 1) + i, b = a + 1, c = a + nx + 2, d = a + nx + 1;
                    if ((i + j) & 1) t.insert(t.end(), {a, b, c, a, c, d});
                    else t.insert(t.end(), {a, b, d, b, c, d});
                }
        }
        add(p, t);
    }
};

uva_arap_opts arap_opts(int iters, int mode, int precision, int line_search) {
    uva_arap_opts o{};
    o.iters = iters;
    o.line_search = line_search;
    o.mode = mode;
    o.precision = precision;
    o.tol = 0.0;
    o.cg_tol = 1e-8;
    o.cg_max_iters = 0;
    return o;
}

// The emulated tasks on one batch and team shape, refine conventions.
void refine_tasks(const Batch& b, int members, int waves, uint32_t extra, bool every_precision) {
    void* h = uva_emu_create((int64_t)b.vo.size() - 1, b.vo.data(), b.to.data(), b.P.data(), b.T.data(), nullptr,
                             extra, members, waves);
    check(h != nullptr, "uva_emu_create (refine)");
    if (!h) return;
    const size_t n2 = 2 * (size_t)b.n();
    std::vector<double> U(n2, 0.0);
    check(uva_emu_lscm(h, U.data(), 1e-8, 0, UVA_PREC_F64) == UVA_OK && finite(U), "lscm f64");
    std::vector<double> U2(n2, 0.0);
    check(uva_emu_lscm(h, U2.data(), 1e-6, 0, UVA_PREC_MIXED) == UVA_OK && finite(U2), "lscm mixed");
    for (double& x : U) x += 0.02 * (urand() - 0.5);
    for (int prec : {UVA_PREC_MIXED, UVA_PREC_F32, UVA_PREC_F64}) {
        if (prec != UVA_PREC_MIXED && !every_precision) continue;
        std::vector<double> V = U;
        const uva_arap_opts o = arap_opts(3, UVA_MODE_REFINE, prec, 1);
        check(uva_emu_arap(h, V.data(), &o) == UVA_OK && finite(V), "arap refine");
    }
    {
        std::vector<double> V = U;
        uva_arap_opts o = arap_opts(2, UVA_MODE_REFINE, UVA_PREC_MIXED, 1);
        o.iters_folded = 4;
        check(uva_emu_arap(h, V.data(), &o) == UVA_OK && finite(V), "arap refine, folded starts");
        uva_slim_opts s{2, 0, 1e-5, 1e-10};
        check(uva_emu_slim(h, V.data(), &s) == UVA_OK && finite(V), "slim");
        std::vector<double> rhs(n2, 0.0), x(n2, 0.0);
        check(uva_emu_local_rhs(h, V.data(), rhs.data()) == UVA_OK && finite(rhs), "local rhs");
        check(uva_emu_cg(h, rhs.data(), x.data(), 1e-8, 0, UVA_PREC_MIXED) == UVA_OK && finite(x), "cg");
    }
    uva_emu_destroy(h);
}

// Batch conventions: Batch.arap, per-row pins, the plan PCG and the step.
void batch_tasks(const Batch& b, int members, int waves) {
    std::vector<int64_t> pins(b.vo.begin(), b.vo.end() - 1);
    void* h = uva_emu_create((int64_t)b.vo.size() - 1, b.vo.data(), b.to.data(), b.P.data(), b.T.data(),
                             pins.data(), UVA_PLAN_BATCH, members, waves);
    check(h != nullptr, "uva_emu_create (batch)");
    if (!h) return;
    const size_t n2 = 2 * (size_t)b.n();
    std::vector<double> U(n2);
    for (int64_t v = 0; v < b.n(); ++v) {
        U[2 * v] = b.P[3 * v] + 0.05 * (urand() - 0.5);
        U[2 * v + 1] = b.P[3 * v + 1] + 0.05 * (urand() - 0.5);
    }
    for (int mode : {UVA_MODE_BATCH, UVA_MODE_HIER}) {
        std::vector<double> V = U;
        const uva_arap_opts o = arap_opts(2, mode, UVA_PREC_MIXED, 1);
        check(uva_emu_arap(h, V.data(), &o) == UVA_OK && finite(V), "arap batch");
    }
    check(uva_emu_set_pins(h, pins.data(), (int64_t)pins.size(), 0.0) == UVA_OK, "set pins");
    std::vector<double> rhs(n2, 0.0), x(n2, 0.0);
    check(uva_emu_local_rhs(h, U.data(), rhs.data()) == UVA_OK, "local rhs (batch)");
    const int its = uva_emu_plan_pcg(h, rhs.data(), x.data(), 200, 1, 1e-12, 1);
    check(its > 0 && finite(x), "plan pcg");
    std::vector<double> V = U;
    for (int ls : {0, 1})
        check(uva_emu_step_apply(h, V.data(), x.data(), UVA_MODE_BATCH, ls) == UVA_OK && finite(V), "step apply");
    uva_emu_destroy(h);
}

// The device plan builder with the product's schedule (256 slots on 8 XCDs,
// teams), both schedules, and with per-island iteration estimates.
void plan_builder(const Batch& b) {
    for (const char* sched : {"claim", "static"}) {
        setenv("UVART_SCHED", sched, 1);
        for (bool batch : {false, true}) {
            uva::HostPlan hp;
            const std::string err =
                uva::build_host_plan(hp, (int64_t)b.vo.size() - 1, b.vo.data(), b.to.data(), b.P.data(), b.T.data(),
                                     nullptr, nullptr, nullptr, nullptr, batch, nullptr, uva::kWaveMaxRows, 256, 8,
                                     true);
            check(err.empty(), "build_host_plan");
            if (!err.empty()) continue;
            std::vector<double> est(b.vo.size() - 1);
            for (double& e : est) e = 5.0 + 200.0 * urand();
            uva::plan_teams(hp, hp.group_isl, 256, 8, true, est.data());
            uva::plan_teams(hp, hp.group_isl, 256, 8, false, nullptr, uva::kResSlicesF64);
        }
    }
    unsetenv("UVART_SCHED");
    // an island with a vertex outside it is an error, not a crash
    Batch bad;
    bad.add({0, 0, 0, 1, 0, 0, 0, 1, 0}, {0, 1, 5});
    uva::HostPlan hp;
    check(!uva::build_host_plan(hp, 1, bad.vo.data(), bad.to.data(), bad.P.data(), bad.T.data(), nullptr, nullptr,
                                nullptr, nullptr, false, nullptr)
               .empty(),
          "build_host_plan rejects a vertex outside its island");
    // a general symmetric matrix: the 1-D Laplacian of 300 rows
    const int64_t m = 300;
    std::vector<int64_t> ip{0}, ix;
    std::vector<double> dv;
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = i - 1; j <= i + 1; ++j)
            if (j >= 0 && j < m) {
                ix.push_back(j);
                dv.push_back(i == j ? 2.0 : -1.0);
            }
        ip.push_back((int64_t)ix.size());
    }
    uva::HostPlan mp;
    check(uva::build_matrix_plan(mp, m, ip.data(), ix.data(), dv.data()).empty(), "build_matrix_plan");
}

void regroup_and_repair(const Batch& b) {
    // every face of the batch as one mesh, the charts' faces in reverse order
    const int64_t nf = b.to.back(), nv = b.n();
    std::vector<int64_t> F(3 * nf), chart(nf), faces(nf), CV(3 * nf);
    for (size_t c = 0; c + 1 < b.vo.size(); ++c)
        for (int64_t f = b.to[c]; f < b.to[c + 1]; ++f) {
            for (int k = 0; k < 3; ++k) F[3 * f + k] = b.T[3 * f + k] + b.vo[c];
            chart[nf - 1 - f] = (int64_t)c;
            faces[nf - 1 - f] = f;
        }
    for (int64_t i = 0; i < nf; ++i)
        for (int k = 0; k < 3; ++k) CV[3 * i + k] = F[3 * faces[i] + k];
    const int64_t nc = (int64_t)b.vo.size() - 1;
    std::vector<int64_t> keep(nc), vo(nc + 1), to(nc + 1), T(3 * nf), vorder(nv), forder(nf);
    std::vector<double> P(3 * nv);
    int64_t nk = 0;
    check(uva::chart_regroup(nf, faces.data(), CV.data(), nv, chart.data(), nc, F.data(), nf, b.P.data(), nv,
                             keep.data(), &nk, vo.data(), to.data(), P.data(), T.data(), vorder.data(), forder.data())
              .empty() &&
              nk == nc,
          "chart_regroup");
    chart[0] = nc;   // a chart id out of range
    check(!uva::chart_regroup(nf, faces.data(), CV.data(), nv, chart.data(), nc, F.data(), nf, b.P.data(), nv,
                              keep.data(), &nk, vo.data(), to.data(), P.data(), T.data(), vorder.data(),
                              forder.data())
               .empty(),
          "chart_regroup rejects a chart id out of range");
    // sequential flip repair of the largest chart's noisy layout
    const size_t c = b.vo.size() - 2;
    const int64_t n = b.vo[c + 1] - b.vo[c], t = b.to[c + 1] - b.to[c];
    std::vector<double> U(2 * n);
    for (int64_t v = 0; v < n; ++v) {
        U[2 * v] = b.P[3 * (b.vo[c] + v)] + 0.8 * (urand() - 0.5);
        U[2 * v + 1] = b.P[3 * (b.vo[c] + v) + 1] + 0.8 * (urand() - 0.5);
    }
    uva::repair_sequential(U.data(), n, b.T.data() + 3 * b.to[c], t, 30);
    check(finite(U), "repair_sequential");
}

}  // namespace

int main() {
    stage("start");
    // one wavefront per island on charts of 1 to 384 triangles; workgroups of
    // four wavefronts and teams of three two-wavefront members (one thread
    // per lane each) on charts of one and of 60 triangles
    Batch tiny, small;
    tiny.grid(0, 0);
    small.grid(0, 0);
    small.grid(1, 1);
    small.grid(3, 2);
    small.grid(9, 7);
    small.grid(16, 12);
    tiny.grid(6, 5);
    refine_tasks(small, 1, 1, 0, true);
    stage("wavefront: refine conventions");
    batch_tasks(small, 1, 1);
    stage("wavefront: batch conventions");
    refine_tasks(small, 1, 1, 1u << 8, false);   // streaming (non-resident) CG
    stage("wavefront: streaming CG");
    for (auto [m, w] : {std::pair<int, int>{1, 4}, {3, 2}}) {
        refine_tasks(tiny, m, w, 0, false);
        batch_tasks(tiny, m, w);
        std::printf("team of %d x %d wavefronts: ", m, w);
        stage("refine and batch conventions");
    }
    Batch large = small;
    for (int k = 0; k < 30; ++k) large.grid(5 + (int)(60 * urand()), 5 + (int)(60 * urand()));
    large.grid(150, 120);
    std::printf("plan builder over %lld islands\n", (long long)large.vo.size() - 1);
    plan_builder(large);
    stage("plan builder");
    regroup_and_repair(large);
    stage("regroup and repair");
    if (failures) {
        std::printf("host_check: %d failure(s)\n", failures);
        return 1;
    }
    std::printf("host_check ok\n");
    return 0;
}
