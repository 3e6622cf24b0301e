// C ABI of libuvart: plan lifetime, device memory, launch sequencing and
// hipGraph capture of the ARAP iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/uvart.h"
#include "kernels.hpp"
#include "setup.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) return fail(UVA_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

constexpr int kMaxSlots = 2048;  // queue counter pairs per call sequence
constexpr int kDefaultMaxIters = 20000;  // CG iterations per global step when the caller gives none

}  // namespace

namespace uva {
int set_error(int code, const char* msg) {
    g_err = msg;
    return code;
}
}  // namespace uva

struct uva_plan {
    uva::DevPlan dp{};
    std::vector<void*> allocs;
    int64_t device_bytes = 0;
    int64_t nnz_pad = 0, nnz_true = 0;
    int cus = 1;                          // compute units of the plan's device
    int clock_khz = 0;                    // device wall-clock rate (wall_clock64 ticks per ms)
    std::vector<double> P;                 // host positions (LSCM pins given later)
    std::vector<int> lpin0, lpin1, v0;     // host copies
    std::vector<double> ldist;             // host copy of the default pins' distances
    std::vector<int64_t> lcofs;            // host copy of the LSCM coarse offsets (the plan's own pins)
    bool lpins_custom = false;             // the device holds caller-given LSCM pins
    // Task schedule of the LSCM solve.  Its FP64 vectors keep fewer rows per
    // workgroup resident (kResSlicesF64) than the FP32 ARAP CG, and its CG
    // iterations grow faster with island size, so its teams and slot lists
    // differ from the ARAP schedule's; built at the first uva_lscm call.
    struct TaskList {
        const int* order_group = nullptr;
        const int* grp_mem = nullptr;
        const int* grp_m = nullptr;
        const int* grp_team = nullptr;
        const int* slot_ptr = nullptr;
        int n_group = 0, n_teams = 0, max_m = 1, slots_per_xcd = 1, claim_tasks = 0;
    } lsched;
    bool have_lsched = false;
    std::vector<int> lmembers;             // per island: workgroups of its LSCM team
    bool last_lscm = false;                // the latest solve call was uva_lscm
    double* pcg_scratch = nullptr;         // block partials and state of uva_plan_pcg
    int pcg_grid = 0;
    std::vector<int> vinv;                 // caller's vertex id -> plan vertex
    uva::HostPlan sched;                   // island sizes and schedule inputs (for uva_plan_tune)
    int team_alloc = 0;                    // teams the barrier arrays hold
    std::vector<int> members;              // per island: workgroups of its team (0 = one wavefront)
    std::vector<double> est;               // CG iterations per global step measured by uva_plan_tune
    double2* ubuf[3] = {nullptr, nullptr, nullptr};  // plan-order copies of the caller's vertex arrays
    int* d_lpin0 = nullptr;
    int* d_lpin1 = nullptr;
    double* d_ldist = nullptr;
    // graph cache of uva_arap
    hipStream_t cap = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<char> gkey;
    // Device memory comes from the device's stream-ordered pool, allocated
    // and released in the order of the stream of the plan's latest call:
    // creating and destroying plans (one per chart for the reference's
    // per-chart calls) then never waits for the whole device.
    hipStream_t mem = nullptr;

    // (mem is always a caller's stream, never the private capture stream:
    // the public entry points set it, the helpers they call do not.)
    ~uva_plan() {
        if (gexec) {
            (void)hipStreamSynchronize(mem);      // a replay may still run
            (void)hipGraphExecDestroy(gexec);
        }
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : allocs) (void)hipFreeAsync(p, mem);
        if (cap) (void)hipStreamDestroy(cap);
    }

    template <class T>
    hipError_t alloc(T** out, size_t n) {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMallocAsync(&p, bytes, mem);
        if (e != hipSuccess) return e;
        allocs.push_back(p);
        device_bytes += (int64_t)bytes;
        *out = (T*)p;
        return hipSuccess;
    }
    template <class T>
    hipError_t upload(const T** out, const std::vector<T>& v, hipStream_t s) {
        T* p = nullptr;
        hipError_t e = alloc(&p, v.size());
        if (e != hipSuccess) return e;
        if (!v.empty()) e = hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
        *out = p;
        return e;
    }
};

namespace {

// The plan's device view with the LSCM task schedule in place of the ARAP one.
uva::DevPlan lscm_view(const uva_plan* pl) {
    uva::DevPlan d = pl->dp;
    const auto& l = pl->lsched;
    d.order_group = l.order_group;
    d.grp_mem = l.grp_mem;
    d.grp_m = l.grp_m;
    d.grp_team = l.grp_team;
    d.slot_ptr = l.slot_ptr;
    d.n_group = l.n_group;
    d.n_teams = l.n_teams;
    d.max_m = l.max_m;
    d.slots_per_xcd = l.slots_per_xcd;
    d.claim_tasks = l.claim_tasks;
    return d;
}

int launch(uva_plan* pl, int kind, int prec, const uva::TaskArgs& a, int slot, hipStream_t s) {
    pl->last_lscm = kind == uva::TASK_LSCM;
    if (kind == uva::TASK_LSCM) HIP_TRY(uva::launch_island_tasks(kind, prec, lscm_view(pl), a, slot, pl->cus, s));
    else HIP_TRY(uva::launch_island_tasks(kind, prec, pl->dp, a, slot, pl->cus, s));
    return UVA_OK;
}

// The plan keeps its vertices in its own order (Morton order inside each
// island); every entry point copies the caller's vertex arrays (n x 2
// float64) into plan order on the stream and the results back.
double2* to_plan(uva_plan* pl, const double* ext, int k, hipStream_t s, int& rc) {
    if (rc) return nullptr;
    hipError_t e = uva::launch_permute(pl->dp, (const double2*)ext, pl->ubuf[k], true, s);
    if (e != hipSuccess) rc = fail(UVA_EHIP, std::string("permute: ") + hipGetErrorString(e));
    return pl->ubuf[k];
}

int from_plan(uva_plan* pl, int k, double* ext, hipStream_t s, int rc) {
    if (rc) return rc;
    HIP_TRY(uva::launch_permute(pl->dp, pl->ubuf[k], (double2*)ext, false, s));
    return UVA_OK;
}

// Task counters and team barrier counters of a call sequence start at zero
// (fill kernels, not memsets, so the sequence can be captured in a graph);
// the first reset of a call also clears the team fault flag.
// (n_teams: teams of the schedule the sequence runs, -1 for the ARAP one)
int reset_queue(uva_plan* pl, int slots, hipStream_t s, bool first = true, int n_teams = -1) {
    if (n_teams < 0) n_teams = pl->dp.n_teams;
    HIP_TRY(uva::launch_fill_int(pl->dp.queue, uva::kQueueStride * slots, 0, s));
    if (n_teams > 0) {
        HIP_TRY(uva::launch_fill_int(pl->dp.team_cnt, slots * n_teams, 0, s));
        if (first) HIP_TRY(uva::launch_fill_int(pl->dp.fault, 1, 0, s));
    }
    return UVA_OK;
}

// A team whose workgroups were not all resident together gives up waiting at
// a barrier after the hang guard and sets the plan's fault flag: the call's
// results are then invalid.  Every call that ran island tasks on a plan with
// multi-workgroup teams reads the flag back (waiting for the stream) and
// reports it; the flag is cleared again for the next call.
int check_fault(uva_plan* pl, hipStream_t s, int rc, int n_teams = -1) {
    if (n_teams < 0) n_teams = pl->dp.n_teams;
    if (rc || n_teams <= 0) return rc;
    int f = 0;
    HIP_TRY(hipMemcpyAsync(&f, pl->dp.fault, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!f) return UVA_OK;
    HIP_TRY(uva::launch_fill_int(pl->dp.fault, 1, 0, s));
    return fail(UVA_EHIP, "a multi-workgroup island team gave up waiting at a barrier (its workgroups were not "
                          "resident together); the results of this call are invalid");
}

uva::TaskArgs base_args(double* U) {
    uva::TaskArgs a{};
    a.U = (double2*)U;
    a.cg_tol = 1e-6;
    a.cg_max = kDefaultMaxIters;
    return a;
}

// One sequence of launches (queue reset, prologue, iterations) with slots
// reused in chunks of kMaxSlots.  With ev (2 events per launch), ev[2k] and
// ev[2k+1] bracket launch k (0 = prologue) on the stream.
int arap_sequence(uva_plan* pl, double* Uext, const uva_arap_opts* o, hipStream_t s,
                  const std::vector<hipEvent_t>* ev = nullptr) {
    int prc = UVA_OK;
    double2* U = to_plan(pl, Uext, 0, s, prc);
    if (prc) return prc;
    uva::TaskArgs a = base_args((double*)U);
    a.cg_tol = o->cg_tol;
    a.cg_max = o->cg_max_iters > 0 ? o->cg_max_iters : kDefaultMaxIters;
    a.tol = o->tol;
    a.line_search = o->line_search;
    a.batch = o->mode == UVA_MODE_BATCH || o->mode == UVA_MODE_HIER;
    a.hier = o->mode == UVA_MODE_HIER;
    a.iters_plain = o->iters_folded > 0 ? std::max(o->iters, 1) : 0;
    a.iters_folded = o->iters_folded;
    int total = 1 + std::max(o->iters, o->iters_folded);
    int done = 0;
    while (done < total) {
        int chunk = std::min(total - done, kMaxSlots);
        int rc = reset_queue(pl, chunk, s, done == 0);
        if (rc) return rc;
        for (int k = 0; k < chunk; ++k) {
            int step = done + k;
            if (ev) HIP_TRY(hipEventRecord((*ev)[2 * step], s));
            rc = launch(pl, step == 0 ? uva::TASK_PROLOGUE : uva::TASK_ARAP_ITER, o->precision, a, k, s);
            if (rc) return rc;
            if (ev) HIP_TRY(hipEventRecord((*ev)[2 * step + 1], s));
        }
        done += chunk;
    }
    return from_plan(pl, 0, Uext, s, UVA_OK);
}

bool valid_plan(const uva_plan* p) { return p != nullptr; }

}  // namespace

namespace {

// XCDs of the current device as the dispatcher uses them: the distinct
// hardware ids seen by the blocks of a probe launch, in ascending order
// (probed once per device; n = 1 if the probe fails or sees one XCD).
struct XccInfo {
    int n = 1;
    int map[16];
};

XccInfo xcc_info() {
    static std::mutex mu;
    static std::map<int, XccInfo> cache;
    int dev = 0;
    XccInfo xi;
    for (int k = 0; k < 16; ++k) xi.map[k] = -1;
    if (hipGetDevice(&dev) != hipSuccess) return xi;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    constexpr int kBlocks = 1024;
    int* d = nullptr;
    std::vector<int> h(kBlocks, -1);
    bool ok = hipMalloc(&d, kBlocks * sizeof(int)) == hipSuccess;
    ok = ok && uva::launch_xcc_probe(d, kBlocks, nullptr) == hipSuccess;
    ok = ok && hipMemcpy(h.data(), d, kBlocks * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    if (d) (void)hipFree(d);
    if (ok) {
        bool seen[16] = {false};
        for (int v : h) {
            if (v < 0 || v > 15) { ok = false; break; }
            seen[v] = true;
        }
        int n = 0;
        for (int k = 0; k < 16 && ok; ++k)
            if (seen[k]) xi.map[k] = n++;
        if (ok && n >= 1 && n <= uva::kMaxXcd) xi.n = n;
        else for (int k = 0; k < 16; ++k) xi.map[k] = -1;
    }
    if (xi.n == 1) for (int k = 0; k < 16; ++k) xi.map[k] = 0;
    cache[dev] = xi;
    return xi;
}

// The current device's default memory pool keeps what plans release for
// reuse instead of returning it at every synchronisation (once per device).
void keep_pool_memory() {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return;
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    done[dev] = true;
}

// Upload a host plan and allocate the scratch of a new device plan.
// Per island: workgroups of its team in the schedule (0 = a wavefront task).
std::vector<int> island_members(const uva::HostPlan& hp) {
    std::vector<int> m((size_t)std::max(hp.n_isl, 0), 0);
    for (size_t k = 0; k < hp.order_group.size(); ++k) m[hp.order_group[k]] = hp.grp_m[k];
    return m;
}

// Host-to-device copies of large plans through two pinned staging chunks
// per device (allocated once per process with hipHostMalloc): host threads
// fill one chunk while the other is copied.  Copies from pageable memory go
// through the runtime's own staging, at about 24 GB/s for a 600 MB plan.
class Staging {
  public:
    static constexpr size_t kChunk = size_t(32) << 20;
    static constexpr size_t kMin = size_t(1) << 20;   // smaller copies go straight from the source
    // Copies each item's (source, length) to its device address on stream s;
    // returns once every staged chunk's copy has completed.
    static hipError_t copy(const std::vector<std::pair<char*, std::pair<const char*, size_t>>>& items,
                           hipStream_t s) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        static std::mutex mu;
        static std::map<int, Staging> per_dev;
        std::lock_guard<std::mutex> lock(mu);
        Staging& st = per_dev[dev];
        if (!st.buf_[0]) {
            for (int k = 0; k < 2; ++k) {
                e = hipHostMalloc((void**)&st.buf_[k], kChunk, hipHostMallocDefault);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&st.ev_[k], hipEventDisableTiming);
                if (e != hipSuccess) {
                    st.release();
                    return e;
                }
            }
        }
        bool used[2] = {false, false};
        int k = 0;
        for (const auto& it : items) {
            char* dst = it.first;
            const char* src = it.second.first;
            size_t len = it.second.second;
            if (len < kMin) {
                e = hipMemcpyAsync(dst, src, len, hipMemcpyHostToDevice, s);
                if (e != hipSuccess) break;
                continue;
            }
            for (size_t o = 0; o < len && e == hipSuccess; o += kChunk) {
                const size_t n = std::min(kChunk, len - o);
                if (used[k]) e = hipEventSynchronize(st.ev_[k]);     // its last copy has finished
                if (e != hipSuccess) break;
                fill(st.buf_[k], src + o, n);
                e = hipMemcpyAsync(dst + o, st.buf_[k], n, hipMemcpyHostToDevice, s);
                if (e == hipSuccess) e = hipEventRecord(st.ev_[k], s);
                used[k] = true;
                k ^= 1;
            }
            if (e != hipSuccess) break;
        }
        for (int j = 0; j < 2; ++j)
            if (used[j]) {
                const hipError_t e2 = hipEventSynchronize(st.ev_[j]);
                if (e == hipSuccess) e = e2;
            }
        return e;
    }

  private:
    char* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    void release() {
        for (int k = 0; k < 2; ++k) {
            if (buf_[k]) (void)hipHostFree(buf_[k]);
            if (ev_[k]) (void)hipEventDestroy(ev_[k]);
            buf_[k] = nullptr;
            ev_[k] = nullptr;
        }
    }
    // memcpy of n bytes with up to 8 threads (slices of at least 2 MB)
    static void fill(char* dst, const char* src, size_t n) {
        const int nth = (int)std::max<size_t>(1, std::min<size_t>(8, n >> 21));
        if (nth == 1) {
            std::memcpy(dst, src, n);
            return;
        }
        const size_t part = (n + nth - 1) / nth;
        std::vector<std::thread> th;
        for (int t = 1; t < nth; ++t) {
            const size_t a = t * part, b = std::min(n, a + part);
            if (a < b) th.emplace_back([=] { std::memcpy(dst + a, src + a, b - a); });
        }
        std::memcpy(dst, src, std::min(n, part));
        for (auto& x : th) x.join();
    }
};

// Every array a plan build creates is a slice of one device block: one pool
// allocation and one or a few host-to-device copies per plan instead of one
// of each per array (the reference's per-chart calls build and drop a plan
// for every chart, so these calls are the cost of a small chart).
// Layout: uploaded arrays, then zero-filled arrays, then scratch whose
// contents the kernels define; slices start at 256-byte boundaries.
class Arena {
  public:
    template <class T, class P>
    void put(P** out, const T* src, size_t n) {          // upload n values of src
        static_assert(sizeof(T) % sizeof(P) == 0 || sizeof(P) % sizeof(T) == 0, "slice type");
        up_items_.push_back({up_, slice(n * sizeof(T)), src, n * sizeof(T), [out](char* b) { *out = (P*)b; }});
        up_ += up_items_.back().bytes;
    }
    template <class T, class A>
    void put(const T** out, const std::vector<T, A>& v) { put<T, const T>(out, v.data(), v.size()); }
    template <class T>
    void zero(T** out, size_t n) { zero_items_.push_back({0, slice(n * sizeof(T)), nullptr, 0, setter(out)}); }
    template <class T>
    void raw(T** out, size_t n) { raw_items_.push_back({0, slice(n * sizeof(T)), nullptr, 0, setter(out)}); }

    // One allocation from the plan's pool, the copies and the zero fill, all
    // on stream s; the host sources must live until s has passed the copies.
    hipError_t commit(uva_plan* pl, hipStream_t s) {
        size_t off = up_;
        for (Item& it : zero_items_) { it.off = off; off += it.bytes; }
        const size_t zero_end = off;
        for (Item& it : raw_items_) { it.off = off; off += it.bytes; }
        char* base = nullptr;
        hipError_t e = pl->alloc(&base, off);
        if (e != hipSuccess) return e;
        if (up_ > 0 && up_ <= kImageBytes) {              // small plans: one copy of a packed image
            image_.assign(up_, 0);
            for (const Item& it : up_items_)
                if (it.len) std::memcpy(image_.data() + it.off, it.src, it.len);
            e = hipMemcpyAsync(base, image_.data(), up_, hipMemcpyHostToDevice, s);
        } else {                                          // large plans: staged from the host arrays
            std::vector<std::pair<char*, std::pair<const char*, size_t>>> items;
            for (const Item& it : up_items_)
                if (it.len) items.push_back({base + it.off, {(const char*)it.src, it.len}});
            e = Staging::copy(items, s);
        }
        if (e == hipSuccess && zero_end > up_) e = hipMemsetAsync(base + up_, 0, zero_end - up_, s);
        for (const std::vector<Item>* v : {&up_items_, &zero_items_, &raw_items_})
            for (const Item& it : *v) it.set(base + it.off);
        return e;
    }

  private:
    static constexpr size_t kImageBytes = size_t(8) << 20;
    struct Item {
        size_t off, bytes;                                // slice
        const void* src;                                  // uploads: host source of len bytes
        size_t len;
        std::function<void(char*)> set;                   // stores the slice's address
    };
    static size_t slice(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~size_t(255); }
    template <class T>
    static std::function<void(char*)> setter(T** out) { return [out](char* b) { *out = (T*)b; }; }
    size_t up_ = 0;
    std::vector<Item> up_items_, zero_items_, raw_items_;
    std::vector<char> image_;
};

int upload_plan(uva::HostPlan& hp, const double* P, hipStream_t s, uva_plan** out) {
    std::unique_ptr<uva_plan> hold(new (std::nothrow) uva_plan());
    uva_plan* pl = hold.get();
    if (!pl) return fail(UVA_ENOMEM, "plan allocation failed");
    pl->mem = s;
    keep_pool_memory();
    auto bail = [&](int code, const std::string& m) { return fail(code, m); };
    uva::DevPlan& d = pl->dp;
    d.n_isl = hp.n_isl;
    d.n_vert = hp.n_vert;
    d.n_tri = hp.n_tri;
    d.mode_batch = hp.batch;
    d.n_slices = hp.sl_ptr.empty() ? 0 : (int)hp.sl_ptr.size() - 1;
    d.m_zero = hp.sl_ptr.empty() ? 0 : hp.sl_ptr.back();
    Arena ar;
#define UP(field, vec) ar.put(&d.field, hp.vec)
    UP(isl_v0, isl_v0); UP(isl_nv, isl_nv); UP(isl_t0, isl_t0); UP(isl_nt, isl_nt);
    UP(isl_s0, isl_s0); UP(isl_ns, isl_ns); UP(isl_pin, isl_pin); UP(isl_wpin, isl_wpin); UP(isl_bw, isl_bw);
    UP(tv0, tv0); UP(tv1, tv1); UP(tv2, tv2);
    UP(fx1, fx1); UP(fx2, fx2); UP(fy2, fy2); UP(farea, farea); UP(fok, fok);
    UP(inc_ptr, inc_ptr); UP(inc_tri, inc_tri);
    UP(sl_ptr, sl_ptr); UP(sl_row, sl_row); UP(m_col, m_col); UP(m_hi, m_hi); UP(m_lo, m_lo); UP(m_b, m_b);
    UP(diag, diag); UP(dinv32, dinv32);
    UP(adj_ptr, adj_ptr); UP(adj_idx, adj_idx); UP(vbnd, vbnd);
    UP(order_group, order_group); UP(order_wave, order_wave);
    UP(grp_mem, grp_mem); UP(grp_m, grp_m); UP(grp_team, grp_team); UP(slot_ptr, slot_ptr);
    UP(vperm, vperm); UP(isl_cofs, isl_cofs); UP(isl_cg, isl_cg); UP(isl_lcg, isl_lcg); UP(cinv, cinv);
    UP(isl_lcofs, isl_lcofs); UP(lcinv, lcinv);
#undef UP
    ar.put<double, const double2>(&d.inc_g, hp.inc_g.data(), hp.inc_g.size());
    ar.put<double, const double2>(&d.tri_g, hp.tri_g.data(), hp.tri_g.size());
    // LSCM pins (mutable through uva_lscm)
    ar.put<int, int>(&pl->d_lpin0, hp.isl_lpin0.data(), hp.isl_lpin0.size());
    ar.put<int, int>(&pl->d_lpin1, hp.isl_lpin1.data(), hp.isl_lpin1.size());
    ar.put<double, double>(&pl->d_ldist, hp.isl_ldist.data(), hp.isl_ldist.size());
    const size_t nv = hp.n_vert, nt = hp.n_tri;
    const std::vector<double> ones(nt, 1.0);
    ar.put<double, double>(&d.fscale, ones.data(), nt);
    // zero-filled: the CG vectors (defined contents from the start), the
    // island states and the fault flag
    void** vecs[6] = {&d.vr, &d.vp, &d.vAp, &d.ve, &d.vu, &d.vw};
    for (void** v : vecs) ar.zero((double2**)v, nv);
    ar.zero(&d.st, (size_t)hp.n_isl);
    ar.zero(&d.fault, 1);
    // scratch the kernels write before reading
    ar.raw(&d.crc, hp.sl_ptr.empty() ? 1 : hp.sl_ptr.size());
    for (int k = 0; k < 3; ++k) ar.raw(&pl->ubuf[k], nv);
    ar.raw(&d.team_cnt, (size_t)kMaxSlots * std::max(1, hp.n_teams));
    ar.raw(&d.team_part, (size_t)std::max(1, hp.n_teams) * 2 * uva::kMaxTeam * 8);
    ar.raw(&d.R, nt);
    ar.raw(&d.b64, nv);
    ar.raw(&d.X64, nv);
    ar.raw(&d.sy, 3 * nt);               // per-triangle corner values of the matrix-free passes (SLIM, backward)
    ar.raw(&d.red, 4 * (size_t)hp.n_isl);
    ar.raw(&d.queue, uva::kQueueStride * (size_t)kMaxSlots);
    {
        hipError_t e = ar.commit(pl, s);
        if (e != hipSuccess)
            return bail(e == hipErrorOutOfMemory ? UVA_ENOMEM : UVA_EHIP,
                        std::string("plan upload: ") + hipGetErrorString(e));
    }
    d.isl_lpin0 = pl->d_lpin0;
    d.isl_lpin1 = pl->d_lpin1;
    d.isl_ldist = pl->d_ldist;
    pl->team_alloc = std::max(1, hp.n_teams);
    pl->vinv.assign(hp.n_vert, 0);
    for (int v = 0; v < hp.n_vert; ++v) pl->vinv[hp.vperm[v]] = v;
    pl->lpin0 = hp.isl_lpin0;
    pl->lpin1 = hp.isl_lpin1;
    pl->ldist = hp.isl_ldist;
    pl->lcofs = hp.isl_lcofs;
    pl->v0 = hp.isl_v0;
    if (P) pl->P.assign(P, P + 3 * (size_t)hp.n_vert);
    d.n_group = (int)hp.order_group.size();
    d.n_wave = (int)hp.order_wave.size();
    d.n_teams = hp.n_teams;
    pl->members = island_members(hp);
    {
        uva::HostPlan& sc = pl->sched;
        sc.n_isl = hp.n_isl;
        sc.isl_nv = hp.isl_nv;
        sc.isl_ns = hp.isl_ns;
        sc.group_isl = hp.group_isl;
        sc.sched_slots = hp.sched_slots;
        sc.sched_teams = hp.sched_teams;
        sc.n_xcd = hp.n_xcd;
    }
    // XCD lists: the probed hardware ids map to lists 0..n_xcd-1 when the plan
    // was built for them, else everything maps to the single list
    d.n_xcd = hp.n_xcd;
    d.max_m = hp.max_m;
    d.slots_per_xcd = hp.slots_per_xcd;
    d.claim_tasks = hp.claim_tasks ? 1 : 0;
    const XccInfo xi = xcc_info();
    d.xcd_local = (hp.n_xcd > 1 && hp.n_xcd == xi.n) ? 1 : 0;
    // an XCD the probe never saw has no list (its workgroups only take wavefront tasks)
    for (int k = 0; k < 16; ++k) d.xcc_map[k] = d.xcd_local ? xi.map[k] : 0;
    pl->nnz_pad = hp.sl_ptr.empty() ? 0 : hp.sl_ptr.back();
    pl->nnz_true = hp.nnz_true;
    // grid: enough workgroups to hold every group task and the wave tasks,
    // capped at what the device keeps resident
    int dev = 0, cus = 1;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipDeviceGetAttribute(&pl->clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) pl->clock_khz = 0;
    pl->cus = cus;
    // a team barrier gives up after 20 s (hang guard only)
    d.spin_ticks = (long long)(pl->clock_khz > 0 ? pl->clock_khz : 100000) * 20000;
    hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess) return bail(UVA_EHIP, std::string("plan upload: ") + hipGetErrorString(se));
    *out = hold.release();
    return UVA_OK;
}

}  // namespace

extern "C" {

int uva_version(void) { return 10000; }

const char* uva_last_error(void) { return g_err.c_str(); }

int uva_struct_size(int32_t which) {
    switch (which) {
        case 0: return (int)sizeof(uva_arap_opts);
        case 1: return (int)sizeof(uva_slim_opts);
        case 2: return (int)sizeof(uva_island_stats);
        case 3: return (int)sizeof(uva_plan_info);
        default: return -1;
    }
}

int uva_plan_create(uva_plan** out, int64_t n_islands, const int64_t* vert_off, const int64_t* tri_off,
                    const double* P, const int64_t* T, const int64_t* pins, const double* X, const double* area,
                    const double* cots, uint32_t flags, void* stream) {
    if (!out || !vert_off || !tri_off || !P || !T) return fail(UVA_EINVAL, "null argument");
    *out = nullptr;
    if ((X || area || cots) && !(X && area && cots)) return fail(UVA_EINVAL, "X, area and cots go together");
    hipStream_t s = (hipStream_t)stream;
    // one CG workgroup per CU is resident: the schedule has one workgroup
    // slot per CU
    int slots = 1;
    {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0)
            slots = c;
    }
    const int n_xcd = xcc_info().n;
    uva::HostPlan hp;
    std::string err;
    try {
        err = uva::build_host_plan(hp, n_islands, vert_off, tri_off, P, T, pins, X, area, cots,
                                   (flags & UVA_PLAN_BATCH) != 0, nullptr,
                                   (flags & UVA_PLAN_NO_WAVE) ? -1 : uva::kWaveMaxRows,
                                   slots, n_xcd, (flags & UVA_PLAN_NO_TEAMS) == 0,
                                   (flags & UVA_PLAN_KEEP_ORDER) == 0, (flags & UVA_PLAN_NO_COARSE) == 0);
    } catch (const std::bad_alloc&) {
        return fail(UVA_ENOMEM, "host allocation failed while building the plan");
    }
    if (!err.empty()) return fail(UVA_EINVAL, err);
    return upload_plan(hp, P, s, out);
}

int uva_plan_create_matrix(uva_plan** out, int64_t n, const int64_t* indptr, const int64_t* indices,
                           const double* data, void* stream) {
    if (!out || !indptr || (!indices && n > 0) || (!data && n > 0)) return fail(UVA_EINVAL, "null argument");
    *out = nullptr;
    uva::HostPlan hp;
    std::string err;
    try {
        err = uva::build_matrix_plan(hp, n, indptr, indices, data);
    } catch (const std::bad_alloc&) {
        return fail(UVA_ENOMEM, "host allocation failed while building the plan");
    }
    if (!err.empty()) return fail(UVA_EINVAL, err);
    return upload_plan(hp, nullptr, (hipStream_t)stream, out);
}

int uva_plan_destroy(uva_plan* plan) {
    if (!plan) return UVA_OK;
    delete plan;   // memory released in the order of the plan's latest stream
    return UVA_OK;
}

int uva_plan_get_info(const uva_plan* plan, uva_plan_info* info) {
    if (!valid_plan(plan) || !info) return fail(UVA_EINVAL, "null argument");
    info->n_islands = plan->dp.n_isl;
    info->n_vertices = plan->dp.n_vert;
    info->n_triangles = plan->dp.n_tri;
    info->nnz = plan->nnz_pad;
    info->nnz_true = plan->nnz_true;
    info->device_bytes = plan->device_bytes;
    info->n_group_tasks = plan->dp.n_group;
    info->n_wave_tasks = plan->dp.n_wave;
    info->n_teams = plan->dp.n_teams;
    info->reserved = 0;
    return UVA_OK;
}

int uva_plan_get_stats(uva_plan* plan, uva_island_stats* stats, void* stream) {
    if (!valid_plan(plan) || !stats) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    hipStream_t s = (hipStream_t)stream;
    std::vector<uva::IslandState> st(plan->dp.n_isl);
    int fault = 0;
    HIP_TRY(hipMemcpyAsync(st.data(), plan->dp.st, st.size() * sizeof(uva::IslandState), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&fault, plan->dp.fault, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (fault) return fail(UVA_EHIP, "a multi-workgroup island team gave up waiting at a barrier; results are invalid");
    for (size_t i = 0; i < st.size(); ++i) {
        uva_island_stats& o = stats[i];
        o.cg_iters = st[i].cg_iters;
        o.cg_iters_max = st[i].cg_iters_max;
        o.arap_iters = st[i].arap_iters;
        o.converged = st[i].converged;
        o.flips0 = st[i].flips0;
        o.use_ls = st[i].use_ls;
        o.status = st[i].status;
        const std::vector<int>& mem = plan->last_lscm ? plan->lmembers : plan->members;
        o.members = i < mem.size() ? mem[i] : 0;
        o.rel_resid = st[i].rel_resid;
        o.energy = st[i].e_prev;
        o.step = st[i].step;
        o.busy_ms = plan->clock_khz > 0 ? (double)st[i].ticks / (double)plan->clock_khz : 0.0;
    }
    return UVA_OK;
}

}  // extern "C"

namespace {

// Re-plan team sizes and the workgroup schedule from per-island CG
// iterations per global step (est, or NULL for the size-based guess) with the
// plan's cost model, and upload the new task arrays.
int replan(uva_plan* plan, const double* est, hipStream_t s) {
    uva::DevPlan& d = plan->dp;
    plan->mem = s;
    uva::HostPlan& sc = plan->sched;
    try {
        uva::plan_teams(sc, sc.group_isl, sc.sched_slots, sc.n_xcd, sc.sched_teams, est);
    } catch (const std::bad_alloc&) {
        return fail(UVA_ENOMEM, "host allocation failed while re-planning");
    }
    // new task arrays (the old ones stay allocated until the plan goes away:
    // a cached graph or work in flight on another stream may still read them)
#define UPS(field, vec)                                                                          \
    do {                                                                                         \
        hipError_t _e = plan->upload(&d.field, sc.vec, s);                                        \
        if (_e != hipSuccess) return fail(UVA_EHIP, std::string("upload " #vec ": ") + hipGetErrorString(_e)); \
    } while (0)
    UPS(order_group, order_group); UPS(grp_mem, grp_mem); UPS(grp_m, grp_m); UPS(grp_team, grp_team);
    UPS(slot_ptr, slot_ptr);
#undef UPS
    if (sc.n_teams > plan->team_alloc) {
        if (plan->alloc(&d.team_cnt, (size_t)kMaxSlots * sc.n_teams) != hipSuccess ||
            plan->alloc(&d.team_part, (size_t)sc.n_teams * 2 * uva::kMaxTeam * 8) != hipSuccess)
            return fail(UVA_ENOMEM, "team arrays");
        plan->team_alloc = sc.n_teams;
    }
    plan->members = island_members(sc);
    d.n_group = (int)sc.order_group.size();
    d.n_teams = sc.n_teams;
    d.max_m = sc.max_m;
    d.slots_per_xcd = sc.slots_per_xcd;
    d.claim_tasks = sc.claim_tasks ? 1 : 0;
    // a cached graph holds the old launch parameters
    if (plan->gexec) { (void)hipGraphExecDestroy(plan->gexec); plan->gexec = nullptr; }
    if (plan->graph) { (void)hipGraphDestroy(plan->graph); plan->graph = nullptr; }
    plan->gkey.clear();
    HIP_TRY(hipStreamSynchronize(s));
    return UVA_OK;
}

// The LSCM task schedule from per-island CG iterations (est, or NULL for a
// size-based guess: about 1.3 n^0.8 Jacobi-PCG iterations to 1e-8 on the
// conformal operator of an n-vertex disk) with the plan's cost model.
int plan_lscm(uva_plan* plan, const double* est, hipStream_t s) {
    uva::DevPlan& d = plan->dp;
    plan->mem = s;
    const uva::HostPlan& ms = plan->sched;
    uva::HostPlan sc;
    sc.n_isl = ms.n_isl;
    sc.isl_nv = ms.isl_nv;
    sc.isl_ns = ms.isl_ns;
    for (int k = 0; k < 3; ++k) sc.cost[k] = ms.cost[k];
    std::vector<double> guess;
    if (!est) {
        guess.resize((size_t)std::max(ms.n_isl, 1));
        for (int i = 0; i < ms.n_isl; ++i) guess[i] = 1.3 * std::pow((double)ms.isl_nv[i], 0.8);
        est = guess.data();
    }
    try {
        uva::plan_teams(sc, ms.group_isl, ms.sched_slots, ms.n_xcd, ms.sched_teams, est, uva::kResSlicesF64);
    } catch (const std::bad_alloc&) {
        return fail(UVA_ENOMEM, "host allocation failed while planning the LSCM schedule");
    }
    auto& l = plan->lsched;
#define UPL(field)                                                                               \
    do {                                                                                         \
        hipError_t _e = plan->upload(&l.field, sc.field, s);                                      \
        if (_e != hipSuccess) return fail(UVA_EHIP, std::string("upload " #field ": ") + hipGetErrorString(_e)); \
    } while (0)
    UPL(order_group); UPL(grp_mem); UPL(grp_m); UPL(grp_team); UPL(slot_ptr);
#undef UPL
    if (sc.n_teams > plan->team_alloc) {
        if (plan->alloc(&d.team_cnt, (size_t)kMaxSlots * sc.n_teams) != hipSuccess ||
            plan->alloc(&d.team_part, (size_t)sc.n_teams * 2 * uva::kMaxTeam * 8) != hipSuccess)
            return fail(UVA_ENOMEM, "team arrays");
        plan->team_alloc = sc.n_teams;
    }
    l.n_group = (int)sc.order_group.size();
    l.n_teams = sc.n_teams;
    l.max_m = sc.max_m;
    l.slots_per_xcd = sc.slots_per_xcd;
    l.claim_tasks = sc.claim_tasks ? 1 : 0;
    plan->lmembers = island_members(sc);
    plan->have_lsched = true;
    HIP_TRY(hipStreamSynchronize(s));
    return UVA_OK;
}

// Weighted least squares fit of the schedule's cost model to the measured
// island times: microseconds per CG iteration = c0 (one workgroup) or c1
// (team member) + c2 x slices per member, weighted by each island's CG
// iterations.  A class without islands keeps its constant; a fit that is not
// positive keeps the old model.
void fit_costs(const std::vector<uva::IslandState>& st, const std::vector<int>& members,
               const std::vector<int>& ns, int clock_khz, double cost[3]) {
    if (clock_khz <= 0) return;
    double A[3][3] = {{0}}, r[3] = {0};
    bool have[2] = {false, false};
    for (size_t i = 0; i < st.size(); ++i) {
        const int m = i < members.size() ? members[i] : 0;
        if (m < 1 || st[i].cg_iters <= 0 || st[i].ticks <= 0) continue;
        const double w = st[i].cg_iters;
        const double y = (double)st[i].ticks / clock_khz * 1e3 / st[i].cg_iters;
        const double x[3] = {m == 1 ? 1.0 : 0.0, m > 1 ? 1.0 : 0.0, (double)((ns[i] + m - 1) / m)};
        have[m > 1 ? 1 : 0] = true;
        for (int a = 0; a < 3; ++a) {
            r[a] += w * x[a] * y;
            for (int b = 0; b < 3; ++b) A[a][b] += w * x[a] * x[b];
        }
    }
    // classes without data: their constant moves to the right-hand side
    double c[3] = {cost[0], cost[1], cost[2]};
    int act[3], na = 0;
    for (int k = 0; k < 2; ++k)
        if (have[k]) act[na++] = k;
    if (na == 0) return;
    act[na++] = 2;
    double M[3][4];
    for (int a = 0; a < na; ++a) {
        double rhs = r[act[a]];
        for (int k = 0; k < 2; ++k)
            if (!have[k]) rhs -= A[act[a]][k] * c[k];
        for (int b = 0; b < na; ++b) M[a][b] = A[act[a]][act[b]];
        M[a][na] = rhs;
    }
    for (int k = 0; k < na; ++k) {             // Gaussian elimination with partial pivoting
        int piv = k;
        for (int a = k + 1; a < na; ++a)
            if (std::fabs(M[a][k]) > std::fabs(M[piv][k])) piv = a;
        if (std::fabs(M[piv][k]) < 1e-12) return;
        for (int b = 0; b <= na; ++b) std::swap(M[k][b], M[piv][b]);
        for (int a = 0; a < na; ++a) {
            if (a == k) continue;
            const double f = M[a][k] / M[k][k];
            for (int b = k; b <= na; ++b) M[a][b] -= f * M[k][b];
        }
    }
    double out[3] = {c[0], c[1], c[2]};
    for (int a = 0; a < na; ++a) out[act[a]] = M[a][na] / M[a][a];
    for (int k = 0; k < 3; ++k)
        if (!(out[k] > 0.0) || !std::isfinite(out[k])) return;
    for (int k = 0; k < 3; ++k) cost[k] = out[k];
}

}  // namespace

extern "C" {

int uva_plan_tune(uva_plan* plan, void* stream) {
    if (!valid_plan(plan)) return fail(UVA_EINVAL, "null argument");
    hipStream_t s = (hipStream_t)stream;
    uva::DevPlan& d = plan->dp;
    std::vector<uva::IslandState> st(d.n_isl);
    HIP_TRY(hipMemcpyAsync(st.data(), d.st, st.size() * sizeof(uva::IslandState), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<double> est(d.n_isl);
    bool any = false;
    for (int i = 0; i < d.n_isl; ++i) {
        est[i] = (double)st[i].cg_iters / std::max(1, st[i].arap_iters);
        any = any || st[i].cg_iters > 0;
    }
    if (!any) return fail(UVA_EINVAL, "uva_plan_tune needs a finished solve on the plan (no CG iterations recorded)");
    // after an LSCM solve: its own schedule from the measured iterations
    if (plan->last_lscm) return plan_lscm(plan, est.data(), s);
    fit_costs(st, plan->members, plan->sched.isl_ns, plan->clock_khz, plan->sched.cost);
    plan->est.assign(est.begin(), est.end());
    return replan(plan, est.data(), s);
}

int uva_plan_get_tuning(const uva_plan* plan, double* cost) {
    if (!valid_plan(plan) || !cost) return fail(UVA_EINVAL, "null argument");
    for (int k = 0; k < 3; ++k) cost[k] = plan->sched.cost[k];
    return UVA_OK;
}

int uva_plan_set_tuning(uva_plan* plan, const double* cost, void* stream) {
    if (!valid_plan(plan) || !cost) return fail(UVA_EINVAL, "null argument");
    for (int k = 0; k < 3; ++k)
        if (!(cost[k] > 0.0) || !std::isfinite(cost[k])) return fail(UVA_EINVAL, "costs must be positive and finite");
    for (int k = 0; k < 3; ++k) plan->sched.cost[k] = cost[k];
    return replan(plan, plan->est.empty() ? nullptr : plan->est.data(), (hipStream_t)stream);
}

int uva_plan_pcg(uva_plan* plan, const double* b, double* x, int32_t max_iters, int32_t check_every, double tol2,
                 int32_t strict, int32_t* iters_out, void* stream) {
    if (!valid_plan(plan) || !b || !x) return fail(UVA_EINVAL, "null argument");
    if (max_iters < 0 || check_every < 1 || !(tol2 >= 0.0))
        return fail(UVA_EINVAL, "max_iters must be >= 0, check_every >= 1 and tol2 >= 0");
    uva::DevPlan& d = plan->dp;
    if (!d.rpin) return fail(UVA_EINVAL, "uva_plan_pcg needs per-row pins (uva_plan_set_pins)");
    if (iters_out) *iters_out = 0;
    if (max_iters == 0 || d.n_vert == 0) return UVA_OK;
    hipStream_t s = (hipStream_t)stream;
    plan->mem = s;
    if (!plan->pcg_scratch) {
        plan->pcg_grid = uva::plan_pcg_grid(d, plan->cus);
        if (plan->alloc(&plan->pcg_scratch, uva::plan_pcg_scratch_doubles(plan->pcg_grid)) != hipSuccess)
            return fail(UVA_ENOMEM, "plan PCG scratch");
    }
    int rc = UVA_OK;
    const double2* bp = to_plan(plan, b, 1, s, rc);
    double2* xp = to_plan(plan, x, 0, s, rc);
    if (rc) return rc;
    HIP_TRY(uva::launch_plan_pcg(d, bp, xp, (double2*)d.vr, (double2*)d.vu, (double2*)d.vp, (double2*)d.vAp,
                                 plan->pcg_scratch, plan->pcg_grid, max_iters, check_every, tol2, strict, s));
    rc = from_plan(plan, 0, x, s, UVA_OK);
    if (rc || !iters_out) return rc;
    int its = 0;
    HIP_TRY(hipMemcpyAsync(&its, uva::plan_pcg_iters_ptr(plan->pcg_scratch, plan->pcg_grid), sizeof(int),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *iters_out = its;
    return UVA_OK;
}

int uva_plan_set_scale(uva_plan* plan, const double* scale_dev, void* stream) {
    if (!valid_plan(plan) || !scale_dev) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (plan->dp.n_tri)
        HIP_TRY(hipMemcpyAsync(plan->dp.fscale, scale_dev, plan->dp.n_tri * sizeof(double), hipMemcpyDeviceToDevice,
                               (hipStream_t)stream));
    return UVA_OK;
}

int uva_plan_set_pins(uva_plan* plan, const int64_t* pins, int64_t n_pins, double w, void* stream) {
    if (!valid_plan(plan) || (n_pins > 0 && !pins)) return fail(UVA_EINVAL, "null argument");
    uva::DevPlan& d = plan->dp;
    if (!d.mode_batch) return fail(UVA_EINVAL, "per-vertex pins need a plan with Batch conventions");
    if (!(w >= 0.0) || !std::isfinite(w)) return fail(UVA_EINVAL, "pin weight must be >= 0 and finite");
    if (n_pins < 0) return fail(UVA_EINVAL, "n_pins must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    plan->mem = s;
    const int n = d.n_vert;
    // w = 0: each pin takes the soft-pin weight of its island
    std::vector<double> wisl((size_t)std::max(d.n_isl, 1), w);
    if (w == 0.0 && d.n_isl) {
        HIP_TRY(hipMemcpyAsync(wisl.data(), d.isl_wpin, d.n_isl * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    // the pinned rows (weights of a row pinned more than once add up, in
    // the pins' order), then every row's weight, pinned diagonal and its
    // FP32 inverse on the device
    std::vector<std::pair<int, double>> pr;
    pr.reserve((size_t)n_pins);
    for (int64_t i = 0; i < n_pins; ++i) {
        if (pins[i] < 0 || pins[i] >= n) return fail(UVA_EINVAL, "pin outside the plan's vertices");
        const int r = plan->vinv[pins[i]];
        const int isl = (int)(std::upper_bound(plan->v0.begin(), plan->v0.end(), r) - plan->v0.begin()) - 1;
        pr.push_back({r, wisl[std::max(isl, 0)]});
    }
    std::stable_sort(pr.begin(), pr.end(),
                     [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
    std::vector<int> rows;
    std::vector<double> rw;
    for (const auto& e : pr) {
        if (!rows.empty() && rows.back() == e.first) {
            rw.back() += e.second;
        } else {
            rows.push_back(e.first);
            rw.push_back(0.0 + e.second);
        }
    }
    if (!d.rpin) {
        double *p = nullptr, *q = nullptr;
        if (plan->alloc(&p, (size_t)std::max(n, 1)) != hipSuccess || plan->alloc(&q, (size_t)std::max(n, 1)) != hipSuccess)
            return fail(UVA_ENOMEM, "pin weights");
        d.rpin = p;
        d.pdiag = q;
    }
    int* drows = nullptr;
    double* dw = nullptr;
    const int m = (int)rows.size();
    if (plan->alloc(&drows, (size_t)std::max(m, 1)) != hipSuccess || plan->alloc(&dw, (size_t)std::max(m, 1)) != hipSuccess)
        return fail(UVA_ENOMEM, "pin list");
    if (m) {
        HIP_TRY(hipMemcpyAsync(drows, rows.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(dw, rw.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
    }
    HIP_TRY(uva::launch_set_pins(d.diag, drows, dw, m, (double*)d.rpin, (double*)d.pdiag, (float*)d.dinv32, n, s));
    // the coarse inverses hold the islands' own pins: Jacobi only from here on
    std::vector<int64_t> none((size_t)std::max(d.n_isl, 1), -1);
    if (d.n_isl)
        HIP_TRY(hipMemcpyAsync((void*)d.isl_cofs, none.data(), d.n_isl * sizeof(int64_t), hipMemcpyHostToDevice, s));
    if (plan->gexec) { (void)hipGraphExecDestroy(plan->gexec); plan->gexec = nullptr; }
    if (plan->graph) { (void)hipGraphDestroy(plan->graph); plan->graph = nullptr; }
    plan->gkey.clear();
    HIP_TRY(hipStreamSynchronize(s));
    return UVA_OK;
}

int uva_lscm(uva_plan* plan, double* U, const int64_t* pins, double tol, int32_t max_iters, void* stream) {
    if (!valid_plan(plan) || !U) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (!(tol > 0.0)) return fail(UVA_EINVAL, "tol must be positive");
    if (plan->dp.mode_batch) return fail(UVA_EINVAL, "LSCM needs a plan built with refine conventions");
    hipStream_t s = (hipStream_t)stream;
    uva::DevPlan& d = plan->dp;
    // The coarse level of an island's LSCM operator leaves out the plan's own
    // pins; an island given other pins solves with the Jacobi preconditioner.
    auto put_pins = [&](const std::vector<int>& a, const std::vector<int>& b, const std::vector<double>& dist,
                        const std::vector<int64_t>& lco) -> int {
        if (d.n_isl <= 0) return UVA_OK;
        HIP_TRY(hipMemcpyAsync(plan->d_lpin0, a.data(), a.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(plan->d_lpin1, b.data(), b.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(plan->d_ldist, dist.data(), dist.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync((void*)d.isl_lcofs, lco.data(), lco.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                               s));
        HIP_TRY(hipStreamSynchronize(s));
        return UVA_OK;
    };
    if (pins) {
        std::vector<int> a(d.n_isl), b(d.n_isl);
        std::vector<double> dist(d.n_isl);
        std::vector<int64_t> lco(plan->lcofs);
        for (int i = 0; i < d.n_isl; ++i) {
            int nv = (i + 1 < d.n_isl ? plan->v0[i + 1] : d.n_vert) - plan->v0[i];
            if (pins[2 * i] < 0 || pins[2 * i] >= nv || pins[2 * i + 1] < 0 || pins[2 * i + 1] >= nv)
                return fail(UVA_EINVAL, "LSCM pin outside its island");
            const int ca = plan->v0[i] + (int)pins[2 * i], cb = plan->v0[i] + (int)pins[2 * i + 1];
            a[i] = plan->vinv[ca];
            b[i] = plan->vinv[cb];
            const double* p0 = &plan->P[3 * (size_t)ca];
            const double* p1 = &plan->P[3 * (size_t)cb];
            dist[i] = std::sqrt((p1[0] - p0[0]) * (p1[0] - p0[0]) + (p1[1] - p0[1]) * (p1[1] - p0[1]) +
                                (p1[2] - p0[2]) * (p1[2] - p0[2]));
            if (a[i] != plan->lpin0[i] || b[i] != plan->lpin1[i]) lco[i] = -1;
        }
        const int rc = put_pins(a, b, dist, lco);
        if (rc) return rc;
        plan->lpins_custom = true;
    } else {
        for (int i = 0; i < d.n_isl; ++i)
            if (plan->lpin0[i] < 0)
                return fail(UVA_ENOBOUNDARY, "island " + std::to_string(i) +
                                                 " has no boundary (closed surface): needs a cut");
        if (plan->lpins_custom) {
            // back to the plan's own pins
            const int rc = put_pins(plan->lpin0, plan->lpin1, plan->ldist, plan->lcofs);
            if (rc) return rc;
            plan->lpins_custom = false;
        }
    }
    if (!plan->have_lsched) {
        const int rc = plan_lscm(plan, nullptr, s);
        if (rc) return rc;
    }
    uva::TaskArgs a = base_args((double*)plan->ubuf[0]);
    a.cg_tol = tol;
    a.cg_max = max_iters > 0 ? max_iters : kDefaultMaxIters;
    const int nt = plan->lsched.n_teams;
    int rc = reset_queue(plan, 1, s, true, nt);
    if (!rc) rc = launch(plan, uva::TASK_LSCM, UVA_PREC_F64, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 0, U, s, rc), nt);
}

int uva_arap(uva_plan* plan, double* U, const uva_arap_opts* o, void* stream) {
    if (!valid_plan(plan) || !U || !o) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (o->iters < 0 || o->iters_folded < 0) return fail(UVA_EINVAL, "iters must be >= 0");
    if (!(o->cg_tol >= 0.0) || (o->cg_tol == 0.0 && o->cg_max_iters <= 0))
        return fail(UVA_EINVAL, "cg_tol must be positive, or 0 with cg_max_iters > 0");
    if (o->precision < UVA_PREC_MIXED || o->precision > UVA_PREC_F64) return fail(UVA_EINVAL, "bad precision");
    if (o->mode != UVA_MODE_REFINE && o->mode != UVA_MODE_BATCH && o->mode != UVA_MODE_HIER)
        return fail(UVA_EINVAL, "bad mode");
    hipStream_t s = (hipStream_t)stream;
    if (!o->use_graph) return check_fault(plan, s, arap_sequence(plan, U, o, s));
    // graph of the whole sequence, cached per (U, options)
    std::vector<char> key(sizeof(uva_arap_opts) + sizeof(double*));
    std::memcpy(key.data(), o, sizeof(uva_arap_opts));
    std::memcpy(key.data() + sizeof(uva_arap_opts), &U, sizeof(double*));
    if (!plan->gexec || key != plan->gkey) {
        if (plan->gexec) { (void)hipGraphExecDestroy(plan->gexec); plan->gexec = nullptr; }
        if (plan->graph) { (void)hipGraphDestroy(plan->graph); plan->graph = nullptr; }
        if (!plan->cap) HIP_TRY(hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
        int rc = arap_sequence(plan, U, o, plan->cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(plan->cap, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(UVA_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        plan->graph = g;
        HIP_TRY(hipGraphInstantiate(&plan->gexec, g, nullptr, nullptr, 0));
        plan->gkey = key;
    }
    HIP_TRY(hipGraphLaunch(plan->gexec, s));
    return check_fault(plan, s, UVA_OK);
}

int uva_arap_profile(uva_plan* plan, double* U, const uva_arap_opts* o, double* launch_ms, int32_t n_launch,
                     void* stream) {
    if (!valid_plan(plan) || !U || !o || !launch_ms) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (o->iters < 0 || o->iters_folded < 0) return fail(UVA_EINVAL, "iters must be >= 0");
    if (n_launch != 1 + std::max(o->iters, o->iters_folded))
        return fail(UVA_EINVAL, "n_launch must be 1 + max(iters, iters_folded)");
    if (!(o->cg_tol >= 0.0) || (o->cg_tol == 0.0 && o->cg_max_iters <= 0))
        return fail(UVA_EINVAL, "cg_tol must be positive, or 0 with cg_max_iters > 0");
    if (o->precision < UVA_PREC_MIXED || o->precision > UVA_PREC_F64) return fail(UVA_EINVAL, "bad precision");
    if (o->mode != UVA_MODE_REFINE && o->mode != UVA_MODE_BATCH && o->mode != UVA_MODE_HIER)
        return fail(UVA_EINVAL, "bad mode");
    hipStream_t s = (hipStream_t)stream;
    std::vector<hipEvent_t> ev(2 * (size_t)n_launch, nullptr);
    int rc = UVA_OK;
    for (auto& e : ev) {
        if (hipEventCreate(&e) != hipSuccess) { rc = fail(UVA_EHIP, "hipEventCreate"); break; }
    }
    if (rc == UVA_OK) rc = check_fault(plan, s, arap_sequence(plan, U, o, s, &ev));
    if (rc == UVA_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(UVA_EHIP, "hipStreamSynchronize");
    for (int k = 0; rc == UVA_OK && k < n_launch; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) rc = fail(UVA_EHIP, "hipEventElapsedTime");
        launch_ms[k] = ms;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int uva_slim(uva_plan* plan, double* U, const uva_slim_opts* o, void* stream) {
    if (!valid_plan(plan) || !U || !o) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (o->iters < 0) return fail(UVA_EINVAL, "iters must be >= 0");
    if (!(o->cg_tol > 0.0)) return fail(UVA_EINVAL, "cg_tol must be positive");
    if (plan->dp.mode_batch) return fail(UVA_EINVAL, "SLIM needs a plan built with refine conventions");
    if (plan->dp.n_tri == 0) return fail(UVA_EINVAL, "SLIM needs triangles");
    hipStream_t s = (hipStream_t)stream;
    uva::DevPlan& d = plan->dp;
    plan->mem = s;
    if (!d.sw) {
        if (plan->alloc(&d.sw, d.n_tri) != hipSuccess || plan->alloc(&d.sdi, d.n_vert) != hipSuccess)
            return fail(UVA_ENOMEM, "SLIM scratch");
    }
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.cg_tol = o->cg_tol;
    a.cg_max = o->cg_max_iters > 0 ? o->cg_max_iters : kDefaultMaxIters;
    a.tol = o->tol;
    const int total = 1 + o->iters;
    for (int done = 0; done < total && !rc;) {
        const int chunk = std::min(total - done, kMaxSlots);
        rc = reset_queue(plan, chunk, s, done == 0);
        for (int k = 0; k < chunk && !rc; ++k)
            rc = launch(plan, done + k == 0 ? uva::TASK_SLIM_PROLOGUE : uva::TASK_SLIM_ITER, UVA_PREC_F64, a, k, s);
        done += chunk;
    }
    return check_fault(plan, s, from_plan(plan, 0, U, s, rc));
}

int uva_local_rhs(uva_plan* plan, const double* U, double* b, int32_t mode, void* stream) {
    if (!valid_plan(plan) || !U || !b) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.out = plan->ubuf[1];
    a.batch = mode == UVA_MODE_BATCH;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_LOCAL_RHS, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 1, b, s, rc));
}

int uva_cg_solve(uva_plan* plan, const double* b, double* x, double tol, int32_t max_iters, int32_t precision,
                 void* stream) {
    if (!valid_plan(plan) || !b || !x) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (!(tol >= 0.0) || (tol == 0.0 && max_iters <= 0))
        return fail(UVA_EINVAL, "tol must be positive, or 0 with max_iters > 0");
    if (precision < UVA_PREC_MIXED || precision > UVA_PREC_F64) return fail(UVA_EINVAL, "bad precision");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, x, 0, s, rc));
    a.in = to_plan(plan, b, 1, s, rc);
    a.cg_tol = tol;
    a.cg_max = max_iters > 0 ? max_iters : kDefaultMaxIters;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_CG, precision, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 0, x, s, rc));
}

int uva_max_step(uva_plan* plan, const double* U, const double* D, double* tmax, int32_t mode, void* stream) {
    if (!valid_plan(plan) || !U || !D || !tmax) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.in = to_plan(plan, D, 1, s, rc);
    a.out_isl = tmax;
    a.batch = mode == UVA_MODE_BATCH || mode == UVA_MODE_HIER;
    a.hier = mode == UVA_MODE_HIER;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_MAX_STEP, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, rc);
}

int uva_step_apply(uva_plan* plan, double* U, const double* Un, int32_t mode, int32_t line_search, void* stream) {
    if (!valid_plan(plan) || !U || !Un) return fail(UVA_EINVAL, "null argument");
    if (mode != UVA_MODE_BATCH && mode != UVA_MODE_HIER) return fail(UVA_EINVAL, "mode must be BATCH or HIER");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.in = to_plan(plan, Un, 1, s, rc);
    a.batch = 1;
    a.hier = mode == UVA_MODE_HIER;
    a.line_search = line_search != 0;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_STEP_APPLY, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 0, U, s, rc));
}

int uva_energy(uva_plan* plan, const double* U, int32_t kind, double* e, void* stream) {
    if (!valid_plan(plan) || !U || !e) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (kind != 0 && kind != 1) return fail(UVA_EINVAL, "kind must be 0 (ARAP) or 1 (symmetric Dirichlet)");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.out_isl = e;
    a.kind = kind;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_ENERGY, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, rc);
}

int uva_energy_fwd(uva_plan* plan, const double* U, const double* w, int32_t kind, double* e, void* stream) {
    if (!valid_plan(plan) || !U || !e) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (kind != 0 && kind != 1) return fail(UVA_EINVAL, "kind must be 0 (ARAP) or 1 (symmetric Dirichlet)");
    if (plan->dp.mode_batch) return fail(UVA_EINVAL, "the energy gradients need a plan built with refine conventions");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.out_isl = e;
    a.kind = kind;
    a.w = w;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_ENERGY, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, rc);
}

int uva_energy_bwd(uva_plan* plan, const double* U, const double* w, int32_t kind, const double* ge, double* gU,
                   double* gw, void* stream) {
    if (!valid_plan(plan) || !U || !ge || !gU) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (kind != 0 && kind != 1) return fail(UVA_EINVAL, "kind must be 0 (ARAP) or 1 (symmetric Dirichlet)");
    if (plan->dp.mode_batch) return fail(UVA_EINVAL, "the energy gradients need a plan built with refine conventions");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.kind = kind;
    a.w = w;
    a.gin = ge;
    a.out = plan->ubuf[1];
    a.gw = gw;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_ENERGY_BWD, UVA_PREC_F64, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 1, gU, s, rc));
}

int uva_adjoint_cg(uva_plan* plan, const double* g, double* lam, double tol, int32_t max_iters, int32_t precision,
                   void* stream) {
    if (!valid_plan(plan) || !g || !lam) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (!(tol > 0.0)) return fail(UVA_EINVAL, "tol must be positive");
    if (precision < UVA_PREC_MIXED || precision > UVA_PREC_F64) return fail(UVA_EINVAL, "bad precision");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)plan->ubuf[0]);
    a.in = to_plan(plan, g, 1, s, rc);
    a.cg_tol = tol;
    a.cg_max = max_iters > 0 ? max_iters : kDefaultMaxIters;
    a.zero_x = 1;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_CG, precision, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 0, lam, s, rc));
}

int uva_arap_step_bwd(uva_plan* plan, const double* U, const double* lam, double* gU, void* stream) {
    if (!valid_plan(plan) || !U || !lam || !gU) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (plan->dp.mode_batch) return fail(UVA_EINVAL, "the step backward needs a plan built with refine conventions");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.in = to_plan(plan, lam, 1, s, rc);
    a.out = plan->ubuf[2];
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_STEP_BWD, UVA_PREC_F64, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 2, gU, s, rc));
}

int uva_signed_areas(uva_plan* plan, const double* U, double* sa, void* stream) {
    if (!valid_plan(plan) || !U || !sa) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    int rc = UVA_OK;
    const double2* Up = to_plan(plan, U, 0, (hipStream_t)stream, rc);
    if (rc) return rc;
    HIP_TRY(uva::launch_signed_areas(plan->dp, Up, sa, (hipStream_t)stream));
    return UVA_OK;
}

int uva_flips_per_chart(uva_plan* plan, const double* U, double* counts, void* stream) {
    if (!valid_plan(plan) || !U || !counts) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.out_isl = counts;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_FLIPS, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, rc);
}

int uva_repair_jacobi(uva_plan* plan, double* U, int32_t rounds, void* stream) {
    if (!valid_plan(plan) || !U) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    if (rounds < 0) return fail(UVA_EINVAL, "rounds must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    a.rounds = rounds;
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_REPAIR, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 0, U, s, rc));
}

int uva_orient_positive(uva_plan* plan, double* U, void* stream) {
    if (!valid_plan(plan) || !U) return fail(UVA_EINVAL, "null argument");
    plan->mem = (hipStream_t)stream;   // allocations and the final release follow this stream
    hipStream_t s = (hipStream_t)stream;
    int rc = UVA_OK;
    uva::TaskArgs a = base_args((double*)to_plan(plan, U, 0, s, rc));
    if (!rc) rc = reset_queue(plan, 1, s);
    if (!rc) rc = launch(plan, uva::TASK_ORIENT, UVA_PREC_MIXED, a, 0, s);
    return check_fault(plan, s, from_plan(plan, 0, U, s, rc));
}

int uva_chart_regroup(int64_t n, const int64_t* faces, const int64_t* CV, int64_t nv, const int64_t* chart,
                      int64_t nchart, const int64_t* mesh_F, int64_t n_mesh_f, const double* mesh_V, int64_t n_mesh_v,
                      int64_t* keep, int64_t* nk, int64_t* vert_off, int64_t* tri_off, double* P, int64_t* T,
                      int64_t* vorder, int64_t* forder) {
    if (!nk || !keep || !vert_off || !tri_off || (n > 0 && (!faces || !CV || !chart || !mesh_F || !T || !forder)) ||
        (nv > 0 && (!P || !vorder || !mesh_V)))
        return fail(UVA_EINVAL, "null argument");
    std::string err;
    try {
        err = uva::chart_regroup(n, faces, CV, nv, chart, nchart, mesh_F, n_mesh_f, mesh_V, n_mesh_v, keep, nk,
                                 vert_off, tri_off, P, T, vorder, forder);
    } catch (const std::bad_alloc&) {
        return fail(UVA_ENOMEM, "host allocation failed");
    }
    return err.empty() ? UVA_OK : fail(UVA_EINVAL, err);
}

int uva_repair_seq(double* U, int64_t n, const int64_t* T, int64_t t, int32_t rounds) {
    if (!U || (!T && t > 0) || n < 0 || t < 0) return fail(UVA_EINVAL, "bad argument");
    for (int64_t i = 0; i < 3 * t; ++i)
        if (T[i] < 0 || T[i] >= n) return fail(UVA_EINVAL, "triangle vertex out of range");
    try {
        uva::repair_sequential(U, n, T, t, rounds);
    } catch (const std::bad_alloc&) {
        return fail(UVA_ENOMEM, "host allocation failed");
    }
    return UVA_OK;
}

}  // extern "C"
