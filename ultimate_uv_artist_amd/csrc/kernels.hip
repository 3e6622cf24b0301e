// gfx950 kernels of the UV parameterization hot path.
//
// One persistent kernel serves every per-island phase.  Its workgroups pull
// islands from a device-side queue ordered most-expensive-first: large
// islands are solved by the whole workgroup (block barriers only), small ones
// by a single wavefront (no barriers at all), so no island ever needs a
// grid-wide synchronisation and every island carries its own convergence
// state.  Rows are mapped slice-by-slice (64 rows = one wavefront, lane =
// row) so matrix loads are contiguous and every phase touches the rows owned
// by the same lane.
//
// All reductions are in FP64 with a fixed order (butterfly inside a
// wavefront, then wavefront slots in order), so results do not depend on
// which workgroup picked an island.
#include "kernels.hpp"

namespace uva {

// ---------------------------------------------------------------- uniform values
// Every value that steers control flow (task ids, reduction results, island
// flags) goes through readfirstlane: it lands in a scalar register, so the
// compiler emits scalar branches and never splits a wavefront's lanes over
// different paths of a team loop.  (A task id broadcast with a lane shuffle
// is a vector value to the compiler; it restructured the task loop so lanes
// 1-63 re-entered it without lane 0's fresh task id.)
// The per-island code (UVA_HD functions) also compiles for the host, where
// the CPU emulator (emu.inc) runs it with one thread per lane; there a value
// is already uniform and the device clock reads 0.
#define UVA_HD __host__ __device__

// Pointers into LDS carry the LDS address space on the device, so their
// accesses compile to LDS instructions (the CPU emulator has no such space).
#if defined(__HIP_DEVICE_COMPILE__)
#define UVA_LDS __attribute__((address_space(3)))
#else
#define UVA_LDS
#endif

UVA_HD __forceinline__ int uniform_i(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

UVA_HD __forceinline__ double uniform_d(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
#else
    return v;
#endif
}

UVA_HD __forceinline__ double4 uniform_d4(double4 v) {
    return make_double4(uniform_d(v.x), uniform_d(v.y), uniform_d(v.z), uniform_d(v.w));
}

UVA_HD __forceinline__ long long uva_clock() {
#if defined(__HIP_DEVICE_COMPILE__)
    return wall_clock64();
#else
    return 0;
#endif
}

// ---------------------------------------------------------------- reductions
// Cross-lane reductions of a whole wavefront (every lane active), built from
// register-to-register lane moves rather than LDS permutes: DPP quad
// permutations swap lanes 1 and 2 apart, the half-row and row mirrors pair
// the two halves of each 8- and 16-lane row, a swizzle pairs lanes 16 apart,
// and the two 32-lane halves meet through lanes 0 and 32 read as scalars.
// After each stage a lane holds the combination of two equal-sized groups
// whose values are uniform within each group, and a + b == b + a bit for bit,
// so every lane ends with the same value in a fixed order:
//   ((((q0 + q1) + ...) per 4, 8, 16, 32 lanes), then lanes 0..31 + 32..63.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double swizzle_xor16(double v) {
    // ds_swizzle bit mode: and 0x1f, or 0, xor 0x10 (lane l reads l ^ 16 of its 32)
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffLL), 0x401F);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

constexpr int kDppQuad1032 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppQuad2301 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<kDppQuad1032>(v);
    v += dpp_move<kDppQuad2301>(v);
    v += dpp_move<kDppRowHalfMirror>(v);
    v += dpp_move<kDppRowMirror>(v);
    v += swizzle_xor16(v);
    return read_lane(v, 0) + read_lane(v, 32);
}

__device__ __forceinline__ double wave_min(double v) {
    v = fmin(v, dpp_move<kDppQuad1032>(v));
    v = fmin(v, dpp_move<kDppQuad2301>(v));
    v = fmin(v, dpp_move<kDppRowHalfMirror>(v));
    v = fmin(v, dpp_move<kDppRowMirror>(v));
    v = fmin(v, swizzle_xor16(v));
    return fmin(read_lane(v, 0), read_lane(v, 32));
}

// A team solves one island.  Its interface: lane()/wave()/nwaves() map rows
// (wavefront w of the team takes slices lo+w, lo+w+W, ... of its slice range
// [sl_lo, sl_hi)), tid()/tstride() and [tri_lo, tri_hi) map triangles,
// leader() is the one thread that writes island state, sync() makes the
// team's global-memory writes visible to the whole team, and the reductions
// return identical (scalar) values in every thread.
struct WaveTeam {
    char* lv;   // this wavefront's share of the LDS vector buffer (or NULL)
    int lvb;    // its size in bytes
    __device__ WaveTeam(char* l, int b) : lv(l), lvb(b) {}
    __device__ void* lvec() const { return lv; }
    __device__ int lvec_bytes() const { return lvb; }
    __device__ bool multi() const { return false; }
    __device__ int lane() const { return threadIdx.x & 63; }
    __device__ int wave() const { return 0; }
    __device__ int nwaves() const { return 1; }
    __device__ int tid() const { return threadIdx.x & 63; }
    __device__ int tstride() const { return 64; }
    __device__ bool leader() const { return (threadIdx.x & 63) == 0; }
    __device__ int members() const { return 1; }
    __device__ int sl_lo(int) const { return 0; }
    __device__ int sl_hi(int ns) const { return ns; }
    __device__ int tri_lo(int) const { return 0; }
    __device__ int tri_hi(int nt) const { return nt; }
    // LDS writes of the workgroup (here: wavefront) become visible to it.
    __device__ void bsync() const { sync(); }
    // Global-memory writes of this wavefront become visible to its other lanes.
    __device__ void sync() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    __device__ double4 sum4(double4 v) const {
        v.x = wave_sum(v.x); v.y = wave_sum(v.y); v.z = wave_sum(v.z); v.w = wave_sum(v.w);
        sync();
        return uniform_d4(v);
    }
    __device__ double min1(double v) const {
        v = wave_min(v);
        sync();
        return uniform_d(v);
    }
    // sum over the wavefront's lanes (every lane gets it; no barrier)
    __device__ double2 lane_sum2(double2 v) const { return make_double2(wave_sum(v.x), wave_sum(v.y)); }
    __device__ void sum8(double4 a, double4 b, double4& sa, double4& sb) const {
        sa = sum4(a);
        b.x = wave_sum(b.x); b.y = wave_sum(b.y); b.z = wave_sum(b.z); b.w = wave_sum(b.w);
        sb = uniform_d4(b);
    }
};

template <int BS>
struct BlockTeam {
    static constexpr int kWaves = BS / 64;
    static constexpr int kSlots = 8;   // doubles per wavefront slot
    double UVA_LDS* lds;  // 2 * kWaves * kSlots doubles
    int par;
    char* lv;     // LDS vector buffer of the workgroup (or NULL)
    int lvb;
    __device__ BlockTeam(double UVA_LDS* l, char* v, int vb) : lds(l), par(0), lv(v), lvb(vb) {}
    __device__ void* lvec() const { return lv; }
    __device__ int lvec_bytes() const { return lvb; }
    __device__ bool multi() const { return false; }
    __device__ int lane() const { return threadIdx.x & 63; }
    __device__ int wave() const { return threadIdx.x >> 6; }
    __device__ int nwaves() const { return kWaves; }
    __device__ int tid() const { return threadIdx.x; }
    __device__ int tstride() const { return BS; }
    __device__ bool leader() const { return threadIdx.x == 0; }
    __device__ int members() const { return 1; }
    __device__ int sl_lo(int) const { return 0; }
    __device__ int sl_hi(int ns) const { return ns; }
    __device__ int tri_lo(int) const { return 0; }
    __device__ int tri_hi(int nt) const { return nt; }
    __device__ void sync() const { __syncthreads(); }
    __device__ void bsync() const { __syncthreads(); }
    // The two slot buffers alternate; a buffer is rewritten only after the
    // barrier of the following reduction, when every thread has read it.
    __device__ double4 sum4(double4 v) {
        double4 s, unused;
        sum8(v, make_double4(0.0, 0.0, 0.0, 0.0), s, unused);
        return s;
    }
    __device__ void sum8(double4 a, double4 b, double4& sa, double4& sb) {
        a.x = wave_sum(a.x); a.y = wave_sum(a.y); a.z = wave_sum(a.z); a.w = wave_sum(a.w);
        b.x = wave_sum(b.x); b.y = wave_sum(b.y); b.z = wave_sum(b.z); b.w = wave_sum(b.w);
        double UVA_LDS* buf = lds + par * kWaves * kSlots;
        par ^= 1;
        if (lane() == 0) {
            double UVA_LDS* q = buf + kSlots * wave();
            q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w;
            q[4] = b.x; q[5] = b.y; q[6] = b.z; q[7] = b.w;
        }
        __syncthreads();
        double4 s = make_double4(0.0, 0.0, 0.0, 0.0), t = make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const double UVA_LDS* q = buf + kSlots * w;
            s.x += q[0]; s.y += q[1]; s.z += q[2]; s.w += q[3];
            t.x += q[4]; t.y += q[5]; t.z += q[6]; t.w += q[7];
        }
        sa = uniform_d4(s);
        sb = uniform_d4(t);
    }
    __device__ double2 lane_sum2(double2 v) const { return make_double2(wave_sum(v.x), wave_sum(v.y)); }
    __device__ double min1(double v) {
        v = wave_min(v);
        double UVA_LDS* buf = lds + par * kWaves * kSlots;
        par ^= 1;
        if (lane() == 0) buf[kSlots * wave()] = v;
        __syncthreads();
        double m = buf[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) m = fmin(m, buf[kSlots * w]);
        return uniform_d(m);
    }
};

// Wait of one thread at a team barrier: until the arrival count reaches
// target (returns 0), the team's fault flag is set (returns 2: a member gave
// up at an earlier barrier, so the team's results are already invalid and
// waiting again would only stall the launch), or more than spin ticks pass
// (returns 1: the caller sets the flag).  The count and the flag are read
// together each round.  Env supplies count(), fault(), clock() and nap(); the
// CPU emulator runs the same loop with host atomics and a counting clock.
template <class Env>
UVA_HD __forceinline__ int team_wait(Env& e, int target, long long spin) {
    const long long t0 = e.clock();
    for (;;) {
        const int c = e.count();
        const int f = e.fault();
        if (c >= target) return 0;
        if (f) return 2;
        if (e.clock() - t0 > spin) return 1;
        e.nap();
    }
}

// The device side of team_wait: relaxed agent-scope loads of the counter and
// the flag (the caller orders memory around the wait).
struct DevWait {
    const int* cnt;
    const int* flt;
    __device__ int count() const { return __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ int fault() const { return __hip_atomic_load(flt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ long long clock() const { return wall_clock64(); }
    __device__ void nap() const { __builtin_amdgcn_s_sleep(1); }
};

// Hardware id of the XCD this wavefront runs on (HW_REG_XCC_ID).
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((15 << 11) | 20) & 15; }

// m workgroups on one island.  Member k owns a contiguous block of the
// island's slices and triangles (halo traffic only at block edges).  sync()
// is a counter barrier; a reduction combines each member's partial sums,
// written to global memory, in member order after one barrier, so the result
// is identical in every member.  After the barrier the m x 8 partials are
// fetched with one load per thread (all in flight together) into LDS, and
// eight threads sum them member by member.  All members are resident (the grid never
// exceeds the resident capacity) and take their member tasks from the queue
// in order, so every member eventually arrives.
//
// Barrier visibility.  When all members run on one XCD (xcd_local), they
// share its L2: every wavefront waits until its own stores have reached L2
// (vmcnt 0), the counter is an L2 atomic, and after the wait the L1 of the CU
// is invalidated so loads see the other members' data.  Otherwise the counter
// uses agent-scope release/acquire, which writes back and invalidates L2.
// A wait longer than spin_ticks of the device clock gives up and sets *fault
// (a hang guard: the island's result is then invalid and the host reports it);
// once the flag is set, no barrier of the launch waits any more, so every
// workgroup still runs to its end.
template <int BS>
struct TeamWG {
    static constexpr int kWaves = BS / 64;
    static constexpr int kSlots = 8;
    // LDS doubles a team reduction needs: wavefront slots, the members'
    // partials and the eight results
    static constexpr int kLdsDoubles = kWaves * kSlots + kMaxTeam * kSlots + kSlots;
    double UVA_LDS* lds;        // kLdsDoubles doubles
    int* cnt;           // arrival counter of this team (zero at launch)
    double* part;       // 2 * kMaxTeam * kSlots partial sums
    int m, mem;
    int rounds;         // barriers passed
    int nred;           // reductions done (selects the partial buffer)
    int local;          // members share one XCD
    long long spin;
    int* fault;
    char* lv;           // LDS vector buffer of this member (or NULL)
    int lvb;
    __device__ TeamWG(double UVA_LDS* l, int* c, double* pt, int m_, int k_, int local_, long long spin_, int* fault_,
                      char* v, int vb)
        : lds(l), cnt(c), part(pt), m(m_), mem(k_), rounds(0), nred(0), local(local_), spin(spin_),
          fault(fault_), lv(v), lvb(vb) {}
    __device__ void* lvec() const { return lv; }
    __device__ int lvec_bytes() const { return lvb; }
    __device__ bool multi() const { return true; }
    __device__ int lane() const { return threadIdx.x & 63; }
    __device__ int wave() const { return threadIdx.x >> 6; }
    __device__ int nwaves() const { return kWaves; }
    __device__ int tid() const { return threadIdx.x; }
    __device__ int tstride() const { return BS; }
    __device__ bool leader() const { return mem == 0 && threadIdx.x == 0; }
    __device__ int members() const { return m; }
    __device__ int sl_lo(int ns) const { return (int)((long long)ns * mem / m); }
    __device__ int sl_hi(int ns) const { return (int)((long long)ns * (mem + 1) / m); }
    __device__ int tri_lo(int nt) const { return (int)((long long)nt * mem / m); }
    __device__ int tri_hi(int nt) const { return (int)((long long)nt * (mem + 1) / m); }
    // LDS writes of this member become visible to the member's threads.
    __device__ void bsync() const { __syncthreads(); }
    __device__ void sync() {
        if (local) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        ++rounds;
        if (threadIdx.x == 0) {
            const int target = rounds * m;
            DevWait w{cnt, fault};
            int r;
            if (local) {
                __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                r = team_wait(w, target, spin);
                asm volatile("buffer_inv sc0\n\ts_waitcnt vmcnt(0)" ::: "memory");
            } else {
                __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                r = team_wait(w, target, spin);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            if (r == 1) atomicOr(fault, 1);
        }
        __syncthreads();
    }
    __device__ void sum8(double4 a, double4 b, double4& sa, double4& sb) {
        a.x = wave_sum(a.x); a.y = wave_sum(a.y); a.z = wave_sum(a.z); a.w = wave_sum(a.w);
        b.x = wave_sum(b.x); b.y = wave_sum(b.y); b.z = wave_sum(b.z); b.w = wave_sum(b.w);
        if (lane() == 0) {
            double UVA_LDS* q = lds + kSlots * wave();
            q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w;
            q[4] = b.x; q[5] = b.y; q[6] = b.z; q[7] = b.w;
        }
        __syncthreads();
        double* pb = part + (nred & 1) * kMaxTeam * kSlots;
        ++nred;
        if (threadIdx.x < kSlots) {
            double v = 0.0;
            for (int w = 0; w < kWaves; ++w) v += lds[kSlots * w + threadIdx.x];
            pb[kSlots * mem + threadIdx.x] = v;
        }
        sync();
        double UVA_LDS* mp = lds + kWaves * kSlots;          // members' partials
        double UVA_LDS* res = mp + kMaxTeam * kSlots;        // the eight sums
        if (threadIdx.x < m * kSlots) mp[threadIdx.x] = pb[threadIdx.x];
        __syncthreads();
        if (threadIdx.x < kSlots) {
            double v = 0.0;
            for (int k = 0; k < m; ++k) v += mp[kSlots * k + threadIdx.x];
            res[threadIdx.x] = v;
        }
        __syncthreads();
        sa = uniform_d4(make_double4(res[0], res[1], res[2], res[3]));
        sb = uniform_d4(make_double4(res[4], res[5], res[6], res[7]));
    }
    __device__ double4 sum4(double4 v) {
        double4 s, unused;
        sum8(v, make_double4(0.0, 0.0, 0.0, 0.0), s, unused);
        return s;
    }
    __device__ double2 lane_sum2(double2 v) const { return make_double2(wave_sum(v.x), wave_sum(v.y)); }
    __device__ double min1(double v) {
        v = wave_min(v);
        if (lane() == 0) lds[kSlots * wave()] = v;
        __syncthreads();
        double* pb = part + (nred & 1) * kMaxTeam * kSlots;
        ++nred;
        if (threadIdx.x == 0) {
            double mn = lds[0];
            for (int w = 1; w < kWaves; ++w) mn = fmin(mn, lds[kSlots * w]);
            pb[kSlots * mem] = mn;
        }
        sync();
        double UVA_LDS* mp = lds + kWaves * kSlots;
        double UVA_LDS* res = mp + kMaxTeam * kSlots;
        if (threadIdx.x < m) mp[threadIdx.x] = pb[kSlots * threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0) {
            double mn = mp[0];
            for (int k = 1; k < m; ++k) mn = fmin(mn, mp[k]);
            res[0] = mn;
        }
        __syncthreads();
        return uniform_d(res[0]);
    }
};

// ---------------------------------------------------------------- island view
struct Isl {
    int v0, nv, t0, nt, s0, ns, pin, bw;
    int cg;               // slices per coarse aggregate
    int lcg;              // the same for the LSCM operator's coarse level
    long long cofs;       // offset of the inverse coarse operator, -1 without one
    long long lcofs;      // the same for the LSCM operator
    double wpin;
};

UVA_HD __forceinline__ long long uniform_ll(long long v) {
    // the halves are joined as unsigned bits (a negative v's high half shifted
    // as a signed value would be undefined)
    const unsigned long long hi = (unsigned)uniform_i((int)(v >> 32));
    const unsigned long long lo = (unsigned)uniform_i((int)(v & 0xffffffffLL));
    return (long long)(hi << 32 | lo);
}

UVA_HD __forceinline__ Isl load_isl(const DevPlan& P, int i) {
    Isl d;
    d.v0 = uniform_i(P.isl_v0[i]); d.nv = uniform_i(P.isl_nv[i]);
    d.t0 = uniform_i(P.isl_t0[i]); d.nt = uniform_i(P.isl_nt[i]);
    d.s0 = uniform_i(P.isl_s0[i]); d.ns = uniform_i(P.isl_ns[i]);
    d.pin = uniform_i(P.isl_pin[i]); d.wpin = uniform_d(P.isl_wpin[i]);
    d.bw = uniform_i(P.isl_bw[i]);
    d.cg = uniform_i(P.isl_cg[i]);
    d.lcg = P.isl_lcg ? uniform_i(P.isl_lcg[i]) : d.cg;
    d.cofs = uniform_ll(P.isl_cofs[i]);
    d.lcofs = P.isl_lcofs ? uniform_ll(P.isl_lcofs[i]) : -1;
    return d;
}

// Iterate the rows owned by this lane: wavefront w of the team takes slices
// lo+w, lo+w+W, ... of its slice range; lane l takes row l of the slice.
#define FOR_ROWS(tm, d, row, sg)                                                                   \
    for (int _s = (tm).sl_lo((d).ns) + (tm).wave(), _se = (tm).sl_hi((d).ns); _s < _se;            \
         _s += (tm).nwaves())                                                                      \
        for (int row = (d).v0 + _s * kSlice + (tm).lane(), sg = (d).s0 + _s, _once = 1;             \
             _once && (_s * kSlice + (tm).lane() < (d).nv); _once = 0)

// Iterate the triangles owned by this thread.
#define FOR_TRIS(tm, d, f)                                                                         \
    for (int f = (d).t0 + (tm).tri_lo((d).nt) + (tm).tid(), _fe = (d).t0 + (tm).tri_hi((d).nt);   \
         f < _fe; f += (tm).tstride())

template <class T>
struct Vec2;
template <>
struct Vec2<float> { using type = float2; };
template <>
struct Vec2<double> { using type = double2; };

// ---------------------------------------------------------------- sparse rows
enum { OP_ARAP = 0, OP_LSCM = 1 };

constexpr int kGather = 8;   // matrix entries whose gathers are issued together

// y = A x for one row: FP32 values (hi) for float vectors, hi + lo for double.
// (o is an entry offset, always within the matrix: unsigned, so the load
// takes a scalar base and one 32-bit offset register)
template <typename V>
UVA_HD __forceinline__ V mval(const DevPlan& P, unsigned o);
template <>
UVA_HD __forceinline__ float mval<float>(const DevPlan& P, unsigned o) { return P.m_hi[o]; }
template <>
UVA_HD __forceinline__ double mval<double>(const DevPlan& P, unsigned o) {
    return (double)P.m_hi[o] + (double)P.m_lo[o];
}

template <typename V, int OP>
UVA_HD __forceinline__ typename Vec2<V>::type spmv_row(const DevPlan& P, int row, int sg, int lane,
                                                           const typename Vec2<V>::type* __restrict__ x, V dg) {
    const int base = P.sl_ptr[sg];
    const int width = (P.sl_ptr[sg + 1] - base) >> 6;
    typename Vec2<V>::type xr = x[row];
    V yx = dg * xr.x, yy = dg * xr.y;
    int o = base + lane;
    // Entries go in groups of kGather: all column indices and values of a
    // group are loaded first, then all gathers are issued, so a row costs two
    // dependent memory round trips per group instead of two per entry.
    for (int k0 = 0; k0 < width; k0 += kGather, o += kGather * kSlice) {
        int c[kGather];
        V v[kGather];
        V bv[kGather];
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            const bool in = k0 + u < width;
            c[u] = in ? P.m_col[o + u * kSlice] : row;
            v[u] = in ? mval<V>(P, o + u * kSlice) : (V)0;
            bv[u] = (OP == OP_LSCM && in) ? (V)P.m_b[o + u * kSlice] : (V)0;
        }
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            const bool in = k0 + u < width;          // padding entries are left out by select
            const typename Vec2<V>::type xc = x[c[u]];
            if (OP == OP_ARAP) {
                yx = in ? fma(v[u], xc.x, yx) : yx;
                yy = in ? fma(v[u], xc.y, yy) : yy;
            } else {
                yx = in ? fma(v[u], xc.x, fma(bv[u], xc.y, yx)) : yx;
                yy = in ? fma(v[u], xc.y, -fma(bv[u], xc.x, -yy)) : yy;
            }
        }
    }
    typename Vec2<V>::type y;
    y.x = yx;
    y.y = yy;
    return y;
}

// Soft-pin weight of a row: the island's pin, or the per-row weights that
// replace it when the plan's islands are pinned chart by chart.
UVA_HD __forceinline__ double pin_weight(const DevPlan& P, const Isl& d, int row) {
    if (P.rpin) return P.rpin[row];
    return row == d.pin ? d.wpin : 0.0;
}

template <int OP>
UVA_HD __forceinline__ double diag_of(const DevPlan& P, const Isl& d, int row) {
    double g = P.diag[row];
    if (OP == OP_ARAP) g += pin_weight(P, d, row);
    return g;
}

template <int OP>
UVA_HD __forceinline__ double dinv_of(const DevPlan& P, const Isl& d, int row) {
    if (OP == OP_ARAP) return (double)P.dinv32[row];
    double g = P.diag[row];
    return g > 1e-30 ? 1.0 / g : 1.0;
}

// ---------------------------------------------------------------- precision
template <int PREC>
struct Prec;
template <>
struct Prec<UVA_PREC_MIXED> {
    using V = float;
    static constexpr int kMaxOuter = 8;
    static constexpr double kInnerFrac2 = 0.25;  // inner target (0.5 tol)^2 leaves room for the true residual
    static constexpr bool kResid64 = true;
    static constexpr bool kRoundX = false;
};
template <>
struct Prec<UVA_PREC_F32> {
    using V = float;
    static constexpr int kMaxOuter = 1;
    static constexpr double kInnerFrac2 = 1.0;
    static constexpr bool kResid64 = false;
    static constexpr bool kRoundX = true;
};
template <>
struct Prec<UVA_PREC_F64> {
    using V = double;
    static constexpr int kMaxOuter = 8;
    static constexpr double kInnerFrac2 = 1.0;
    static constexpr bool kResid64 = true;
    static constexpr bool kRoundX = false;
};

// FP32-only accumulation for the all-FP32 variant, FP64 otherwise.
template <int PREC>
UVA_HD __forceinline__ double acc_mul(float a, float b) {
    if (PREC == UVA_PREC_F32) return (double)(a * b);
    return (double)a * (double)b;
}
template <int PREC>
UVA_HD __forceinline__ double acc_mul(double a, double b) { return a * b; }

// CG step length alpha = gamma / (p, A p).  The curvature (p, A p) may be
// negative: the reference clips cotangent weights to +-1e4, and a clipped
// triangle's quadratic form can be indefinite, which makes the whole operator
// indefinite (a needle island with one clipped weight has an eigenvalue near
// -38).  CG still converges on such systems as long as it keeps the signed
// step, as the direct solve the reference uses would.  A curvature that the
// vectors' precision cannot resolve (|den| below that precision times the
// uncorrected term nd) stops the step instead: alpha = 0.
template <int PREC>
UVA_HD __forceinline__ double cg_alpha(double g, double den, double nd) {
    constexpr double kRes = sizeof(typename Prec<PREC>::V) == 4 ? 1e-6 : 1e-13;
    if (!(fabs(den) > kRes * fabs(nd)) || !__builtin_isfinite(den)) return 0.0;
    return g / den;
}
template <int PREC>
UVA_HD __forceinline__ double cg_alpha(double g, double den) {
    return cg_alpha<PREC>(g, den, den);
}

#include "cg_island.hpp"
#include "cg_resident.hpp"

// ---------------------------------------------------------------- per-triangle pieces
struct Jac {
    double j00, j01, j10, j11;
};

UVA_HD __forceinline__ Jac jacobian(const DevPlan& P, int f, const double2* __restrict__ U) {
    const double2 ua = U[P.tv0[f]], ub = U[P.tv1[f]], uc = U[P.tv2[f]];
    const double x1 = P.fx1[f], x2 = P.fx2[f], y2 = P.fy2[f];
    Jac J;
    J.j00 = (ub.x - ua.x) / x1;
    J.j10 = (ub.y - ua.y) / x1;
    J.j01 = ((uc.x - ua.x) - J.j00 * x2) / y2;
    J.j11 = ((uc.y - ua.y) - J.j10 * x2) / y2;
    return J;
}

// ARAP energy of one triangle (param.py:306-311): signed singular values q +- r.
UVA_HD __forceinline__ double arap_energy_tri(const Jac& J, double area) {
    const double qe = 0.5 * (J.j00 + J.j11), qh = 0.5 * (J.j10 - J.j01);
    const double rf = 0.5 * (J.j00 - J.j11), rg = 0.5 * (J.j10 + J.j01);
    const double q = sqrt(qe * qe + qh * qh), r = sqrt(rf * rf + rg * rg);
    const double a = q + r - 1.0, b = q - r - 1.0;
    return area * (a * a + b * b);
}

// Symmetric Dirichlet energy of one triangle (param.py:313-318).
UVA_HD __forceinline__ double sd_energy_tri(const Jac& J, double area) {
    const double dj = J.j00 * J.j11 - J.j01 * J.j10;
    if (!(dj > 0.0)) return __builtin_huge_val();  // +inf (also for NaN)
    const double fro = J.j00 * J.j00 + J.j01 * J.j01 + J.j10 * J.j10 + J.j11 * J.j11;
    return area * (fro + fro / (dj * dj) - 4.0);
}

// Local step: per-triangle target rotation (cos, sin) into P.R, and the ARAP
// energy of U.  refine: normalised conformal part of J, identity when the
// frame is degenerate or |.| < 1e-300 (param.py:214-233).  batch: the same
// rotation (= angle of atan2) for thick finite triangles, identity otherwise,
// times the triangle's target scale (gpu.py:166-170).
template <class Team>
UVA_HD __forceinline__ double local_step(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U, bool batch,
                             bool want_energy) {
    double en = 0.0;
    FOR_TRIS(tm, d, f) {
        const bool ok = P.fok[f];
        double ca = 1.0, cb = 0.0;
        if (ok || batch) {
            const Jac J = jacobian(P, f, U);
            const double A = J.j00 + J.j11, B = J.j10 - J.j01;
            const double h = sqrt(A * A + B * B);
            if (batch) {
                if (ok && __builtin_isfinite(h) && h > 0.0) { ca = A / h; cb = B / h; }
                const double sc = P.fscale[f];
                ca *= sc;
                cb *= sc;
            } else if (!(h < 1e-300)) {
                ca = A / h;
                cb = B / h;
            }
            if (want_energy && ok) en += arap_energy_tri(J, P.farea[f]);
        }
        P.R[f] = make_double2(ca, cb);
    }
    if (!want_energy) { tm.sync(); return 0.0; }
    return tm.sum4(make_double4(en, 0.0, 0.0, 0.0)).x;
}

// rhs: b[v] = sum over incident corners of R(tri) g (param.py:457-468, gpu.py:171-181)
// plus the soft-pin term; X64 = U (warm start).  Returns |b|^2 per column.
template <class Team>
UVA_HD __forceinline__ double2 rhs_step(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U) {
    double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
    FOR_ROWS(tm, d, row, sg) {
        (void)sg;
        double bx = 0.0, by = 0.0;
        for (int q = P.inc_ptr[row]; q < P.inc_ptr[row + 1]; ++q) {
            const double2 R = P.R[P.inc_tri[q]];
            const double2 g = P.inc_g[q];
            bx += R.x * g.x - R.y * g.y;
            by += R.y * g.x + R.x * g.y;
        }
        const double2 u = U[row];
        const double pw = pin_weight(P, d, row);
        if (pw != 0.0) {
            bx += pw * u.x;
            by += pw * u.y;
        }
        P.b64[row] = make_double2(bx, by);
        P.X64[row] = u;
        acc.x += bx * bx;
        acc.y += by * by;
    }
    double4 s = tm.sum4(acc);
    return make_double2(s.x, s.y);
}

// Flip-preventing step along D = X64 - U (param.py:259-291 for refine;
// gpu.py:204-227 for batch), minimum over the island's triangles.
template <class Team>
UVA_HD __forceinline__ double max_step_island(Team& tm, const DevPlan& P, const Isl& d, const double2* __restrict__ U,
                                  const double2* __restrict__ Xn, bool batch, bool usable_only = false) {
    double best = 1e30;
    FOR_TRIS(tm, d, f) {
        const int a = P.tv0[f], b = P.tv1[f], c = P.tv2[f];
        const double2 ua = U[a], ub = U[b], uc = U[c];
        const double2 xa = Xn[a], xb = Xn[b], xc = Xn[c];
        const double ux1 = ub.x - ua.x, uy1 = ub.y - ua.y, ux2 = uc.x - ua.x, uy2 = uc.y - ua.y;
        const double dx1 = (xb.x - ub.x) - (xa.x - ua.x), dy1 = (xb.y - ub.y) - (xa.y - ua.y);
        const double dx2 = (xc.x - uc.x) - (xa.x - ua.x), dy2 = (xc.y - uc.y) - (xa.y - ua.y);
        const double C = ux1 * uy2 - ux2 * uy1;
        const double B = ux1 * dy2 + dx1 * uy2 - ux2 * dy1 - dx2 * uy1;
        const double A = dx1 * dy2 - dx2 * dy1;
        if (!batch) {
            if (C <= 0.0) continue;
            double bm = fabs(B);
            if (!(bm >= 1e-300)) bm = (bm != bm) ? bm : 1e-300;  // Python max(|B|, 1e-300), NaN kept
            if (A == 0.0 || fabs(A) < 1e-30 * bm) {
                if (B < 0.0) {
                    const double rt = -C / B;
                    if (rt < best) best = rt;
                }
                continue;
            }
            const double disc = B * B - 4.0 * A * C;
            if (disc < 0.0) continue;
            const double sq = sqrt(disc);
            const double r1 = (-B - sq) / (2.0 * A), r2 = (-B + sq) / (2.0 * A);
            if (r1 > 0.0 && r1 < best) best = r1;
            if (r2 > 0.0 && r2 < best) best = r2;
        } else {
            const double big = 1e30;
            const double lin = (fabs(A) < 1e-30 && B < 0.0) ? -C / B : big;
            const double disc = B * B - 4.0 * A * C;
            const double sq = sqrt(fmax(disc, 0.0));
            const double A2 = fabs(A) < 1e-30 ? 1.0 : 2.0 * A;
            const double r1 = (-B - sq) / A2, r2 = (-B + sq) / A2;
            double quad = fmin(r1 > 0.0 ? r1 : big, r2 > 0.0 ? r2 : big);
            quad = (disc >= 0.0 && fabs(A) >= 1e-30) ? quad : big;
            // hier.py:494 also leaves out triangles without a usable frame
            const double t = (C > 0.0 && (!usable_only || P.fok[f])) ? fmin(lin, quad) : big;
            best = fmin(best, t);
        }
    }
    return tm.min1(best);
}

UVA_HD __forceinline__ double signed_area(const DevPlan& P, int f, const double2* __restrict__ U) {
    const double2 a = U[P.tv0[f]], b = U[P.tv1[f]], c = U[P.tv2[f]];
    return 0.5 * ((b.x - a.x) * (c.y - a.y) - (c.x - a.x) * (b.y - a.y));
}

#include "slim_island.hpp"
#include "diff_island.hpp"

// ---------------------------------------------------------------- island tasks
template <int KIND, int PREC, class Team>
UVA_HD __forceinline__ void run_task(Team& tm, const DevPlan& P, int isl, const TaskArgs& a) {
    IslandState* st = P.st + isl;
    const Isl d = load_isl(P, isl);
    double2* U = a.U;
    if (KIND == TASK_PROLOGUE) {
        // flips of the start, line-search decision, rotations and energy of U
        double nflip = 0.0;
        FOR_TRIS(tm, d, f) nflip += signed_area(P, f, U) <= 0.0 ? 1.0 : 0.0;
        const double fl = tm.sum4(make_double4(nflip, 0.0, 0.0, 0.0)).x;
        const double en = local_step(tm, P, d, U, a.batch, !a.batch);
        if (tm.leader()) {
            st->flips0 = (int)fl;
            st->use_ls = a.line_search && (a.batch || fl == 0.0);
            st->iters_cap = fl > 0.0 ? a.iters_folded : a.iters_plain;
            st->e_prev = en;
            st->done = 0;
            st->cg_iters = 0;
            st->cg_iters_max = 0;
            st->arap_iters = 0;
            st->converged = 0;
            st->rel_resid = 0.0;
            st->step = 0.0;
            st->tmax = 1e30;
            st->status = 0;
            st->ticks = 0;
        }
    } else if (KIND == TASK_ARAP_ITER) {
        if (uniform_i(st->done)) return;
        const int cap = uniform_i(st->iters_cap);
        if (cap > 0 && uniform_i(st->arap_iters) >= cap) return;
        const long long t_start = uva_clock();
        const double2 b2 = rhs_step(tm, P, d, U);
        // Batch and hier solves cap the CG iterations (40 and 20 by default), so
        // their result depends on the preconditioner: they keep the
        // reference's Jacobi one (gpu.py:183-201, hier.py:458-467)
        const CGOut cg = cg_arap<PREC>(tm, P, d, b2, a.cg_tol, a.cg_max, !a.batch);
        if (a.batch) {
            // islands whose solve went non-finite keep their layout (gpu.py:235-240)
            double bad = 0.0;
            FOR_ROWS(tm, d, row, sg) {
                (void)sg;
                const double2 x = P.X64[row];
                bad += (__builtin_isfinite(x.x) && __builtin_isfinite(x.y)) ? 0.0 : 1.0;
            }
            if (tm.sum4(make_double4(bad, 0.0, 0.0, 0.0)).x > 0.0) {
                FOR_ROWS(tm, d, row, sg) {
                    (void)sg;
                    P.X64[row] = U[row];
                }
                tm.sync();
            }
        }
        int use_ls = uniform_i(st->use_ls);
        if (a.hier && use_ls) {
            // hier.py:500-505: a chart with a folded usable triangle takes the
            // plain ARAP step this iteration
            double nf = 0.0;
            FOR_TRIS(tm, d, f) nf += (P.fok[f] && signed_area(P, f, U) < 0.0) ? 1.0 : 0.0;
            if (tm.sum4(make_double4(nf, 0.0, 0.0, 0.0)).x > 0.0) use_ls = 0;
        }
        double step = 1.0, tmax = 1e30;
        if (use_ls) {
            tmax = max_step_island(tm, P, d, U, P.X64, a.batch, a.hier != 0);
            step = fmin(1.0, 0.9 * tmax);
        }
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 x = P.X64[row];
            if (use_ls) {
                const double2 u = U[row];
                U[row] = make_double2(u.x + step * (x.x - u.x), u.y + step * (x.y - u.y));
            } else {
                U[row] = x;
            }
        }
        tm.sync();
        const double en = local_step(tm, P, d, U, a.batch, !a.batch);
        if (tm.leader()) {
            st->cg_iters += cg.iters;
            st->cg_iters_max = max(st->cg_iters_max, cg.iters);
            st->arap_iters += 1;
            st->converged = cg.conv;
            st->rel_resid = cg.rel;
            st->step = step;
            st->tmax = tmax;
            st->ticks += uva_clock() - t_start;
            if (!a.batch) {
                const double prev = st->e_prev;
                if (fabs(prev - en) <= a.tol * fmax(prev, 1e-300)) st->done = 1;
                st->e_prev = en;
            }
        }
    } else if (KIND == TASK_LSCM) {
        const long long t_start = uva_clock();
        const int p0 = uniform_i(P.isl_lpin0[isl]), p1 = uniform_i(P.isl_lpin1[isl]);
        if (p0 < 0) {
            if (tm.leader()) st->status = -5;
            return;
        }
        const double dist = uniform_d(P.isl_ldist[isl]);
        // rhs = -H x_pin on the free rows (param.py:153-165)
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            P.X64[row] = make_double2(row == p1 ? dist : 0.0, 0.0);
        }
        tm.sync();
        double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
        FOR_ROWS(tm, d, row, sg) {
            double2 y = spmv_row<double, OP_LSCM>(P, row, sg, tm.lane(), P.X64, P.diag[row]);
            double2 bb = (row == p0 || row == p1) ? make_double2(0.0, 0.0) : make_double2(-y.x, -y.y);
            P.b64[row] = bb;
            acc.x += bb.x * bb.x;
            acc.y += bb.y * bb.y;
        }
        const double4 s = tm.sum4(acc);
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            P.X64[row] = make_double2(0.0, 0.0);
        }
        tm.sync();
        const CGOut cg = cg_solve<PREC, OP_LSCM>(tm, P, d, make_double2(s.x, s.y), a.cg_tol, a.cg_max, true, p0, p1);
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            double2 x = P.X64[row];
            if (row == p0) x = make_double2(0.0, 0.0);
            if (row == p1) x = make_double2(dist, 0.0);
            U[row] = x;
        }
        if (tm.leader()) {
            st->cg_iters = cg.iters;
            st->cg_iters_max = cg.iters;
            st->converged = cg.conv;
            st->rel_resid = cg.rel;
            st->status = 0;
            st->ticks = uva_clock() - t_start;
        }
        tm.sync();
    } else if (KIND == TASK_LOCAL_RHS) {
        local_step(tm, P, d, U, a.batch, false);
        const double2 b2 = rhs_step(tm, P, d, U);
        (void)b2;
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            a.out[row] = P.b64[row];
        }
    } else if (KIND == TASK_CG) {
        // b in a.in, warm start (or zero start) and result in U
        double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 bb = a.in[row];
            P.b64[row] = bb;
            P.X64[row] = a.zero_x ? make_double2(0.0, 0.0) : U[row];
            acc.x += bb.x * bb.x;
            acc.y += bb.y * bb.y;
        }
        const double4 s = tm.sum4(acc);
        const CGOut cg = cg_arap<PREC>(tm, P, d, make_double2(s.x, s.y), a.cg_tol, a.cg_max, true);
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            U[row] = P.X64[row];
        }
        if (tm.leader()) {
            st->cg_iters = cg.iters;
            st->cg_iters_max = cg.iters;
            st->converged = cg.conv;
            st->rel_resid = cg.rel;
        }
        tm.sync();
    } else if (KIND == TASK_MAX_STEP) {
        // D in a.in: step limit along D from U; X64 = U + D
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 u = U[row], dd = a.in[row];
            P.X64[row] = make_double2(u.x + dd.x, u.y + dd.y);
        }
        tm.sync();
        const double t = max_step_island(tm, P, d, U, P.X64, a.batch, a.hier != 0);
        if (tm.leader()) a.out_isl[isl] = t;
    } else if (KIND == TASK_STEP_APPLY) {
        // The rest of one Batch / hier ARAP iteration after its coupled
        // global step (gpu.py:235-247, hier.py:498-509): a.in holds the
        // step's result Un.  An island with a non-finite row keeps its
        // layout; with line search (hier: unless a usable triangle of U is
        // folded) the step is min(1, 0.9 x the flip-safe limit along Un - U),
        // else 1; then U += step (Un - U).
        double bad = 0.0;
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 x = a.in[row];
            bad += (__builtin_isfinite(x.x) && __builtin_isfinite(x.y)) ? 0.0 : 1.0;
            P.X64[row] = x;
        }
        if (tm.sum4(make_double4(bad, 0.0, 0.0, 0.0)).x > 0.0) {
            FOR_ROWS(tm, d, row, sg) {
                (void)sg;
                P.X64[row] = U[row];
            }
            tm.sync();
        }
        int use_ls = a.line_search;
        if (a.hier && use_ls) {
            double nf = 0.0;
            FOR_TRIS(tm, d, f) nf += (P.fok[f] && signed_area(P, f, U) < 0.0) ? 1.0 : 0.0;
            if (tm.sum4(make_double4(nf, 0.0, 0.0, 0.0)).x > 0.0) use_ls = 0;
        }
        double step = 1.0, tmax = 1e30;
        if (use_ls) {
            tmax = max_step_island(tm, P, d, U, P.X64, true, a.hier != 0);
            step = fmin(1.0, 0.9 * tmax);
        }
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            const double2 u = U[row], x = P.X64[row];
            U[row] = make_double2(u.x + (x.x - u.x) * step, u.y + (x.y - u.y) * step);
        }
        if (tm.leader()) {
            st->step = step;
            st->tmax = tmax;
        }
        tm.sync();
    } else if (KIND == TASK_ENERGY) {
        const double s = energy_fwd_island(tm, P, d, U, a.w, a.kind);
        if (tm.leader()) a.out_isl[isl] = s;
    } else if (KIND == TASK_ENERGY_BWD) {
        energy_bwd_island(tm, P, d, U, a.w, a.kind, uniform_d(a.gin[isl]), a.out, a.gw);
    } else if (KIND == TASK_STEP_BWD) {
        step_bwd_island(tm, P, d, U, a.in, a.out);
    } else if (KIND == TASK_REPAIR) {
        // Jacobi 1-ring relaxation of the free vertices of folded thick triangles
        // (gpu.py:251-273); the batch-wide "no fold left" stop is per island here,
        // which gives the same layout.
        int* mark = (int*)P.vr;
        FOR_ROWS(tm, d, row, sg) {
            (void)sg;
            mark[row] = 0;
        }
        tm.sync();
        for (int round = 0; round < a.rounds; ++round) {
            double nbad = 0.0;
            FOR_TRIS(tm, d, f) {
                if (signed_area(P, f, U) < 0.0 && P.fok[f]) {
                    nbad += 1.0;
                    mark[P.tv0[f]] = 1;
                    mark[P.tv1[f]] = 1;
                    mark[P.tv2[f]] = 1;
                }
            }
            if (tm.sum4(make_double4(nbad, 0.0, 0.0, 0.0)).x == 0.0) break;
            FOR_ROWS(tm, d, row, sg) {
                (void)sg;
                double2 u = U[row];
                if (mark[row] && !P.vbnd[row]) {
                    double sx = 0.0, sy = 0.0;
                    const int q0 = P.adj_ptr[row], q1 = P.adj_ptr[row + 1];
                    for (int q = q0; q < q1; ++q) {
                        const double2 w = U[P.adj_idx[q]];
                        sx += w.x;
                        sy += w.y;
                    }
                    const double dg = fmax((double)(q1 - q0), 1.0);
                    u = make_double2(sx / dg, sy / dg);
                }
                P.X64[row] = u;
                mark[row] = 0;
            }
            tm.sync();
            FOR_ROWS(tm, d, row, sg) {
                (void)sg;
                U[row] = P.X64[row];
            }
            tm.sync();
        }
    } else if (KIND == TASK_ORIENT) {
        double s = 0.0;
        FOR_TRIS(tm, d, f) {
            const double sa = signed_area(P, f, U);
            const double sg = sa > 0.0 ? 1.0 : (sa < 0.0 ? -1.0 : 0.0);
            s += sg * P.farea[f];
        }
        const double tot = tm.sum4(make_double4(s, 0.0, 0.0, 0.0)).x;
        if (tot < 0.0) {
            FOR_ROWS(tm, d, row, sg) {
                (void)sg;
                double2 u = U[row];
                U[row] = make_double2(u.x, -u.y);
            }
        }
        tm.sync();
    } else if (KIND == TASK_FLIPS) {
        double c = 0.0;
        FOR_TRIS(tm, d, f)
            c += (signed_area(P, f, U) < 0.0 && P.fok[f]) ? 1.0 : 0.0;
        const double s = tm.sum4(make_double4(c, 0.0, 0.0, 0.0)).x;
        if (tm.leader()) a.out_isl[isl] = s;
    } else if (KIND == TASK_SLIM_PROLOGUE) {
        // param.py:418-421: symmetric Dirichlet energy of the start (a
        // non-finite start leaves U as it is), regularisation 1e-12 * area
        double en = 0.0, ar = 0.0;
        FOR_TRIS(tm, d, f) {
            ar += P.farea[f];
            if (P.fok[f]) en += sd_energy_tri(jacobian(P, f, U), P.farea[f]);
        }
        const double4 s = tm.sum4(make_double4(en, ar, 0.0, 0.0));
        if (tm.leader()) {
            st->e_prev = s.x;
            st->reg = 1e-6 * s.y * 1e-6;
            st->done = __builtin_isfinite(s.x) ? 0 : 1;
            st->cg_iters = 0;
            st->cg_iters_max = 0;
            st->arap_iters = 0;
            st->converged = 1;
            st->rel_resid = 0.0;
            st->step = 0.0;
            st->tmax = 1e30;
            st->status = 0;
            st->flips0 = 0;
            st->use_ls = 1;
            st->ticks = 0;
        }
        tm.sync();
    } else if (KIND == TASK_SLIM_ITER) {
        // one iteration of param.py:423-453
        if (uniform_i(st->done)) return;
        const long long t_start = uva_clock();
        const double reg = uniform_d(st->reg), prev = uniform_d(st->e_prev);
        int conv = 0;
        double rel = 0.0;
        const int its = slim_solve(tm, P, d, U, reg, a.cg_tol, a.cg_max, conv, rel);
        const double tm_ = max_step_island(tm, P, d, U, P.X64, false);
        double step = fmin(1.0, 0.9 * tm_);
        double e = slim_energy_at(tm, P, d, U, step);
        for (int k = 0; k < 12 && (!__builtin_isfinite(e) || e > prev); ++k) {
            step *= 0.5;
            e = slim_energy_at(tm, P, d, U, step);
        }
        const bool accept = __builtin_isfinite(e) && !(e > prev);
        if (accept) {
            FOR_ROWS(tm, d, row, sg) {
                (void)sg;
                const double2 u = U[row], xn = P.X64[row];
                const double dx = xn.x - u.x, dy = xn.y - u.y;
                U[row] = make_double2(u.x + step * dx, u.y + step * dy);
            }
        }
        if (tm.leader()) {
            st->cg_iters += its;
            st->cg_iters_max = its > st->cg_iters_max ? its : st->cg_iters_max;
            st->converged = st->converged && conv;
            st->rel_resid = rel;
            st->tmax = tm_;
            if (accept) {
                st->arap_iters += 1;
                st->step = step;
                st->done = fabs(prev - e) <= a.tol * prev ? 1 : 0;
                st->e_prev = e;
            } else {
                st->done = 1;
            }
            st->ticks += uva_clock() - t_start;
        }
        tm.sync();
    }
}

#if !defined(UVA_EMULATOR)
template <int BS, int KIND, int PREC>
__global__ __launch_bounds__(BS) void island_kernel(DevPlan P, TaskArgs a, int slot) {
    constexpr int kRed = 2 * (BS / 64) * BlockTeam<BS>::kSlots > TeamWG<BS>::kLdsDoubles
                             ? 2 * (BS / 64) * BlockTeam<BS>::kSlots : TeamWG<BS>::kLdsDoubles;
    __shared__ double red[kRed];
    __shared__ int s_task;
    // vectors of the resident CG (kinds that solve the ARAP or LSCM system only)
    constexpr int kLv = lds_vec_bytes(KIND);
    __shared__ double lvec[kLv > 0 ? kLv / 8 : 1];
    char* const lv = kLv > 0 ? (char*)lvec : nullptr;
    BlockTeam<BS> bt((double UVA_LDS*)red, lv, kLv);
    // the workgroup slots of this XCD (each slot's tasks in order), then
    // wavefront tasks of any island
    const int list = uniform_i(P.xcc_map[xcc_id()]);
    int* const q = P.queue + kQueueStride * slot;
    for (int t_cur = 0, t_end = 0; list >= 0;) {
        if (t_cur == t_end) {
            // claim the next unclaimed slot (or task) of the XCD
            if (threadIdx.x == 0) s_task = atomicAdd(q + list, 1);
            __syncthreads();
            const int k = uniform_i(s_task);
            __syncthreads();
            if (P.claim_tasks) {
                // the next task of the XCD's list
                const int b = uniform_i(P.slot_ptr[list]), e = uniform_i(P.slot_ptr[list + 1]);
                if (k >= e - b) break;
                t_cur = b + k;
                t_end = t_cur + 1;
            } else {
                if (k >= P.slots_per_xcd) break;
                t_cur = uniform_i(P.slot_ptr[list * P.slots_per_xcd + k]);
                t_end = uniform_i(P.slot_ptr[list * P.slots_per_xcd + k + 1]);
            }
            continue;
        }
        const int task = t_cur++;
        const int isl = uniform_i(P.order_group[task]);
        const int m = uniform_i(P.grp_m[task]);
        if (m <= 1) {
            run_task<KIND, PREC>(bt, P, isl, a);
        } else {
            const int team = uniform_i(P.grp_team[task]);
            TeamWG<BS> tw((double UVA_LDS*)red, P.team_cnt + slot * P.n_teams + team,
                          P.team_part + (size_t)team * 2 * kMaxTeam * TeamWG<BS>::kSlots, m,
                          uniform_i(P.grp_mem[task]), P.xcd_local, P.spin_ticks, P.fault, lv, kLv);
            run_task<KIND, PREC>(tw, P, isl, a);
        }
        __syncthreads();
    }
    constexpr int kLvWave = kLv / (BS / 64);
    WaveTeam wt(kLv > 0 ? lv + (threadIdx.x >> 6) * kLvWave : nullptr, kLvWave);
    for (;;) {
        // the first active lane takes the next task for its wavefront; every
        // lane is active here, so that is lane 0
        int task = 0;
        if (__lane_id() == 0) task = atomicAdd(q + kMaxXcd, 1);
        task = uniform_i(task);
        if (task >= P.n_wave) break;
        run_task<KIND, PREC>(wt, P, uniform_i(P.order_wave[task]), a);
    }
}

__global__ void xcc_probe_kernel(int* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

__global__ void fill_int_kernel(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void permute_kernel(const int* __restrict__ vperm, const double2* __restrict__ src,
                               double2* __restrict__ dst, int n, int to_plan) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (to_plan) dst[v] = src[vperm[v]];
    else dst[vperm[v]] = src[v];
}

__global__ void signed_areas_kernel(DevPlan P, const double2* __restrict__ U, double* __restrict__ sa) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < P.n_tri) sa[f] = signed_area(P, f, U);
}

// Per-row soft pins: rpin[r] = the weight of row r in the m pinned rows
// (ascending, distinct; 0 elsewhere), pdiag = diag + rpin and its FP32
// inverse (1 where the pinned diagonal is not above 1e-30).
__global__ void set_pins_kernel(const double* __restrict__ diag, const int* __restrict__ rows,
                                const double* __restrict__ w, int m, double* __restrict__ rpin,
                                double* __restrict__ pdiag, float* __restrict__ dinv32, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int lo = 0, hi = m;               // first pinned row >= r
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rows[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    const double rp = (lo < m && rows[lo] == r) ? w[lo] : 0.0;
    const double g = diag[r] + rp;
    rpin[r] = rp;
    pdiag[r] = g;
    dinv32[r] = (float)(g > 1e-30 ? 1.0 / g : 1.0);
}

hipError_t launch_set_pins(const double* diag, const int* rows, const double* w, int m, double* rpin, double* pdiag,
                           float* dinv32, int n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(set_pins_kernel, dim3((n + 255) / 256), dim3(256), 0, s, diag, rows, w, m, rpin, pdiag,
                           dinv32, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
template <int KIND, int PREC>
static hipError_t launch_kind(const DevPlan& P, const TaskArgs& a, int slot, int cus, hipStream_t s) {
    constexpr int BS = block_size(KIND, PREC);
    // resident workgroups per CU of this instantiation (the same on every
    // call; a function-local static is initialised once, thread-safely)
    static const int occ = [] {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, island_kernel<BS, KIND, PREC>, BS, 0) != hipSuccess)
            nb = 1;
        return nb > 0 ? nb : 1;
    }();
    // enough workgroups for every workgroup task and the wavefront tasks, at
    // most what the device keeps resident; blocks go to the XCDs round-robin,
    // so every XCD list gets at least as many workgroups as its largest team
    const int cap = cus * occ;
    int grid = P.n_group + (P.n_wave + BS / 64 - 1) / (BS / 64);
    if (P.n_group > 0) grid = grid > P.n_xcd * P.slots_per_xcd ? grid : P.n_xcd * P.slots_per_xcd;
    grid = (grid + P.n_xcd - 1) / P.n_xcd * P.n_xcd;
    grid = grid < cap ? (grid > 0 ? grid : 1) : cap;
    hipLaunchKernelGGL((island_kernel<BS, KIND, PREC>), dim3(grid), dim3(BS), 0, s, P, a, slot);
    return hipGetLastError();
}

hipError_t launch_island_tasks(int kind, int prec, const DevPlan& P, const TaskArgs& a, int slot, int cus,
                               hipStream_t s) {
#define UVA_K(K)                                                                 \
    case K:                                                                      \
        switch (prec) {                                                          \
            case UVA_PREC_F32: return launch_kind<K, UVA_PREC_F32>(P, a, slot, cus, s); \
            case UVA_PREC_F64: return launch_kind<K, UVA_PREC_F64>(P, a, slot, cus, s); \
            default: return launch_kind<K, UVA_PREC_MIXED>(P, a, slot, cus, s);   \
        }
    switch (kind) {
        UVA_K(TASK_ARAP_ITER)
        UVA_K(TASK_CG)
        case TASK_PROLOGUE: return launch_kind<TASK_PROLOGUE, UVA_PREC_MIXED>(P, a, slot, cus, s);
        // FP64 vectors: FP32 inner passes restarted from the FP64 residual need
        // 1.6-3x the iterations on the conformal operator
        case TASK_LSCM: return launch_kind<TASK_LSCM, UVA_PREC_F64>(P, a, slot, cus, s);
        case TASK_LOCAL_RHS: return launch_kind<TASK_LOCAL_RHS, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_MAX_STEP: return launch_kind<TASK_MAX_STEP, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_STEP_APPLY: return launch_kind<TASK_STEP_APPLY, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_ENERGY: return launch_kind<TASK_ENERGY, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_REPAIR: return launch_kind<TASK_REPAIR, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_ORIENT: return launch_kind<TASK_ORIENT, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_FLIPS: return launch_kind<TASK_FLIPS, UVA_PREC_MIXED>(P, a, slot, cus, s);
        case TASK_SLIM_PROLOGUE: return launch_kind<TASK_SLIM_PROLOGUE, UVA_PREC_F64>(P, a, slot, cus, s);
        case TASK_SLIM_ITER: return launch_kind<TASK_SLIM_ITER, UVA_PREC_F64>(P, a, slot, cus, s);
        case TASK_ENERGY_BWD: return launch_kind<TASK_ENERGY_BWD, UVA_PREC_F64>(P, a, slot, cus, s);
        case TASK_STEP_BWD: return launch_kind<TASK_STEP_BWD, UVA_PREC_F64>(P, a, slot, cus, s);
        default: return hipErrorInvalidValue;
    }
#undef UVA_K
}

hipError_t launch_xcc_probe(int* out, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(xcc_probe_kernel, dim3(blocks), dim3(64), 0, s, out);
    return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_permute(const DevPlan& P, const double2* src, double2* dst, bool to_plan, hipStream_t s) {
    if (P.n_vert <= 0) return hipSuccess;
    hipLaunchKernelGGL(permute_kernel, dim3((P.n_vert + 255) / 256), dim3(256), 0, s, P.vperm, src, dst, P.n_vert,
                       to_plan ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_signed_areas(const DevPlan& P, const double2* U, double* sa, hipStream_t s) {
    if (P.n_tri <= 0) return hipSuccess;
    hipLaunchKernelGGL(signed_areas_kernel, dim3((P.n_tri + 255) / 256), dim3(256), 0, s, P, U, sa);
    return hipGetLastError();
}
#endif  // !UVA_EMULATOR

}  // namespace uva

#if defined(UVA_EMULATOR) && !defined(__HIP_DEVICE_COMPILE__)
#include "emu.inc"
#endif
