// qb_kernels.hip -- gfx950 kernels of the simulated-annealing QUBO engine.
//
// Float path (ray x voxel QUBO, src/ising_recon.hpp:105-166):
//   ray_team_kernel     persistent; one team of workgroups per XCD, each team
//                       annealing its own groups of BW branches. Every colour class
//                       is one step: BW lanes evaluate one bit for the BW branches
//                       of a group -- dE from the bit's CSC column (ray, a, w) and the
//                       group's residual line r[ray][0..BW), the neighbour term from
//                       spin masks, one Philox draw per lane, and on acceptance the
//                       residual scatter. Bits of one class share no ray and are not
//                       neighbours, so a step needs no atomics and its result does not
//                       depend on the order inside it. Steps are separated by a team
//                       barrier in the XCD's L2.
//   ray_outputs_kernel  branch agreement: p_void, tri-state, certainty, branches.
//   ray_energy_*        FP64 binary_energy of every branch, from scratch.
// Integer path (src/ising_weighted.hpp:193-267, src/ising_gpu.cuh:149-272):
//   int_sweep_kernel    one workgroup per replica, state in LDS (or global memory
//                       when it does not fit), fields updated with integer atomics.
//   int_energy_kernel   -cut over the edge list, one block per state.
//
// Every file of the engine is compiled with -ffp-contract=off: each floating-point
// expression is evaluated as written, in the order of the CPU oracle.
#include <algorithm>
#include <type_traits>

#include "qb_internal.h"
#include "qb_rng.h"

namespace qb {

// ---------------------------------------------------------------------------------
// Grid barrier for the persistent sweep. Monotone counter: the k-th barrier of a
// launch waits until the counter reaches k * gridDim.x. Every wave first waits for
// its own global stores to complete (s_barrier does not wait for other waves'
// stores on this target), so when thread 0's agent-scope release writes the L2
// back, the whole block's residual and spin stores of the step are in it; the
// acquire after the wait makes every block's stores visible before the next step
// reads them. A wait longer than ~2 s (s_memrealtime runs at 100 MHz) sets *err and
// every workgroup leaves its loops, so the grid drains.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ bool grid_barrier(unsigned* bar, unsigned target, int* err) {
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        if (ok) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            unsigned polls = 0;
            while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(1);
                if ((++polls & 255u) == 0) {
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                        __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = 0;
                        break;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <typename T>
__device__ __forceinline__ T residual_add(T r, float a, double delta);
template <>
__device__ __forceinline__ double residual_add<double>(double r, float a, double delta) {
    return r + double(a) * delta;                       // r[ray] += a * delta (ising_recon.hpp:139)
}
template <>
__device__ __forceinline__ float residual_add<float>(float r, float a, double delta) {
    return r + (delta > 0.0 ? a : -a);                  // the FP32 carry: one float add
}

// ---------------------------------------------------------------------------------
// XCD teams. The sweep's workgroups form one team per XCD (the hardware XCC_ID), and
// a team anneals its own branch groups alone. Inside an XCD the L2 is the point of
// coherence: stores write through the CU's L1 to it, so a team shares its mutable
// state (residuals, spin masks) by draining stores before each barrier and reading
// that state with loads that skip L1 and hit L2 (scope bit sc1). No cache is written
// back or invalidated between steps, and teams never touch each other's state.
// ---------------------------------------------------------------------------------


// Team barrier: every member has its own arrival word in the team's 64-word area of the
// barrier array, words[rank], and the k-th barrier completes when every member's word
// has reached k * team size (`target`). Every wave drains its stores, then thread 0
// publishes the member's arrival with one store; wave 0 then polls all members' words,
// one per lane in one load, until every one has arrived. No member waits on an atomic's
// return or on a second hop through a release word. While wave 0 polls, every thread
// first runs `after_arrival` (work that must not delay the arrival, such as the
// caller's task cursor). The error flag is read only while polling; a wait longer than
// ~2 s sets *err, as in grid_barrier.
static_assert(kMaxTeamSize <= 64, "wave 0 polls one member per lane");
template <typename F>
__device__ __forceinline__ bool team_barrier(unsigned* words, int rank, int team_size, unsigned target, int* err,
                                             F&& after_arrival) {
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(words + rank, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    after_arrival();
    if (threadIdx.x < 64) {
        const int lane = int(threadIdx.x);
        int ok = 1;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        unsigned polls = 0;
        while (true) {
            const unsigned v = lane < team_size ? ld_l2(words + lane) : target;
            if (__ballot(v < target) == 0) break;
            if ((++polls & 1023u) == 0) [[unlikely]] {
                if (ld_l2(err) != 0 || __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                    if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        if (lane == 0) s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

__device__ __forceinline__ bool team_barrier(unsigned* words, int rank, int team_size, unsigned target, int* err) {
    return team_barrier(words, rank, team_size, target, err, [] {});
}

// The dE of a bit sums its CSR entries in chains of kSeg: the first chain starts
// from lin * delta, later chains start from their first term; the chain sums are
// added in runs of kChainRun (in order inside a run) and the run sums in order; the
// neighbour terms follow in CSR order. For a bit with at most kSeg entries this is
// local_field_dE's order exactly (ising_recon.hpp:105-118).

// Loads of one chain: the entries [ka, kb) (1 <= kb - ka <= kSeg) and the group's
// residuals on their rays. Every load is unconditional -- past kb the index is
// clamped to kb - 1 -- so all of them issue back to back and the chain costs two
// dependent memory round trips.
template <typename T>
struct Chain {
    RayEnt e[kSeg];
    T rv[kSeg];
};

template <typename T, int BW>
__device__ __forceinline__ void chain_load(const RayEnt* __restrict__ ent, int ka, int kb, const T* rg, int b,
                                           Chain<T>& ch) {
#pragma unroll
    for (int u = 0; u < kSeg; ++u) ch.e[u] = ent[min(ka + u, kb - 1)];
#pragma unroll
    for (int u = 0; u < kSeg; ++u) ch.rv[u] = ld_l2(rg + size_t(ch.e[u].ray) * BW + b);
}

// The chain's sum: seeded with `seed` (the first chain of a bit, seed = lin * delta)
// or started from its first term; terms w (r a + a^2 / 2), a = a_k delta, in order.
template <typename T>
__device__ __forceinline__ double chain_sum(const RayEnt* e, const T* rv, int len, double delta, double seed,
                                            bool seeded) {
    double acc = seed;
#pragma unroll
    for (int u = 0; u < kSeg; ++u) {
        const double ad = double(e[u].a) * delta;
        const double t = e[u].w * (double(rv[u]) * ad + 0.5 * ad * ad);
        const double nxt = (u == 0 && !seeded) ? t : acc + t;
        acc = u < len ? nxt : acc;
    }
    return acc;
}
template <typename T>
__device__ __forceinline__ double chain_sum(const Chain<T>& ch, int len, double delta, double seed, bool seeded) {
    return chain_sum<T>(ch.e, ch.rv, len, delta, seed, seeded);
}
// The same sum from the entries' shares and weights held apart.
template <typename T>
__device__ __forceinline__ double chain_sum_aw(const float* ea, const double* ew, const T* rv, int len, double delta,
                                               double seed, bool seeded) {
    double acc = seed;
#pragma unroll
    for (int u = 0; u < kSeg; ++u) {
        const double ad = double(ea[u]) * delta;
        const double t = ew[u] * (double(rv[u]) * ad + 0.5 * ad * ad);
        const double nxt = (u == 0 && !seeded) ? t : acc + t;
        acc = u < len ? nxt : acc;
    }
    return acc;
}

// On acceptance: r[ray] += a * delta for the chain's entries, from the loaded values
// (no other bit of the class touches these rays).
template <typename T, int BW>
__device__ __forceinline__ void chain_scatter(const RayEnt* e, const T* rv, int len, T* rg, int b, double delta) {
#pragma unroll
    for (int u = 0; u < kSeg; ++u)
        if (u < len) rg[size_t(e[u].ray) * BW + b] = residual_add<T>(rv[u], e[u].a, delta);
}
template <typename T, int BW>
__device__ __forceinline__ void chain_scatter(const Chain<T>& ch, int len, T* rg, int b, double delta) {
    chain_scatter<T, BW>(ch.e, ch.rv, len, rg, b, delta);
}

// lambda * (+-1) per neighbour, in CSR order; the first 8 neighbours' masks are loaded
// together (the 6-neighbour grid never needs the tail loop).
__device__ __forceinline__ double neighbour_terms(const DevProblem& p, const uint16_t* xg, int n0, int n1, int b,
                                                  int xb, double acc) {
    if (n1 <= n0) return acc;
    constexpr int kNbL = 8;
    int q[kNbL];
    uint16_t m[kNbL];
#pragma unroll
    for (int u = 0; u < kNbL; ++u) q[u] = p.nb[min(n0 + u, n1 - 1)];
#pragma unroll
    for (int u = 0; u < kNbL; ++u) m[u] = ld_l2(xg + q[u]);
#pragma unroll
    for (int u = 0; u < kNbL; ++u) {
        const int before = ((m[u] >> b) & 1) != xb;
        const double nxt = acc + p.lambda * (before ? -1.0 : 1.0);
        acc = n0 + u < n1 ? nxt : acc;
    }
    for (int k = n0 + kNbL; k < n1; ++k) {
        const int before = ((ld_l2(xg + p.nb[k]) >> b) & 1) != xb;
        acc += p.lambda * (before ? -1.0 : 1.0);
    }
    return acc;
}

// The Metropolis test of one branch with its draw u = draw_u01(seed, sweep, bit)
// already made: accept iff dE <= 0 or u < qexp(-dE * beta), decided through the float
// exponential with a guard band (uphill_accept: qexp's decision, bit for bit).
template <typename T>
__device__ __forceinline__ bool decide_drawn(const DevState<T>& s, int branch, double u, double beta, double dE) {
    if (branch >= s.R) return false;
    return dE <= 0.0 || uphill_accept(-dE * beta, u);
}

// The accepted lanes of this lane's BW-wide segment, as a mask of the group's branches.
template <int BW>
__device__ __forceinline__ uint16_t segment_flips(bool acc) {
    const uint64_t bal = __ballot(acc);
    const unsigned base = threadIdx.x & 63u & ~unsigned(BW - 1);
    return uint16_t((bal >> base) & ((uint64_t(1) << BW) - 1));
}

// The sweep and temperature one bit's proposal uses. In the plain form every bit is at
// the launch's sweep. In the staged form (DevState::clock) the bit's own sweep is the
// launch's plus its clock offset, beta is read at it, and the proposal is live only while
// that sweep is inside the schedule and the branch is one of its request's.
template <bool STAGED, typename T>
__device__ __forceinline__ bool bit_time(const DevState<T>& s, uint32_t clk, int branch, int sweep, double beta,
                                         int& own, double& bt) {
    if constexpr (!STAGED) {
        (void)s;
        (void)clk;
        (void)branch;
        own = sweep;
        bt = beta;
        return true;
    } else {
        (void)beta;
        own = sweep + int(clk & kClockOffsetMask);
        const int rq = int(clk >> 24);
        const bool live = own < s.total_sweeps && (rq == 0 || branch < rq);
        bt = live ? s.beta[own] : 0.0;
        return live;
    }
}

// ---------------------------------------------------------------------------------
// Packed steps. A batch's static operands -- every slot's chain entries and head --
// form one block (DevBatches::blk), which the workgroup copies straight into LDS a
// step ahead: every wave moves 64 granules of 16 bytes per instruction, in a few
// instructions, while the previous step decides, scatters and waits at the barrier.
// Once the barrier opens a step issues only its state loads (the spin masks, the
// chains' residuals), with their addresses read from LDS.
// ---------------------------------------------------------------------------------

__device__ __forceinline__ int slot_len(uint32_t i) { return int(i & 31u); }
__device__ __forceinline__ int slot_j(uint32_t i) { return int((i >> 5) & 255u); }
__device__ __forceinline__ int slot_nch(uint32_t i) { return int((i >> 13) & 255u); }
__device__ __forceinline__ int slot_nnb(uint32_t i) { return int((i >> 21) & 1023u); }
__device__ __forceinline__ bool slot_seq(uint32_t i) { return (i >> 31) != 0u; }

// Copies `n` granules from global `src` into LDS at `dst`, every wave of the workgroup
// 64 granules per instruction (lanes past the end repeat the last granule, so `dst`
// must hold n rounded up to 64 granules).
template <int BLOCK>
__device__ __forceinline__ void copy_to_lds(const uint4* src, int n, unsigned char* dst) {
    const int tid = int(threadIdx.x);
    for (int v0 = (tid >> 6) * 64; v0 < n; v0 += BLOCK) {
        const uint4* g = src + min(v0 + (tid & 63), n - 1);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)g,
                                         (__attribute__((address_space(3))) void*)(dst + size_t(v0) * 16), 16, 0, 0);
    }
}

// Issues the copy of a batch's block ({first granule, slots used} = meta) into the LDS
// buffer pb: the entries to pb, the heads to pb + KSLOTS * kSeg granules. The copy
// lands asynchronously: a wave's vmcnt covers it, so a wait for every memory operation
// and a workgroup barrier (as the team barrier makes) precede any read of the buffer.
template <int BLOCK, int KSLOTS>
__device__ __forceinline__ void issue_block(const uint4* blk, int2 meta, unsigned char* pb) {
    static_assert((KSLOTS * kSeg) % 64 == 0 && (KSLOTS * kHeadGranules) % 64 == 0,
                  "the buffer's parts hold whole 64-granule copies");
    const int first = __builtin_amdgcn_readfirstlane(meta.x), n = __builtin_amdgcn_readfirstlane(meta.y);
    if (n <= 0) return;
    copy_to_lds<BLOCK>(blk + first, n * kSeg, pb);
    copy_to_lds<BLOCK>(blk + first + n * kSeg, n * kHeadGranules, pb + size_t(KSLOTS) * kSeg * 16);
}

// The per-lane operands of a packed task that depend only on its block, its sweep and
// its group: the slot's bit and chain, the rays of the entries the lane loads, the
// neighbour positions, the draw and, with pairs, each loaded entry's share and bit. A
// task's prep is made from its block in LDS before its step -- in the barrier wait before
// it when a barrier separates it from the previous task -- so once the barrier opens the
// step only issues its state loads.
template <typename T, int BW>
struct PackedPrep {
    static constexpr bool kPair = BW == 2 && sizeof(T) == 4;
    static constexpr int kLoads = kPair ? kSeg / 2 : kSeg;
    int pos;
    uint32_t info;
    int ray[kLoads];
    int q[kNb];
    int pbs[kPair ? kLoads : 1];
    float pa[kPair ? kLoads : 1];
    double u, bt, lin;
    bool live;
};

// Residual loads across the wave, in whole chains (whose rays share few lines) rather
// than one entry of 64 / BW unrelated chains. With two FP32 branches a lane loads an
// entry's pair of residuals as one 8-byte word: load u of lane l is entry q % kSeg of the
// wave's chain q / kSeg, q = 64 u + l, so a wave issues kSeg / 2 loads. Otherwise a lane
// loads one branch's residual: q = (64 u + l) / BW, kSeg loads.
template <typename T, int BW, int BLOCK, bool STAGED>
__device__ __forceinline__ void packed_prep(const DevState<T>& s, const unsigned char* pb, int n_used, int g,
                                            int sweep, double beta, PackedPrep<T, BW>& P) {
    constexpr int kSlots = BLOCK / BW;
    constexpr int kSpw = 64 / BW;
    constexpr bool kPair = PackedPrep<T, BW>::kPair;
    constexpr int kLoads = PackedPrep<T, BW>::kLoads;
    const RayEnt* E = reinterpret_cast<const RayEnt*>(pb);
    const SlotHead* Hd = reinterpret_cast<const SlotHead*>(pb + size_t(kSlots) * kSeg * sizeof(RayEnt));
    const int tid = int(threadIdx.x);
    const int b = tid % BW, slot = tid / BW;
    const int lane = tid & 63, wave = tid >> 6;
    // every LDS read first, unconditionally (slots past n_used hold stale data, which the
    // masks below discard), so they are in flight together
    const bool used = slot < n_used;
    const int hpos = Hd[slot].pos;
    const uint32_t hinfo = Hd[slot].info;
    const uint64_t hkey = Hd[slot].key;
    const int hbit = Hd[slot].bit;
    const double hlin = Hd[slot].lin;
    int hq[kNb];
#pragma unroll
    for (int v = 0; v < kNb; ++v) hq[v] = Hd[slot].q[v];
    int ray[kLoads];
    uint32_t cinfo[kPair ? kLoads : 1];
    float ca[kPair ? kLoads : 1];
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int q = kPair ? u * 64 + lane : (u * 64 + lane) / BW;
        const int c = wave * kSpw + q / kSeg;
        const RayEnt* e = E + blk_entry(c, q % kSeg);
        ray[u] = e->ray;
        if constexpr (kPair) {
            ca[u] = e->a;
            cinfo[u] = Hd[c].info;
        }
    }
    P.pos = used ? hpos : -1;
    P.info = used ? hinfo : 0u;
    const int j = slot_j(P.info), nn = slot_nnb(P.info);
    const bool first = P.pos >= 0 && j == 0;
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int q = kPair ? u * 64 + lane : (u * 64 + lane) / BW;
        const int c = wave * kSpw + q / kSeg;
        P.ray[u] = c < n_used ? ray[u] : -1;
    }
#pragma unroll
    for (int v = 0; v < kNb; ++v) P.q[v] = first && nn > 0 ? hq[v] : 0;
    if constexpr (kPair) {
        // each loaded entry's share and its bit's first slot (-1: no entry, or a column
        // whose rays repeat, which that slot scatters itself)
#pragma unroll
        for (int w = 0; w < kLoads; ++w) {
            const int q = w * 64 + lane;
            const int c = wave * kSpw + q / kSeg;
            const bool has = P.ray[w] >= 0;
            P.pbs[w] = has && !slot_seq(cinfo[w]) ? c - slot_j(cinfo[w]) : -1;
            P.pa[w] = has ? ca[w] : 0.0f;
        }
    }
    // the draw depends only on (branch, sweep, bit)
    P.u = 0.0;
    P.bt = beta;
    P.lin = 0.0;
    P.live = false;
    if (first) {
        const uint32_t clk = STAGED ? s.clock[P.pos] : 0u;
        int own = sweep;
        P.live = bit_time<STAGED>(s, clk, g * BW + b, sweep, beta, own, P.bt);
        P.u = draw_u01(s.seed_offset + hkey + uint64_t(g * BW + b) + 1u, uint32_t(own), uint32_t(hbit));
        P.lin = hlin;
    }
}

// One packed batch for the BW branches of group g, its block in LDS at pb and its prep
// made: every slot evaluates one chain of one bit; the chain sums meet in LDS, the first
// slot of each run of kChainRun chains adds the run, the bit's first slot adds the runs
// and the neighbour terms and decides; then the accepted chains are scattered (with two
// FP32 branches by the lanes that loaded each entry's pair, else by each slot).
// `after_sync` runs once every slot's chain sum is in LDS (the caller's copy of the next
// block). `trow` is the residual transposition area (kSweepBlock * kSeg values of T),
// `part`, `accepted` and `fx` the workgroup's chain sums, decisions and per-bit flip
// records.
template <typename T, int BW, int BLOCK, bool STAGED, typename F2>
__device__ __forceinline__ void packed_step(const DevProblem& p, const DevState<T>& s, const unsigned char* pb,
                                            const PackedPrep<T, BW>& P, T* trow, double (*part)[BW],
                                            int (*accepted)[BW], uint32_t* fx, int g, F2&& after_sync) {
    constexpr int kSlots = BLOCK / BW;
    constexpr int kSpw = 64 / BW;   // slots per wave
    constexpr bool kPair = PackedPrep<T, BW>::kPair;
    constexpr int kLoads = PackedPrep<T, BW>::kLoads;
    const RayEnt* E = reinterpret_cast<const RayEnt*>(pb);
    const SlotHead* Hd = reinterpret_cast<const SlotHead*>(pb + size_t(kSlots) * kSeg * sizeof(RayEnt));
    const int tid = int(threadIdx.x);
    const int b = tid % BW, slot = tid / BW;
    const int lane = tid & 63, wave = tid >> 6;
    const int pos = P.pos;
    const uint32_t info = P.info;
    const int len = slot_len(info), j = slot_j(info), nn = slot_nnb(info);
    const bool first = pos >= 0 && j == 0;
    uint16_t* xg = s.x + size_t(g) * size_t(p.n);
    T* rg = s.r + size_t(g) * size_t(p.n_rays) * BW;
    // The residuals come back through this wave's rows of trow in chain order (a wave's
    // LDS operations complete in order, so no barrier is needed).
    T* tw = trow + wave * kSpw * BW * kSeg;
    const int ls = slot - wave * kSpw;
    T rvt[kPair ? 1 : kSeg];
    uint64_t rpair[kPair ? kLoads : 1];
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int ray = P.ray[u];
        if constexpr (kPair)
            rpair[u] = ray >= 0 ? ld_l2(reinterpret_cast<const uint64_t*>(rg + size_t(ray) * BW)) : 0ull;
        else
            rvt[u] = ray >= 0 ? ld_l2(rg + size_t(ray) * BW + b) : T(0);
    }
    uint16_t m[kNb];
    uint16_t xm = 0;
    if (pos >= 0) {
        xm = ld_l2(xg + pos);
        if (first && nn > 0) {
            if constexpr (BW == 2) {
                // the slot's two lanes load half of the masks each (swapped below)
#pragma unroll
                for (int u = 0; u < kNb / 2; ++u) m[u] = ld_l2(xg + (b ? P.q[u + kNb / 2] : P.q[u]));
            } else {
#pragma unroll
                for (int u = 0; u < kNb; ++u) m[u] = ld_l2(xg + P.q[u]);
            }
        }
    }
    // the chain's shares and weights, from LDS while the loads are in flight
    float ea[kSeg];
    double ew[kSeg];
    if (pos >= 0) {
#pragma unroll
        for (int v = 0; v < kSeg; ++v) {
            const RayEnt ev = E[blk_entry(slot, v)];
            ea[v] = ev.a;
            ew[v] = ev.w;
        }
    }
    const double u = P.u, bt = P.bt, lin = P.lin;
    const bool live = P.live;
    // the residuals back in chain order
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        if constexpr (kPair) {
            const int q = u * 64 + lane;
            T* row = tw + (q / kSeg) * BW * kSeg + q % kSeg;
            row[0] = __uint_as_float(uint32_t(rpair[u]));
            row[kSeg] = __uint_as_float(uint32_t(rpair[u] >> 32));
        } else {
            const int q = (u * 64 + lane) / BW;
            tw[((q / kSeg) * BW + b) * kSeg + q % kSeg] = rvt[u];
        }
    }
    T rv[kSeg];
#pragma unroll
    for (int u = 0; u < kSeg; ++u) rv[u] = tw[(ls * BW + b) * kSeg + u];
    const int xb = (xm >> b) & 1;
    const double delta = xb ? -1.0 : 1.0;
    // the first kNb neighbours' disagreement with this branch's spin, one bit each
    uint32_t unlike = 0;
    if (first && nn > 0) {
        if constexpr (BW == 2) {
            // both lanes of the slot are here: each takes the other's half (DPP quad_perm 1,0,3,2)
            uint16_t mm[kNb];
#pragma unroll
            for (int u = 0; u < kNb / 2; ++u) {
                const uint16_t o = uint16_t(__builtin_amdgcn_mov_dpp(int(m[u]), 0xB1, 0xF, 0xF, false));
                mm[u] = b ? o : m[u];
                mm[u + kNb / 2] = b ? m[u] : o;
            }
#pragma unroll
            for (int v = 0; v < kNb; ++v) unlike |= uint32_t(((mm[v] >> b) & 1) != xb) << v;
        } else {
#pragma unroll
            for (int v = 0; v < kNb; ++v) unlike |= uint32_t(((m[v] >> b) & 1) != xb) << v;
        }
    }
    if (pos >= 0) {
        double partial = j == 0 ? lin * delta : 0.0;
        if (len > 0) partial = chain_sum_aw<T>(ea, ew, rv, len, delta, partial, j == 0);
        part[slot][b] = partial;
    }
    __syncthreads();
    after_sync();
    // the first slot of every run of kChainRun chains adds the run's sums in order, in
    // place (no other slot reads them)
    if (pos >= 0 && j % kChainRun == 0) {
        const int nch = slot_nch(info);
        if (j + 1 < nch) {
            double v[kChainRun];
#pragma unroll
            for (int w = 1; w < kChainRun; ++w) v[w] = part[min(slot + w, kSlots - 1)][b];
            double gsum = part[slot][b];
#pragma unroll
            for (int w = 1; w < kChainRun; ++w) gsum = j + w < nch ? gsum + v[w] : gsum;
            part[slot][b] = gsum;
        }
    }
    __syncthreads();
    if (first) {
        const int nch = slot_nch(info);
        double dE = part[slot][b];
        for (int q0 = kChainRun; q0 < nch; q0 += 8 * kChainRun) {
            double v[8];
#pragma unroll
            for (int w = 0; w < 8; ++w) v[w] = part[min(slot + q0 + w * kChainRun, kSlots - 1)][b];
#pragma unroll
            for (int w = 0; w < 8; ++w) dE = q0 + w * kChainRun < nch ? dE + v[w] : dE;
        }
        // lambda * (+-1) per neighbour in CSR order
#pragma unroll
        for (int v = 0; v < kNb; ++v) {
            const double nxt = dE + p.lambda * (((unlike >> v) & 1u) ? -1.0 : 1.0);
            dE = v < nn ? nxt : dE;
        }
        if (nn > kNb) [[unlikely]] {
            const int nb0 = Hd[slot].nb0;
            for (int k = nb0 + kNb; k < nb0 + nn; ++k) {
                const int before = ((ld_l2(xg + p.nb[k]) >> b) & 1) != xb;
                dE += p.lambda * (before ? -1.0 : 1.0);
            }
        }
        const bool acc = live && decide_drawn<T>(s, g * BW + b, u, bt, dE);
        accepted[slot][b] = acc ? 1 : 0;
        const uint16_t flips = segment_flips<BW>(acc);
        if (b == 0 && flips) xg[pos] = uint16_t(xm ^ flips);
        if constexpr (kPair)
            if (b == 0) fx[slot] = uint32_t(flips) | (uint32_t(xm) << 16);
    }
    __syncthreads();
    if constexpr (kPair) {
        // every entry's pair from the loading lane: both branches written back, the
        // flipped ones updated (rays repeating in a column are left to the bit's first slot)
        uint32_t f[kLoads];
#pragma unroll
        for (int w = 0; w < kLoads; ++w) f[w] = P.pbs[w] >= 0 ? fx[P.pbs[w]] : 0u;
#pragma unroll
        for (int w = 0; w < kLoads; ++w) {
            if ((f[w] & 3u) == 0) continue;
            float r0 = __uint_as_float(uint32_t(rpair[w])), r1 = __uint_as_float(uint32_t(rpair[w] >> 32));
            if (f[w] & 1u) r0 = residual_add<float>(r0, P.pa[w], (f[w] & 0x10000u) ? -1.0 : 1.0);
            if (f[w] & 2u) r1 = residual_add<float>(r1, P.pa[w], (f[w] & 0x20000u) ? -1.0 : 1.0);
            *reinterpret_cast<uint64_t*>(rg + size_t(P.ray[w]) * BW) =
                uint64_t(__float_as_uint(r0)) | (uint64_t(__float_as_uint(r1)) << 32);
        }
    }
    if (pos >= 0 && accepted[slot - j][b]) {
        if (slot_seq(info)) [[unlikely]] {
            // a ray repeats in this column: the bit's first slot scatters it in CSR order
            if (j == 0)
                for (int k = p.col_ptr[pos]; k < p.col_ptr[pos + 1]; ++k) {
                    const RayEnt ek = p.ent[k];
                    T* q = rg + size_t(ek.ray) * BW + b;
                    *q = residual_add<T>(ld_l2(q), ek.a, delta);
                    __builtin_amdgcn_s_waitcnt(0);
                }
        } else if (!kPair && len > 0) {
#pragma unroll
            for (int v = 0; v < kSeg; ++v)
                if (v < len) rg[size_t(E[blk_entry(slot, v)].ray) * BW + b] = residual_add<T>(rv[v], ea[v], delta);
        }
    }
}

// A multi-pass batch's bit: slot descriptors 0 and 1 of the batch carry
// {-1, pos, k0, k1} and {-1, n0, n1, seq} (pos -1, so both read as empty slots).
struct HeavyBit {
    int pos, k0, k1, n0, n1, seq;
};

__device__ __forceinline__ HeavyBit heavy_bit(const DevBatches& bt, int batch, int slots) {
    const int4* d = reinterpret_cast<const int4*>(bt.desc + size_t(batch) * size_t(slots));
    const int4 a = d[0], c = d[1];
    return HeavyBit{a.y, a.z, a.w, c.y, c.z, c.w};
}

// A bit with more chains than a workgroup has slots, by a whole workgroup in several
// passes: the chains are spread over the block's BW-lane slots, their sums meet in
// LDS, slot 0 adds them (in runs of kChainRun) and decides, and the residual scatter is spread
// over the block again.
template <typename T, int BW, int BLOCK, bool STAGED, typename F>
__device__ __forceinline__ void heavy_update(const DevProblem& p, const DevState<T>& s, const HeavyBit& h, int g,
                                             int sweep, double beta, double (*part)[BW], int* accepted, F&& after) {
    const int pos = h.pos;
    constexpr int kSlots = BLOCK / BW;
    const int b = threadIdx.x % BW;
    const int slot = threadIdx.x / BW;
    uint16_t* xg = s.x + size_t(g) * size_t(p.n);
    T* rg = s.r + size_t(g) * size_t(p.n_rays) * BW;
    const uint16_t xm = ld_l2(xg + pos);
    const int k0 = h.k0, k1 = h.k1;
    const double lin = p.lin[pos];
    const int xb = (xm >> b) & 1;
    const double delta = xb ? -1.0 : 1.0;
    const int n_chain = (k1 - k0 + kSeg - 1) / kSeg;
    double dE = 0.0, run = 0.0;
    Chain<T> ch;
    for (int j0 = 0; j0 < n_chain; j0 += kSlots) {
        const int j = j0 + slot;
        if (j < n_chain) {
            const int ka = k0 + j * kSeg;
            const int kb = min(ka + kSeg, k1);
            chain_load<T, BW>(p.ent, ka, kb, rg, b, ch);
            part[slot][b] = chain_sum<T>(ch, kb - ka, delta, lin * delta, j == 0);
        }
        __syncthreads();
        if (slot == 0) {
            const int m = min(n_chain - j0, kSlots);
            for (int q = 0; q < m; ++q) {
                const int jq = j0 + q;
                run = jq % kChainRun == 0 ? part[q][b] : run + part[q][b];
                if (jq % kChainRun == kChainRun - 1 || jq == n_chain - 1) dE = jq < kChainRun ? run : dE + run;
            }
        }
        __syncthreads();
    }
    if (slot == 0) {
        dE = neighbour_terms(p, xg, h.n0, h.n1, b, xb, dE);
        int own = sweep;
        double bt = beta;
        const bool live = bit_time<STAGED>(s, STAGED ? s.clock[pos] : 0u, g * BW + b, sweep, beta, own, bt);
        const double u = draw_u01(s.seed_offset + p.draw_key[pos] + uint64_t(g * BW + b) + 1u, uint32_t(own),
                                  uint32_t(p.draw_bit[pos]));
        const bool acc = live && decide_drawn<T>(s, g * BW + b, u, bt, dE);
        accepted[b] = acc ? 1 : 0;
        const uint16_t flips = segment_flips<BW>(acc);
        if (b == 0 && flips) xg[pos] = uint16_t(xm ^ flips);
    }
    __syncthreads();
    if (accepted[b]) {
        if (h.seq) {
            // a ray repeats in this column: its updates must land in CSR order
            if (slot == 0)
                for (int k = k0; k < k1; ++k) {
                    const RayEnt e = p.ent[k];
                    T* q = rg + size_t(e.ray) * BW + b;
                    *q = residual_add<T>(ld_l2(q), e.a, delta);
                    __builtin_amdgcn_s_waitcnt(0);
                }
        } else {
            // the chain this slot evaluated in the last pass is still in registers:
            // scatter it first, then load and scatter the slot's earlier chains
            const int last_pass = ((n_chain - 1) / kSlots) * kSlots;
            if (last_pass + slot < n_chain) {
                const int ka = k0 + (last_pass + slot) * kSeg;
                chain_scatter<T, BW>(ch, min(ka + kSeg, k1) - ka, rg, b, delta);
            }
            for (int j = slot; j < last_pass; j += kSlots) {
                const int ka = k0 + j * kSeg;
                const int kb = min(ka + kSeg, k1);
                chain_load<T, BW>(p.ent, ka, kb, rg, b, ch);
                chain_scatter<T, BW>(ch, kb - ka, rg, b, delta);
            }
        }
    }
    __syncthreads();
    after();
}

// A bit with more chains than a workgroup has slots, up to kHeavyPasses chains per
// slot: its column is staged in LDS by the whole workgroup (every load of it in flight
// at once), then every slot issues the residual loads of all its chains at once, sums
// each chain, and after the decision scatters them from registers. The neighbour
// positions and masks are loaded by the deciding slot at the start, beside the column.
// A step of such a bit costs the staging plus one round of residual loads, where the
// pass-by-pass form costs two dependent round trips per pass and again per pass to
// scatter. The staging area is the step's own block buffer, h0 (empty for a multi-pass
// batch), for the first kSlots chains and the sweep's work area hx for the rest, which
// also holds the chain sums. `after` runs once the decision is made (the caller's copy
// of the next block into the other buffer).

template <typename T, int BW, int BLOCK, bool STAGED, typename F>
__device__ __forceinline__ void heavy_lds_update(const DevProblem& p, const DevState<T>& s, const HeavyBit& h, int g,
                                                 int sweep, double beta, RayEnt* h0, RayEnt* hx, F&& after) {
    constexpr int kSlots = BLOCK / BW;
    constexpr int kE0 = kSlots * kSeg;   // entries staged in h0; the rest go to hx
    constexpr int kMaxEnt = kHeavyPasses * kE0;
    constexpr int kNbHeavy = 8;
    static_assert(kE0 % 64 == 0 && kMaxEnt % 64 == 0, "a wave stages 64 entries at a time");
    static_assert(BLOCK == kSweepBlock, "the staging area is sized for the sweep workgroup");
    double (*hpart)[BW] = reinterpret_cast<double (*)[BW]>(hx + (kHeavyPasses - 1) * kE0);
    __shared__ int hacc[BW];
    // chain j's entries: the first kSlots chains in h0, the others in hx
    auto chain_at = [&](int j) -> const RayEnt* { return j < kSlots ? h0 + j * kSeg : hx + (j - kSlots) * kSeg; };
    const int tid = int(threadIdx.x);
    const int b = tid % BW, slot = tid / BW;
    uint16_t* xg = s.x + size_t(g) * size_t(p.n);
    T* rg = s.r + size_t(g) * size_t(p.n_rays) * BW;
    const int pos = h.pos, k0 = h.k0, deg = h.k1 - h.k0;   // kSlots * kSeg < deg <= kMaxEnt
    const int n_chain = (deg + kSeg - 1) / kSeg;
    const uint16_t xm = ld_l2(xg + pos);
    // stage the column straight into LDS: wave w copies entries [64 v, 64 v + 64) for
    // v = w, w + 4, ... (lane i to row 64 v + i; rows past deg repeat the last entry and
    // stay inside the area, as deg <= kMaxEnt and kE0 is a multiple of 64)
    {
        for (int v0 = (tid >> 6) * 64; v0 < deg; v0 += BLOCK) {
            const RayEnt* src = p.ent + k0 + min(v0 + (tid & 63), deg - 1);
            RayEnt* dst = v0 < kE0 ? h0 + v0 : hx + (v0 - kE0);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
        }
        // the deciding slot's neighbour positions, then their masks (neighbours are in
        // other classes, final since the barrier)
        int q[kNbHeavy];
        uint16_t m[kNbHeavy];
        if (slot == 0 && h.n1 > h.n0) {
#pragma unroll
            for (int u = 0; u < kNbHeavy; ++u) q[u] = p.nb[min(h.n0 + u, h.n1 - 1)];
#pragma unroll
            for (int u = 0; u < kNbHeavy; ++u) m[u] = ld_l2(xg + q[u]);
        }
        __builtin_amdgcn_s_waitcnt(0);   // the LDS copies (and the neighbour loads) have landed
        __syncthreads();
        const int xb = (xm >> b) & 1;
        const double delta = xb ? -1.0 : 1.0;
        // the residuals of every chain of this slot, all in flight together
        T rv[kHeavyPasses][kSeg];
#pragma unroll
        for (int qq = 0; qq < kHeavyPasses; ++qq) {
            const int j = slot + qq * kSlots;
            const int len = j < n_chain ? min(kSeg, deg - j * kSeg) : 0;
            const RayEnt* e = chain_at(min(j, n_chain - 1));
#pragma unroll
            for (int u = 0; u < kSeg; ++u) rv[qq][u] = T(0);
            if (len > 0) {
#pragma unroll
                for (int u = 0; u < 4; ++u) rv[qq][u] = ld_l2(rg + size_t(e[min(u, len - 1)].ray) * BW + b);
            }
            if (len > 4) {
#pragma unroll
                for (int u = 4; u < 8; ++u) rv[qq][u] = ld_l2(rg + size_t(e[min(u, len - 1)].ray) * BW + b);
            }
            if (len > 8) {
#pragma unroll
                for (int u = 8; u < 12; ++u) rv[qq][u] = ld_l2(rg + size_t(e[min(u, len - 1)].ray) * BW + b);
            }
            if (len > 12) {
#pragma unroll
                for (int u = 12; u < 16; ++u) rv[qq][u] = ld_l2(rg + size_t(e[min(u, len - 1)].ray) * BW + b);
            }
        }
        double u01 = 0.0, bt = beta;
        bool live = false;
        if (slot == 0) {
            int own = sweep;
            live = bit_time<STAGED>(s, STAGED ? s.clock[pos] : 0u, g * BW + b, sweep, beta, own, bt);
            u01 = draw_u01(s.seed_offset + p.draw_key[pos] + uint64_t(g * BW + b) + 1u, uint32_t(own),
                           uint32_t(p.draw_bit[pos]));
        }
        const double lin = p.lin[pos];
#pragma unroll
        for (int qq = 0; qq < kHeavyPasses; ++qq) {
            const int j = slot + qq * kSlots;
            if (j < n_chain) {
                const int len = min(kSeg, deg - j * kSeg);
                RayEnt e[kSeg];
                const RayEnt* ej = chain_at(j);
#pragma unroll
                for (int u = 0; u < kSeg; ++u) e[u] = ej[min(u, len - 1)];
                hpart[j][b] = chain_sum<T>(e, rv[qq], len, delta, lin * delta, j == 0);
            }
        }
        __syncthreads();
        // run r of kChainRun chain sums, added in order by slot r, in place
        const int n_run = (n_chain + kChainRun - 1) / kChainRun;   // <= kHeavyPasses * kSlots / kChainRun
        if (slot < n_run) {
            const int j0 = slot * kChainRun;
            double v[kChainRun];
#pragma unroll
            for (int u = 1; u < kChainRun; ++u) v[u] = hpart[min(j0 + u, kHeavyPasses * kSlots - 1)][b];
            double g = hpart[j0][b];
#pragma unroll
            for (int u = 1; u < kChainRun; ++u) g = j0 + u < n_chain ? g + v[u] : g;
            hpart[j0][b] = g;
        }
        __syncthreads();
        if (slot == 0) {
            double dE = hpart[0][b];
            for (int r0 = 1; r0 < n_run; r0 += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = hpart[min((r0 + u) * kChainRun, kHeavyPasses * kSlots - 1)][b];
#pragma unroll
                for (int u = 0; u < 8; ++u) dE = r0 + u < n_run ? dE + v[u] : dE;
            }
            // lambda * (+-1) per neighbour in CSR order
#pragma unroll
            for (int u = 0; u < kNbHeavy; ++u) {
                const int before = ((m[u] >> b) & 1) != xb;
                const double nxt = dE + p.lambda * (before ? -1.0 : 1.0);
                dE = h.n0 + u < h.n1 ? nxt : dE;
            }
            for (int k = h.n0 + kNbHeavy; k < h.n1; ++k) {
                const int before = ((ld_l2(xg + p.nb[k]) >> b) & 1) != xb;
                dE += p.lambda * (before ? -1.0 : 1.0);
            }
            const bool acc = live && decide_drawn<T>(s, g * BW + b, u01, bt, dE);
            hacc[b] = acc ? 1 : 0;
            const uint16_t flips = segment_flips<BW>(acc);
            if (b == 0 && flips) xg[pos] = uint16_t(xm ^ flips);
        }
        __syncthreads();
        after();
        if (hacc[b]) {
            if (h.seq) {
                // a ray repeats in this column: its updates must land in CSR order
                if (slot == 0)
                    for (int k = 0; k < deg; ++k) {
                        const RayEnt e = k < kE0 ? h0[k] : hx[k - kE0];
                        T* qp = rg + size_t(e.ray) * BW + b;
                        *qp = residual_add<T>(ld_l2(qp), e.a, delta);
                        __builtin_amdgcn_s_waitcnt(0);
                    }
            } else {
#pragma unroll
                for (int qq = 0; qq < kHeavyPasses; ++qq) {
                    const int j = slot + qq * kSlots;
                    if (j < n_chain) {
                        const int len = min(kSeg, deg - j * kSeg);
                        const RayEnt* ej = chain_at(j);
#pragma unroll
                        for (int u = 0; u < kSeg; ++u)
                            if (u < len) {
                                const RayEnt e = ej[u];
                                rg[size_t(e.ray) * BW + b] = residual_add<T>(rv[qq][u], e.a, delta);
                            }
                    }
                }
            }
        }
    }
    __syncthreads();   // the staging area, hpart and hacc are reused by the next step
}

// Dispatch of a multi-pass batch: staged in LDS when its chains fit kHeavyPasses
// rounds of the workgroup's slots (and the launch reserved the staging area), else
// pass by pass.
template <typename T, int BW, int BLOCK, bool STAGED, typename F>
__device__ __forceinline__ void heavy_step(const DevProblem& p, const DevState<T>& s, const DevBatches& bt, int batch,
                                           int g, int sweep, double beta, RayEnt* h0, RayEnt* hx, double (*part)[BW],
                                           int (*accepted)[BW], F&& after) {
    constexpr int kSlots = BLOCK / BW;
    const HeavyBit h = heavy_bit(bt, batch, kSlots);
    if (bt.heavy_lds && h.k1 - h.k0 <= kHeavyPasses * kSlots * kSeg)
        heavy_lds_update<T, BW, BLOCK, STAGED>(p, s, h, g, sweep, beta, h0, hx, after);
    else
        heavy_update<T, BW, BLOCK, STAGED>(p, s, h, g, sweep, beta, part, accepted[0], after);
}

// The light bits of one class, [lb, le): no ray entry, at most kLightNb neighbours,
// so dE = lin * delta + lambda * (+-1) per neighbour in CSR order -- the same sum the
// packed step makes for such a bit. A lane takes a (bit, group) task and decides every
// branch of the group, so each load serves BW branches. Waves claim chunks of
// 64 * kLightU tasks from the team's counter (DevState::tickets[64 t + 32]), so ranks
// that spent the class on multi-pass or packed tasks take fewer. A wave's claims end with
// its first claim past the class's last chunk, so a class adds exactly n_chunks + n_waves
// to the counter and every wave knows the next class's base without a reset. Bits of one
// class are independent, so which lane takes which bit does not matter. No LDS and no
// workgroup sync.
constexpr int kLightU = 2;
constexpr int kLightCounter = 32;   // word of a team's 64-word ticket area (its own 128-byte line)

template <typename T, int BW, bool STAGED>
__device__ __forceinline__ void light_sweep(const DevProblem& p, const DevState<T>& s, int lb, int le, int my_groups,
                                            int team, int n_waves, unsigned& base, int sweep, double beta) {
    const int n_light = le - lb;
    if (n_light <= 0) return;
    constexpr int kChunk = 64 * kLightU;
    const int tasks = n_light * my_groups;
    const unsigned n_chunks = unsigned((tasks + kChunk - 1) / kChunk);
    unsigned* ctr = s.tickets + team * 64 + kLightCounter;
    const int lane = int(threadIdx.x) & 63;
    auto claim = [&] {
        unsigned v = 0;
        if (lane == 0) v = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return unsigned(__builtin_amdgcn_readfirstlane(int(v)));
    };
    unsigned got = claim();
    while (got - base < n_chunks) {
        const int c0 = int(got - base) * kChunk;   // the chunk's first task: < tasks
        const int t0 = c0 + lane;
        int pos[kLightU], grp[kLightU];
        bool valid[kLightU];
        int4 qa[kLightU], qb[kLightU];
        double lin[kLightU];
        int bit[kLightU];
        uint64_t key[kLightU];
        uint32_t clk[kLightU];
        uint16_t xm[kLightU];
#pragma unroll
        for (int u = 0; u < kLightU; ++u) {
            const int t = t0 + u * 64;
            valid[u] = t < tasks;
            const int tt = valid[u] ? t : c0;   // lanes past the end load a real task's operands
            const int gi = tt / n_light;
            pos[u] = lb + (tt - gi * n_light);
            grp[u] = team + gi * kTeams;
            const int4* q4 = reinterpret_cast<const int4*>(p.light_nb + size_t(pos[u]) * kLightNb);
            qa[u] = q4[0];
            qb[u] = q4[1];
            lin[u] = p.lin[pos[u]];
            bit[u] = p.draw_bit[pos[u]];
            key[u] = p.draw_key[pos[u]];
            clk[u] = STAGED ? s.clock[pos[u]] : 0u;
            xm[u] = ld_l2(s.x + size_t(grp[u]) * size_t(p.n) + pos[u]);
        }
        got = claim();   // the next chunk, in flight while this one is decided
        uint16_t m[kLightU][kLightNb];
#pragma unroll
        for (int u = 0; u < kLightU; ++u) {
            const uint16_t* xg = s.x + size_t(grp[u]) * size_t(p.n);
            const int q[kLightNb] = {qa[u].x, qa[u].y, qa[u].z, qa[u].w, qb[u].x, qb[u].y, qb[u].z, qb[u].w};
#pragma unroll
            for (int v = 0; v < kLightNb; ++v) m[u][v] = ld_l2(xg + (q[v] >= 0 ? q[v] : pos[u]));
        }
#pragma unroll
        for (int u = 0; u < kLightU; ++u) {
            const int q[kLightNb] = {qa[u].x, qa[u].y, qa[u].z, qa[u].w, qb[u].x, qb[u].y, qb[u].z, qb[u].w};
            uint16_t flips = 0;
#pragma unroll
            for (int b = 0; b < BW; ++b) {
                const int xb = (xm[u] >> b) & 1;
                const double delta = xb ? -1.0 : 1.0;
                double dE = lin[u] * delta;
#pragma unroll
                for (int v = 0; v < kLightNb; ++v) {
                    const int before = ((m[u][v] >> b) & 1) != xb;
                    const double nxt = dE + p.lambda * (before ? -1.0 : 1.0);
                    dE = q[v] >= 0 ? nxt : dE;
                }
                const int branch = grp[u] * BW + b;
                int own = sweep;
                double bt = beta;
                const bool live = bit_time<STAGED>(s, clk[u], branch, sweep, beta, own, bt);
                const double draw =
                    draw_u01(s.seed_offset + key[u] + uint64_t(branch) + 1u, uint32_t(own), uint32_t(bit[u]));
                const bool acc = live && decide_drawn<T>(s, branch, draw, bt, dE);
                flips |= uint16_t(acc ? 1u << b : 0u);
            }
            if (valid[u] && flips) s.x[size_t(grp[u]) * size_t(p.n) + pos[u]] = uint16_t(xm[u] ^ flips);
        }
    }
    base += n_chunks + unsigned(n_waves);
}

// Residuals of ray `ray` for branch lane b, from scratch: -d + sum of a over the
// ray's set bits in increasing original bit index (binary_energy's order).
template <typename T, int BW>
__device__ __forceinline__ void ray_recompute(const DevProblem& p, const DevState<T>& s, int ray, int g, int b) {
    const uint16_t* xg = s.x + size_t(g) * size_t(p.n);
    double r = -p.d[ray];
    const int q0 = p.rt_ptr[ray], q1 = p.rt_ptr[ray + 1];
    for (int qa = q0; qa < q1; qa += kSeg) {
        BitEnt e[kSeg];
        uint16_t m[kSeg];
#pragma unroll
        for (int u = 0; u < kSeg; ++u) e[u] = p.rt[min(qa + u, q1 - 1)];
#pragma unroll
        for (int u = 0; u < kSeg; ++u) m[u] = ld_l2(xg + e[u].pos);
#pragma unroll
        for (int u = 0; u < kSeg; ++u) {
            const double nxt = r + double(e[u].a);
            r = (qa + u < q1 && ((m[u] >> b) & 1)) ? nxt : r;
        }
    }
    s.r[(size_t(g) * size_t(p.n_rays) + size_t(ray)) * BW + b] = T(r);
}

// The persistent sweep. Block b runs on XCD b mod 8 (hardware round-robin); the
// blocks read their XCC_ID to form the teams, take a rank by ticket, and check
// once, behind an agent-scope grid barrier, that every team is complete. Team t
// anneals groups t, t + 8, ... through every class of every sweep.

template <typename T, int BW, int BLOCK, bool STAGED>
__global__ void __launch_bounds__(BLOCK)
ray_team_kernel(DevProblem p, DevBatches bt, DevState<T> s, int sweep_begin, int sweep_end, int recompute_every,
                int team_size) {
    constexpr int kSlots = BLOCK / BW;
    __shared__ int s_team, s_rank, s_ok;
    if (threadIdx.x == 0) {
        s_team = int(xcc_id() & (kTeams - 1));
        s_rank = int(__hip_atomic_fetch_add(s.tickets + s_team * 64, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    if (!grid_barrier(s.bar, gridDim.x, s.err)) return;
    if (threadIdx.x == 0) {
        const unsigned members = __hip_atomic_load(s.tickets + s_team * 64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = members >= unsigned(team_size) ? 1 : 0;
        if (!s_ok) __hip_atomic_store(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int team = s_team, rank = s_rank;
    if (!s_ok || rank >= team_size) return;   // extra blocks on a crowded XCD leave
    unsigned* bar = s.bar + 64 + team * 64;
    const int my_groups = s.n_groups > team ? (s.n_groups - team + kTeams - 1) / kTeams : 0;
    if (my_groups == 0) return;
    const int b = threadIdx.x % BW;
    const int slot = int(threadIdx.x) / BW;
    const int n_slots = team_size * kSlots;
    const int my_slot = rank * kSlots + slot;
    if (p.n_classes == 0) return;
    // class -> batch table and heavy-batch bits, copied to LDS once; then the two block
    // buffers and the work area (sweep_dynamic_lds)
    extern __shared__ __align__(16) int32_t tables[];
    constexpr size_t kBlockBytes = size_t(kSlots) * (kSeg * sizeof(RayEnt) + sizeof(SlotHead));
    unsigned char* pb0 = reinterpret_cast<unsigned char*>(tables) + sweep_tables_bytes_aligned(p.n_classes);
    unsigned char* xa = pb0 + 2 * kBlockBytes;
    __shared__ double s_part[kSlots][BW];
    __shared__ uint32_t s_fx[kSlots];   // a deciding slot's flips | spin mask << 16 (two FP32 branches)
    __shared__ int s_acc[kSlots][BW];
    int32_t* cls = tables;                          // n_classes + 1 batch offsets
    int32_t* nheavy = tables + p.n_classes + 1;    // n_classes multi-pass batch counts
    int32_t* lbeg = nheavy + p.n_classes;           // n_classes first light positions
    int32_t* cend = lbeg + p.n_classes;             // n_classes + 1 class position offsets
    for (int i = int(threadIdx.x); i <= p.n_classes; i += BLOCK) cls[i] = bt.cls_batch[i];
    for (int i = int(threadIdx.x); i < p.n_classes; i += BLOCK) nheavy[i] = bt.cls_heavy[i];
    for (int i = int(threadIdx.x); i < p.n_classes; i += BLOCK) lbeg[i] = p.light_begin[i];
    for (int i = int(threadIdx.x); i <= p.n_classes; i += BLOCK) cend[i] = p.cls_ptr[i];
    __syncthreads();
    // Class c has H = nh * G multi-pass tasks and P = (nb - nh) * G packed tasks, each a
    // (batch, group) pair; Hr = heavy_ranks(H, team) ranks take the multi-pass tasks
    // round-robin and the other ranks the packed ones (with Hr = 0 all tasks, multi-pass
    // first, go round-robin over the team). A rank's k-th task in the class:
    struct Split {
        int H, P, Hr, nh, np;
    };
    auto split_of = [&](int c) {
        const int nh = nheavy[c], np = cls[c + 1] - cls[c] - nh;
        const int H = nh * my_groups, P = np * my_groups;
        return Split{H, P, heavy_ranks(H, team_size), nh, np};
    };
    auto n_tasks = [&](int c) {
        const Split q = split_of(c);
        if (q.Hr == 0) {
            const int total = q.H + q.P;
            return rank < total ? (total - rank + team_size - 1) / team_size : 0;
        }
        if (rank < q.Hr) return (q.H - rank + q.Hr - 1) / q.Hr;
        const int r = rank - q.Hr, n = team_size - q.Hr;
        return r < q.P ? (q.P - r + n - 1) / n : 0;
    };
    // With one group per team (Hr = 0 always) a rank's k-th task in class c is the class's
    // batch rank + k * team_size, so the cursor needs no division.
    const bool one_group = my_groups == 1 && heavy_ranks(1, team_size) == 0;
    // the batch and group of this rank's k-th task in class c; heavy = multi-pass
    auto task_of = [&](int c, int k, int& g, bool& heavy) {
        if (one_group) {
            const int t = rank + k * team_size;
            heavy = t < nheavy[c];
            g = team;
            return cls[c] + t;
        }
        const Split q = split_of(c);
        int h = -1, pk = -1;
        if (q.Hr == 0) {
            const int t = rank + k * team_size;
            if (t < q.H) h = t;
            else pk = t - q.H;
        } else if (rank < q.Hr) {
            h = rank + k * q.Hr;
        } else {
            pk = (rank - q.Hr) + k * (team_size - q.Hr);
        }
        heavy = h >= 0;
        if (heavy) {
            const int gi = h / q.nh;
            g = team + gi * kTeams;
            return cls[c] + (h - gi * q.nh);
        }
        const int gi = pk / q.np;
        g = team + gi * kTeams;
        return cls[c] + q.nh + (pk - gi * q.np);
    };
    // Two lookahead cursors: (s1, c1, k1) is the task after the one being processed,
    // (s2, c2, k2) the one after that. A step issues the copy of task 1's block into the
    // other buffer once its own chain sums are made (its state loads have returned, so
    // the copy does not delay them), and advances cursor 2 while its state loads are in
    // flight; the copy lands while the step decides, scatters and waits at the barrier.
    auto advance = [&](int& sw, int& c, int& k) {
        ++k;
        if (one_group) {
            while (rank + k * team_size >= cls[c + 1] - cls[c]) {
                k = 0;
                if (++c == p.n_classes) {
                    c = 0;
                    if (++sw >= sweep_end) return false;
                }
            }
            return true;
        }
        while (k >= n_tasks(c)) {
            k = 0;
            if (++c == p.n_classes) {
                c = 0;
                if (++sw >= sweep_end) return false;
            }
        }
        return true;
    };
    // a batch's {first granule, slots used}, by a vector load (a scalar one would hold up
    // the step's LDS waits until it returned)
    auto meta_of = [&](int batch) {
        const uint64_t v = ld_l2(reinterpret_cast<const uint64_t*>(bt.blk_meta) + batch);
        return int2{int(uint32_t(v)), int(uint32_t(v >> 32))};
    };
    int s1 = sweep_begin, c1 = 0, k1 = -1;
    bool a1 = advance(s1, c1, k1);
    int s2 = s1, c2 = c1, k2 = k1;
    bool a2 = a1 && advance(s2, c2, k2);
    int b1 = 0, g1 = 0, b2 = 0, g2 = 0;
    bool h1 = false, h2 = false;
    if (a1) b1 = task_of(c1, k1, g1, h1);
    if (a2) b2 = task_of(c2, k2, g2, h2);
    // advances cursor 2 once per task: in the barrier wait after the task, or, when the
    // next task follows in the same class, before it
    bool cursor_due = false;
    auto advance_cursor = [&] {
        if (cursor_due && a2) {
            a2 = advance(s2, c2, k2);
            if (a2) b2 = task_of(c2, k2, g2, h2);
        }
        cursor_due = false;
    };
    int2 meta_cur = int2{0, 0};
    if (a1 && !h1) {
        meta_cur = meta_of(b1);
        issue_block<BLOCK, kSlots>(bt.blk, meta_cur, pb0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int buf = 0;
    bool synced = true;   // the current block is visible to every wave (a barrier since its copy)
    // The next packed task's prep: made in the barrier wait before it (its block has
    // landed and every wave has passed the barrier's drain), else at its start.
    PackedPrep<T, BW> prep;
    bool prepped = false;
    double beta_now = 0.0, beta_next = 0.0;
    int sweep_now = sweep_begin;
    auto at_arrival = [&] {
        if (a1 && !h1 && !prepped) {
            const double b1beta = STAGED ? 0.0 : s1 == sweep_now ? beta_now : s1 == sweep_now + 1 ? beta_next
                                                                                            : s.beta[s1];
            packed_prep<T, BW, BLOCK, STAGED>(s, pb0 + size_t(buf) * kBlockBytes,
                                              __builtin_amdgcn_readfirstlane(meta_cur.y), g1, s1, b1beta, prep);
            prepped = true;
        }
        advance_cursor();
    };
    unsigned target = 0;
    unsigned light_base = 0;   // the team's light-task counter at this class's first claim
    const int n_waves = team_size * (BLOCK / 64);
    for (int sweep = sweep_begin; sweep < sweep_end; ++sweep) {
        // the staged form reads beta per bit, at its own sweep; the next sweep's is at hand
        // for a prep made in this sweep's last barrier
        const double beta = STAGED ? 0.0 : s.beta[sweep];
        beta_now = beta;
        beta_next = STAGED || sweep + 1 >= sweep_end ? 0.0 : s.beta[sweep + 1];
        sweep_now = sweep;
        for (int c = 0; c < p.n_classes; ++c) {
            while (a1 && s1 == sweep && c1 == c) {
                if (!synced) {
                    // a second task of this class: its block was copied during the last step
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                }
                synced = false;
                advance_cursor();
                const int batch = b1, g = g1;
                const bool heavy = h1;
                const int n_used = __builtin_amdgcn_readfirstlane(meta_cur.y);
                if (!heavy && !prepped)
                    packed_prep<T, BW, BLOCK, STAGED>(s, pb0 + size_t(buf) * kBlockBytes, n_used, g, sweep, beta, prep);
                prepped = false;
                // shift the cursors: the next task is the old task 2
                s1 = s2;
                c1 = c2;
                k1 = k2;
                a1 = a2;
                b1 = b2;
                g1 = g2;
                h1 = h2;
                cursor_due = true;
                const bool pf = a1 && !h1;
                const int2 meta_nxt = pf ? meta_of(b1) : int2{0, 0};
                unsigned char* pc = pb0 + size_t(buf) * kBlockBytes;
                unsigned char* pn = pb0 + size_t(buf ^ 1) * kBlockBytes;
                auto next_block = [&] {
                    if (pf) issue_block<BLOCK, kSlots>(bt.blk, meta_nxt, pn);
                };
                if (heavy) [[unlikely]] {
                    heavy_step<T, BW, BLOCK, STAGED>(p, s, bt, batch, g, sweep, beta, reinterpret_cast<RayEnt*>(pc),
                                                     reinterpret_cast<RayEnt*>(xa), s_part, s_acc, next_block);
                } else {
                    packed_step<T, BW, BLOCK, STAGED>(p, s, pc, prep, reinterpret_cast<T*>(xa), s_part, s_acc, s_fx, g,
                                                      next_block);
                }
                meta_cur = meta_nxt;
                buf ^= 1;
            }
            if (lbeg[c] < cend[c + 1]) [[unlikely]]   // few classes hold light bits: keep their code out of line
                light_sweep<T, BW, STAGED>(p, s, lbeg[c], cend[c + 1], my_groups, team, n_waves, light_base, sweep, beta);
            target += unsigned(team_size);
            if (!team_barrier(bar, rank, team_size, target, s.err, at_arrival)) return;
            synced = true;
        }
        if (recompute_every > 0 && (sweep + 1) % recompute_every == 0) [[unlikely]] {
            const int tasks = p.n_rays * my_groups;
            for (int t = my_slot; t < tasks; t += n_slots) {
                const int gi = t / p.n_rays;
                ray_recompute<T, BW>(p, s, t - gi * p.n_rays, team + gi * kTeams, b);
            }
            target += unsigned(team_size);
            if (!team_barrier(bar, rank, team_size, target, s.err)) return;
        }
    }
}

// One thread per (batch, slot) of the packed batches: writes a used slot's chain
// (entries past its length: ray -1, a 0, w 0) and head into its batch's block.
__global__ void batch_blocks_kernel(DevProblem p, const SlotDesc* desc, const int2* meta, int n_batches, int slots,
                                    uint4* blk) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(n_batches) * slots) return;
    const int bi = int(t / slots), sl = int(t - int64_t(bi) * slots);
    const int2 m = meta[bi];
    if (sl >= m.y) return;
    const SlotDesc d = desc[t];
    RayEnt* E = reinterpret_cast<RayEnt*>(blk + m.x);
    const int len = slot_len(d.info), nn = slot_nnb(d.info);
    for (int u = 0; u < kSeg; ++u) E[blk_entry(sl, u)] = u < len ? p.ent[d.ka + u] : RayEnt{-1, 0.0f, 0.0};
    SlotHead h;
    h.pos = d.pos;
    h.info = d.info;
    h.nb0 = d.nb0;
    h.bit = p.draw_bit[d.pos];
    for (int u = 0; u < kNb; ++u) h.q[u] = nn > 0 ? p.nb[d.nb0 + min(u, nn - 1)] : d.pos;
    h.lin = p.lin[d.pos];
    h.key = p.draw_key[d.pos];
    h.pad[0] = 0;
    h.pad[1] = 0;
    reinterpret_cast<SlotHead*>(blk + m.x + int64_t(m.y) * kSeg)[sl] = h;
}

hipError_t launch_batch_blocks(const DevProblem& p, const SlotDesc* desc, const int2* meta, int n_batches, int slots,
                               uint4* blk, hipStream_t st) {
    const int64_t items = int64_t(n_batches) * slots;
    if (items <= 0) return hipSuccess;
    batch_blocks_kernel<<<unsigned((items + 255) / 256), 256, 0, st>>>(p, desc, meta, n_batches, slots, blk);
    return hipGetLastError();
}

template <typename T, int BW>
__global__ void ray_init_kernel(DevProblem p, uint16_t* x, T* r, int n_groups) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t nx = size_t(n_groups) * size_t(p.n);
    const size_t nr = size_t(n_groups) * size_t(p.n_rays) * BW;
    if (i < nx) x[i] = 0;
    if (i < nr) r[i] = T(-p.d[(i / BW) % size_t(p.n_rays)]);
}

// Branch agreement (branch_fractions' p_v, ising_recon.hpp:160-163, and classify,
// :305-309) plus certainty max(p, 1 - p); outputs in original bit order.
__global__ void ray_outputs_kernel(DevProblem p, const uint16_t* x, int n_groups, int bw, int R, float* p_void,
                                   uint8_t* tri, float* certainty, uint8_t* branches) {
    const int pos = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (pos >= p.n) return;
    const int bit = p.pos_bit[pos];
    int cnt = 0;
    for (int g = 0; g < n_groups; ++g) {
        const int live = R - g * bw < bw ? R - g * bw : bw;
        const unsigned mask = (1u << live) - 1u;
        const unsigned m = unsigned(x[size_t(g) * size_t(p.n) + pos]) & mask;
        cnt += __popc(m);
        if (branches)
            for (int b = 0; b < live; ++b)
                branches[size_t(g * bw + b) * size_t(p.n) + size_t(bit)] = uint8_t((m >> b) & 1u);
    }
    const float pv = float(cnt) / float(R);
    if (p_void) p_void[bit] = pv;
    if (tri) tri[bit] = pv >= 0.9f ? 1 : (pv <= 0.1f ? 0 : 2);
    if (certainty) certainty[bit] = pv > 1.0f - pv ? pv : 1.0f - pv;
}

// ---- FP64 energies -------------------------------------------------------------------
constexpr int kEnergyBlock = 256;

// Data term per (ray, branch lane): 1/2 w r^2 with r from scratch, summed over the
// block's rays per lane; also max |r_carried - r_scratch|.
template <typename T, int BW>
__global__ void __launch_bounds__(kEnergyBlock)
ray_energy_data_kernel(DevProblem p, const uint16_t* x, const T* r_carried, double* part, int part_stride,
                       float* resid_err) {
    constexpr int kItems = kEnergyBlock / BW;
    __shared__ double red[kEnergyBlock];
    const int b = threadIdx.x % BW;
    const int g = blockIdx.y;
    const int ray = blockIdx.x * kItems + int(threadIdx.x / BW);
    double v = 0.0;
    float err = 0.0f;
    if (ray < p.n_rays) {
        const uint16_t* xg = x + size_t(g) * size_t(p.n);
        double r = -p.d[ray];
        for (int k = p.rt_ptr[ray]; k < p.rt_ptr[ray + 1]; ++k) {
            const BitEnt e = p.rt[k];
            if ((xg[e.pos] >> b) & 1) r += double(e.a);
        }
        v = 0.5 * p.w[ray] * r * r;
        if (r_carried) {
            const double rc = double(r_carried[(size_t(g) * size_t(p.n_rays) + size_t(ray)) * BW + b]);
            err = float(fabs(rc - r));
        }
    }
    red[threadIdx.x] = v;
    if (r_carried && err > 0.0f) atomicMax(reinterpret_cast<unsigned*>(resid_err), __float_as_uint(err));
    __syncthreads();
    for (int stride = kEnergyBlock / 2; stride >= BW; stride >>= 1) {
        if (int(threadIdx.x) < stride) red[threadIdx.x] += red[threadIdx.x + stride];
        __syncthreads();
    }
    if (int(threadIdx.x) < BW) part[size_t(blockIdx.x) * part_stride + size_t(g) * BW + threadIdx.x] = red[threadIdx.x];
}

// Prior term per (position, branch lane): lin x + lambda per unlike neighbour j > i.
template <int BW>
__global__ void __launch_bounds__(kEnergyBlock)
ray_energy_prior_kernel(DevProblem p, const uint16_t* x, double* part, int part_stride) {
    constexpr int kItems = kEnergyBlock / BW;
    __shared__ double red[kEnergyBlock];
    const int b = threadIdx.x % BW;
    const int g = blockIdx.y;
    const int pos = blockIdx.x * kItems + int(threadIdx.x / BW);
    double v = 0.0;
    if (pos < p.n) {
        const uint16_t* xg = x + size_t(g) * size_t(p.n);
        const int xi = (xg[pos] >> b) & 1;
        const int bit = p.pos_bit[pos];
        v = p.lin[pos] * double(xi);
        for (int k = p.nb_ptr[pos]; k < p.nb_ptr[pos + 1]; ++k)
            if (p.nb_bit[k] > bit && ((xg[p.nb[k]] >> b) & 1) != xi) v += p.lambda;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int stride = kEnergyBlock / 2; stride >= BW; stride >>= 1) {
        if (int(threadIdx.x) < stride) red[threadIdx.x] += red[threadIdx.x + stride];
        __syncthreads();
    }
    if (int(threadIdx.x) < BW) part[size_t(blockIdx.x) * part_stride + size_t(g) * BW + threadIdx.x] = red[threadIdx.x];
}

// Sums the partials of each branch in block order (deterministic).
__global__ void ray_energy_final_kernel(const double* part_data, int blocks_data, const double* part_prior,
                                        int blocks_prior, int stride, int R, double* data, double* prior,
                                        double* total) {
    const int br = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (br >= R) return;
    double d = 0.0, q = 0.0;
    for (int k = 0; k < blocks_data; ++k) d += part_data[size_t(k) * stride + br];
    for (int k = 0; k < blocks_prior; ++k) q += part_prior[size_t(k) * stride + br];
    data[br] = d;
    prior[br] = q;
    if (total) total[br] = d + q;
}

// K states (original bit order, one byte per bit) into 16-bit group masks.
__global__ void pack_states_kernel(const uint8_t* X, int K, int n, const int32_t* pos_bit, uint16_t* x) {
    const int pos = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int g = blockIdx.y;
    if (pos >= n) return;
    const int bit = pos_bit[pos];
    unsigned m = 0;
    for (int b = 0; b < kGroup; ++b) {
        const int k = g * kGroup + b;
        if (k < K && X[size_t(k) * size_t(n) + size_t(bit)]) m |= 1u << b;
    }
    x[size_t(g) * size_t(n) + pos] = uint16_t(m);
}

// ---------------------------------------------------------------------------------
// Integer path
// ---------------------------------------------------------------------------------
template <int BLOCK, bool LDS>
__global__ void __launch_bounds__(BLOCK)
int_sweep_kernel(DevGraph g, uint64_t seed_base, const double* beta, int sweep_begin, int sweep_end, int init,
                 int8_t* s_state, int32_t* f_state, long long* energy) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ long long red[BLOCK / 64];
    const int r = blockIdx.x;
    const int n = g.n, np = g.n_pad;
    int8_t* gs = s_state + size_t(r) * size_t(np);
    int32_t* gf = f_state + size_t(r) * size_t(np);
    int32_t* f = LDS ? reinterpret_cast<int32_t*>(smem) : gf;
    int8_t* s = LDS ? reinterpret_cast<int8_t*>(smem + size_t(np) * 4) : gs;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    if (init) {
        for (int i = tid; i < n; i += BLOCK) { f[i] = g.f0[i]; s[i] = -1; }
    } else if (LDS) {
        for (int i = tid; i < n; i += BLOCK) { f[i] = gf[i]; s[i] = gs[i]; }
    }
    __syncthreads();
    const uint64_t seed = seed_base + uint64_t(r);
    long long dsum = 0;
    for (int sweep = sweep_begin; sweep < sweep_end; ++sweep) {
        const double bt = beta[sweep];
        for (int c = 0; c < g.n_classes; ++c) {
            const int c0 = g.cls_ptr[c], c1 = g.cls_ptr[c + 1];
            for (int base = c0; base < c1; base += BLOCK) {
                const int pos = base + tid;
                int si = 0;
                bool acc = false;
                long long dE = 0;
                if (pos < c1) {
                    int fi;
                    if (LDS) {
                        si = s[pos];
                        fi = f[pos];
                    } else {
                        si = __hip_atomic_load(&s[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        fi = __hip_atomic_load(&f[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    dE = -(long long)si * (long long)fi;          // State::dE, ising_weighted.hpp:209
                    acc = dE <= 0 || metropolis_uphill(double(dE), bt, seed, uint32_t(sweep),
                                                       uint32_t(g.pos_bit[pos]));
                    if (acc) dsum += dE;
                }
                uint64_t m = __ballot(acc);
                while (m) {                                      // State::flip, :210-217
                    const int L = __builtin_ctzll(m);
                    m &= m - 1;
                    const int pj = __shfl(pos, L);
                    const int sj = __shfl(si, L);
                    for (int k = g.nb_ptr[pj] + lane; k < g.nb_ptr[pj + 1]; k += 64)
                        atomicAdd(&f[g.nb[k]], -2 * g.nw[k] * sj);
                    if (lane == 0) s[pj] = int8_t(-sj);
                }
            }
            // global state: this wave's field atomics and spin stores complete before
            // the barrier (s_barrier alone does not wait for them)
            if (!LDS) __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
    if (LDS)
        for (int i = tid; i < n; i += BLOCK) { gf[i] = f[i]; gs[i] = s[i]; }
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_down(dsum, o);
    if (lane == 0) red[tid >> 6] = dsum;
    __syncthreads();
    if (tid == 0) {
        long long t = 0;
        for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
        energy[r] = (init ? 0 : energy[r]) + t;
    }
}

__global__ void int_unpermute_kernel(DevGraph g, int R, const int8_t* s_state, int8_t* s_out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = size_t(R) * size_t(g.n_pad);
    if (i >= total) return;
    const int r = int(i / size_t(g.n_pad));
    const int pos = int(i % size_t(g.n_pad));
    if (pos < g.n)
        s_out[size_t(r) * g.n_pad + size_t(g.pos_bit[pos])] = s_state[size_t(r) * g.n_pad + pos];
    else
        s_out[size_t(r) * g.n_pad + size_t(pos)] = 0;
}

// -cut of state blockIdx.x over the edge list (oracle_kernel, ising_gpu.cuh:257-272).
__global__ void __launch_bounds__(256) int_energy_kernel(DevGraph g, const uint8_t* X, long long* out) {
    __shared__ long long red[4];
    const uint8_t* x = X + size_t(blockIdx.x) * size_t(g.n);
    long long acc = 0;
    for (int64_t e = threadIdx.x; e < g.m; e += 256)
        if (x[g.eu[e]] != x[g.ev[e]]) acc += g.ew[e];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = -(red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------
// Dispatch on the group width: fn.template operator()<BW>().
template <typename F>
static hipError_t with_bw(int bw, F&& fn) {
    switch (bw) {
        case 2: return fn.template operator()<2>();
        case 4: return fn.template operator()<4>();
        case 8: return fn.template operator()<8>();
        case 16: return fn.template operator()<16>();
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ray_init(const DevProblem& p, int precision, int bw, void* x, void* r, int n_groups, hipStream_t st) {
    const size_t nx = size_t(n_groups) * size_t(p.n);
    const size_t nr = size_t(n_groups) * size_t(p.n_rays) * size_t(bw);
    const size_t total = nx > nr ? nx : nr;
    const unsigned blocks = unsigned((total + 255) / 256);
    return with_bw(bw, [&]<int BW>() {
        if (precision == 1)
            ray_init_kernel<float, BW><<<blocks, 256, 0, st>>>(p, static_cast<uint16_t*>(x), static_cast<float*>(r), n_groups);
        else
            ray_init_kernel<double, BW><<<blocks, 256, 0, st>>>(p, static_cast<uint16_t*>(x), static_cast<double*>(r), n_groups);
        return hipGetLastError();
    });
}

hipError_t launch_ray_sweeps(const DevProblem& p, const DevBatches& bt, int precision, int bw, void* x, void* r,
                             int n_groups, int R,
                             uint64_t seed_offset, const double* beta, int sb, int se, int recompute_every,
                             unsigned* bar, unsigned* tickets, int* err, int team_size, hipStream_t st,
                             const uint32_t* clock, int total_sweeps) {
    return with_bw(bw, [&]<int BW>() {
        auto go = [&](auto zero, auto staged) -> hipError_t {
            using T = decltype(zero);
            constexpr bool STAGED = decltype(staged)::value;
            DevState<T> s;
            s.n_groups = n_groups;
            s.R = R;
            s.seed_offset = seed_offset;
            s.x = static_cast<uint16_t*>(x);
            s.r = static_cast<T*>(r);
            s.beta = beta;
            s.bar = bar;
            s.tickets = tickets;
            s.err = err;
            s.clock = clock;
            s.total_sweeps = total_sweeps;
            // A plain launch of at most the co-resident workgroup count (the host checks it
            // against the occupancy with this dynamic LDS). Not a cooperative launch: the
            // kernel's barriers are its own atomics, and a cooperative launch takes the
            // device's global wave sync resource, which one process holds at a time, so a
            // second process on the GPU would be refused. A workgroup that is not resident
            // leaves its team short or a barrier waiting: both end in the error flag.
            const size_t lds = sweep_dynamic_lds(p.n_classes, BW, bt.heavy_lds != 0, precision);
            const hipError_t e =
                hipFuncSetAttribute(reinterpret_cast<const void*>(&ray_team_kernel<T, BW, kSweepBlock, STAGED>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (e != hipSuccess) return e;
            ray_team_kernel<T, BW, kSweepBlock, STAGED>
                <<<dim3(unsigned(team_size * kTeams)), dim3(kSweepBlock), lds, st>>>(p, bt, s, sb, se, recompute_every,
                                                                                   team_size);
            return hipGetLastError();
        };
        if (clock)
            return precision == 1 ? go(0.0f, std::true_type{}) : go(0.0, std::true_type{});
        return precision == 1 ? go(0.0f, std::false_type{}) : go(0.0, std::false_type{});
    });
}

int ray_sweep_blocks_per_cu(int precision, int bw, size_t table_bytes, bool staged) {
    int per_cu = 0;
    const hipError_t e = with_bw(bw, [&]<int BW>() {
        const void* k =
            staged ? (precision == 1 ? reinterpret_cast<const void*>(&ray_team_kernel<float, BW, kSweepBlock, true>)
                                     : reinterpret_cast<const void*>(&ray_team_kernel<double, BW, kSweepBlock, true>))
                   : (precision == 1 ? reinterpret_cast<const void*>(&ray_team_kernel<float, BW, kSweepBlock, false>)
                                     : reinterpret_cast<const void*>(&ray_team_kernel<double, BW, kSweepBlock, false>));
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kSweepBlock, table_bytes);
    });
    return e == hipSuccess ? per_cu : 0;
}

hipError_t launch_ray_outputs(const DevProblem& p, const uint16_t* x, int n_groups, int bw, int R, float* p_void,
                              uint8_t* tri, float* certainty, uint8_t* branches, hipStream_t st) {
    ray_outputs_kernel<<<unsigned((p.n + 255) / 256), 256, 0, st>>>(p, x, n_groups, bw, R, p_void, tri, certainty,
                                                                     branches);
    return hipGetLastError();
}

// Partial-sum blocks of the energy kernels at group width bw.
int ray_energy_blocks(const DevProblem& p, int bw) {
    const int items = kEnergyBlock / bw;
    return (p.n_rays + items - 1) / items + (p.n + items - 1) / items;
}

hipError_t launch_ray_energy(const DevProblem& p, const uint16_t* x, int n_groups, int bw, int R, const void* r_carried,
                             int precision, double* part, int n_part_blocks, double* data, double* prior,
                             float* resid_err, hipStream_t st, double* total) {
    const int items = kEnergyBlock / bw;
    const int bd = (p.n_rays + items - 1) / items;
    const int bp = (p.n + items - 1) / items;
    if (bd + bp > n_part_blocks) return hipErrorInvalidValue;
    const int stride = n_groups * bw;
    double* part_prior = part + size_t(bd) * stride;
    return with_bw(bw, [&]<int BW>() {
        if (bd > 0) {
            const dim3 grid{unsigned(bd), unsigned(n_groups), 1u};
            if (precision == 1)
                ray_energy_data_kernel<float, BW><<<grid, kEnergyBlock, 0, st>>>(
                    p, x, static_cast<const float*>(r_carried), part, stride, resid_err);
            else
                ray_energy_data_kernel<double, BW><<<grid, kEnergyBlock, 0, st>>>(
                    p, x, static_cast<const double*>(r_carried), part, stride, resid_err);
        }
        if (bp > 0)
            ray_energy_prior_kernel<BW><<<dim3(unsigned(bp), unsigned(n_groups)), kEnergyBlock, 0, st>>>(p, x, part_prior,
                                                                                                      stride);
        ray_energy_final_kernel<<<unsigned((R + 63) / 64), 64, 0, st>>>(part, bd, part_prior, bp, stride, R, data, prior,
                                                                              total);
        return hipGetLastError();
    });
}

hipError_t launch_pack_states(const uint8_t* X, int K, int n, const int32_t* pos_bit, uint16_t* x, hipStream_t st) {
    const int n_groups = (K + kGroup - 1) / kGroup;
    pack_states_kernel<<<dim3(unsigned((n + 255) / 256), unsigned(n_groups)), 256, 0, st>>>(X, K, n, pos_bit, x);
    return hipGetLastError();
}

constexpr int kIntLdsBytes = 64 * 1024;

int int_lds_capacity_spins() { return kIntLdsBytes / 5 - 16; }

hipError_t launch_int_sweeps(const DevGraph& g, int R, uint64_t seed_base, const double* beta, int sb, int se,
                             int init, int8_t* s_state, int32_t* f_state, long long* energy, bool lds_state,
                             hipStream_t st) {
    const int max_class_block = 256;
    (void)max_class_block;
    if (lds_state) {
        const size_t smem = size_t(g.n_pad) * 5;
        if (smem > size_t(kIntLdsBytes)) return hipErrorInvalidValue;
        int_sweep_kernel<64, true><<<unsigned(R), 64, smem, st>>>(g, seed_base, beta, sb, se, init, s_state, f_state,
                                                                  energy);
    } else {
        int_sweep_kernel<256, false><<<unsigned(R), 256, 0, st>>>(g, seed_base, beta, sb, se, init, s_state,
                                                                  f_state, energy);
    }
    return hipGetLastError();
}

// ---- the reference's own trajectory (wising::anneal, src/ising_weighted.hpp:250-267) ----
// Replica r anneals from s = -1 with seed_base + r: bits in index order, one splitmix64
// stream (state seed * 0x2545F4914F6CDD1D + 1) drawn only for uphill proposals, a flip
// taken iff dE <= 0 or u < exp(-dE / T) with T per sweep (the host's expression). One wave
// per replica, its spins and integer fields in LDS, positions of the colour layout:
// the wave evaluates the next 64 bits at once against the current state. The k-th uphill
// proposal of the window draws from state + k * gamma (splitmix64 is a counter: the n-th
// draw mixes state_0 + n * gamma), so every lane decides its own bit as the sequential
// loop would if no earlier bit of the window flipped. The first accepted bit is the one
// the loop flips: the draws up to it are consumed, it flips (its row's fields updated
// by the wave's lanes), and the later lanes keep their decisions up to the first lane
// whose field the flip changed, where the next window starts. With none accepted the
// whole window is consumed.
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// u < exp(-dE / T) for u = u53 2^-53 and dE > 0, with x = -dE / T the host's quotient.
// A float exponential of xa = dE * (-1 / T) decides outside a band of 2^-14 around it:
// xa is within 1e-14 of x, and the float exponential's error, the rounding of xa to float
// included, is below 2e-5 for -79 < xa <= 0. Inside the band, and for the rare u = 0,
// the exact quotient and the FP64 qexp decide. qexp and the host's libm exp differ by at
// most an ulp or two, so a decision can differ from the host's only for a u within that
// distance of the exponential.
__device__ __forceinline__ bool ref_accept(uint64_t u53, long long dE, double T, double nrT) {
    const double xa = double(dE) * nrT;
    const double u = double(u53) * 0x1.0p-53;
    if (!(xa > -79.0)) return u < qexp(-double(dE) / T);   // only u = 0 can pass below e^-79
    constexpr double kBand = 0x1.0p-14;
    const double e = double(__builtin_amdgcn_exp2f(float(xa) * 1.44269504f));
    if (u < e * (1.0 - kBand)) return true;
    if (u > e * (1.0 + kBand)) return false;
    return u < qexp(-double(dE) / T);
}

// IN_LDS: a replica's spins and fields are copied into LDS for the launch (n_pad * 5 bytes
// a replica); otherwise they are worked on where they lie in global memory, read through
// the CU's L1, which a wave's own completed stores update (larger graphs).
template <bool IN_LDS>
__device__ __forceinline__ void int_ref_sync() {
    if constexpr (IN_LDS) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

template <int WAVES, bool IN_LDS>
__global__ void __launch_bounds__(64 * WAVES) int_ref_kernel(DevGraph g, int R, uint64_t seed_base, const double* temp,
                                                              int sweep_begin, int sweep_end, int init,
                                                              int8_t* s_state, int32_t* f_state, uint64_t* rng,
                                                              long long* energy) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = int(threadIdx.x) / 64, lane = int(threadIdx.x) % 64;
    const int r = int(blockIdx.x) * WAVES + w;
    if (r >= R) return;   // the whole wave leaves; waves never sync with each other
    const int n = g.n, np = g.n_pad;
    int8_t* gs = s_state + size_t(r) * size_t(np);
    int32_t* gf = f_state + size_t(r) * size_t(np);
    int32_t* f = IN_LDS ? reinterpret_cast<int32_t*>(smem + size_t(w) * size_t(np) * 5) : gf;
    int8_t* s = IN_LDS ? reinterpret_cast<int8_t*>(f + np) : gs;
    uint64_t st;
    long long e;
    if (init) {
        for (int p = lane; p < np; p += 64) {
            s[p] = -1;
            f[p] = p < n ? g.f0[p] : 0;
        }
        st = (seed_base + uint64_t(r)) * 0x2545F4914F6CDD1Dull + 1;
        e = 0;
    } else {
        if constexpr (IN_LDS) {
            for (int p = lane; p < np; p += 64) {
                s[p] = gs[p];
                f[p] = gf[p];
            }
        }
        st = rng[r];
        e = energy[r];
    }
    int_ref_sync<IN_LDS>();
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    for (int sweep = sweep_begin; sweep < sweep_end; ++sweep) {
        const double T = temp[sweep];
        const double nrT = -1.0 / T;
        int i0 = 0;
        while (i0 < n) {
            const int i = i0 + lane;
            const bool valid = i < n;
            const int p = valid ? g.bit_pos[i] : 0;
            // each lane's adjacency row bounds, loaded with the window so that a flip does not
            // wait on them
            const int row0 = valid ? g.nb_ptr[p] : 0, row1 = valid ? g.nb_ptr[p + 1] : 0;
            const int32_t fp = f[p];
            const long long dE = valid ? -(long long)s[p] * (long long)fp : 0;
            const bool up = valid && dE > 0;
            const uint64_t m_up = __ballot(up);
            bool acc = valid && dE <= 0;
            if (up) acc = ref_accept(mix64(st + uint64_t(__popcll(m_up & below) + 1) * kGamma) >> 11, dE, T, nrT);
            const uint64_t m_acc = __ballot(acc);
            // Lanes in [lo, lim) hold decisions made on the current state: a flip changes
            // no s but its own and only its neighbours' f, so a later lane whose f is
            // unchanged keeps its dE, its uphill status and (all earlier lanes being
            // unchanged too) its draw. The first lane whose f changed bounds lim.
            uint64_t todo = ~uint64_t(0);   // lanes >= lo
            uint64_t ok = ~uint64_t(0);     // lanes < lim
            for (;;) {
                const uint64_t cand = m_acc & todo & ok;
                if (cand == 0) {
                    st += uint64_t(__popcll(m_up & todo & ok)) * kGamma;
                    i0 += ok == ~uint64_t(0) ? 64 : __builtin_ctzll(~ok);
                    break;
                }
                const int js = __builtin_ctzll(cand);
                const uint64_t upto = js == 63 ? ~uint64_t(0) : (uint64_t(2) << js) - 1u;
                st += uint64_t(__popcll(m_up & todo & upto)) * kGamma;
                const int pf = __shfl(p, js), q0 = __shfl(row0, js), q1 = __shfl(row1, js);
                e += __shfl(dE, js);
                const int si = s[pf];
                for (int q = q0 + lane; q < q1; q += 64) f[g.nb[q]] -= 2 * g.nw[q] * si;
                if (lane == 0) s[pf] = int8_t(-si);
                int_ref_sync<IN_LDS>();
                todo = ~upto;
                const uint64_t m_dirty = __ballot(valid && f[p] != fp) & todo;
                if (m_dirty) ok &= (uint64_t(1) << __builtin_ctzll(m_dirty)) - 1u;
            }
        }
    }
    if constexpr (IN_LDS) {
        for (int p = lane; p < np; p += 64) {
            gs[p] = s[p];
            gf[p] = f[p];
        }
    }
    if (lane == 0) {
        rng[r] = st;
        energy[r] = e;
    }
}

// Where a replica's state lives: LDS, four replicas per workgroup while their states fit
// the LDS limit, else one; global memory beyond one replica's fit (four per workgroup).
static int int_ref_waves(const DevGraph& g) { return size_t(4) * size_t(g.n_pad) * 5 <= kRefLdsLimit ? 4 : 1; }

bool int_ref_in_lds(const DevGraph& g) { return size_t(g.n_pad) * 5 <= kRefLdsLimit; }

size_t int_ref_lds_bytes(const DevGraph& g) {
    return int_ref_in_lds(g) ? size_t(int_ref_waves(g)) * size_t(g.n_pad) * 5 : 0;
}

hipError_t launch_int_ref(const DevGraph& g, int R, uint64_t seed_base, const double* temp, int sweep_begin,
                          int sweep_end, int init, int8_t* s_state, int32_t* f_state, uint64_t* rng,
                          long long* energy, hipStream_t st) {
    if (!int_ref_in_lds(g)) {
        int_ref_kernel<4, false><<<unsigned((R + 3) / 4), 256, 0, st>>>(g, R, seed_base, temp, sweep_begin, sweep_end,
                                                                         init, s_state, f_state, rng, energy);
        return hipGetLastError();
    }
    const size_t lds = int_ref_lds_bytes(g);
    const int waves = int_ref_waves(g);
    const void* k = waves == 4 ? reinterpret_cast<const void*>(&int_ref_kernel<4, true>)
                               : reinterpret_cast<const void*>(&int_ref_kernel<1, true>);
    const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    if (waves == 4)
        int_ref_kernel<4, true><<<unsigned((R + 3) / 4), 256, lds, st>>>(g, R, seed_base, temp, sweep_begin, sweep_end,
                                                                         init, s_state, f_state, rng, energy);
    else
        int_ref_kernel<1, true><<<unsigned(R), 64, lds, st>>>(g, R, seed_base, temp, sweep_begin, sweep_end, init,
                                                              s_state, f_state, rng, energy);
    return hipGetLastError();
}

// ---- the reference's own float trajectory (anneal_branch, src/ising_recon.hpp:119-143) ----
// Branch b anneals from x = 0, r = -d with seed b + 1 + seed_offset: bits in index order,
// dE = local_field_dE (lin delta, then w (r a + a^2 / 2) per entry with a = a_k delta in CSR
// order, then lambda (+-1) per neighbour in CSR order), one splitmix64 stream drawn only
// for uphill proposals, a flip iff dE <= 0 or u < exp(-dE / T), and on a flip r += a_k delta
// per entry. One wave per branch; x (bytes, bit order) and r (FP64) of a branch live in
// global memory, read through the CU's L1 (a wave sees its own completed stores). As in the
// integer trajectory, the wave evaluates the next 64 bits at once -- each lane sums its
// bit's dE alone, in the reference's order -- and the first accepted bit is the loop's next
// flip. A flip of a bit with rays ends the window; after a ray-free bit's flip the later
// lanes only sum their neighbour terms again (the last part of the reference's sum) and
// decide anew. Draw keys and counters of unions (draw_bit, draw_key) do not apply here.
__device__ __forceinline__ bool ref_accept_f(uint64_t u53, double dE, double T) {
    const double x = -dE / T;
    const double u = double(u53) * 0x1.0p-53;
    if (!(x > -80.0)) return u < qexp(x);
    constexpr double kBand = 0x1.0p-14;
    const double e = double(__builtin_amdgcn_exp2f(float(x) * 1.44269504f));
    if (u < e * (1.0 - kBand)) return true;
    if (u > e * (1.0 + kBand)) return false;
    return u < qexp(x);
}

// A launch runs sweeps [sweep_begin, sweep_end); the first one starts every branch from
// x = 0, r = -d and its seed's splitmix64 state, a later one continues from the x, r and
// state (sm[b]) the previous launch left, so the split never changes the trajectory.
__global__ void __launch_bounds__(256) ray_ref_kernel(DevProblem p, int R, uint64_t seed_offset, const double* temp,
                                                       int sweep_begin, int sweep_end, uint8_t* xs, double* rs,
                                                       uint64_t* sm) {
    const int w = int(threadIdx.x) / 64, lane = int(threadIdx.x) % 64;
    const int b = int(blockIdx.x) * 4 + w;
    if (b >= R) return;   // the whole wave leaves; waves never sync with each other
    const int n = p.n;
    uint8_t* x = xs + size_t(b) * size_t(n);
    double* r = rs + size_t(b) * size_t(p.n_rays);
    uint64_t st;
    if (sweep_begin == 0) {
        for (int i = lane; i < n; i += 64) x[i] = 0;
        for (int k = lane; k < p.n_rays; k += 64) r[k] = -p.d[k];
        st = (uint64_t(b) + 1u + seed_offset) * 0x2545F4914F6CDD1Dull + 1;
    } else {
        st = sm[b];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    for (int sweep = sweep_begin; sweep < sweep_end; ++sweep) {
        const double T = temp[sweep];
        int i0 = 0;
        while (i0 < n) {
            const int i = i0 + lane;
            const bool valid = i < n;
            const int pos = valid ? p.bit_pos[i] : 0;
            double dE = 0.0, base = 0.0;   // base: the sum up to the neighbour terms
            const int xi = valid ? int(x[i]) : 0;
            // each lane's column bounds and update order, kept for the flip
            const int c0 = valid ? p.col_ptr[pos] : 0, c1 = valid ? p.col_ptr[pos + 1] : 0;
            const int seq = valid ? int(p.seq_update[pos]) : 0;
            constexpr int kChunk = 8;
            if (valid) {
                // eight entries' loads, then their residuals, in flight together; the sum
                // itself stays in CSR order
                const double delta = xi ? -1.0 : 1.0;
                dE = p.lin[pos] * delta;
                const int k1 = c1;
                for (int kb = c0; kb < k1; kb += kChunk) {
                    RayEnt e[kChunk];
                    double rv[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) e[u] = p.ent[min(kb + u, k1 - 1)];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) rv[u] = r[e[u].ray];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) {
                        if (kb + u >= k1) break;
                        const double a = double(e[u].a) * delta;
                        dE += e[u].w * (rv[u] * a + 0.5 * a * a);
                    }
                }
                base = dE;
                const int q1 = p.nb_ptr[pos + 1];
                for (int qb = p.nb_ptr[pos]; qb < q1; qb += kChunk) {
                    int xv[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) xv[u] = x[p.nb_bit[min(qb + u, q1 - 1)]];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) {
                        if (qb + u >= q1) break;
                        dE += p.lambda * (xv[u] != xi ? -1.0 : 1.0);
                    }
                }
            }
            // decisions of the lanes in todo (lanes >= lo) on the current state, the k-th
            // uphill one drawing st + k gamma
            uint64_t todo = ~uint64_t(0), m_up = 0, m_acc = 0;
            auto decide = [&]() {
                const bool mine = valid && ((todo >> lane) & 1u);
                const bool up = mine && dE > 0.0;
                m_up = __ballot(up) & todo;
                bool acc = mine && dE <= 0.0;
                if (up) acc = ref_accept_f(mix64(st + uint64_t(__popcll(m_up & below) + 1) * kGamma) >> 11, dE, T);
                m_acc = __ballot(acc) & todo;
            };
            decide();
            for (;;) {
                if (m_acc == 0) {
                    st += uint64_t(__popcll(m_up)) * kGamma;
                    i0 += 64;
                    break;
                }
                const int js = __builtin_ctzll(m_acc);
                const uint64_t upto = js == 63 ? ~uint64_t(0) : (uint64_t(2) << js) - 1u;
                st += uint64_t(__popcll(m_up & upto)) * kGamma;
                const int fi = i0 + js;
                const int k0 = __shfl(c0, js), k1 = __shfl(c1, js), seq_f = __shfl(seq, js);
                const int xf = __shfl(xi, js);
                const double delta = xf ? -1.0 : 1.0;
                if (!seq_f) {
                    for (int k = k0 + lane; k < k1; k += 64) r[p.ent[k].ray] += double(p.ent[k].a) * delta;
                } else if (lane == 0) {   // a ray repeats in the column: its additions in CSR order
                    for (int k = k0; k < k1; ++k) r[p.ent[k].ray] += double(p.ent[k].a) * delta;
                }
                if (lane == 0) x[fi] = uint8_t(xf ^ 1);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                todo = ~upto;
                if (k1 > k0 || todo == 0) {   // residuals changed (or the window is done)
                    i0 = fi + 1;
                    break;
                }
                // A ray-free bit flipped: no residual changed, so a later lane's sum up to its
                // neighbour terms stands; only those terms are summed again, in order.
                if (valid && ((todo >> lane) & 1u)) {
                    dE = base;
                    const int q1 = p.nb_ptr[pos + 1];
                    for (int qb = p.nb_ptr[pos]; qb < q1; qb += kChunk) {
                        int xv[kChunk];
#pragma unroll
                        for (int u = 0; u < kChunk; ++u) xv[u] = x[p.nb_bit[min(qb + u, q1 - 1)]];
#pragma unroll
                        for (int u = 0; u < kChunk; ++u) {
                            if (qb + u >= q1) break;
                            dE += p.lambda * (xv[u] != xi ? -1.0 : 1.0);
                        }
                    }
                }
                decide();
            }
        }
    }
    if (lane == 0) sm[b] = st;
}

hipError_t launch_ray_ref(const DevProblem& p, int R, uint64_t seed_offset, const double* temp, int sweep_begin,
                          int sweep_end, uint8_t* xs, double* rs, uint64_t* sm, hipStream_t st) {
    ray_ref_kernel<<<unsigned((R + 3) / 4), 256, 0, st>>>(p, R, seed_offset, temp, sweep_begin, sweep_end, xs, rs,
                                                          sm);
    return hipGetLastError();
}

// ---- the team barrier's visibility, checked on the device ---------------------------------
// The sweep relies on a team's plain stores, drained before its barrier, being seen by the
// L2-hitting loads of every member after it. This kernel exercises exactly that: each
// step, every thread of every member stores the step number into its own word, the team
// passes team_barrier, and every thread then reads its word of each member with ld_l2; a
// word below the step is a stale read and is counted. Words only grow, so a member that
// runs ahead is never miscounted. Teams form as in the sweep (XCC_ID, tickets).
__global__ void __launch_bounds__(256) team_barrier_probe_kernel(unsigned* words, unsigned* bar, unsigned* tickets,
                                                                 int* err, int steps, int team_size,
                                                                 unsigned long long* stale) {
    __shared__ int s_team, s_rank, s_ok;
    if (threadIdx.x == 0) {
        s_team = int(xcc_id() & (kTeams - 1));
        s_rank = int(__hip_atomic_fetch_add(tickets + s_team * 64, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    if (!grid_barrier(bar, gridDim.x, err)) return;
    if (threadIdx.x == 0) {
        const unsigned members = __hip_atomic_load(tickets + s_team * 64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = members >= unsigned(team_size) ? 1 : 0;
        if (!s_ok) __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int team = s_team, rank = s_rank;
    if (!s_ok || rank >= team_size) return;
    unsigned* tbar = bar + 64 + team * 64;
    unsigned* mine = words + (size_t(team) * team_size + size_t(rank)) * 256;
    const unsigned* team_words = words + size_t(team) * team_size * 256;
    unsigned long long bad = 0;
    unsigned target = 0;
    for (int s = 1; s <= steps; ++s) {
        mine[threadIdx.x] = unsigned(s);
        target += unsigned(team_size);
        if (!team_barrier(tbar, rank, team_size, target, err)) break;
        for (int r = 0; r < team_size; ++r) bad += ld_l2(team_words + size_t(r) * 256 + threadIdx.x) < unsigned(s);
    }
    if (bad) atomicAdd(stale, bad);
}

hipError_t launch_team_barrier_probe(unsigned* words, unsigned* bar, unsigned* tickets, int* err, int steps,
                                     int team_size, unsigned long long* stale, hipStream_t st) {
    team_barrier_probe_kernel<<<unsigned(team_size * kTeams), 256, 0, st>>>(words, bar, tickets, err, steps, team_size,
                                                                          stale);
    return hipGetLastError();
}

// ---- staged solves: bit clocks, state carried between stages, per-request outputs ---------
// A request occupies original bits [bit0, bit0 + n_bits) and rays [ray0, ray0 + n_rays) of a
// union. Its clock words (DevState::clock) say where its bits are in their own schedule and
// how many branches it anneals; its state moves between unions span by span.

__global__ void stage_clock_kernel(const int32_t* bit_pos, int bit0, int n_bits, uint32_t word, uint32_t* clock) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n_bits) clock[bit_pos[bit0 + i]] = word;
}

// Spins and clocks of bits src [sb0, sb0 + n_bits) -> dst [db0, ...), residuals of rays src
// [sr0, sr0 + n_rays) -> dst [dr0, ...), groups [0, groups). A bit's own sweep is the same
// before and after: its destination offset is its source offset plus the source stage's
// sweeps (the destination stage has run none).
template <typename T, int BW>
__global__ void stage_import_kernel(const int32_t* sbit_pos, int sn, int sn_rays, const uint16_t* sx, const T* sr,
                                    const uint32_t* sclock, uint32_t src_sweeps, const int32_t* dbit_pos, int dn,
                                    int dn_rays, uint16_t* dx, T* dr, uint32_t* dclock, int sb0, int db0, int n_bits,
                                    int sr0, int dr0, int n_rays, int groups) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int g = int(blockIdx.y);
    if (g >= groups) return;
    if (i < size_t(n_bits)) {
        const int sp = sbit_pos[sb0 + int(i)], dp = dbit_pos[db0 + int(i)];
        dx[size_t(g) * size_t(dn) + size_t(dp)] = sx[size_t(g) * size_t(sn) + size_t(sp)];
        if (g == 0) {
            const uint32_t c = sclock[sp];
            dclock[dp] = (c & ~kClockOffsetMask) | ((c & kClockOffsetMask) + src_sweeps);
        }
    }
    if (i < size_t(n_rays) * BW) {
        const size_t ray = i / BW, b = i % BW;
        dr[(size_t(g) * size_t(dn_rays) + size_t(dr0) + ray) * BW + b] =
            sr[(size_t(g) * size_t(sn_rays) + size_t(sr0) + ray) * BW + b];
    }
}

// Branch agreement of one request over its own first rq branches (branch_fractions' p_v
// and classify, src/ising_recon.hpp:160-163, 302-309), outputs in the span's bit order.
__global__ void stage_outputs_kernel(DevProblem p, const uint16_t* x, int bw, int bit0, int n_bits, int rq,
                                     float* p_void, uint8_t* tri, float* certainty, uint8_t* branches) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n_bits) return;
    const int pos = p.bit_pos[bit0 + i];
    int cnt = 0;
    for (int g = 0; g * bw < rq; ++g) {
        const int live = rq - g * bw < bw ? rq - g * bw : bw;
        const unsigned m = unsigned(x[size_t(g) * size_t(p.n) + pos]) & ((1u << live) - 1u);
        cnt += __popc(m);
        if (branches)
            for (int b = 0; b < live; ++b)
                branches[size_t(g * bw + b) * size_t(n_bits) + size_t(i)] = uint8_t((m >> b) & 1u);
    }
    const float pv = float(cnt) / float(rq);
    p_void[i] = pv;
    tri[i] = pv >= 0.9f ? 1 : (pv <= 0.1f ? 0 : 2);
    certainty[i] = pv > 1.0f - pv ? pv : 1.0f - pv;
}

hipError_t launch_stage_clock(const DevProblem& p, int bit0, int n_bits, uint32_t word, uint32_t* clock,
                              hipStream_t st) {
    if (n_bits > 0)
        stage_clock_kernel<<<unsigned((n_bits + 255) / 256), 256, 0, st>>>(p.bit_pos, bit0, n_bits, word, clock);
    return hipGetLastError();
}

hipError_t launch_stage_import(const DevProblem& src, const void* sx, const void* sr, const uint32_t* sclock,
                               int src_sweeps, const DevProblem& dst, void* dx, void* dr, uint32_t* dclock,
                               int precision, int bw, int sb0, int db0, int n_bits, int sr0, int dr0, int n_rays,
                               int groups, hipStream_t st) {
    const size_t items = std::max(size_t(n_bits), size_t(n_rays) * size_t(bw));
    if (items == 0 || groups <= 0) return hipSuccess;
    const dim3 grid{unsigned((items + 255) / 256), unsigned(groups), 1u};
    return with_bw(bw, [&]<int BW>() {
        if (precision == 1)
            stage_import_kernel<float, BW><<<grid, 256, 0, st>>>(
                src.bit_pos, src.n, src.n_rays, static_cast<const uint16_t*>(sx), static_cast<const float*>(sr), sclock,
                uint32_t(src_sweeps), dst.bit_pos, dst.n, dst.n_rays, static_cast<uint16_t*>(dx),
                static_cast<float*>(dr), dclock, sb0, db0, n_bits, sr0, dr0, n_rays, groups);
        else
            stage_import_kernel<double, BW><<<grid, 256, 0, st>>>(
                src.bit_pos, src.n, src.n_rays, static_cast<const uint16_t*>(sx), static_cast<const double*>(sr),
                sclock, uint32_t(src_sweeps), dst.bit_pos, dst.n, dst.n_rays, static_cast<uint16_t*>(dx),
                static_cast<double*>(dr), dclock, sb0, db0, n_bits, sr0, dr0, n_rays, groups);
        return hipGetLastError();
    });
}

hipError_t launch_stage_outputs(const DevProblem& p, const uint16_t* x, int bw, int bit0, int n_bits, int rq,
                                float* p_void, uint8_t* tri, float* certainty, uint8_t* branches, hipStream_t st) {
    if (n_bits > 0)
        stage_outputs_kernel<<<unsigned((n_bits + 255) / 256), 256, 0, st>>>(p, x, bw, bit0, n_bits, rq, p_void, tri,
                                                                            certainty, branches);
    return hipGetLastError();
}

// ---- freeze_provable_rock's fixpoint (src/ising_recon.hpp:187-205) ------------------------
// The data bound of a 0 -> 1 flip, sum_k w_b a_k (a_k / 2 - d_b) in CSR order, one thread per
// position; then rounds in which every unfrozen bit freezes iff lb + lin + lambda (rock - free)
// >= margin against the previous round's frozen set. With lambda >= 0 a bound only grows as
// neighbours freeze, so these rounds reach the reference's in-order fixpoint.
__global__ void freeze_bound_kernel(DevProblem p, double* lb) {
    const int pos = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (pos >= p.n) return;
    double s = 0.0;
    for (int k = p.col_ptr[pos]; k < p.col_ptr[pos + 1]; ++k) {
        const RayEnt e = p.ent[k];
        s += e.w * double(e.a) * (0.5 * double(e.a) - p.d[e.ray]);
    }
    lb[pos] = s;
}

__global__ void freeze_round_kernel(DevProblem p, const double* lb, double margin, const uint8_t* frozen,
                                    uint8_t* next, unsigned* changed) {
    const int pos = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (pos >= p.n) return;
    if (frozen[pos]) {
        next[pos] = 1;
        return;
    }
    int rock = 0, free_nb = 0;
    for (int q = p.nb_ptr[pos]; q < p.nb_ptr[pos + 1]; ++q) (frozen[p.nb[q]] ? rock : free_nb)++;
    const bool f = lb[pos] + p.lin[pos] + p.lambda * (rock - free_nb) >= margin;
    next[pos] = f ? 1 : 0;
    if (f) atomicAdd(changed, 1u);
}

hipError_t launch_freeze_bound(const DevProblem& p, double* lb, hipStream_t st) {
    if (p.n > 0) freeze_bound_kernel<<<unsigned((p.n + 255) / 256), 256, 0, st>>>(p, lb);
    return hipGetLastError();
}

hipError_t launch_freeze_round(const DevProblem& p, const double* lb, double margin, const uint8_t* frozen,
                               uint8_t* next, unsigned* changed, hipStream_t st) {
    if (p.n > 0) freeze_round_kernel<<<unsigned((p.n + 255) / 256), 256, 0, st>>>(p, lb, margin, frozen, next, changed);
    return hipGetLastError();
}

hipError_t launch_int_unpermute(const DevGraph& g, int R, const int8_t* s_state, int8_t* s_out, hipStream_t st) {
    const size_t total = size_t(R) * size_t(g.n_pad);
    int_unpermute_kernel<<<unsigned((total + 255) / 256), 256, 0, st>>>(g, R, s_state, s_out);
    return hipGetLastError();
}

hipError_t launch_int_energies(const DevGraph& g, const uint8_t* X, int K, long long* out, hipStream_t st) {
    int_energy_kernel<<<unsigned(K), 256, 0, st>>>(g, X, out);
    return hipGetLastError();
}

// A plain streaming copy, 16 bytes per lane per step: the reference point for HBM byte
// counters and for the bandwidth the chip sustains on a pure read + write stream.
__global__ void __launch_bounds__(256) copy_probe_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         size_t n16) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_probe(const void* src, void* dst, size_t n16, hipStream_t st) {
    copy_probe_kernel<<<2048, 256, 0, st>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
    return hipGetLastError();
}

}  // namespace qb
