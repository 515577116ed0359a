// qb_dense.hip -- gfx950 kernel of the dense-QUBO field engine (the precision study's
// dense instance).
//
// E(x) = sum_i h_i x_i + sum_{i<j} Q_ij x_i x_j with Q dense, symmetric, zero diagonal.
// Every bit couples to every other, so a sweep is sequential: bits in index order. The
// engine keeps the local fields f_i = h_i + sum_j Q_ij x_j and updates them on every
// flip: f += delta_t Q_t (delta_t = +1 for 0 -> 1, -1 for 1 -> 0). One workgroup anneals
// one replica with x and f in LDS:
//
//   - a window of 64 bits is decided by one wave, one bit per lane: the window's 64 x 64
//     block of Q and its draws were staged in LDS by the other waves during the previous
//     window; bit t decides on dE = (1 - 2 x_t) f_t and, on a flip, every lane adds
//     Q_{t,lane} delta_t to its field -- so later bits of the window see earlier flips;
//   - the whole workgroup then adds the window's flipped rows, in flip order, to every field
//     outside the window (a rank-k update, k = flips in the window).
//
// Every field therefore receives the same additions in the same order as the sequential
// sweep. Fields are FP64, or FP32 carried incrementally (one float add of +-Q per flip) with
// an FP64 recompute f = h + Q x every k sweeps. A flip is accepted iff dE <= 0 or
// dE * beta < -qlog(u), u from Philox keyed by (replica seed, sweep, bit): the Metropolis
// test written with the log, so the sequential steps need no exponential.
#include "qb_internal.h"
#include "qb_rng.h"

namespace qb {

constexpr int kDenseWin = 64;
constexpr int kRankChunk = 8;
constexpr int kFieldGroup = 4;   // fields a thread updates together (i, i + block, ...)

// Fields i = g0 + q STRIDE (q < kFieldGroup, i < n) from scratch:
// h_i + sum_j x_j Q_ji in FP64, j ascending (Q_ji = Q_ij), additions of unset bits
// skipped; every field's loads of a 16-column chunk in flight together.
template <int STRIDE = kDenseBlock>
__device__ __forceinline__ void scratch_fields(const DevDense& d, const uint8_t* x, int g0, double (&s)[kFieldGroup]) {
    constexpr int kChunk = 16;
    const int n = d.n;
    int col[kFieldGroup];
#pragma unroll
    for (int q = 0; q < kFieldGroup; ++q) {
        col[q] = min(g0 + q * STRIDE, n - 1);
        s[q] = double(d.h[col[q]]);
    }
    for (int j0 = 0; j0 < n; j0 += kChunk) {
        float v[kFieldGroup][kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
#pragma unroll
            for (int q = 0; q < kFieldGroup; ++q) v[q][u] = d.Q[size_t(min(j0 + u, n - 1)) * size_t(n) + size_t(col[q])];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            const bool set = j0 + u < n && x[min(j0 + u, n - 1)];
#pragma unroll
            for (int q = 0; q < kFieldGroup; ++q) {
                const double nxt = s[q] + double(v[q][u]);
                s[q] = set ? nxt : s[q];
            }
        }
    }
}

// Staging of the window at i0n (in sweep sw) into one of two buffers while wave 0 decides
// the current window: waves 1.. of the workgroup copy its block, one row per direct-to-LDS
// load, bq[t * 64 + l] = Q[i0n + min(t, wl - 1)][min(i0n + l, n - 1)] (lanes past the
// window hold columns nobody reads), and wave 1 makes its -log draws L[l]. The copy is
// covered by the issuing waves' vmcnt: they drain it before the barrier that makes the
// buffer visible.
template <int FIRST, int NW>
__device__ __forceinline__ void stage_window(const DevDense& d, uint64_t key, int i0n, int sw, float* bq, double* L) {
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    const int n = d.n, wln = min(kDenseWin, n - i0n);
    const int col = min(i0n + lane, n - 1);
    for (int t = wave - FIRST; t < kDenseWin; t += NW) {
        const float* src = d.Q + size_t(i0n + min(t, wln - 1)) * size_t(n) + size_t(col);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                         (__attribute__((address_space(3))) void*)(bq + t * kDenseWin), 4, 0, 0);
    }
    if (wave == FIRST) L[lane] = -qlog(draw_u01(key, uint32_t(sw), uint32_t(i0n + (lane < wln ? lane : 0))));
}

// One window's sequential decisions by one wave, lane l holding bit i0 + l: its field fl,
// spin xl and -log draw L, the window's block in LDS (blk[t * 64 + l] = Q[i0 + t][i0 + l]).
// Bit t decides on dE = (1 - 2 x_t) f_t -- accept iff dE <= 0 or dE beta < L_t -- with every
// earlier flip of the window in f_t. Every lane from the first undecided bit on tests its
// own bit against its current field; the first that accepts is the next flip, and the bits
// before it decide as they would one at a time, since no flip lies between them. Every lane
// then adds Q[i0 + t][i0 + l] delta_t and the search goes on after t: a window costs one
// round per flip plus one. Lane 0 records the flips (s_t, s_delta); returns their count.
template <typename T>
__device__ __forceinline__ int decide_window(int wl, double b, double L, const float* blk, T& fl, int& xl, int* s_t,
                                             float* s_delta) {
    const int lane = int(threadIdx.x) & 63;
    int k = 0;
    int from = 0;
    while (true) {
        const double dE = xl ? -double(fl) : double(fl);
        const bool accept = lane >= from && lane < wl && (dE <= 0.0 || dE * b < L);
        const uint64_t m = __ballot(accept);
        if (m == 0) break;
        const int t = __builtin_ctzll(m);
        const float dl = __builtin_amdgcn_readlane(xl, t) ? -1.0f : 1.0f;
        fl = fl + T(blk[t * kDenseWin + lane]) * T(dl);
        if (lane == t) xl ^= 1;
        if (lane == 0) {
            s_t[k] = t;
            s_delta[k] = dl;
        }
        ++k;
        from = t + 1;
    }
    return k;
}

// Fields i = g0 + q kDenseBlock (q < kFieldGroup, i < n, outside the window [i0, i0 + w))
// plus the window's k flipped rows in flip order, CHUNK rows' loads at a time (past the last
// flip the index repeats it; its value is unused).
template <typename T, int CHUNK>
__device__ __forceinline__ void add_rows_group(const DevDense& d, int i0, int w, int g0, int k, const int* s_t,
                                               const float* s_delta, T* f) {
    const int n = d.n;
    int col[kFieldGroup];
    bool upd[kFieldGroup];
    T acc[kFieldGroup];
#pragma unroll
    for (int q = 0; q < kFieldGroup; ++q) {
        const int i = g0 + q * kDenseBlock;
        upd[q] = i < n && !(i >= i0 && i < i0 + w);
        col[q] = min(i, n - 1);
        acc[q] = f[col[q]];
    }
    for (int m0 = 0; m0 < k; m0 += CHUNK) {
        float v[kFieldGroup][CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const size_t row = size_t(i0 + s_t[min(m0 + u, k - 1)]) * size_t(n);
#pragma unroll
            for (int q = 0; q < kFieldGroup; ++q) v[q][u] = d.Q[row + size_t(col[q])];
        }
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const T dl = T(s_delta[min(m0 + u, k - 1)]);
#pragma unroll
            for (int q = 0; q < kFieldGroup; ++q) {
                const T nxt = acc[q] + T(v[q][u]) * dl;
                acc[q] = m0 + u < k ? nxt : acc[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kFieldGroup; ++q)
        if (upd[q]) f[g0 + q * kDenseBlock] = acc[q];
}

template <typename T>
__global__ void __launch_bounds__(kDenseBlock) dense_anneal_kernel(DevDense d, const double* beta, int n_sweeps,
                                                                    int recompute_every, uint64_t seed_offset,
                                                                    uint8_t* x_out, double* energy, double* field_err) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = d.n;
    float* blk = reinterpret_cast<float*>(smem);   // two staging buffers: the next windows' Q blocks
    T* f = reinterpret_cast<T*>(blk + 2 * kDenseWin * kDenseWin);
    uint8_t* x = reinterpret_cast<uint8_t*>(f + n);
    __shared__ double s_L[2][kDenseWin];   // and their -log draws
    __shared__ int s_k;
    __shared__ int s_t[kDenseWin];
    __shared__ float s_delta[kDenseWin];
    __shared__ double s_red[kDenseBlock / 64][2];
    const int tid = int(threadIdx.x);
    const uint64_t key = seed_offset + uint64_t(blockIdx.x) + 1u;
    for (int i = tid; i < n; i += kDenseBlock) {
        x[i] = 0;
        f[i] = T(d.h[i]);
    }
    if (tid >= 64 && n_sweeps > 0) stage_window<1, kDenseBlock / 64 - 1>(d, key, 0, 0, blk, s_L[0]);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int nbuf = 0;   // the buffer of the current window
    for (int sweep = 0; sweep < n_sweeps; ++sweep) {
        const double b = beta[sweep];
        for (int i0 = 0; i0 < n; i0 += kDenseWin) {
            const int w = min(kDenseWin, n - i0);
            // the next window (after the last, the next sweep's first), staged during this one
            const int n0 = i0 + kDenseWin < n ? i0 + kDenseWin : 0;
            const bool staging = n0 > 0 || sweep + 1 < n_sweeps;
            if (tid >= 64 && staging)
                stage_window<1, kDenseBlock / 64 - 1>(d, key, n0, n0 > 0 ? sweep : sweep + 1,
                                                      blk + (nbuf ^ 1) * kDenseWin * kDenseWin, s_L[nbuf ^ 1]);
            if (tid < 64) {
                const int lane = tid;
                const bool live = lane < w;
                const int i = i0 + (live ? lane : 0);
                const double L = s_L[nbuf][lane];
                T fl = f[i];
                int xl = x[i];
                const int k = decide_window<T>(w, b, L, blk + nbuf * kDenseWin * kDenseWin, fl, xl, s_t, s_delta);
                if (live) {
                    f[i] = fl;
                    x[i] = uint8_t(xl);
                }
                if (lane == 0) s_k = k;
            }
            __syncthreads();
            const int k = s_k;
            if (k > 0) {
                // up to kRankChunk flipped rows at a time (fewer when fewer flipped): their
                // loads in flight together, the additions in flip order
                for (int g0 = tid; g0 < n; g0 += kFieldGroup * kDenseBlock) {
                    if (k <= 1) add_rows_group<T, 1>(d, i0, w, g0, k, s_t, s_delta, f);
                    else if (k <= 2) add_rows_group<T, 2>(d, i0, w, g0, k, s_t, s_delta, f);
                    else if (k <= 4) add_rows_group<T, 4>(d, i0, w, g0, k, s_t, s_delta, f);
                    else add_rows_group<T, kRankChunk>(d, i0, w, g0, k, s_t, s_delta, f);
                }
            }
            if (staging) __builtin_amdgcn_s_waitcnt(0);   // the staged block has landed before the barrier
            __syncthreads();
            nbuf ^= 1;
        }
        if (recompute_every > 0 && (sweep + 1) % recompute_every == 0) {
            // f = h + Q x from scratch in FP64, j ascending (column i of Q = row i)
            for (int g0 = tid; g0 < n; g0 += kFieldGroup * kDenseBlock) {
                double sc[kFieldGroup];
                scratch_fields(d, x, g0, sc);   // reads x and Q only, so fields may change meanwhile
#pragma unroll
                for (int q = 0; q < kFieldGroup; ++q)
                    if (g0 + q * kDenseBlock < n) f[g0 + q * kDenseBlock] = T(sc[q]);
            }
            __syncthreads();
        }
    }
    // outputs: spins, the FP64 energy and max |f_carried - f_scratch|
    double e = 0.0, err = 0.0;
    for (int g0 = tid; g0 < n; g0 += kFieldGroup * kDenseBlock) {
        double sc[kFieldGroup];
        scratch_fields(d, x, g0, sc);
#pragma unroll
        for (int q = 0; q < kFieldGroup; ++q) {
            const int i = g0 + q * kDenseBlock;
            if (i >= n) continue;
            err = fmax(err, fabs(double(f[i]) - sc[q]));
            if (x[i]) e += double(d.h[i]) + 0.5 * (sc[q] - double(d.h[i]));
            x_out[size_t(blockIdx.x) * size_t(n) + size_t(i)] = x[i];
        }
    }
    for (int o = 32; o > 0; o /= 2) {
        e += __shfl_down(e, o);
        err = fmax(err, __shfl_down(err, o));
    }
    if ((tid & 63) == 0) {
        s_red[tid / 64][0] = e;
        s_red[tid / 64][1] = err;
    }
    __syncthreads();
    if (tid == 0) {
        double es = 0.0, em = 0.0;
        for (int wv = 0; wv < kDenseBlock / 64; ++wv) {
            es += s_red[wv][0];
            em = fmax(em, s_red[wv][1]);
        }
        energy[blockIdx.x] = es;
        field_err[blockIdx.x] = em;
    }
}

// ---- several workgroups per replica ----------------------------------------------------
// kDenseParts workgroups on one XCD anneal one replica together. Part p owns a contiguous
// run of windows and the fields of their bits; every part keeps the whole spin vector.
// Windows go in order: the window's owner decides it as above (its fields are its own) and
// publishes the window's flips -- a count and one byte per flip, t | 0x80 for 1 -> 0 -- in a
// slot of the replica's ring (two sweeps deep). Every word of the slot carries the window's
// tag, so the owner goes on without waiting for its stores; every other part polls the slot
// until each word has the tag, reads the flips and applies them to its spins. Then every
// part adds the flipped rows, in flip order, to its fields. Each field therefore receives the
// same additions in the same order as with one workgroup, and the spins are the same bit for
// bit; only the final energy's partial sums are grouped by part. The parts share state
// through the XCD's L2 (sc1 loads), like the ray sweep's teams.
//
// The ring's two sweeps suffice: the owner of a window in sweep s + 2 decides only after
// every part has published its windows of sweep s + 1, which each did after applying the
// same window of sweep s.
constexpr int kDenseSlotWords = 1 + kDenseWin / 2;   // a header word, then two flips a word
constexpr int kTeamRankChunk = 16;
constexpr int kDenseTeamBlock = 512;

struct DenseTeamWork {
    unsigned* tickets;   // [8 * 64]: per-XCD tickets
    unsigned* done;      // [R * 32]: parts finished, at done[32 r]
    double* part_e;      // [R * kDenseParts]
    double* part_err;    // [R * kDenseParts]
    uint32_t* slots;     // [R][2][n_win][kDenseSlotWords]
};

__host__ __device__ inline int dense_part_fields_max(int n) {
    const int n_win = (n + kDenseWin - 1) / kDenseWin;
    return ((n_win + kDenseParts - 1) / kDenseParts) * kDenseWin;
}

size_t dense_team_work_bytes(int n, int R) {
    const size_t n_win = size_t((n + kDenseWin - 1) / kDenseWin);
    return sizeof(unsigned) * (8 * 64 + 32 * size_t(R)) + sizeof(double) * 2 * size_t(R) * kDenseParts +
           sizeof(uint32_t) * size_t(R) * 2 * n_win * kDenseSlotWords;
}

static DenseTeamWork dense_team_work(void* work, int n, int R) {
    unsigned char* p = static_cast<unsigned char*>(work);
    DenseTeamWork w;
    w.tickets = reinterpret_cast<unsigned*>(p);
    w.done = w.tickets + 8 * 64;
    w.part_e = reinterpret_cast<double*>(w.done + 32 * size_t(R));
    w.part_err = w.part_e + size_t(R) * kDenseParts;
    w.slots = reinterpret_cast<uint32_t*>(w.part_err + size_t(R) * kDenseParts);
    (void)n;
    return w;
}

size_t dense_team_lds_bytes(int n, int precision) {
    return 3 * sizeof(float) * kDenseWin * kDenseWin +
           size_t(dense_part_fields_max(n)) * (precision == 1 ? sizeof(float) : sizeof(double)) + size_t(n);
}

// A slot word's tag for the idx-th window (idx from 1), in its low 16 bits: never 0, the
// zeroed ring, and different from the slot's previous use, 2 n_win windows earlier, while
// n_win < 16384 (n is far smaller: its fields fit LDS).
__device__ __forceinline__ uint32_t dense_tag(unsigned idx) { return (idx & 0x7FFFu) | 0x8000u; }

// Waits (one wave, lane q reading word q) until every word of the slot carries `tag`; the
// lane's word is left in `word`. False (and *err set) after ~2 s or when another workgroup
// has set *err.
__device__ __forceinline__ bool dense_wait_slot(const uint32_t* slot, uint32_t tag, uint32_t& word, int* err) {
    const int lane = int(threadIdx.x) & 63;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    unsigned polls = 0;
    while (true) {
        word = lane < kDenseSlotWords ? ld_l2(slot + lane) : tag;
        if (__ballot((word & 0xFFFFu) != tag) == 0) return true;
        if ((++polls & 1023u) == 0 &&
            (ld_l2(err) != 0 || __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull)) {
            if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
}

// Field i plus the window's k flipped rows, acc + Q[i0 + s_t[m]][i] s_delta[m] in flip order,
// CHUNK rows' loads at a time (past the last flip the index repeats it; its value is unused).
template <typename T, int CHUNK>
__device__ __forceinline__ T add_rows(const DevDense& d, int i0, int i, int k, const int* s_t, const float* s_delta,
                                      T acc) {
    for (int m0 = 0; m0 < k; m0 += CHUNK) {
        float v[CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) v[u] = d.Q[size_t(i0 + s_t[min(m0 + u, k - 1)]) * size_t(d.n) + size_t(i)];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const T nxt = acc + T(v[u]) * T(s_delta[min(m0 + u, k - 1)]);
            acc = m0 + u < k ? nxt : acc;
        }
    }
    return acc;
}

// Fields col[j] (j < NF) plus the window's k flipped rows, acc[j] + Q[i0 + s_t[m]][col[j]]
// s_delta[m] in flip order, the CHUNK rows' loads of every field in flight together (past the
// last flip the index repeats it; its value is unused).
template <typename T, int CHUNK, int NF>
__device__ __forceinline__ void add_rows_fields(const DevDense& d, int i0, const int (&col)[NF], int k, const int* s_t,
                                                const float* s_delta, T (&acc)[NF]) {
    for (int m0 = 0; m0 < k; m0 += CHUNK) {
        float v[NF][CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const size_t row = size_t(i0 + s_t[min(m0 + u, k - 1)]) * size_t(d.n);
#pragma unroll
            for (int j = 0; j < NF; ++j) v[j][u] = d.Q[row + size_t(col[j])];
        }
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const T dl = T(s_delta[min(m0 + u, k - 1)]);
#pragma unroll
            for (int j = 0; j < NF; ++j) {
                const T nxt = acc[j] + T(v[j][u]) * dl;
                acc[j] = m0 + u < k ? nxt : acc[j];
            }
        }
    }
}

// Field i plus the window's k flipped rows in flip order, as add_rows, with flip m's bit and
// sign in lane m of tl and dl (read across the wave), CHUNK rows' loads at a time.
template <typename T, int CHUNK>
__device__ __forceinline__ T add_rows_lanes(const DevDense& d, int i0, int i, int k, int tl, float dl, T acc) {
    for (int m0 = 0; m0 < k; m0 += CHUNK) {
        float v[CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
            v[u] = d.Q[size_t(i0 + __builtin_amdgcn_readlane(tl, min(m0 + u, k - 1))) * size_t(d.n) + size_t(i)];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const float dm = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dl), min(m0 + u, k - 1)));
            const T nxt = acc + T(v[u]) * T(dm);
            acc = m0 + u < k ? nxt : acc;
        }
    }
    return acc;
}

template <typename T>
__global__ void __launch_bounds__(kDenseTeamBlock) dense_team_kernel(DevDense d, DenseTeamWork tw, int R,
                                                                 const double* beta, int n_sweeps,
                                                                 int recompute_every, uint64_t seed_offset,
                                                                 uint8_t* x_out, double* energy, double* field_err,
                                                                 int* err) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = d.n;
    const int n_win = (n + kDenseWin - 1) / kDenseWin;
    __shared__ int s_xcd, s_rank;
    __shared__ double s_red[kDenseTeamBlock / 64][2];
    const int tid = int(threadIdx.x);
    if (tid == 0) {
        s_xcd = int(xcc_id() & 7u);
        s_rank = int(__hip_atomic_fetch_add(tw.tickets + s_xcd * 64, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    __syncthreads();
    const int rank = s_rank;
    if (rank >= kDenseTeamSlots * kDenseParts) return;
    const int r = s_xcd + 8 * (rank / kDenseParts);
    if (r >= R) return;
    const int part = rank % kDenseParts;
    const int wb = part * n_win / kDenseParts, we = (part + 1) * n_win / kDenseParts;
    const int ib = wb * kDenseWin, ie = min(we * kDenseWin, n);
    // three staging buffers (window G's block and draws in buffer G % 3)
    float* blk = reinterpret_cast<float*>(smem);
    T* f = reinterpret_cast<T*>(blk + 3 * kDenseWin * kDenseWin);   // fields of bits [ib, ie)
    uint8_t* x = reinterpret_cast<uint8_t*>(f + dense_part_fields_max(n));
    __shared__ double s_L[3][kDenseWin];
    // the flip ring: window G's flips (t, delta) and count in slot G % 4
    __shared__ int s_rt[4][kDenseWin];
    __shared__ float s_rd[4][kDenseWin];
    __shared__ int s_rk[4];
    // progress words: windows whose flips are in the ring; windows whose next-window update is
    // done (wave 1); windows whose bulk update and staging are done (waves 2..7, one word each)
    __shared__ int s_flips, s_aprog, s_abort;
    __shared__ int s_bprog[kDenseTeamBlock / 64 - 2];
    constexpr int kNB = kDenseTeamBlock / 64 - 2;
    constexpr int kBulk = kDenseTeamBlock - 128;   // threads of the bulk update
    constexpr int kBulkFields = 3;                 // fields a bulk thread updates together
    const uint64_t key = seed_offset + uint64_t(r) + 1u;
    const int wave = tid >> 6, lane = tid & 63;
    const int T_total = n_sweeps * n_win;
    auto own_g = [&](int G) { const int w = G % n_win; return w >= wb && w < we; };
    auto stage = [&](int G) {
        stage_window<2, kNB>(d, key, (G % n_win) * kDenseWin, G / n_win, blk + (G % 3) * kDenseWin * kDenseWin,
                             s_L[G % 3]);
    };
    // Waits (one wave) until *p >= v; false once the workgroup aborts or after ~2 s (then it
    // aborts, with *err set).
    auto wait_ge = [&](int* p, int v) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        unsigned polls = 0;
        while (__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < v) {
            if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
            __builtin_amdgcn_s_sleep(1);
            if ((++polls & 1023u) == 0 && __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                if (lane == 0) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                return false;
            }
        }
        return true;
    };
    // Waits (one wave) until s_aprog >= va, every s_bprog[j] >= vb and s_flips >= vf: lane
    // j < kNB reads s_bprog[j], lane kNB s_aprog and lane kNB + 1 s_flips, all in one load a
    // poll. False as for wait_ge.
    auto wait_ab = [&](int va, int vb, int vf) {
        int* w = lane < kNB ? &s_bprog[lane] : lane == kNB ? &s_aprog : &s_flips;
        const int need = lane < kNB ? vb : lane == kNB ? va : lane == kNB + 1 ? vf : INT_MIN;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        unsigned polls = 0;
        while (__ballot(__hip_atomic_load(w, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need) != 0) {
            if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
            __builtin_amdgcn_s_sleep(1);
            if ((++polls & 1023u) == 0 && __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                if (lane == 0) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                return false;
            }
        }
        return true;
    };
    auto release = [&](int* p, int v) {
        if (lane == 0) __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    for (int i = tid; i < n; i += kDenseTeamBlock) x[i] = 0;
    for (int i = ib + tid; i < ie; i += kDenseTeamBlock) f[i - ib] = T(d.h[i]);
    if (tid == 0) {
        s_flips = 0;
        s_aprog = 0;
        s_abort = 0;
    }
    if (tid < kNB) s_bprog[tid] = 0;
    if (wave >= 2) {
        for (int G = 0; G < 2 && G < T_total; ++G)
            if (own_g(G)) stage(G);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int sweep = 0; sweep < n_sweeps; ++sweep) {
        const double b = beta[sweep];
        for (int w = 0; w < n_win; ++w) {
            const int G = sweep * n_win + w;
            const int i0 = w * kDenseWin;
            const int wl = min(kDenseWin, n - i0);
            const bool own = w >= wb && w < we;
            const int q = G & 3;
            const bool next_own = G + 1 < T_total && own_g(G + 1);
            if (wave == 0) {
                // the ring slot is free once window G - 4 is fully applied
                // (an owned window's fields hold every earlier window's flips once wave 1 has
                // added window G - 1's, after the bulk of window G - 2)
                if (!wait_ab(own ? G : G - 3, G - 3, INT_MIN)) return;
                const unsigned idx = unsigned(G) + 1u;
                uint32_t* slot = tw.slots + ((size_t(r) * 2 + size_t(sweep & 1)) * size_t(n_win) + size_t(w)) * kDenseSlotWords;
                if (own) {
                    const bool live = lane < wl;
                    const int i = i0 + (live ? lane : 0);
                    const double L = s_L[G % 3][lane];
                    T fl = f[i - ib];
                    int xl = x[i];
                    const int k = decide_window<T>(wl, b, L, blk + (G % 3) * kDenseWin * kDenseWin, fl, xl, s_rt[q],
                                                   s_rd[q]);
                    if (live) {
                        f[i - ib] = fl;
                        x[i] = uint8_t(xl);
                    }
                    if (lane == 0) s_rk[q] = k;
                    // the workgroup's waves need only the ring: raise the word before the slot
                    // stores, so that its release does not wait for them
                    release(&s_flips, G + 1);
                    // publish the flips, every word tagged: lane 0 the count, lane q > 0 flips
                    // 2q - 2 and 2q - 1 in bits 16-23 and 24-31
                    __builtin_amdgcn_s_waitcnt(0);   // the ring written (LDS) before it is read
                    if (lane < kDenseSlotWords) {
                        uint32_t word = dense_tag(idx);
                        if (lane == 0) {
                            word |= uint32_t(k) << 16;
                        } else {
#pragma unroll
                            for (int u = 0; u < 2; ++u) {
                                const int m = 2 * (lane - 1) + u;
                                if (m < k) word |= uint32_t(s_rt[q][m] | (s_rd[q][m] < 0.0f ? 0x80 : 0)) << (16 + 8 * u);
                            }
                        }
                        __hip_atomic_store(slot + lane, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                } else {
                    uint32_t word = 0;
                    if (!dense_wait_slot(slot, dense_tag(idx), word, err)) {
                        if (lane == 0) __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        return;
                    }
                    const int k = int((__builtin_amdgcn_readfirstlane(int(word)) >> 16) & 0xFF);
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const int m = 2 * (lane - 1) + u;
                        if (lane >= 1 && lane < kDenseSlotWords && m < k) {
                            const uint32_t c = (word >> (16 + 8 * u)) & 0xFFu;
                            const int t = int(c & 0x7Fu);
                            s_rt[q][m] = t;
                            s_rd[q][m] = (c & 0x80u) ? -1.0f : 1.0f;
                            x[i0 + t] = (c & 0x80u) ? 0 : 1;
                        }
                    }
                    if (lane == 0) s_rk[q] = k;
                    release(&s_flips, G + 1);
                }
            } else if (wave == 1) {
                // the next window's 64 fields get this window's flips, after the bulk of window
                // G - 1 has given them that window's
                if (!wait_ab(0, next_own ? G : INT_MIN, G + 1)) return;
                if (next_own) {
                    // the count and every lane's flip read together; the branch is uniform, so
                    // every lane holds its flip when the rows are read across the wave
                    const int k = __builtin_amdgcn_readfirstlane(s_rk[q]);
                    const int tl = s_rt[q][lane];
                    const float dl = s_rd[q][lane];
                    const int i = ((G + 1) % n_win) * kDenseWin + lane;
                    if (k > 0) {
                        const int ic = min(i, min(ie, n) - 1);
                        const T a = f[ic - ib];
                        const T r = k <= 1   ? add_rows_lanes<T, 1>(d, i0, ic, k, tl, dl, a)
                                    : k <= 2 ? add_rows_lanes<T, 2>(d, i0, ic, k, tl, dl, a)
                                    : k <= 4 ? add_rows_lanes<T, 4>(d, i0, ic, k, tl, dl, a)
                                    : k <= 8 ? add_rows_lanes<T, 8>(d, i0, ic, k, tl, dl, a)
                                             : add_rows_lanes<T, kTeamRankChunk>(d, i0, ic, k, tl, dl, a);
                        if (i < min(ie, n)) f[i - ib] = r;
                    }
                }
                __builtin_amdgcn_s_waitcnt(0);
                release(&s_aprog, G + 1);
            } else {
                // waves 2..: stage window G + 2 (its buffer's last reader, window G - 1, is
                // decided), then add this window's flips to the part's other fields
                if (G + 2 < T_total && own_g(G + 2)) stage(G + 2);
                if (!wait_ge(&s_flips, G + 1)) return;
                const int k = s_rk[q];
                if (k > 0) {
                    // kBulkFields fields a thread, their row loads in flight together; a field
                    // of the own or the next own window takes part in the loads only
                    const int i_next = next_own ? ((G + 1) % n_win) * kDenseWin : -1 - kDenseWin;
                    for (int g = ib + tid - 128; g < ie; g += kBulkFields * kBulk) {
                        int col[kBulkFields];
                        bool use[kBulkFields];
                        T a[kBulkFields];
#pragma unroll
                        for (int j = 0; j < kBulkFields; ++j) {
                            const int i = g + j * kBulk;
                            use[j] = i < ie && !(own && i >= i0 && i < i0 + wl) && !(i >= i_next && i < i_next + kDenseWin);
                            col[j] = i < ie ? i : ib;
                            a[j] = f[col[j] - ib];
                        }
                        if (k <= 1) add_rows_fields<T, 1, kBulkFields>(d, i0, col, k, s_rt[q], s_rd[q], a);
                        else if (k <= 2) add_rows_fields<T, 2, kBulkFields>(d, i0, col, k, s_rt[q], s_rd[q], a);
                        else if (k <= 4) add_rows_fields<T, 4, kBulkFields>(d, i0, col, k, s_rt[q], s_rd[q], a);
                        else if (k <= 8) add_rows_fields<T, 8, kBulkFields>(d, i0, col, k, s_rt[q], s_rd[q], a);
                        else add_rows_fields<T, kTeamRankChunk, kBulkFields>(d, i0, col, k, s_rt[q], s_rd[q], a);
#pragma unroll
                        for (int j = 0; j < kBulkFields; ++j)
                            if (use[j]) f[col[j] - ib] = a[j];
                    }
                }
                __builtin_amdgcn_s_waitcnt(0);   // the staged rows have landed too
                release(&s_bprog[wave - 2], G + 1);
            }
        }
        if (recompute_every > 0 && (sweep + 1) % recompute_every == 0) {
            // every wave is past the sweep's last window: the fields from scratch in FP64
            __syncthreads();
            for (int g0 = ib + tid; g0 < ie; g0 += kFieldGroup * kDenseTeamBlock) {
                double sc[kFieldGroup];
                scratch_fields<kDenseTeamBlock>(d, x, g0, sc);
#pragma unroll
                for (int qq = 0; qq < kFieldGroup; ++qq)
                    if (g0 + qq * kDenseTeamBlock < ie) f[g0 + qq * kDenseTeamBlock - ib] = T(sc[qq]);
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (s_abort) return;
    // outputs: this part's spins, its share of the FP64 energy and of max |f - f_scratch|
    double e = 0.0, er = 0.0;
    for (int g0 = ib + tid; g0 < ie; g0 += kFieldGroup * kDenseTeamBlock) {
        double sc[kFieldGroup];
        scratch_fields<kDenseTeamBlock>(d, x, g0, sc);
#pragma unroll
        for (int q = 0; q < kFieldGroup; ++q) {
            const int i = g0 + q * kDenseTeamBlock;
            if (i >= ie) continue;
            er = fmax(er, fabs(double(f[i - ib]) - sc[q]));
            if (x[i]) e += double(d.h[i]) + 0.5 * (sc[q] - double(d.h[i]));
            x_out[size_t(r) * size_t(n) + size_t(i)] = x[i];
        }
    }
    for (int o = 32; o > 0; o /= 2) {
        e += __shfl_down(e, o);
        er = fmax(er, __shfl_down(er, o));
    }
    if ((tid & 63) == 0) {
        s_red[tid / 64][0] = e;
        s_red[tid / 64][1] = er;
    }
    __syncthreads();
    if (tid == 0) {
        double es = 0.0, em = 0.0;
        for (int wv = 0; wv < kDenseTeamBlock / 64; ++wv) {
            es += s_red[wv][0];
            em = fmax(em, s_red[wv][1]);
        }
        tw.part_e[size_t(r) * kDenseParts + part] = es;
        tw.part_err[size_t(r) * kDenseParts + part] = em;
        __builtin_amdgcn_s_waitcnt(0);
        // the last part to finish adds the parts' shares in part order
        const unsigned before =
            __hip_atomic_fetch_add(tw.done + 32 * size_t(r), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (before + 1u == unsigned(kDenseParts)) {
            double et = 0.0, mt = 0.0;
            for (int q = 0; q < kDenseParts; ++q) {
                et += ld_l2(tw.part_e + size_t(r) * kDenseParts + q);
                mt = fmax(mt, ld_l2(tw.part_err + size_t(r) * kDenseParts + q));
            }
            energy[r] = et;
            field_err[r] = mt;
        }
    }
}

size_t dense_lds_bytes(int n, int precision) {
    return 2 * sizeof(float) * kDenseWin * kDenseWin + size_t(n) * (precision == 1 ? sizeof(float) : sizeof(double)) +
           size_t(n);
}

hipError_t launch_dense(const DevDense& d, int precision, int R, const double* beta, int n_sweeps,
                        int recompute_every, uint64_t seed_offset, uint8_t* x_out, double* energy, double* field_err,
                        hipStream_t st) {
    const size_t lds = dense_lds_bytes(d.n, precision);
    if (precision == 1) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dense_anneal_kernel<float>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        dense_anneal_kernel<float><<<R, kDenseBlock, lds, st>>>(d, beta, n_sweeps, recompute_every, seed_offset, x_out,
                                                                energy, field_err);
    } else {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dense_anneal_kernel<double>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        dense_anneal_kernel<double><<<R, kDenseBlock, lds, st>>>(d, beta, n_sweeps, recompute_every, seed_offset, x_out,
                                                                 energy, field_err);
    }
    return hipGetLastError();
}

hipError_t launch_dense_team(const DevDense& d, int precision, int R, const double* beta, int n_sweeps,
                             int recompute_every, uint64_t seed_offset, uint8_t* x_out, double* energy,
                             double* field_err, void* work, int* err, hipStream_t st) {
    if (R < 1 || R > kTeams * kDenseTeamSlots || d.n < kDenseWin * kDenseParts) return hipErrorInvalidValue;
    const size_t lds = dense_team_lds_bytes(d.n, precision);
    const DenseTeamWork tw = dense_team_work(work, d.n, R);
    // every XCD gets kDenseParts workgroups per replica slot it holds (round-robin dispatch)
    const unsigned grid = unsigned(kTeams * kDenseParts * ((R + kTeams - 1) / kTeams));
    if (precision == 1) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dense_team_kernel<float>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        dense_team_kernel<float><<<grid, kDenseTeamBlock, lds, st>>>(d, tw, R, beta, n_sweeps, recompute_every, seed_offset,
                                                                 x_out, energy, field_err, err);
    } else {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dense_team_kernel<double>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        dense_team_kernel<double><<<grid, kDenseTeamBlock, lds, st>>>(d, tw, R, beta, n_sweeps, recompute_every,
                                                                  seed_offset, x_out, energy, field_err, err);
    }
    return hipGetLastError();
}

}  // namespace qb
