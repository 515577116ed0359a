// qb_host.cpp -- host side of libqublar_hip.so: argument checks, the colouring and
// colour-class layout, device residency, and the C ABI of include/qublar.h.
#include "qublar.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "qb_internal.h"


using namespace qb;

// ---- errors -------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define QB_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(QB_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                \
    } while (0)

extern "C" const char* qb_last_error(void) { return g_err.c_str(); }

namespace qb {
// A device builder's failure (qb_build.hip), recorded for qb_last_error.
int build_fail(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace qb
extern "C" int qb_version(void) { return 100; }

// Device buffer that frees itself.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int reserve(size_t n) {
        if (n <= bytes) return QB_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, n ? n : 1) != hipSuccess) { p = nullptr; return fail(QB_E_NOMEM, "hipMalloc(%zu) failed", n); }
        bytes = n;
        return QB_OK;
    }
    template <typename T>
    int upload(const std::vector<T>& v, hipStream_t st) {
        int rc = reserve(sizeof(T) * v.size());
        if (rc) return rc;
        if (!v.empty() && hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(QB_E_HIP, "hipMemcpyAsync H2D failed");
        return QB_OK;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

static int device_props(hipDeviceProp_t& prop) {
    int dev = 0;
    QB_HIP(hipGetDevice(&dev));
    QB_HIP(hipGetDeviceProperties(&prop, dev));
    return QB_OK;
}

extern "C" int qb_get_device_info(qb_device_info* out) {
    if (!out) return fail(QB_E_ARG, "qb_get_device_info: null output");
    hipDeviceProp_t prop;
    int rc = device_props(prop);
    if (rc) return rc;
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->name, sizeof out->name, "%s", prop.name);
    std::snprintf(out->arch, sizeof out->arch, "%s", prop.gcnArchName);
    out->compute_units = prop.multiProcessorCount;
    out->xcds = prop.multiProcessorCount >= 256 ? 8 : 1;
    out->hbm_bytes = int64_t(prop.totalGlobalMem);
    out->lds_bytes_per_cu = int(prop.maxSharedMemoryPerMultiProcessor);
    return QB_OK;
}

extern "C" int qb_copy_probe(int64_t bytes, int reps, double* ms_per_copy) {
    if (bytes < 16 || bytes % 16 != 0 || reps < 1 || !ms_per_copy)
        return fail(QB_E_ARG, "qb_copy_probe: bytes must be a positive multiple of 16, reps >= 1");
    DevBuf src, dst;
    int rc = src.reserve(size_t(bytes));
    if (!rc) rc = dst.reserve(size_t(bytes));
    if (rc) return rc;
    hipStream_t st = nullptr;
    QB_HIP(hipMemsetAsync(src.p, 1, size_t(bytes), st));
    const size_t n16 = size_t(bytes) / 16;
    QB_HIP(launch_copy_probe(src.p, dst.p, n16, st));   // first touch, untimed
    hipEvent_t e0, e1;
    QB_HIP(hipEventCreate(&e0));
    QB_HIP(hipEventCreate(&e1));
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < reps; ++i) (void)launch_copy_probe(src.p, dst.p, n16, st);
    (void)hipEventRecord(e1, st);
    const hipError_t se = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (se != hipSuccess) return fail(QB_E_HIP, "qb_copy_probe: %s", hipGetErrorString(se));
    QB_HIP(hipGetLastError());
    *ms_per_copy = double(ms) / reps;
    return QB_OK;
}

// ---- schedule -----------------------------------------------------------------------
// T(s) = t_hot (t_cold / t_hot)^min(1, s / max(1, A - 1)) for s < A, then t_cold
// (src/ising_recon.hpp:129-131); the engine uses beta = 1 / T.
static std::vector<double> betas_of(const qb_schedule& s) {
    const int total = s.anneal_sweeps + s.hold_sweeps;
    std::vector<double> beta(size_t(std::max(total, 0)));
    for (int sweep = 0; sweep < total; ++sweep) {
        const double frac = std::min(1.0, double(sweep) / double(std::max(1, s.anneal_sweeps - 1)));
        const double temp = sweep < s.anneal_sweeps ? s.t_hot * std::pow(s.t_cold / s.t_hot, frac) : s.t_cold;
        beta[size_t(sweep)] = 1.0 / temp;
    }
    return beta;
}

static int check_schedule(const qb_schedule* s) {
    if (!s) return fail(QB_E_ARG, "null schedule");
    if (s->anneal_sweeps < 0 || s->hold_sweeps < 0 || s->anneal_sweeps + s->hold_sweeps <= 0)
        return fail(QB_E_ARG, "schedule needs at least one sweep (anneal %d, hold %d)", s->anneal_sweeps,
                    s->hold_sweeps);
    if (!(s->t_hot > 0.0) || !(s->t_cold > 0.0) || !std::isfinite(s->t_hot) || !std::isfinite(s->t_cold))
        return fail(QB_E_ARG, "temperatures must be positive and finite (t_hot %g, t_cold %g)", s->t_hot, s->t_cold);
    return QB_OK;
}

// ---- greedy colouring ---------------------------------------------------------------
// Bits in `visit` order take the smallest colour not used by any bit they conflict
// with; conflicts(i, mark) calls mark(j) for every conflicting bit j.
template <typename Conflicts>
static int greedy_colour(int n, const std::vector<int>& visit, Conflicts conflicts, std::vector<int>& colour) {
    colour.assign(size_t(n), -1);
    std::vector<int> stamp;
    int n_colours = 0;
    for (int step = 0; step < n; ++step) {
        const int i = visit[size_t(step)];
        conflicts(i, [&](int j) {
            const int c = colour[size_t(j)];
            if (c >= 0 && j != i) {
                if (size_t(c) >= stamp.size()) stamp.resize(size_t(c) + 64, -1);
                stamp[size_t(c)] = i;
            }
        });
        int c = 0;
        while (size_t(c) < stamp.size() && stamp[size_t(c)] == i) ++c;
        colour[size_t(i)] = c;
        n_colours = std::max(n_colours, c + 1);
    }
    return n_colours;
}

// ---- saturation colouring ------------------------------------------------------------
// DSatur (Brelaz, CACM 22(4), 1979): the next bit coloured is the uncoloured one whose
// conflicting bits already carry the most distinct colours (its saturation), ties by
// conflict weight, then by index; it takes the smallest colour none of them carries.
// On ray x voxel problems it needs markedly fewer classes -- fewer dependent sweep
// steps -- than one greedy pass in weight order. The queue holds every uncoloured bit
// in the rank bitmap of its saturation level (rank: position in weight order), so the
// next bit is the highest rank of the highest non-empty level.
namespace {
struct RankSet {
    std::vector<uint64_t> w, s;   // rank bitmap and its summary (one bit per non-zero word)
    int count = 0;
    void init(int n) {
        w.assign(size_t(n + 63) / 64, 0);
        s.assign((w.size() + 63) / 64, 0);
    }
    void insert(int r) {
        uint64_t& x = w[size_t(r >> 6)];
        if (!x) s[size_t(r >> 12)] |= uint64_t(1) << ((r >> 6) & 63);
        x |= uint64_t(1) << (r & 63);
        ++count;
    }
    void erase(int r) {
        uint64_t& x = w[size_t(r >> 6)];
        x &= ~(uint64_t(1) << (r & 63));
        if (!x) s[size_t(r >> 12)] &= ~(uint64_t(1) << ((r >> 6) & 63));
        --count;
    }
    int top() const {
        for (size_t k = s.size(); k-- > 0;)
            if (s[k]) {
                const size_t wi = k * 64 + size_t(63 - __builtin_clzll(s[k]));
                return int(wi * 64 + size_t(63 - __builtin_clzll(w[wi])));
            }
        return -1;
    }
};
}  // namespace

// `first`: bits coloured before the others, greedily in the given order (they see only
// each other). conflicts(i, mark) calls mark(j) for every conflicting bit j.
template <typename Conflicts>
static int saturation_colour(int n, const std::vector<int>& by_weight, const std::vector<int>& first,
                             Conflicts conflicts, std::vector<int>& colour) {
    colour.assign(size_t(n), -1);
    std::vector<int> rank(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) rank[size_t(by_weight[size_t(k)])] = n - 1 - k;
    int W = 2;                                    // words per saturation set
    std::vector<uint64_t> sat(size_t(n) * size_t(W), 0);
    std::vector<int> level(static_cast<size_t>(n), 0);
    std::vector<RankSet> queue;
    bool queued = false;
    int K = 0, top = 0;
    auto first_free = [&](int i) {
        const uint64_t* st = &sat[size_t(i) * size_t(W)];
        int k = 0;
        while (k < W && st[k] == ~uint64_t(0)) ++k;
        return k == W ? 64 * W : 64 * k + __builtin_ctzll(~st[k]);
    };
    auto enqueue = [&](int j) {
        const int l = level[size_t(j)];
        while (int(queue.size()) <= l) {
            queue.emplace_back();
            queue.back().init(n);
        }
        queue[size_t(l)].insert(rank[size_t(j)]);
        top = std::max(top, l);
    };
    auto assign = [&](int i, int c) {
        if (c >= 64 * W) {                        // widen every set
            std::vector<uint64_t> wider(size_t(n) * size_t(2 * W), 0);
            for (int j = 0; j < n; ++j)
                std::memcpy(&wider[size_t(j) * size_t(2 * W)], &sat[size_t(j) * size_t(W)], sizeof(uint64_t) * size_t(W));
            sat.swap(wider);
            W *= 2;
        }
        colour[size_t(i)] = c;
        K = std::max(K, c + 1);
        conflicts(i, [&](int j) {
            if (colour[size_t(j)] >= 0) return;
            uint64_t& word = sat[size_t(j) * size_t(W) + size_t(c >> 6)];
            const uint64_t m = uint64_t(1) << (c & 63);
            if (word & m) return;
            word |= m;
            if (queued) queue[size_t(level[size_t(j)])].erase(rank[size_t(j)]);
            ++level[size_t(j)];
            if (queued) enqueue(j);
        });
    };
    for (int i : first) assign(i, first_free(i));
    for (int i = 0; i < n; ++i)
        if (colour[size_t(i)] < 0) enqueue(i);
    queued = true;
    for (;;) {
        while (top >= 0 && queue[size_t(top)].count == 0) --top;
        if (top < 0) break;
        const int r = queue[size_t(top)].top();
        queue[size_t(top)].erase(r);
        const int i = by_weight[size_t(n - 1 - r)];
        assign(i, first_free(i));
    }
    return K;
}

// Positions: bits sorted by (colour, index).
static void order_by_colour(const std::vector<int>& colour, int n_colours, std::vector<int32_t>& pos_bit,
                            std::vector<int32_t>& cls_ptr) {
    cls_ptr.assign(size_t(n_colours) + 1, 0);
    for (int c : colour) cls_ptr[size_t(c) + 1]++;
    for (int c = 0; c < n_colours; ++c) cls_ptr[size_t(c) + 1] += cls_ptr[size_t(c)];
    std::vector<int32_t> fill(cls_ptr.begin(), cls_ptr.end() - 1);
    pos_bit.assign(colour.size(), 0);
    for (size_t i = 0; i < colour.size(); ++i) pos_bit[size_t(fill[size_t(colour[i])]++)] = int32_t(i);
}

// ==== float path =======================================================================

struct qb_problem {
    hipStream_t stream = nullptr;
    DevProblem dev{};
    int n = 0, n_rays = 0, n_classes = 0;
    int64_t nnz = 0, n_nbr = 0;
    int max_class = 0, min_class = 0, max_entries = 0, max_bits_per_ray = 0;
    int64_t device_bytes = 0;
    std::vector<int32_t> pos_bit, cls_ptr;
    DevBuf b_cls, b_pos, b_col, b_ent, b_nbp, b_nb, b_nbbit, b_lin, b_d, b_w, b_rtp, b_rt, b_seq, b_dbit, b_dkey,
        b_lbeg, b_lnb, b_bitpos;
    std::vector<int32_t> light_begin;   // host copy: each class's first light position
    std::vector<int32_t> col_ptr, nb_ptr;   // host copies, position space
    std::vector<uint8_t> seq_update;
    struct Batches {
        DevBuf cls, cls_heavy, first, count, pos, slot0, desc, blk, blk_meta;
        DevBatches dev{};
    };
    std::map<int, std::unique_ptr<Batches>> batches;   // by group width and team size
    // workspace
    DevBuf w_x, w_r, w_beta, w_bar, w_err, w_part, w_pv, w_tri, w_cert, w_br, w_data, w_prior, w_rerr, w_X, w_rref,
        w_rng, w_energy;
    int device = 0;   // the device current at creation: its stream, buffers and scratch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~qb_problem() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int validate_problem(const qb_problem_desc* d) {
    if (!d) return fail(QB_E_ARG, "null problem description");
    if (d->n_vars < 0 || d->n_rays < 0) return fail(QB_E_ARG, "negative sizes (n_vars %d, n_rays %d)", d->n_vars, d->n_rays);
    if (!d->row_ptr || !d->nbr_ptr) return fail(QB_E_ARG, "row_ptr and nbr_ptr are required");
    if (d->n_rays > 0 && (!d->d || !d->w)) return fail(QB_E_ARG, "d and w are required when n_rays > 0");
    if (d->row_ptr[0] != 0 || d->nbr_ptr[0] != 0) return fail(QB_E_ARG, "row_ptr[0] and nbr_ptr[0] must be 0");
    for (int i = 0; i < d->n_vars; ++i) {
        if (d->row_ptr[i + 1] < d->row_ptr[i]) return fail(QB_E_ARG, "row_ptr decreases at bit %d", i);
        if (d->nbr_ptr[i + 1] < d->nbr_ptr[i]) return fail(QB_E_ARG, "nbr_ptr decreases at bit %d", i);
    }
    const int64_t nnz = d->row_ptr[d->n_vars], nn = d->nbr_ptr[d->n_vars];
    if (nnz > 0 && (!d->ray || !d->a)) return fail(QB_E_ARG, "ray and a are required when entries exist");
    if (nn > 0 && !d->nbr) return fail(QB_E_ARG, "nbr is required when neighbours exist");
    for (int64_t k = 0; k < nnz; ++k)
        if (d->ray[k] < 0 || d->ray[k] >= d->n_rays) return fail(QB_E_ARG, "entry %lld names ray %d of %d", (long long)k, d->ray[k], d->n_rays);
    for (int64_t k = 0; k < nn; ++k)
        if (d->nbr[k] < 0 || d->nbr[k] >= d->n_vars) return fail(QB_E_ARG, "neighbour entry %lld names bit %d of %d", (long long)k, d->nbr[k], d->n_vars);
    if (d->draw_bit)
        for (int i = 0; i < d->n_vars; ++i)
            if (d->draw_bit[i] < 0) return fail(QB_E_ARG, "draw_bit[%d] is negative (%d)", i, d->draw_bit[i]);
    return QB_OK;
}

// Distance-2 colouring of the bit-ray incidence plus the neighbour graph: two bits
// share a colour only if they share no ray and are not neighbours. Conflict weight =
// the bit's rays' lengths plus its neighbours. Bits with more entries than a sweep
// workgroup has chain slots (kHeavyEntries) are coloured first, greedily by decreasing
// weight (ties by index); then every other bit in saturation order.
// Inside a class the order is free; it is used to group the work: bits with more
// entries than one chain first, then the others, and last the light bits -- no entry
// and at most kLightNb neighbours -- which the sweep runs without batches. Each part
// keeps index order.
static bool light_bit(const qb_problem_desc* d, int i) {
    return d->row_ptr[i + 1] == d->row_ptr[i] && d->nbr_ptr[i + 1] - d->nbr_ptr[i] <= kLightNb;
}
static void order_within_classes(const qb_problem_desc* desc, int K, std::vector<int32_t>& pos_bit,
                                 const std::vector<int32_t>& cls_ptr) {
    auto part = [&](int32_t i) { return desc->row_ptr[i + 1] - desc->row_ptr[i] > kSeg ? 0 : (light_bit(desc, i) ? 2 : 1); };
    for (int c = 0; c < K; ++c)
        std::stable_sort(pos_bit.begin() + cls_ptr[size_t(c)], pos_bit.begin() + cls_ptr[size_t(c) + 1],
                         [&](int32_t a, int32_t b) { return part(a) < part(b); });
}

// A caller's colouring: verified at distance 2 (bits of one ray, neighbours), unused
// colours dropped (order kept). Returns the number of classes, or -1 with g_err set.
static int given_colouring(const qb_problem_desc* desc, const std::vector<int32_t>& rt_ptr,
                           const std::vector<int32_t>& rt_bit, std::vector<int>& colour) {
    const int n = desc->n_vars, nr = desc->n_rays;
    int K = 0;
    for (int i = 0; i < n; ++i) {
        if (desc->colour[i] < 0) return -fail(QB_E_ARG, "colour[%d] is negative", i);
        K = std::max(K, desc->colour[i] + 1);
    }
    std::vector<int32_t> seen_ray(static_cast<size_t>(K), -1), seen_bit(static_cast<size_t>(K), -1);
    for (int b = 0; b < nr; ++b)
        for (int q = rt_ptr[size_t(b)]; q < rt_ptr[size_t(b) + 1]; ++q) {
            const int i = rt_bit[size_t(q)], c = desc->colour[i];
            if (seen_ray[size_t(c)] == b && seen_bit[size_t(c)] != i)
                return -fail(QB_E_ARG, "colour: bits %d and %d share ray %d and colour %d", seen_bit[size_t(c)], i, b, c);
            seen_ray[size_t(c)] = b;
            seen_bit[size_t(c)] = i;
        }
    for (int i = 0; i < n; ++i)
        for (int k = desc->nbr_ptr[i]; k < desc->nbr_ptr[i + 1]; ++k)
            if (desc->nbr[k] != i && desc->colour[desc->nbr[k]] == desc->colour[i])
                return -fail(QB_E_ARG, "colour: neighbours %d and %d share colour %d", i, desc->nbr[k], desc->colour[i]);
    std::vector<int> remap(static_cast<size_t>(K), -1);
    for (int i = 0; i < n; ++i) remap[size_t(desc->colour[i])] = 0;
    int used = 0;
    for (int c = 0; c < K; ++c)
        if (remap[size_t(c)] == 0) remap[size_t(c)] = used++;
    colour.resize(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) colour[size_t(i)] = remap[size_t(desc->colour[i])];
    return used;
}

static int colour_problem(const qb_problem_desc* desc, const std::vector<int32_t>& rt_ptr,
                          const std::vector<int32_t>& rt_bit, std::vector<int32_t>& pos_bit,
                          std::vector<int32_t>& cls_ptr, int& max_bits_per_ray) {
    const int n = desc->n_vars, nr = desc->n_rays;
    max_bits_per_ray = 0;
    for (int b = 0; b < nr; ++b) max_bits_per_ray = std::max(max_bits_per_ray, rt_ptr[size_t(b) + 1] - rt_ptr[size_t(b)]);
    if (desc->colour) {
        std::vector<int> colour;
        const int K = given_colouring(desc, rt_ptr, rt_bit, colour);
        if (K < 0) return K;
        order_by_colour(colour, K, pos_bit, cls_ptr);
        order_within_classes(desc, K, pos_bit, cls_ptr);
        return K;
    }
    std::vector<int64_t> weight(static_cast<size_t>(n), 0);
    for (int i = 0; i < n; ++i) {
        int64_t wsum = desc->nbr_ptr[i + 1] - desc->nbr_ptr[i];
        for (int k = desc->row_ptr[i]; k < desc->row_ptr[i + 1]; ++k)
            wsum += rt_ptr[size_t(desc->ray[k]) + 1] - rt_ptr[size_t(desc->ray[k])];
        weight[size_t(i)] = wsum;
    }
    std::vector<int> by_weight(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) by_weight[size_t(i)] = i;
    std::stable_sort(by_weight.begin(), by_weight.end(),
                     [&](int a, int b) { return weight[size_t(a)] > weight[size_t(b)]; });
    // the multi-pass bits first, so the few of them share few classes
    std::vector<int> heavy;
    for (int i : by_weight)
        if (desc->row_ptr[i + 1] - desc->row_ptr[i] > kHeavyEntries) heavy.push_back(i);
    std::vector<int> colour;
    const int K = saturation_colour(n, by_weight, heavy, [&](int i, auto mark) {
        for (int k = desc->row_ptr[i]; k < desc->row_ptr[i + 1]; ++k) {
            const int b = desc->ray[k];
            for (int q = rt_ptr[size_t(b)]; q < rt_ptr[size_t(b) + 1]; ++q) mark(rt_bit[size_t(q)]);
        }
        for (int k = desc->nbr_ptr[i]; k < desc->nbr_ptr[i + 1]; ++k) mark(desc->nbr[k]);
    }, colour);
    order_by_colour(colour, K, pos_bit, cls_ptr);
    order_within_classes(desc, K, pos_bit, cls_ptr);
    return K;
}

// ray -> bits in increasing bit index (then CSR order), with each entry's a.
static void ray_lists(const qb_problem_desc* desc, std::vector<int32_t>& rt_ptr, std::vector<int32_t>& rt_bit,
                      std::vector<float>& rt_a) {
    const int n = desc->n_vars, nr = desc->n_rays;
    const int64_t nnz = desc->row_ptr[n];
    rt_ptr.assign(size_t(nr) + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) rt_ptr[size_t(desc->ray[k]) + 1]++;
    for (int b = 0; b < nr; ++b) rt_ptr[size_t(b) + 1] += rt_ptr[size_t(b)];
    rt_bit.assign(size_t(nnz), 0);
    rt_a.assign(size_t(nnz), 0.0f);
    std::vector<int32_t> fill(rt_ptr.begin(), rt_ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int k = desc->row_ptr[i]; k < desc->row_ptr[i + 1]; ++k) {
            const int b = desc->ray[k];
            rt_bit[size_t(fill[size_t(b)])] = i;
            rt_a[size_t(fill[size_t(b)]++)] = desc->a[k];
        }
}

extern "C" int qb_colour_problem(const qb_problem_desc* desc, int32_t* order, int32_t* class_ptr, int* n_classes) {
    if (!order || !class_ptr || !n_classes) return fail(QB_E_ARG, "qb_colour_problem: null output");
    int rc = validate_problem(desc);
    if (rc) return rc;
    std::vector<int32_t> rt_ptr, rt_bit, pos_bit, cls_ptr;
    std::vector<float> rt_a;
    ray_lists(desc, rt_ptr, rt_bit, rt_a);
    int mbr = 0;
    const int K = colour_problem(desc, rt_ptr, rt_bit, pos_bit, cls_ptr, mbr);
    if (K < 0) return -K;
    std::memcpy(order, pos_bit.data(), sizeof(int32_t) * pos_bit.size());
    std::memcpy(class_ptr, cls_ptr.data(), sizeof(int32_t) * cls_ptr.size());
    *n_classes = K;
    return QB_OK;
}

extern "C" int qb_problem_create(const qb_problem_desc* desc, qb_problem** out) {
    if (!out) return fail(QB_E_ARG, "qb_problem_create: null output");
    *out = nullptr;
    int rc = validate_problem(desc);
    if (rc) return rc;
    const int n = desc->n_vars, nr = desc->n_rays;
    const int64_t nnz = desc->row_ptr[n], nn = desc->nbr_ptr[n];
    std::vector<int32_t> rt_ptr, rt_bit;
    std::vector<float> rt_a;
    ray_lists(desc, rt_ptr, rt_bit, rt_a);
    std::vector<int32_t> pos_bit, cls_ptr;
    int max_bits_per_ray = 0;
    const int K = colour_problem(desc, rt_ptr, rt_bit, pos_bit, cls_ptr, max_bits_per_ray);
    if (K < 0) return -K;

    auto P = new (std::nothrow) qb_problem();
    if (!P) return fail(QB_E_NOMEM, "out of host memory");
    std::unique_ptr<qb_problem> guard(P);
    P->n = n;
    P->n_rays = nr;
    P->n_classes = K;
    P->nnz = nnz;
    P->n_nbr = nn;
    P->max_bits_per_ray = max_bits_per_ray;
    P->pos_bit = std::move(pos_bit);
    P->cls_ptr = std::move(cls_ptr);
    P->max_class = 0;
    P->min_class = n;
    for (int c = 0; c < K; ++c) {
        const int sz = P->cls_ptr[size_t(c) + 1] - P->cls_ptr[size_t(c)];
        P->max_class = std::max(P->max_class, sz);
        P->min_class = std::min(P->min_class, sz);
    }
    std::vector<int32_t> pos_of(static_cast<size_t>(n));
    for (int p = 0; p < n; ++p) pos_of[size_t(P->pos_bit[size_t(p)])] = p;
    // colour-major CSC columns and neighbour lists
    std::vector<int32_t> col_ptr(size_t(n) + 1, 0), nb_ptr(size_t(n) + 1, 0);
    std::vector<RayEnt> ent(static_cast<size_t>(nnz));
    std::vector<int32_t> nb(static_cast<size_t>(nn)), nb_bit(static_cast<size_t>(nn));
    std::vector<double> lin(static_cast<size_t>(n));
    for (int p = 0; p < n; ++p) {
        const int i = P->pos_bit[size_t(p)];
        int k0 = col_ptr[size_t(p)];
        for (int k = desc->row_ptr[i]; k < desc->row_ptr[i + 1]; ++k, ++k0) {
            const int b = desc->ray[k];
            ent[size_t(k0)] = RayEnt{b, desc->a[k], desc->w[b]};
        }
        col_ptr[size_t(p) + 1] = k0;
        P->max_entries = std::max(P->max_entries, desc->row_ptr[i + 1] - desc->row_ptr[i]);
        int q0 = nb_ptr[size_t(p)];
        for (int k = desc->nbr_ptr[i]; k < desc->nbr_ptr[i + 1]; ++k, ++q0) {
            nb[size_t(q0)] = pos_of[size_t(desc->nbr[k])];
            nb_bit[size_t(q0)] = desc->nbr[k];
        }
        nb_ptr[size_t(p) + 1] = q0;
        lin[size_t(p)] = desc->field ? desc->field[i] : desc->kappa;
    }
    std::vector<uint8_t> seq_update(static_cast<size_t>(n), 0);
    {
        std::vector<int32_t> seen(static_cast<size_t>(nr), -1);
        for (int p = 0; p < n; ++p)
            for (int k = col_ptr[size_t(p)]; k < col_ptr[size_t(p) + 1]; ++k) {
                const int b = ent[size_t(k)].ray;
                if (seen[size_t(b)] == p) seq_update[size_t(p)] = 1;
                seen[size_t(b)] = p;
            }
    }
    // light bits: each class's last run of positions, neighbour positions at stride kLightNb
    std::vector<int32_t> light_begin(static_cast<size_t>(K), 0);
    std::vector<int32_t> light_nb(static_cast<size_t>(n) * kLightNb, -1);
    for (int c = 0; c < K; ++c) {
        int lb = P->cls_ptr[size_t(c) + 1];
        while (lb > P->cls_ptr[size_t(c)] && light_bit(desc, P->pos_bit[size_t(lb) - 1])) --lb;
        light_begin[size_t(c)] = lb;
        for (int p = lb; p < P->cls_ptr[size_t(c) + 1]; ++p)
            for (int k = nb_ptr[size_t(p)]; k < nb_ptr[size_t(p) + 1]; ++k)
                light_nb[size_t(p) * kLightNb + size_t(k - nb_ptr[size_t(p)])] = nb[size_t(k)];
    }
    // draw keys in position order
    std::vector<int32_t> draw_bit(static_cast<size_t>(n));
    std::vector<uint64_t> draw_key(static_cast<size_t>(n), 0u);
    for (int p = 0; p < n; ++p) {
        const int i = P->pos_bit[size_t(p)];
        draw_bit[size_t(p)] = desc->draw_bit ? desc->draw_bit[i] : i;
        if (desc->draw_key) draw_key[size_t(p)] = desc->draw_key[i];
    }
    std::vector<BitEnt> rt(static_cast<size_t>(nnz));
    for (size_t q = 0; q < rt.size(); ++q) rt[q] = BitEnt{pos_of[size_t(rt_bit[q])], rt_a[q]};
    std::vector<double> dv(desc->d, desc->d + nr), wv(desc->w, desc->w + nr);

    QB_HIP(hipGetDevice(&P->device));
    QB_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    QB_HIP(hipEventCreate(&P->ev0));
    QB_HIP(hipEventCreate(&P->ev1));
    hipStream_t st = P->stream;
    if ((rc = P->b_cls.upload(P->cls_ptr, st)) ||
        (rc = P->b_seq.upload(seq_update, st)) || (rc = P->b_pos.upload(P->pos_bit, st)) ||
        (rc = P->b_dbit.upload(draw_bit, st)) || (rc = P->b_dkey.upload(draw_key, st)) ||
        (rc = P->b_lbeg.upload(light_begin, st)) || (rc = P->b_lnb.upload(light_nb, st)) ||
        (rc = P->b_col.upload(col_ptr, st)) || (rc = P->b_ent.upload(ent, st)) || (rc = P->b_nbp.upload(nb_ptr, st)) ||
        (rc = P->b_nb.upload(nb, st)) || (rc = P->b_nbbit.upload(nb_bit, st)) || (rc = P->b_lin.upload(lin, st)) ||
        (rc = P->b_d.upload(dv, st)) || (rc = P->b_w.upload(wv, st)) || (rc = P->b_rtp.upload(rt_ptr, st)) ||
        (rc = P->b_rt.upload(rt, st)) || (rc = P->b_bitpos.upload(pos_of, st)))
        return rc;
    QB_HIP(hipStreamSynchronize(st));
    P->col_ptr = col_ptr;
    P->nb_ptr = nb_ptr;
    P->seq_update = seq_update;
    P->light_begin = light_begin;
    P->device_bytes = int64_t(P->b_dbit.bytes + P->b_dkey.bytes + P->b_lbeg.bytes + P->b_lnb.bytes +
                              P->b_cls.bytes + P->b_seq.bytes + P->b_pos.bytes + P->b_col.bytes + P->b_ent.bytes + P->b_nbp.bytes +
                              P->b_nb.bytes + P->b_nbbit.bytes + P->b_lin.bytes + P->b_d.bytes + P->b_w.bytes +
                              P->b_rtp.bytes + P->b_rt.bytes);
    DevProblem& D = P->dev;
    D.n = n;
    D.n_rays = nr;
    D.n_classes = K;
    D.nnz = nnz;
    D.n_nbr = nn;
    D.cls_ptr = P->b_cls.as<int32_t>();
    D.seq_update = P->b_seq.as<uint8_t>();
    D.draw_bit = P->b_dbit.as<int32_t>();
    D.draw_key = P->b_dkey.as<uint64_t>();
    D.light_begin = P->b_lbeg.as<int32_t>();
    D.light_nb = P->b_lnb.as<int32_t>();
    D.pos_bit = P->b_pos.as<int32_t>();
    D.bit_pos = P->b_bitpos.as<int32_t>();
    D.col_ptr = P->b_col.as<int32_t>();
    D.ent = P->b_ent.as<RayEnt>();
    D.nb_ptr = P->b_nbp.as<int32_t>();
    D.nb = P->b_nb.as<int32_t>();
    D.nb_bit = P->b_nbbit.as<int32_t>();
    D.lin = P->b_lin.as<double>();
    D.d = P->b_d.as<double>();
    D.w = P->b_w.as<double>();
    D.rt_ptr = P->b_rtp.as<int32_t>();
    D.rt = P->b_rt.as<BitEnt>();
    D.lambda = desc->lambda;
    *out = guard.release();
    return QB_OK;
}

extern "C" int qb_problem_destroy(qb_problem* p) {
    delete p;
    return QB_OK;
}

extern "C" int qb_problem_get_info(const qb_problem* p, qb_problem_info* out) {
    if (!p || !out) return fail(QB_E_ARG, "qb_problem_get_info: null argument");
    out->n_vars = p->n;
    out->n_rays = p->n_rays;
    out->nnz = p->nnz;
    out->n_nbr = p->n_nbr;
    out->n_classes = p->n_classes;
    out->max_class = p->max_class;
    out->min_class = p->min_class;
    out->max_entries_per_bit = p->max_entries;
    out->max_bits_per_ray = p->max_bits_per_ray;
    out->device_bytes = p->device_bytes;
    return QB_OK;
}

extern "C" int qb_problem_colouring(const qb_problem* p, int32_t* order, int32_t* class_ptr) {
    if (!p || !order || !class_ptr) return fail(QB_E_ARG, "qb_problem_colouring: null argument");
    std::memcpy(order, p->pos_bit.data(), sizeof(int32_t) * p->pos_bit.size());
    std::memcpy(class_ptr, p->cls_ptr.data(), sizeof(int32_t) * p->cls_ptr.size());
    return QB_OK;
}

// Workgroups per XCD team: `requested` total (a multiple of 8), or one per CU.
static int team_size_for(int requested, int precision, int bw, size_t table_bytes, bool staged = false) {
    hipDeviceProp_t prop;
    int rc = device_props(prop);
    if (rc) return -rc;
    const int per_cu = ray_sweep_blocks_per_cu(precision, bw, table_bytes, staged);
    if (per_cu <= 0)
        return -fail(QB_E_LIMIT, "the sweep kernel does not fit a CU with %zu bytes of dynamic LDS (class tables, "
                                 "multi-pass staging)", table_bytes);
    const int cap = per_cu * prop.multiProcessorCount;
    const int total = requested > 0 ? requested : prop.multiProcessorCount;
    const int limit = std::min(cap, 8 * kMaxTeamSize);
    if (total % 8 != 0 || total > limit)
        return -fail(QB_E_ARG, "workgroups must be a positive multiple of 8 and at most %d (got %d)", limit, total);
    return total / 8;
}

// True when some bit of the problem has more chains than a sweep workgroup has slots
// at group width bw -- a multi-pass batch of batches_for, whose staging needs LDS.
static bool has_multipass(const qb_problem* P, int bw) {
    const int slots = kSweepBlock / bw;
    for (size_t q = 0; q + 1 < P->col_ptr.size(); ++q)
        if ((P->col_ptr[q + 1] - P->col_ptr[q] + kSeg - 1) / kSeg > slots) return true;
    return false;
}

// Branches per group: the smallest of 2, 4, 8, 16 that spreads the branches over
// the 8 XCD teams in at most 8 groups, unless the caller fixes it; wider while the
// sweep kernel with the problem's dynamic LDS (class tables, multi-pass staging) does
// not fit a CU.
static int group_width_for(const qb_problem* P, int R, int requested, int precision) {
    if (requested) return requested;
    int bw = 2;
    while (bw < 16 && (R + bw - 1) / bw > 8) bw *= 2;
    while (bw < 16 &&
           ray_sweep_blocks_per_cu(precision, bw, sweep_dynamic_lds(P->n_classes, bw, has_multipass(P, bw), precision)) <= 0)
        bw *= 2;
    return bw;
}

// The work batches of every class at group width bw (cached per width): bits sorted
// by chain count, the largest first; a bit needing more chains than a workgroup has
// slots is a batch of its own, the rest are packed into workgroups next-fit.
// Packs each class's bits into batches for a team of `team_size` workgroups. A bit
// with more chains than a workgroup has slots is a multi-pass batch of its own. A
// class whose other bits fit the remaining workgroups in one round is spread over
// them evenly (largest bit first into the least filled batch), so its step costs one
// batch per workgroup; a larger class is cut into full batches, next fit.
static int batches_for(qb_problem* P, int bw, int team_size, const DevBatches** out) {
    const int key = bw * 65536 + team_size;
    auto found = P->batches.find(key);
    if (found != P->batches.end()) {
        *out = &found->second->dev;
        return QB_OK;
    }
    const int slots = kSweepBlock / bw;
    const int K = P->n_classes;
    std::vector<int32_t> cls_batch(static_cast<size_t>(K) + 1, 0), cls_heavy(static_cast<size_t>(K), 0), first,
        count, pos, slot0;
    std::vector<int32_t> items;
    for (int c = 0; c < K; ++c) {
        cls_batch[size_t(c)] = int32_t(first.size());
        items.clear();
        for (int q = P->cls_ptr[size_t(c)]; q < P->light_begin[size_t(c)]; ++q) items.push_back(q);
        auto chains = [&](int q) {
            return std::max(1, (P->col_ptr[size_t(q) + 1] - P->col_ptr[size_t(q)] + kSeg - 1) / kSeg);
        };
        std::stable_sort(items.begin(), items.end(), [&](int a, int b) { return chains(a) > chains(b); });
        // a bit needing more slots than a workgroup has: a multi-pass batch of its own
        std::vector<int32_t> packable;
        int n_heavy = 0;
        int64_t total = 0;
        for (int q : items) {
            const int k = chains(q);
            if (k > slots) {
                first.push_back(int32_t(pos.size()));
                count.push_back(-1);
                pos.push_back(q);
                slot0.push_back(0);
                ++n_heavy;
            } else {
                packable.push_back(q);
                total += k;
            }
        }
        auto emit = [&](const std::vector<int32_t>& bin) {
            first.push_back(int32_t(pos.size()));
            count.push_back(int32_t(bin.size()));
            int used = 0;
            for (int q : bin) {
                pos.push_back(q);
                slot0.push_back(used);
                used += chains(q);
            }
        };
        cls_heavy[size_t(c)] = n_heavy;
        const int room = team_size - n_heavy;   // one task per workgroup, multi-pass included
        if (!packable.empty() && room >= 1 && total <= int64_t(room) * slots) {
            // the class fits the team in one round: largest bit first into the least
            // filled of at most `room` batches, so every workgroup gets one batch
            const int B = std::min<int>(room, int(packable.size()));
            std::vector<std::vector<int32_t>> bins(static_cast<size_t>(B));
            std::vector<int> load(static_cast<size_t>(B), 0);
            for (int q : packable) {
                const int k = chains(q);
                size_t best = 0;
                for (size_t i = 1; i < load.size(); ++i)
                    if (load[i] < load[best]) best = i;
                if (load[best] + k > slots) {
                    bins.emplace_back();
                    load.push_back(0);
                    best = load.size() - 1;
                }
                bins[best].push_back(q);
                load[best] += k;
            }
            for (const auto& bin : bins)
                if (!bin.empty()) emit(bin);
        } else {
            // more than one round: full batches, next fit
            std::vector<int32_t> bin;
            int used = 0;
            for (int q : packable) {
                const int k = chains(q);
                if (!bin.empty() && used + k > slots) {
                    emit(bin);
                    bin.clear();
                    used = 0;
                }
                bin.push_back(q);
                used += k;
            }
            if (!bin.empty()) emit(bin);
        }
    }
    cls_batch[size_t(K)] = int32_t(first.size());
    // slot descriptors: one per chain of every packed bit, empty slots pos = -1
    std::vector<SlotDesc> desc(first.size() * size_t(slots), SlotDesc{-1, 0, 0, 0u});
    for (size_t bi = 0; bi < first.size(); ++bi) {
        if (count[bi] < 0) {
            // a multi-pass batch: its bit in slots 0 and 1, both with pos -1
            const int q = pos[size_t(first[bi])];
            desc[bi * size_t(slots)] = SlotDesc{-1, q, P->col_ptr[size_t(q)], uint32_t(P->col_ptr[size_t(q) + 1])};
            desc[bi * size_t(slots) + 1] = SlotDesc{-1, P->nb_ptr[size_t(q)], P->nb_ptr[size_t(q) + 1],
                                                    uint32_t(P->seq_update[size_t(q)] ? 1 : 0)};
            continue;
        }
        for (int it = 0; it < count[bi]; ++it) {
            const int q = pos[size_t(first[bi] + it)];
            const int k0 = P->col_ptr[size_t(q)], k1 = P->col_ptr[size_t(q) + 1];
            const int nch = std::max(1, (k1 - k0 + kSeg - 1) / kSeg);
            const int nn = P->nb_ptr[size_t(q) + 1] - P->nb_ptr[size_t(q)];
            if (nn > kMaxPackedNeighbours)
                return fail(QB_E_LIMIT, "bit %d has %d neighbours; the sweep's limit is %d", P->pos_bit[size_t(q)], nn,
                            kMaxPackedNeighbours);
            for (int j = 0; j < nch; ++j) {
                const int ka = k0 + j * kSeg;
                const int len = std::max(0, std::min(kSeg, k1 - ka));
                const uint32_t info = uint32_t(len) | (uint32_t(j) << 5) | (uint32_t(nch) << 13) |
                                      (uint32_t(nn) << 21) | (uint32_t(P->seq_update[size_t(q)] ? 1u : 0u) << 31);
                desc[bi * size_t(slots) + size_t(slot0[size_t(first[bi] + it)] + j)] =
                    SlotDesc{q, ka, P->nb_ptr[size_t(q)], info};
            }
        }
    }
    const int n_batches = int(first.size());
    if (sweep_table_bytes(K) > size_t(kSweepTableLimit))
        return fail(QB_E_LIMIT, "%d colour classes exceed the sweep's %d-byte LDS table", K, kSweepTableLimit);
    // every packed batch's block: its used slots (a prefix of the batch) with their
    // chains and heads, built on the device from the descriptors
    std::vector<int2> meta(first.size(), int2{0, 0});
    int64_t granules = 0;
    for (size_t bi = 0; bi < first.size(); ++bi) {
        if (count[bi] < 0) continue;
        int used = 0;
        while (used < slots && desc[bi * size_t(slots) + size_t(used)].pos >= 0) ++used;
        for (int q = used; q < slots; ++q)
            if (desc[bi * size_t(slots) + size_t(q)].pos >= 0)
                return fail(QB_E_ARG, "batch %zu: its used slots are not a prefix", bi);
        meta[bi] = int2{int(granules), used};
        granules += int64_t(used) * (kSeg + kHeadGranules);
        if (granules > int64_t(INT32_MAX))
            return fail(QB_E_LIMIT, "the sweep's batch blocks exceed 2^31 granules of 16 bytes");
    }
    auto B = std::make_unique<qb_problem::Batches>();
    int rc;
    if ((rc = B->cls.upload(cls_batch, P->stream)) || (rc = B->first.upload(first, P->stream)) ||
        (rc = B->count.upload(count, P->stream)) || (rc = B->pos.upload(pos, P->stream)) ||
        (rc = B->slot0.upload(slot0, P->stream)) || (rc = B->desc.upload(desc, P->stream)) ||
        (rc = B->cls_heavy.upload(cls_heavy, P->stream)) || (rc = B->blk_meta.upload(meta, P->stream)) ||
        (rc = B->blk.reserve(size_t(std::max<int64_t>(granules, 1)) * 16)))
        return rc;
    QB_HIP(launch_batch_blocks(P->dev, B->desc.as<SlotDesc>(), B->blk_meta.as<int2>(), n_batches, slots,
                               B->blk.as<uint4>(), P->stream));
    QB_HIP(hipStreamSynchronize(P->stream));
    const bool heavy = std::any_of(count.begin(), count.end(), [](int32_t c) { return c < 0; });
    B->dev = DevBatches{B->cls.as<int32_t>(), B->cls_heavy.as<int32_t>(), B->first.as<int32_t>(),
                        B->count.as<int32_t>(), B->pos.as<int32_t>(), B->slot0.as<int32_t>(),
                        B->desc.as<SlotDesc>(), B->blk.as<uint4>(), B->blk_meta.as<int2>(), n_batches,
                        heavy ? 1 : 0};
    *out = &B->dev;
    P->batches[key] = std::move(B);
    return QB_OK;
}

extern "C" int qb_anneal_branches(qb_problem* P, const qb_schedule* s, int R, uint64_t seed_offset,
                                  const qb_anneal_opts* opts, uint8_t* branches, float* p_void, uint8_t* tri,
                                  float* certainty, double* energy, qb_report* rep) {
    if (!P) return fail(QB_E_ARG, "qb_anneal_branches: null problem");
    int rc = check_schedule(s);
    if (rc) return rc;
    if (R <= 0) return fail(QB_E_ARG, "n_branches must be positive (got %d)", R);
    qb_anneal_opts o{};
    if (opts) o = *opts;
    if (o.precision != QB_FP64 && o.precision != QB_FP32) return fail(QB_E_ARG, "precision must be QB_FP64 or QB_FP32");
    if (o.recompute_every < 0 || o.sweeps_per_launch < 0 || o.workgroups < 0)
        return fail(QB_E_ARG, "negative option");
    if (o.group_width != 0 && o.group_width != 2 && o.group_width != 4 && o.group_width != 8 && o.group_width != 16)
        return fail(QB_E_ARG, "group_width must be 0, 2, 4, 8 or 16 (got %d)", o.group_width);
    const int total = s->anneal_sweeps + s->hold_sweeps;
    const int n = P->n, nr = P->n_rays;
    const int bw = group_width_for(P, R, o.group_width, o.precision);
    const int n_groups = (R + bw - 1) / bw;
    const size_t tsz = o.precision == QB_FP32 ? sizeof(float) : sizeof(double);
    const int team_size = team_size_for(o.workgroups, o.precision, bw, 0);
    if (team_size <= 0) return -team_size;
    const DevBatches* bt = nullptr;
    if ((rc = batches_for(P, bw, team_size, &bt))) return rc;
    {
        // the dynamic LDS must not lower the residency the team size was chosen for
        const int checked =
            team_size_for(o.workgroups, o.precision, bw, sweep_dynamic_lds(P->n_classes, bw, bt->heavy_lds != 0, o.precision));
        if (checked <= 0) return -checked;
    }
    const int n_part = ray_energy_blocks(P->dev, bw);
    if ((rc = P->w_x.reserve(sizeof(uint16_t) * size_t(n_groups) * size_t(std::max(n, 1)))) ||
        (rc = P->w_r.reserve(tsz * size_t(n_groups) * size_t(std::max(nr, 1)) * size_t(bw))) ||
        (rc = P->w_beta.reserve(sizeof(double) * size_t(total))) ||
        (rc = P->w_bar.reserve(sizeof(unsigned) * (kBarWords + kTicketWords))) ||
        (rc = P->w_err.reserve(sizeof(int))) ||
        (rc = P->w_part.reserve(sizeof(double) * size_t(n_part) * size_t(n_groups) * size_t(bw))) ||
        (rc = P->w_pv.reserve(sizeof(float) * size_t(std::max(n, 1)))) || (rc = P->w_tri.reserve(size_t(std::max(n, 1)))) ||
        (rc = P->w_cert.reserve(sizeof(float) * size_t(std::max(n, 1)))) ||
        (rc = P->w_br.reserve(branches ? size_t(R) * size_t(std::max(n, 1)) : 1)) ||
        (rc = P->w_data.reserve(sizeof(double) * size_t(R))) || (rc = P->w_prior.reserve(sizeof(double) * size_t(R))) ||
        (rc = P->w_rerr.reserve(sizeof(float))))
        return rc;
    hipStream_t st = P->stream;
    const std::vector<double> beta = betas_of(*s);
    QB_HIP(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    QB_HIP(hipMemcpyAsync(P->w_beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice, st));
    QB_HIP(launch_ray_init(P->dev, o.precision, bw, P->w_x.p, P->w_r.p, n_groups, st));
    QB_HIP(hipMemsetAsync(P->w_err.p, 0, sizeof(int), st));
    const int chunk = o.sweeps_per_launch > 0 ? std::min(o.sweeps_per_launch, total) : total;
    int launches = 0;
    QB_HIP(hipEventRecord(P->ev0, st));
    for (int sb = 0; sb < total; sb += chunk) {
        QB_HIP(hipMemsetAsync(P->w_bar.p, 0, sizeof(unsigned) * (kBarWords + kTicketWords), st));
        QB_HIP(launch_ray_sweeps(P->dev, *bt, o.precision, bw, P->w_x.p, P->w_r.p, n_groups, R, seed_offset,
                                 P->w_beta.as<double>(), sb, std::min(total, sb + chunk), o.recompute_every,
                                 P->w_bar.as<unsigned>(), P->w_bar.as<unsigned>() + kBarWords, P->w_err.as<int>(),
                                 team_size, st));
        ++launches;
    }
    QB_HIP(hipEventRecord(P->ev1, st));
    const bool want_outputs = p_void || tri || certainty || branches;
    if (want_outputs)
        QB_HIP(launch_ray_outputs(P->dev, P->w_x.as<uint16_t>(), n_groups, bw, R, P->w_pv.as<float>(),
                                  P->w_tri.as<uint8_t>(), P->w_cert.as<float>(),
                                  branches ? P->w_br.as<uint8_t>() : nullptr, st));
    const bool want_energy = energy || rep;
    if (want_energy) {
        QB_HIP(hipMemsetAsync(P->w_rerr.p, 0, sizeof(float), st));
        QB_HIP(launch_ray_energy(P->dev, P->w_x.as<uint16_t>(), n_groups, bw, R, P->w_r.p, o.precision,
                                 P->w_part.as<double>(), n_part, P->w_data.as<double>(), P->w_prior.as<double>(),
                                 P->w_rerr.as<float>(), st));
    }
    if (p_void) QB_HIP(hipMemcpyAsync(p_void, P->w_pv.p, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, st));
    if (tri) QB_HIP(hipMemcpyAsync(tri, P->w_tri.p, size_t(n), hipMemcpyDeviceToHost, st));
    if (certainty) QB_HIP(hipMemcpyAsync(certainty, P->w_cert.p, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, st));
    if (branches) QB_HIP(hipMemcpyAsync(branches, P->w_br.p, size_t(R) * size_t(n), hipMemcpyDeviceToHost, st));
    std::vector<double> dat(static_cast<size_t>(R)), pri(static_cast<size_t>(R));
    float rerr = 0.0f;
    int err_flag = 0;
    if (want_energy) {
        QB_HIP(hipMemcpyAsync(dat.data(), P->w_data.p, sizeof(double) * size_t(R), hipMemcpyDeviceToHost, st));
        QB_HIP(hipMemcpyAsync(pri.data(), P->w_prior.p, sizeof(double) * size_t(R), hipMemcpyDeviceToHost, st));
        QB_HIP(hipMemcpyAsync(&rerr, P->w_rerr.p, sizeof(float), hipMemcpyDeviceToHost, st));
    }
    QB_HIP(hipMemcpyAsync(&err_flag, P->w_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (err_flag == 2)
        return fail(QB_E_TIMEOUT, "an XCD team came up short of %d workgroups (workgroup placement)", team_size);
    if (err_flag) return fail(QB_E_TIMEOUT, "a sweep barrier wait exceeded its limit (%d workgroups)", 8 * team_size);
    if (energy)
        for (int b = 0; b < R; ++b) energy[b] = dat[size_t(b)] + pri[size_t(b)];
    if (rep) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, P->ev0, P->ev1));
        rep->wall_s = wall;
        rep->kernel_ms = ms;
        rep->launches = launches;
        rep->spin_updates = double(R) * double(total) * double(n);
        rep->fp32_residual_err = rerr;
    }
    return QB_OK;
}

// ==== the team barrier's visibility, checked on the device ===================================

extern "C" int qb_team_barrier_check(int steps, int workgroups, uint64_t* stale, double* us_per_step) {
    if (!stale || steps <= 0) return fail(QB_E_ARG, "qb_team_barrier_check: null output or no steps");
    const int team_size = team_size_for(workgroups, QB_FP32, 2, 0);
    if (team_size <= 0) return -team_size;
    DevBuf words, bar, err, cnt;
    int rc;
    if ((rc = words.reserve(sizeof(unsigned) * 256 * size_t(team_size) * kTeams)) ||
        (rc = bar.reserve(sizeof(unsigned) * (kBarWords + kTicketWords))) || (rc = err.reserve(sizeof(int))) ||
        (rc = cnt.reserve(sizeof(unsigned long long))))
        return rc;
    hipStream_t st = nullptr;
    QB_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        hipEvent_t a = nullptr, b = nullptr;
        ~Guard() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            (void)hipStreamDestroy(s);
        }
    } g{st};
    QB_HIP(hipEventCreate(&g.a));
    QB_HIP(hipEventCreate(&g.b));
    QB_HIP(hipMemsetAsync(words.p, 0, words.bytes, st));
    QB_HIP(hipMemsetAsync(bar.p, 0, bar.bytes, st));
    QB_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
    QB_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), st));
    QB_HIP(hipEventRecord(g.a, st));
    QB_HIP(launch_team_barrier_probe(words.as<unsigned>(), bar.as<unsigned>(), bar.as<unsigned>() + kBarWords,
                                     err.as<int>(), steps, team_size, cnt.as<unsigned long long>(), st));
    QB_HIP(hipEventRecord(g.b, st));
    unsigned long long bad = 0;
    int flag = 0;
    QB_HIP(hipMemcpyAsync(&bad, cnt.p, sizeof(bad), hipMemcpyDeviceToHost, st));
    QB_HIP(hipMemcpyAsync(&flag, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    if (flag == 2) return fail(QB_E_TIMEOUT, "an XCD team came up short of %d workgroups (workgroup placement)", team_size);
    if (flag) return fail(QB_E_TIMEOUT, "a team barrier wait exceeded its limit");
    *stale = bad;
    if (us_per_step) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, g.a, g.b));
        *us_per_step = double(ms) * 1e3 / steps;
    }
    return QB_OK;
}

// ==== staged solves ==========================================================================
// The branch state of one problem (in practice a union of requests) kept on the device
// between calls: stages of a few sweeps at a time, requests joining with their state
// imported from the previous union's stage, and leaving with their outputs read.

struct qb_stage {
    qb_problem* P = nullptr;
    int precision = QB_FP64, recompute_every = 0, R = 0, bw = 0, n_groups = 0, team_size = 0, total = 0;
    int sweeps = 0;                  // stage sweeps run so far
    const DevBatches* bt = nullptr;  // owned by P (cached per group width and team size)
    DevBuf x, r, beta, bar, err, clock, out_pv, out_tri, out_cert, out_br;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~qb_stage() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

extern "C" int qb_stage_create(qb_problem* P, const qb_schedule* s, int R, const qb_anneal_opts* opts,
                               qb_stage** out) {
    if (!P || !out) return fail(QB_E_ARG, "qb_stage_create: null argument");
    *out = nullptr;
    int rc = check_schedule(s);
    if (rc) return rc;
    if (R <= 0) return fail(QB_E_ARG, "n_branches must be positive (got %d)", R);
    qb_anneal_opts o{};
    if (opts) o = *opts;
    if (o.precision != QB_FP64 && o.precision != QB_FP32) return fail(QB_E_ARG, "precision must be QB_FP64 or QB_FP32");
    if (o.recompute_every < 0 || o.workgroups < 0) return fail(QB_E_ARG, "negative option");
    if (o.group_width != 0 && o.group_width != 2 && o.group_width != 4 && o.group_width != 8 && o.group_width != 16)
        return fail(QB_E_ARG, "group_width must be 0, 2, 4, 8 or 16 (got %d)", o.group_width);
    const int total = s->anneal_sweeps + s->hold_sweeps;
    if (total > int(kClockOffsetMask)) return fail(QB_E_LIMIT, "a staged schedule has at most %u sweeps", kClockOffsetMask);
    auto S = std::make_unique<qb_stage>();
    S->P = P;
    S->precision = o.precision;
    S->recompute_every = o.precision == QB_FP32 ? o.recompute_every : 0;
    S->R = R;
    S->total = total;
    S->bw = group_width_for(P, R, o.group_width, o.precision);
    S->n_groups = (R + S->bw - 1) / S->bw;
    S->team_size = team_size_for(o.workgroups, o.precision, S->bw, 0, true);
    if (S->team_size <= 0) return -S->team_size;
    if ((rc = batches_for(P, S->bw, S->team_size, &S->bt))) return rc;
    {
        const int checked = team_size_for(o.workgroups, o.precision, S->bw,
                                          sweep_dynamic_lds(P->n_classes, S->bw, S->bt->heavy_lds != 0, o.precision), true);
        if (checked <= 0) return -checked;
    }
    const int n = P->n, nr = P->n_rays;
    const size_t tsz = o.precision == QB_FP32 ? sizeof(float) : sizeof(double);
    if ((rc = S->x.reserve(sizeof(uint16_t) * size_t(S->n_groups) * size_t(std::max(n, 1)))) ||
        (rc = S->r.reserve(tsz * size_t(S->n_groups) * size_t(std::max(nr, 1)) * size_t(S->bw))) ||
        (rc = S->beta.reserve(sizeof(double) * size_t(total))) ||
        (rc = S->bar.reserve(sizeof(unsigned) * (kBarWords + kTicketWords))) || (rc = S->err.reserve(sizeof(int))) ||
        (rc = S->clock.reserve(sizeof(uint32_t) * size_t(std::max(n, 1)))))
        return rc;
    QB_HIP(hipEventCreate(&S->ev0));
    QB_HIP(hipEventCreate(&S->ev1));
    hipStream_t st = P->stream;
    const std::vector<double> beta = betas_of(*s);
    QB_HIP(hipMemcpyAsync(S->beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice, st));
    QB_HIP(launch_ray_init(P->dev, S->precision, S->bw, S->x.p, S->r.p, S->n_groups, st));
    QB_HIP(hipMemsetAsync(S->clock.p, 0, sizeof(uint32_t) * size_t(std::max(n, 1)), st));
    QB_HIP(hipMemsetAsync(S->err.p, 0, sizeof(int), st));
    QB_HIP(hipStreamSynchronize(st));
    *out = S.release();
    return QB_OK;
}

extern "C" int qb_stage_destroy(qb_stage* S) {
    if (S && S->P) (void)hipStreamSynchronize(S->P->stream);
    delete S;
    return QB_OK;
}

// A span of bits of a fresh stage: in range, the stage not yet run.
static int check_span(const qb_stage* S, int bit0, int n_bits, const char* what) {
    if (S->sweeps != 0) return fail(QB_E_ARG, "%s: the stage has run %d sweeps; spans are set before the first", what,
                                    S->sweeps);
    if (bit0 < 0 || n_bits < 0 || int64_t(bit0) + n_bits > S->P->n)
        return fail(QB_E_ARG, "%s: bits [%d, %d) outside [0, %d)", what, bit0, bit0 + n_bits, S->P->n);
    return QB_OK;
}

extern "C" int qb_stage_set_span(qb_stage* S, int bit0, int n_bits, int sweeps_done, int n_branches) {
    if (!S) return fail(QB_E_ARG, "qb_stage_set_span: null stage");
    if (int rc = check_span(S, bit0, n_bits, "qb_stage_set_span")) return rc;
    if (sweeps_done < 0 || sweeps_done > S->total)
        return fail(QB_E_ARG, "sweeps_done %d outside [0, %d]", sweeps_done, S->total);
    if (S->recompute_every > 0 && sweeps_done % S->recompute_every != 0)
        return fail(QB_E_ARG, "sweeps_done %d is not a multiple of the recompute period %d (the recompute after "
                              "every k-th own sweep must fall on the stage's)", sweeps_done, S->recompute_every);
    if (n_branches < 0 || n_branches > S->R || n_branches > kClockMaxBranches)
        return fail(QB_E_ARG, "n_branches %d outside [0, min(%d, %d)]", n_branches, S->R, kClockMaxBranches);
    const uint32_t word = uint32_t(sweeps_done) | (uint32_t(n_branches) << 24);
    QB_HIP(launch_stage_clock(S->P->dev, bit0, n_bits, word, S->clock.as<uint32_t>(), S->P->stream));
    QB_HIP(hipStreamSynchronize(S->P->stream));
    return QB_OK;
}

extern "C" int qb_stage_import(qb_stage* D, const qb_stage* S, int src_bit0, int dst_bit0, int n_bits, int src_ray0,
                               int dst_ray0, int n_rays) {
    if (!D || !S) return fail(QB_E_ARG, "qb_stage_import: null stage");
    if (D == S) return fail(QB_E_ARG, "qb_stage_import: a stage cannot import from itself");
    if (int rc = check_span(D, dst_bit0, n_bits, "qb_stage_import")) return rc;
    if (src_bit0 < 0 || int64_t(src_bit0) + n_bits > S->P->n)
        return fail(QB_E_ARG, "qb_stage_import: source bits [%d, %d) outside [0, %d)", src_bit0, src_bit0 + n_bits,
                    S->P->n);
    if (n_rays < 0 || src_ray0 < 0 || dst_ray0 < 0 || int64_t(src_ray0) + n_rays > S->P->n_rays ||
        int64_t(dst_ray0) + n_rays > D->P->n_rays)
        return fail(QB_E_ARG, "qb_stage_import: ray span outside a problem's rays");
    if (D->precision != S->precision || D->bw != S->bw)
        return fail(QB_E_ARG, "qb_stage_import: stages differ in precision or group width (%d/%d, %d/%d)",
                    D->precision, S->precision, D->bw, S->bw);
    if (D->total != S->total) return fail(QB_E_ARG, "qb_stage_import: stages differ in schedule length");
    if (D->recompute_every != S->recompute_every ||
        (S->recompute_every > 0 && S->sweeps % S->recompute_every != 0))
        return fail(QB_E_ARG, "qb_stage_import: a span moves only at a multiple of the recompute period (source at "
                              "sweep %d, period %d)", S->sweeps, S->recompute_every);
    if (int64_t(S->sweeps) + S->total > int64_t(kClockOffsetMask))
        return fail(QB_E_LIMIT, "qb_stage_import: the source stage has run too many sweeps for a clock offset");
    QB_HIP(hipStreamSynchronize(S->P->stream));
    QB_HIP(launch_stage_import(S->P->dev, S->x.p, S->r.p, S->clock.as<uint32_t>(), S->sweeps, D->P->dev, D->x.p, D->r.p,
                               D->clock.as<uint32_t>(), D->precision, D->bw, src_bit0, dst_bit0, n_bits, src_ray0,
                               dst_ray0, n_rays, std::min(S->n_groups, D->n_groups), D->P->stream));
    QB_HIP(hipStreamSynchronize(D->P->stream));
    return QB_OK;
}

extern "C" int qb_stage_run(qb_stage* S, int n_sweeps, qb_report* rep) {
    if (!S) return fail(QB_E_ARG, "qb_stage_run: null stage");
    if (n_sweeps <= 0) return fail(QB_E_ARG, "n_sweeps must be positive (got %d)", n_sweeps);
    if (int64_t(S->sweeps) + n_sweeps > int64_t(kClockOffsetMask) - S->total)
        return fail(QB_E_LIMIT, "qb_stage_run: the stage has run too many sweeps for its clocks");
    qb_problem* P = S->P;
    hipStream_t st = P->stream;
    const auto t0 = std::chrono::steady_clock::now();
    QB_HIP(hipEventRecord(S->ev0, st));
    QB_HIP(hipMemsetAsync(S->bar.p, 0, sizeof(unsigned) * (kBarWords + kTicketWords), st));
    QB_HIP(launch_ray_sweeps(P->dev, *S->bt, S->precision, S->bw, S->x.p, S->r.p, S->n_groups, S->R, 0,
                             S->beta.as<double>(), S->sweeps, S->sweeps + n_sweeps, S->recompute_every,
                             S->bar.as<unsigned>(), S->bar.as<unsigned>() + kBarWords, S->err.as<int>(), S->team_size,
                             st, S->clock.as<uint32_t>(), S->total));
    QB_HIP(hipEventRecord(S->ev1, st));
    int err_flag = 0;
    QB_HIP(hipMemcpyAsync(&err_flag, S->err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    if (err_flag == 2)
        return fail(QB_E_TIMEOUT, "an XCD team came up short of %d workgroups (workgroup placement)", S->team_size);
    if (err_flag) return fail(QB_E_TIMEOUT, "a sweep barrier wait exceeded its limit (%d workgroups)", 8 * S->team_size);
    S->sweeps += n_sweeps;
    if (rep) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, S->ev0, S->ev1));
        rep->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->kernel_ms = ms;
        rep->launches = 1;
        rep->spin_updates = double(S->R) * double(n_sweeps) * double(P->n);
        rep->fp32_residual_err = 0.0;
    }
    return QB_OK;
}

extern "C" int qb_stage_outputs(qb_stage* S, int bit0, int n_bits, int n_branches, uint8_t* branches, float* p_void,
                                uint8_t* tri, float* certainty) {
    if (!S) return fail(QB_E_ARG, "qb_stage_outputs: null stage");
    if (bit0 < 0 || n_bits < 0 || int64_t(bit0) + n_bits > S->P->n)
        return fail(QB_E_ARG, "qb_stage_outputs: bits [%d, %d) outside [0, %d)", bit0, bit0 + n_bits, S->P->n);
    if (n_branches <= 0 || n_branches > S->R)
        return fail(QB_E_ARG, "qb_stage_outputs: n_branches %d outside [1, %d]", n_branches, S->R);
    const size_t nb = size_t(std::max(n_bits, 1));
    int rc;
    if ((rc = S->out_pv.reserve(sizeof(float) * nb)) || (rc = S->out_tri.reserve(nb)) ||
        (rc = S->out_cert.reserve(sizeof(float) * nb)) ||
        (rc = S->out_br.reserve(branches ? size_t(n_branches) * nb : 1)))
        return rc;
    hipStream_t st = S->P->stream;
    QB_HIP(launch_stage_outputs(S->P->dev, S->x.as<uint16_t>(), S->bw, bit0, n_bits, n_branches, S->out_pv.as<float>(),
                                S->out_tri.as<uint8_t>(), S->out_cert.as<float>(),
                                branches ? S->out_br.as<uint8_t>() : nullptr, st));
    if (n_bits > 0) {
        if (p_void) QB_HIP(hipMemcpyAsync(p_void, S->out_pv.p, sizeof(float) * size_t(n_bits), hipMemcpyDeviceToHost, st));
        if (tri) QB_HIP(hipMemcpyAsync(tri, S->out_tri.p, size_t(n_bits), hipMemcpyDeviceToHost, st));
        if (certainty)
            QB_HIP(hipMemcpyAsync(certainty, S->out_cert.p, sizeof(float) * size_t(n_bits), hipMemcpyDeviceToHost, st));
        if (branches)
            QB_HIP(hipMemcpyAsync(branches, S->out_br.p, size_t(n_branches) * size_t(n_bits), hipMemcpyDeviceToHost, st));
    }
    QB_HIP(hipStreamSynchronize(st));
    return QB_OK;
}

extern "C" int qb_stage_sweeps(const qb_stage* S, int* sweeps) {
    if (!S || !sweeps) return fail(QB_E_ARG, "qb_stage_sweeps: null argument");
    *sweeps = S->sweeps;
    return QB_OK;
}

// Proposals one launch of the reference float trajectory makes at most when the caller
// leaves the split to the library (about a second at a few branches of the muon QUBO).
constexpr double kRefFloatUpdatesPerLaunch = 33554432.0;

extern "C" int qb_anneal_branches_reference(qb_problem* P, const qb_schedule* s, int n_branches, uint64_t seed_offset,
                                           uint8_t* branches, float* p_void, uint8_t* tri, float* certainty,
                                           double* energy, qb_report* rep) {
    return qb_anneal_branches_reference_split(P, s, n_branches, seed_offset, 0, branches, p_void, tri, certainty,
                                              energy, rep);
}

extern "C" int qb_anneal_branches_reference_split(qb_problem* P, const qb_schedule* s, int n_branches,
                                                 uint64_t seed_offset, int sweeps_per_launch, uint8_t* branches,
                                                 float* p_void, uint8_t* tri, float* certainty, double* energy,
                                                 qb_report* rep) {
    if (!P) return fail(QB_E_ARG, "qb_anneal_branches_reference: null problem");
    int rc = check_schedule(s);
    if (rc) return rc;
    const int R = n_branches;
    if (R <= 0) return fail(QB_E_ARG, "n_branches must be positive (got %d)", R);
    if (sweeps_per_launch < 0) return fail(QB_E_ARG, "sweeps_per_launch must be >= 0 (got %d)", sweeps_per_launch);
    const int total = s->anneal_sweeps + s->hold_sweeps;
    const int n = P->n, nr = P->n_rays;
    int per_launch = sweeps_per_launch;
    if (per_launch == 0)
        per_launch = int(std::clamp(kRefFloatUpdatesPerLaunch / (double(R) * double(std::max(n, 1))), 1.0,
                                    double(std::max(total, 1))));
    const int n_groups = (R + kGroup - 1) / kGroup;
    const int n_part = ray_energy_blocks(P->dev, kGroup);
    if ((rc = P->w_X.reserve(size_t(R) * size_t(std::max(n, 1)))) ||
        (rc = P->w_rref.reserve(sizeof(double) * size_t(R) * size_t(std::max(nr, 1)))) ||
        (rc = P->w_beta.reserve(sizeof(double) * size_t(total))) ||
        (rc = P->w_x.reserve(sizeof(uint16_t) * size_t(n_groups) * size_t(std::max(n, 1)))) ||
        (rc = P->w_part.reserve(sizeof(double) * size_t(n_part) * size_t(n_groups) * kGroup)) ||
        (rc = P->w_pv.reserve(sizeof(float) * size_t(std::max(n, 1)))) || (rc = P->w_tri.reserve(size_t(std::max(n, 1)))) ||
        (rc = P->w_cert.reserve(sizeof(float) * size_t(std::max(n, 1)))) ||
        (rc = P->w_br.reserve(branches ? size_t(R) * size_t(std::max(n, 1)) : 1)) ||
        (rc = P->w_data.reserve(sizeof(double) * size_t(R))) || (rc = P->w_prior.reserve(sizeof(double) * size_t(R))) ||
        (rc = P->w_rng.reserve(sizeof(uint64_t) * size_t(R))))
        return rc;
    // the temperature of every sweep by the reference's expression (src/ising_recon.hpp:128-130)
    std::vector<double> temp(static_cast<size_t>(total));
    for (int sweep = 0; sweep < total; ++sweep) {
        const double frac = std::min(1.0, double(sweep) / double(std::max(1, s->anneal_sweeps - 1)));
        temp[size_t(sweep)] = sweep < s->anneal_sweeps ? s->t_hot * std::pow(s->t_cold / s->t_hot, frac) : s->t_cold;
    }
    hipStream_t st = P->stream;
    QB_HIP(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    QB_HIP(hipMemcpyAsync(P->w_beta.p, temp.data(), sizeof(double) * temp.size(), hipMemcpyHostToDevice, st));
    QB_HIP(hipEventRecord(P->ev0, st));
    int launches = 0;
    for (int s0 = 0; s0 < total || launches == 0; s0 += per_launch, ++launches)
        QB_HIP(launch_ray_ref(P->dev, R, seed_offset, P->w_beta.as<double>(), s0, std::min(total, s0 + per_launch),
                              P->w_X.as<uint8_t>(), P->w_rref.as<double>(), P->w_rng.as<uint64_t>(), st));
    QB_HIP(hipEventRecord(P->ev1, st));
    // the branches in the group masks of the read-out and energy kernels
    if (n > 0) QB_HIP(launch_pack_states(P->w_X.as<uint8_t>(), R, n, P->dev.pos_bit, P->w_x.as<uint16_t>(), st));
    if (p_void || tri || certainty || branches)
        QB_HIP(launch_ray_outputs(P->dev, P->w_x.as<uint16_t>(), n_groups, kGroup, R, P->w_pv.as<float>(),
                                  P->w_tri.as<uint8_t>(), P->w_cert.as<float>(),
                                  branches ? P->w_br.as<uint8_t>() : nullptr, st));
    const bool want_energy = energy != nullptr;
    if (want_energy)
        QB_HIP(launch_ray_energy(P->dev, P->w_x.as<uint16_t>(), n_groups, kGroup, R, nullptr, 0, P->w_part.as<double>(),
                                 n_part, P->w_data.as<double>(), P->w_prior.as<double>(), nullptr, st));
    if (p_void) QB_HIP(hipMemcpyAsync(p_void, P->w_pv.p, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, st));
    if (tri) QB_HIP(hipMemcpyAsync(tri, P->w_tri.p, size_t(n), hipMemcpyDeviceToHost, st));
    if (certainty) QB_HIP(hipMemcpyAsync(certainty, P->w_cert.p, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, st));
    if (branches) QB_HIP(hipMemcpyAsync(branches, P->w_br.p, size_t(R) * size_t(n), hipMemcpyDeviceToHost, st));
    std::vector<double> dat(static_cast<size_t>(R)), pri(static_cast<size_t>(R));
    if (want_energy) {
        QB_HIP(hipMemcpyAsync(dat.data(), P->w_data.p, sizeof(double) * size_t(R), hipMemcpyDeviceToHost, st));
        QB_HIP(hipMemcpyAsync(pri.data(), P->w_prior.p, sizeof(double) * size_t(R), hipMemcpyDeviceToHost, st));
    }
    QB_HIP(hipStreamSynchronize(st));
    if (energy)
        for (int b = 0; b < R; ++b) energy[b] = dat[size_t(b)] + pri[size_t(b)];
    if (rep) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, P->ev0, P->ev1));
        rep->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->kernel_ms = ms;
        rep->launches = launches;
        rep->spin_updates = double(R) * double(total) * double(n);
        rep->fp32_residual_err = 0.0;
    }
    return QB_OK;
}

extern "C" int qb_freeze(qb_problem* P, double margin, uint8_t* frozen, int* rounds) {
    if (!P || !frozen) return fail(QB_E_ARG, "qb_freeze: null argument");
    if (!std::isfinite(margin)) return fail(QB_E_ARG, "margin must be finite");
    if (P->dev.lambda < 0.0)
        return fail(QB_E_ARG, "freeze needs lambda >= 0 (then its fixpoint does not depend on the visit order)");
    const int n = P->n;
    int rc;
    // workspace: the two frozen masks, then the change counter on its own 64-byte line
    const size_t mask = size_t(std::max(n, 1));
    const size_t counter_at = (2 * mask + 63) / 64 * 64;
    if ((rc = P->w_part.reserve(sizeof(double) * mask)) || (rc = P->w_X.reserve(counter_at + 64))) return rc;
    hipStream_t st = P->stream;
    double* lb = P->w_part.as<double>();
    uint8_t* a = P->w_X.as<uint8_t>();
    uint8_t* b = a + mask;
    unsigned* changed = reinterpret_cast<unsigned*>(P->w_X.as<uint8_t>() + counter_at);
    QB_HIP(hipMemsetAsync(a, 0, mask, st));
    QB_HIP(launch_freeze_bound(P->dev, lb, st));
    int r = 0;
    for (;;) {
        ++r;
        QB_HIP(hipMemsetAsync(changed, 0, sizeof(unsigned), st));
        QB_HIP(launch_freeze_round(P->dev, lb, margin, a, b, changed, st));
        unsigned nc = 0;
        QB_HIP(hipMemcpyAsync(&nc, changed, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        QB_HIP(hipStreamSynchronize(st));
        std::swap(a, b);
        if (nc == 0 || r > n + 1) break;   // at most n rounds can freeze a bit
    }
    std::vector<uint8_t> byp(static_cast<size_t>(n));
    if (n > 0) QB_HIP(hipMemcpyAsync(byp.data(), a, size_t(n), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    for (int pos = 0; pos < n; ++pos) frozen[P->pos_bit[size_t(pos)]] = byp[size_t(pos)];
    if (rounds) *rounds = r;
    return QB_OK;
}

// Workspace of one energy evaluation of K states: the states as given (K x n bytes), the
// same packed into 16-bit group masks in position order, the per-block partial sums, and
// the data, prior and total energies (K doubles each). Each part starts on a 256-byte
// boundary.
struct EnergyWorkspace {
    size_t X = 0, x = 0, part = 0, data = 0, prior = 0, total = 0, bytes = 0;
    int n_groups = 0, n_part = 0;
};

static EnergyWorkspace energy_workspace(const qb_problem* P, int K) {
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    EnergyWorkspace L;
    const size_t n = size_t(std::max(P->n, 1));
    L.n_groups = (K + kGroup - 1) / kGroup;
    L.n_part = ray_energy_blocks(P->dev, kGroup);
    L.X = 0;
    L.x = L.X + up(size_t(K) * n);
    L.part = L.x + up(sizeof(uint16_t) * size_t(L.n_groups) * n);
    L.data = L.part + up(sizeof(double) * size_t(L.n_part) * size_t(L.n_groups) * kGroup);
    L.prior = L.data + up(sizeof(double) * size_t(K));
    L.total = L.prior + up(sizeof(double) * size_t(K));
    L.bytes = L.total + up(sizeof(double) * size_t(K));
    return L;
}

// binary_energy of K states into the caller's outputs, all device work inside `ws`
// (laid out by energy_workspace). Allocates nothing.
static int energy_into(qb_problem* P, const uint8_t* X, int K, unsigned char* ws, const EnergyWorkspace& L,
                       double* total, double* data, double* prior) {
    const int n = P->n;
    hipStream_t st = P->stream;
    auto at = [ws](size_t off) { return ws + off; };
    if (n > 0) {
        QB_HIP(hipMemcpyAsync(at(L.X), X, size_t(K) * size_t(n), hipMemcpyHostToDevice, st));
        QB_HIP(launch_pack_states(at(L.X), K, n, P->dev.pos_bit, reinterpret_cast<uint16_t*>(at(L.x)), st));
    }
    QB_HIP(launch_ray_energy(P->dev, reinterpret_cast<const uint16_t*>(at(L.x)), L.n_groups, kGroup, K, nullptr, 0,
                             reinterpret_cast<double*>(at(L.part)), L.n_part, reinterpret_cast<double*>(at(L.data)),
                             reinterpret_cast<double*>(at(L.prior)), nullptr, st,
                             reinterpret_cast<double*>(at(L.total))));
    const size_t kb = sizeof(double) * size_t(K);
    if (data) QB_HIP(hipMemcpyAsync(data, at(L.data), kb, hipMemcpyDeviceToHost, st));
    if (prior) QB_HIP(hipMemcpyAsync(prior, at(L.prior), kb, hipMemcpyDeviceToHost, st));
    if (total) QB_HIP(hipMemcpyAsync(total, at(L.total), kb, hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    return QB_OK;
}

extern "C" int qb_energy(qb_problem* P, const uint8_t* X, int K, double* total, double* data, double* prior) {
    if (!P || (!X && K > 0)) return fail(QB_E_ARG, "qb_energy: null argument");
    if (K <= 0) return fail(QB_E_ARG, "qb_energy: K must be positive");
    const EnergyWorkspace L = energy_workspace(P, K);
    if (int rc = P->w_energy.reserve(L.bytes)) return rc;
    return energy_into(P, X, K, P->w_energy.as<unsigned char>(), L, total, data, prior);
}

// ---- the allocation-free energy entry point ---------------------------------------------
// `scratch` is a file-scope device array: every device the library runs on holds its own
// copy until the process ends. In host code its name is only a host-side stand-in, so no
// kernel argument, memset or copy ever names it: its device address is looked up with
// hipGetSymbolAddress once per device, with that device current, and that address is used.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_addr[kMaxDevices] = {};

static int scratch_on(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(QB_E_LIMIT, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        int cur = 0;
        QB_HIP(hipGetDevice(&cur));
        if (cur != device) QB_HIP(hipSetDevice(device));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) QB_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !ptr)
            return fail(QB_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
        g_scratch_addr[device] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_addr[device];
    return QB_OK;
}

extern "C" int qb_energy_noalloc(qb_problem* P, const uint8_t* X, int K, double* total, double* data,
                                 double* prior) {
    if (!P || (!X && K > 0)) return fail(QB_E_ARG, "qb_energy_noalloc: null argument");
    if (K <= 0) return fail(QB_E_ARG, "qb_energy_noalloc: K must be positive");
    const EnergyWorkspace L = energy_workspace(P, K);
    if (L.bytes > kNoAllocScratchBytes)
        return fail(QB_E_LIMIT,
                    "qb_energy_noalloc: %d states of %d bits need %zu bytes of scratch; the static scratch holds %zu "
                    "(call qb_energy, or pass fewer states)",
                    K, P->n, L.bytes, kNoAllocScratchBytes);
    unsigned char* ws = nullptr;
    if (int rc = scratch_on(P->device, &ws)) return rc;
    return energy_into(P, X, K, ws, L, total, data, prior);
}

// ==== integer path =====================================================================

struct qb_graph {
    hipStream_t stream = nullptr;
    DevGraph dev{};
    qb_plan plan{};
    bool lds_state = false;
    std::vector<int32_t> pos_bit, cls_ptr;
    DevBuf b_cls, b_pos, b_nbp, b_nb, b_nw, b_f0, b_eu, b_ev, b_ew, b_bitpos;
    DevBuf w_s, w_f, w_e, w_sout, w_beta, w_X, w_out, w_rng;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~qb_graph() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// make_plan (src/ising_gpu.cuh:74-101): bounds and statistics of the weights.
static void make_plan(int n, int64_t m, const int32_t* ew, const std::vector<int64_t>& row, const std::vector<int32_t>& w,
                      qb_plan& p) {
    std::memset(&p, 0, sizeof p);
    p.n = n;
    p.m = m;
    p.n_pad = (n + 15) / 16 * 16;
    double sum_w2 = 0.0;
    for (int64_t e = 0; e < m; ++e) {
        const int64_t a = ew[e] < 0 ? -int64_t(ew[e]) : int64_t(ew[e]);
        sum_w2 += double(a) * double(a);
        if (a > p.max_w) p.max_w = a;
        if (a != 0 && (p.w_min == 0 || a < p.w_min)) p.w_min = a;
    }
    for (int i = 0; i < n; ++i) {
        int64_t acc = 0;
        for (int64_t k = row[size_t(i)]; k < row[size_t(i) + 1]; ++k) acc += w[size_t(k)] < 0 ? -int64_t(w[size_t(k)]) : int64_t(w[size_t(k)]);
        p.B = std::max(p.B, acc);
    }
    if (p.B > 2147483647LL) std::snprintf(p.refused, sizeof p.refused, "field bound exceeds int32");
    if (p.w_min == 0) std::snprintf(p.refused, sizeof p.refused, "all weights are zero");
    p.field_bytes = p.B <= 127 ? 1 : (p.B <= 32767 ? 2 : 4);
    p.weight_bytes = p.max_w <= 127 ? 1 : 4;
    p.dense = n <= 16384 && double(2 * m) / double(n) >= double(n) / 4.0;
    p.sigma = std::sqrt(2.0 * sum_w2 / double(n));
}

// Symmetric CSR in parse_rudy's fill order (src/ising_weighted.hpp:134-147).
static void graph_csr(const qb_graph_desc* d, std::vector<int64_t>& row, std::vector<int32_t>& col,
                      std::vector<int32_t>& w) {
    const int n = d->n;
    const int64_t m = d->m;
    row.assign(size_t(n) + 1, 0);
    for (int64_t e = 0; e < m; ++e) { row[size_t(d->eu[e]) + 1]++; row[size_t(d->ev[e]) + 1]++; }
    for (int i = 0; i < n; ++i) row[size_t(i) + 1] += row[size_t(i)];
    col.assign(size_t(2 * m), 0);
    w.assign(size_t(2 * m), 0);
    std::vector<int64_t> fill(row.begin(), row.end() - 1);
    for (int64_t e = 0; e < m; ++e) {
        const int32_t a = d->eu[e], b = d->ev[e], wt = d->ew[e];
        col[size_t(fill[size_t(a)])] = b; w[size_t(fill[size_t(a)]++)] = wt;
        col[size_t(fill[size_t(b)])] = a; w[size_t(fill[size_t(b)]++)] = wt;
    }
}

// Distance-1 colouring, vertices by decreasing degree (ties by index).
static int colour_graph(int n, const std::vector<int64_t>& row, const std::vector<int32_t>& col,
                        std::vector<int32_t>& pos_bit, std::vector<int32_t>& cls_ptr) {
    std::vector<int> visit(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) visit[size_t(i)] = i;
    std::stable_sort(visit.begin(), visit.end(), [&](int a, int b) {
        return row[size_t(a) + 1] - row[size_t(a)] > row[size_t(b) + 1] - row[size_t(b)];
    });
    std::vector<int> colour;
    const int K = greedy_colour(n, visit, [&](int i, auto mark) {
        for (int64_t k = row[size_t(i)]; k < row[size_t(i) + 1]; ++k) mark(col[size_t(k)]);
    }, colour);
    order_by_colour(colour, K, pos_bit, cls_ptr);
    return K;
}

static int validate_graph(const qb_graph_desc* d) {
    if (!d) return fail(QB_E_ARG, "null graph description");
    if (d->n <= 0 || d->m < 0) return fail(QB_E_ARG, "bad graph sizes (n %d, m %lld)", d->n, (long long)d->m);
    if (d->m > 0 && (!d->eu || !d->ev || !d->ew)) return fail(QB_E_ARG, "edge arrays are required");
    if (d->m > (int64_t(1) << 30)) return fail(QB_E_LIMIT, "more than 2^30 edges");
    for (int64_t e = 0; e < d->m; ++e) {
        if (d->eu[e] < 0 || d->eu[e] >= d->n || d->ev[e] < 0 || d->ev[e] >= d->n)
            return fail(QB_E_ARG, "edge %lld names a vertex out of range", (long long)e);
        if (d->eu[e] == d->ev[e]) return fail(QB_E_ARG, "edge %lld is a self-loop", (long long)e);
    }
    // A repeated edge (either orientation) would list a neighbour twice in one adjacency row,
    // and the kernels update each neighbour of a flipped vertex once per listing, from
    // different lanes. Bucket the edges by their lower endpoint, then stamp each bucket's
    // upper endpoints.
    std::vector<int64_t> start(size_t(d->n) + 1, 0);
    for (int64_t e = 0; e < d->m; ++e) ++start[size_t(std::min(d->eu[e], d->ev[e])) + 1];
    for (int i = 0; i < d->n; ++i) start[size_t(i) + 1] += start[size_t(i)];
    std::vector<int32_t> upper(size_t(d->m)), edge_of(size_t(d->m));
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t e = 0; e < d->m; ++e) {
        const size_t at = size_t(fill[size_t(std::min(d->eu[e], d->ev[e]))]++);
        upper[at] = std::max(d->eu[e], d->ev[e]);
        edge_of[at] = int32_t(e);
    }
    std::vector<int32_t> stamp(size_t(d->n), -1);
    for (int i = 0; i < d->n; ++i)
        for (int64_t k = start[size_t(i)]; k < start[size_t(i) + 1]; ++k) {
            if (stamp[size_t(upper[size_t(k)])] == i)
                return fail(QB_E_ARG, "edge %d repeats an earlier edge between vertices %d and %d", edge_of[size_t(k)],
                            i, upper[size_t(k)]);
            stamp[size_t(upper[size_t(k)])] = i;
        }
    return QB_OK;
}

extern "C" int qb_colour_graph(const qb_graph_desc* d, int32_t* order, int32_t* class_ptr, int* n_classes) {
    if (!order || !class_ptr || !n_classes) return fail(QB_E_ARG, "qb_colour_graph: null output");
    int rc = validate_graph(d);
    if (rc) return rc;
    std::vector<int64_t> row;
    std::vector<int32_t> col, w, pos_bit, cls_ptr;
    graph_csr(d, row, col, w);
    const int K = colour_graph(d->n, row, col, pos_bit, cls_ptr);
    std::memcpy(order, pos_bit.data(), sizeof(int32_t) * pos_bit.size());
    std::memcpy(class_ptr, cls_ptr.data(), sizeof(int32_t) * cls_ptr.size());
    *n_classes = K;
    return QB_OK;
}

extern "C" int qb_graph_create(const qb_graph_desc* d, int force_global, qb_graph** out) {
    if (!out) return fail(QB_E_ARG, "qb_graph_create: null output");
    *out = nullptr;
    int rc0 = validate_graph(d);
    if (rc0) return rc0;
    const int n = d->n;
    const int64_t m = d->m;
    std::vector<int64_t> row;
    std::vector<int32_t> col, w;
    graph_csr(d, row, col, w);
    auto G = new (std::nothrow) qb_graph();
    if (!G) return fail(QB_E_NOMEM, "out of host memory");
    std::unique_ptr<qb_graph> guard(G);
    make_plan(n, m, d->ew, row, w, G->plan);
    const int K = colour_graph(n, row, col, G->pos_bit, G->cls_ptr);
    G->plan.n_classes = K;
    std::vector<int32_t> pos_of(static_cast<size_t>(n));
    for (int p = 0; p < n; ++p) pos_of[size_t(G->pos_bit[size_t(p)])] = p;
    std::vector<int32_t> nb_ptr(size_t(n) + 1, 0), nb(size_t(2 * m)), nw(size_t(2 * m)), f0(static_cast<size_t>(n));
    for (int p = 0; p < n; ++p) {
        const int i = G->pos_bit[size_t(p)];
        int64_t q = nb_ptr[size_t(p)];
        int64_t acc = 0;
        for (int64_t k = row[size_t(i)]; k < row[size_t(i) + 1]; ++k, ++q) {
            nb[size_t(q)] = pos_of[size_t(col[size_t(k)])];
            nw[size_t(q)] = w[size_t(k)];
            acc -= w[size_t(k)];
        }
        nb_ptr[size_t(p) + 1] = int32_t(q);
        f0[size_t(p)] = int32_t(acc);
    }
    const int64_t lds_need = int64_t(G->plan.n_pad) * 5;
    G->lds_state = !force_global && G->plan.n_pad <= int_lds_capacity_spins() && lds_need <= 64 * 1024;
    G->plan.shared_state = G->lds_state ? 1 : 0;
    QB_HIP(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
    QB_HIP(hipEventCreate(&G->ev0));
    QB_HIP(hipEventCreate(&G->ev1));
    hipStream_t st = G->stream;
    std::vector<int32_t> eu(d->eu, d->eu + m), ev(d->ev, d->ev + m), ewv(d->ew, d->ew + m);
    int rc;
    if ((rc = G->b_cls.upload(G->cls_ptr, st)) || (rc = G->b_pos.upload(G->pos_bit, st)) ||
        (rc = G->b_nbp.upload(nb_ptr, st)) || (rc = G->b_nb.upload(nb, st)) || (rc = G->b_nw.upload(nw, st)) ||
        (rc = G->b_f0.upload(f0, st)) || (rc = G->b_eu.upload(eu, st)) || (rc = G->b_ev.upload(ev, st)) ||
        (rc = G->b_ew.upload(ewv, st)) || (rc = G->b_bitpos.upload(pos_of, st)))
        return rc;
    QB_HIP(hipStreamSynchronize(st));
    DevGraph& D = G->dev;
    D.n = n;
    D.n_pad = G->plan.n_pad;
    D.n_classes = K;
    D.m = m;
    D.cls_ptr = G->b_cls.as<int32_t>();
    D.pos_bit = G->b_pos.as<int32_t>();
    D.nb_ptr = G->b_nbp.as<int32_t>();
    D.nb = G->b_nb.as<int32_t>();
    D.nw = G->b_nw.as<int32_t>();
    D.f0 = G->b_f0.as<int32_t>();
    D.bit_pos = G->b_bitpos.as<int32_t>();
    D.eu = G->b_eu.as<int32_t>();
    D.ev = G->b_ev.as<int32_t>();
    D.ew = G->b_ew.as<int32_t>();
    *out = guard.release();
    return QB_OK;
}

extern "C" int qb_graph_destroy(qb_graph* g) {
    delete g;
    return QB_OK;
}

extern "C" int qb_graph_get_plan(const qb_graph* g, qb_plan* out) {
    if (!g || !out) return fail(QB_E_ARG, "qb_graph_get_plan: null argument");
    *out = g->plan;
    return QB_OK;
}

extern "C" int qb_graph_colouring(const qb_graph* g, int32_t* order, int32_t* class_ptr) {
    if (!g || !order || !class_ptr) return fail(QB_E_ARG, "qb_graph_colouring: null argument");
    std::memcpy(order, g->pos_bit.data(), sizeof(int32_t) * g->pos_bit.size());
    std::memcpy(class_ptr, g->cls_ptr.data(), sizeof(int32_t) * g->cls_ptr.size());
    return QB_OK;
}

extern "C" int qb_run_batch(qb_graph* G, const qb_schedule* b, uint64_t seed_base, int R, double updates_per_launch,
                            int8_t* s_out, int64_t* energy_out, qb_batch_report* rep) {
    if (!G) return fail(QB_E_ARG, "qb_run_batch: null graph");
    int rc = check_schedule(b);
    if (rc) return rc;
    if (R <= 0) return fail(QB_E_ARG, "R must be positive (got %d)", R);
    if (G->plan.refused[0]) return fail(QB_E_REFUSED, "plan refused: %s", G->plan.refused);
    const int total = b->anneal_sweeps + b->hold_sweeps;
    const int np = G->plan.n_pad, n = G->plan.n;
    const size_t cells = size_t(R) * size_t(np);
    if ((rc = G->w_s.reserve(cells)) || (rc = G->w_f.reserve(cells * 4)) || (rc = G->w_e.reserve(sizeof(long long) * size_t(R))) ||
        (rc = G->w_sout.reserve(cells)) || (rc = G->w_beta.reserve(sizeof(double) * size_t(total))))
        return rc;
    const std::vector<double> beta = betas_of(*b);
    hipStream_t st = G->stream;
    int chunk = int(std::floor(updates_per_launch / (double(R) * double(n))));
    chunk = std::max(1, std::min(total, chunk));
    QB_HIP(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    QB_HIP(hipMemcpyAsync(G->w_beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice, st));
    QB_HIP(hipEventRecord(G->ev0, st));
    int launches = 0;
    for (int sb = 0; sb < total; sb += chunk) {
        QB_HIP(launch_int_sweeps(G->dev, R, seed_base, G->w_beta.as<double>(), sb, std::min(total, sb + chunk),
                                 sb == 0 ? 1 : 0, G->w_s.as<int8_t>(), G->w_f.as<int32_t>(), G->w_e.as<long long>(),
                                 G->lds_state, st));
        ++launches;
    }
    QB_HIP(hipEventRecord(G->ev1, st));
    QB_HIP(launch_int_unpermute(G->dev, R, G->w_s.as<int8_t>(), G->w_sout.as<int8_t>(), st));
    if (s_out) QB_HIP(hipMemcpyAsync(s_out, G->w_sout.p, cells, hipMemcpyDeviceToHost, st));
    if (energy_out) QB_HIP(hipMemcpyAsync(energy_out, G->w_e.p, sizeof(int64_t) * size_t(R), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    if (rep) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, G->ev0, G->ev1));
        rep->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->kernel_ms = ms;
        rep->launches = launches;
        rep->sweeps_per_launch = chunk;
    }
    return QB_OK;
}

extern "C" int qb_run_batch_reference(qb_graph* G, const qb_schedule* b, uint64_t seed_base, int R,
                                      double updates_per_launch, int8_t* s_out, int64_t* energy_out,
                                      qb_batch_report* rep) {
    if (!G) return fail(QB_E_ARG, "qb_run_batch_reference: null graph");
    int rc = check_schedule(b);
    if (rc) return rc;
    if (R <= 0) return fail(QB_E_ARG, "R must be positive (got %d)", R);
    if (G->plan.refused[0]) return fail(QB_E_REFUSED, "plan refused: %s", G->plan.refused);
    const int total = b->anneal_sweeps + b->hold_sweeps;
    const int np = G->plan.n_pad, n = G->plan.n;
    const size_t cells = size_t(R) * size_t(np);
    if ((rc = G->w_s.reserve(cells)) || (rc = G->w_f.reserve(cells * 4)) ||
        (rc = G->w_e.reserve(sizeof(long long) * size_t(R))) || (rc = G->w_rng.reserve(sizeof(uint64_t) * size_t(R))) ||
        (rc = G->w_sout.reserve(cells)) || (rc = G->w_beta.reserve(sizeof(double) * size_t(total))))
        return rc;
    // the temperature of every sweep by the reference's expression (src/ising_weighted.hpp:255-258)
    std::vector<double> temp(static_cast<size_t>(total));
    for (int sweep = 0; sweep < total; ++sweep) {
        const double frac = std::min(1.0, double(sweep) / double(std::max(1, b->anneal_sweeps - 1)));
        temp[size_t(sweep)] = sweep < b->anneal_sweeps ? b->t_hot * std::pow(b->t_cold / b->t_hot, frac) : b->t_cold;
    }
    hipStream_t st = G->stream;
    int chunk = int(std::floor(updates_per_launch / (double(R) * double(n))));
    chunk = std::max(1, std::min(total, chunk));
    QB_HIP(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    QB_HIP(hipMemcpyAsync(G->w_beta.p, temp.data(), sizeof(double) * temp.size(), hipMemcpyHostToDevice, st));
    QB_HIP(hipEventRecord(G->ev0, st));
    int launches = 0;
    for (int sb = 0; sb < total; sb += chunk) {
        QB_HIP(launch_int_ref(G->dev, R, seed_base, G->w_beta.as<double>(), sb, std::min(total, sb + chunk),
                              sb == 0 ? 1 : 0, G->w_s.as<int8_t>(), G->w_f.as<int32_t>(), G->w_rng.as<uint64_t>(),
                              G->w_e.as<long long>(), st));
        ++launches;
    }
    QB_HIP(hipEventRecord(G->ev1, st));
    QB_HIP(launch_int_unpermute(G->dev, R, G->w_s.as<int8_t>(), G->w_sout.as<int8_t>(), st));
    if (s_out) QB_HIP(hipMemcpyAsync(s_out, G->w_sout.p, cells, hipMemcpyDeviceToHost, st));
    if (energy_out) QB_HIP(hipMemcpyAsync(energy_out, G->w_e.p, sizeof(int64_t) * size_t(R), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    if (rep) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, G->ev0, G->ev1));
        rep->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->kernel_ms = ms;
        rep->launches = launches;
        rep->sweeps_per_launch = chunk;
    }
    return QB_OK;
}

extern "C" int qb_device_energies(qb_graph* G, const uint8_t* X, int K, int64_t* energy_out) {
    if (!G || !X || !energy_out) return fail(QB_E_ARG, "qb_device_energies: null argument");
    if (K <= 0) return fail(QB_E_ARG, "K must be positive");
    int rc;
    if ((rc = G->w_X.reserve(size_t(K) * size_t(G->plan.n))) || (rc = G->w_out.reserve(sizeof(long long) * size_t(K))))
        return rc;
    hipStream_t st = G->stream;
    QB_HIP(hipMemcpyAsync(G->w_X.p, X, size_t(K) * size_t(G->plan.n), hipMemcpyHostToDevice, st));
    QB_HIP(launch_int_energies(G->dev, G->w_X.as<uint8_t>(), K, G->w_out.as<long long>(), st));
    QB_HIP(hipMemcpyAsync(energy_out, G->w_out.p, sizeof(int64_t) * size_t(K), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    return QB_OK;
}

// ==== batched sparse QUBOs ==============================================================

struct qb_qubo_batch {
    hipStream_t stream = nullptr;
    int n_inst = 0;
    int64_t n_total = 0;
    std::vector<int32_t> n_bits, replicas, inst_pos, inst_cls, cls_ptr, pos_bit;
    std::vector<int64_t> entries;     // per instance
    std::vector<int64_t> x_off;       // per instance: offset of its replicas' bits in the x output
    std::vector<int32_t> e_off;       // per instance: offset of its replicas in the energy output
    std::vector<uint64_t> seed;
    int total_replicas = 0;
    DevBuf b_inst_pos, b_inst_cls, b_cls, b_row, b_col, b_J, b_h, b_bit, b_seed, b_ctile, b_tptr, b_tcol, b_tJ;
    struct Jobs {
        DevBuf by_xcd, by_inst;       // the same jobs, grouped by XCD queue and in instance order
        QuboQueues segs{};            // the XCD queues' ranges in by_xcd
        std::vector<QuboJob> host;    // instance order
        std::vector<int32_t> first;   // per instance: its first job in instance order
        int64_t mask_words = 0;
    };
    std::map<int, std::unique_ptr<Jobs>> jobs;   // by replica block
    DevBuf w_mask, w_beta, w_queue, w_energy;
    std::vector<uint32_t> h_mask;
    // Recorded launch schedules: one per (shape, mode, sweeps, workgroups) and the buffers
    // they were recorded on (a grown buffer makes a new recording).
    struct Recorded {
        int rb, block, n_sweeps, mode, workgroups;
        const void *mask, *queue, *beta;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Recorded> recorded;
    // side streams and events of a recording's parallel branches (one launch per instance)
    static constexpr int kBranches = 4;
    hipStream_t side[kBranches] = {};
    hipEvent_t fork = nullptr, join[kBranches] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~qb_qubo_batch() {
        for (int i = 0; i < kBranches; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        for (Recorded& r : recorded) {
            if (r.exec) (void)hipGraphExecDestroy(r.exec);
            if (r.graph) (void)hipGraphDestroy(r.graph);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int check_qubo_rows(int k, int n, const int32_t* row_ptr, const int32_t* col, const float* J);

// A QUBO instance's colouring: saturation order over its coupling graph (ties by
// degree, then index); pos_bit lists the bits class by class, cp the class bounds.
static int colour_qubo(int n, const int32_t* rp, const int32_t* cl, std::vector<int32_t>& pos_bit,
                       std::vector<int32_t>& cp) {
    std::vector<int> by_weight(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) by_weight[size_t(i)] = i;
    std::stable_sort(by_weight.begin(), by_weight.end(),
                     [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
    std::vector<int> colour;
    const int nc = saturation_colour(n, by_weight, {}, [&](int i, auto mark) {
        for (int e = rp[i]; e < rp[i + 1]; ++e) mark(cl[e]);
    }, colour);
    order_by_colour(colour, nc, pos_bit, cp);
    return nc;
}

extern "C" int qb_colour_qubo(int n, const int32_t* row_ptr, const int32_t* col, const float* J, int32_t* order,
                              int32_t* class_ptr, int* n_classes) {
    if (n <= 0 || !row_ptr || !order || !class_ptr || !n_classes) return fail(QB_E_ARG, "qb_colour_qubo: bad argument");
    int rc = check_qubo_rows(0, n, row_ptr, col, J);
    if (rc) return rc;
    std::vector<int32_t> pos_bit, cp;
    const int nc = colour_qubo(n, row_ptr, col, pos_bit, cp);
    std::memcpy(order, pos_bit.data(), sizeof(int32_t) * size_t(n));
    std::memcpy(class_ptr, cp.data(), sizeof(int32_t) * cp.size());
    *n_classes = nc;
    return QB_OK;
}

// Instance k's rows: symmetric (every J_ij also listed as J_ji, equal), no self
// coupling, columns in range. Returns QB_OK or fails naming the first defect.
static int check_qubo_rows(int k, int n, const int32_t* row_ptr, const int32_t* col, const float* J) {
    if (row_ptr[0] != 0) return fail(QB_E_ARG, "instance %d: row_ptr must start at 0", k);
    struct Pair {
        uint64_t key;
        float w;
    };
    std::vector<Pair> up, down;
    for (int i = 0; i < n; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) return fail(QB_E_ARG, "instance %d: row_ptr decreases at bit %d", k, i);
        for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
            const int j = col[e];
            if (j < 0 || j >= n) return fail(QB_E_ARG, "instance %d: bit %d names column %d of %d", k, i, j, n);
            if (j == i) return fail(QB_E_ARG, "instance %d: bit %d couples to itself", k, i);
            if (!std::isfinite(J[e])) return fail(QB_E_ARG, "instance %d: coupling of bits %d, %d is not finite", k, i, j);
            const uint64_t key = (uint64_t(uint32_t(std::min(i, j))) << 32) | uint32_t(std::max(i, j));
            (i < j ? up : down).push_back(Pair{key, J[e]});
        }
    }
    auto by_key = [](const Pair& a, const Pair& b) { return a.key < b.key; };
    std::sort(up.begin(), up.end(), by_key);
    std::sort(down.begin(), down.end(), by_key);
    if (up.size() != down.size()) return fail(QB_E_ARG, "instance %d: couplings are not symmetric", k);
    for (size_t t = 0; t < up.size(); ++t) {
        if (t > 0 && up[t].key == up[t - 1].key)
            return fail(QB_E_ARG, "instance %d: bits %u and %u are coupled twice", k, unsigned(up[t].key >> 32),
                        unsigned(up[t].key & 0xffffffffu));
        if (up[t].key != down[t].key || up[t].w != down[t].w)
            return fail(QB_E_ARG, "instance %d: coupling of bits %u and %u is not symmetric", k,
                        unsigned(up[t].key >> 32), unsigned(up[t].key & 0xffffffffu));
    }
    return QB_OK;
}

extern "C" int qb_qubo_batch_create(const qb_qubo_batch_desc* d, qb_qubo_batch** out) {
    if (!out) return fail(QB_E_ARG, "qb_qubo_batch_create: null output");
    *out = nullptr;
    if (!d || d->n_instances <= 0 || !d->n_bits || !d->row_ptr || !d->h || !d->replicas || !d->seed)
        return fail(QB_E_ARG, "qb_qubo_batch_create: missing arrays or no instances");
    auto B = std::make_unique<qb_qubo_batch>();
    const int K = d->n_instances;
    B->n_inst = K;
    int64_t rows = 0, ents = 0, xs = 0;
    for (int k = 0; k < K; ++k) {
        const int n = d->n_bits[k];
        if (n <= 0) return fail(QB_E_ARG, "instance %d has %d bits", k, n);
        if (d->replicas[k] < 1 || d->replicas[k] > 32)
            return fail(QB_E_ARG, "instance %d: replicas must be 1..32 (got %d)", k, d->replicas[k]);
        const int32_t* rp = d->row_ptr + rows;
        int rc = check_qubo_rows(k, n, rp, d->col + ents, d->J + ents);
        if (rc) return rc;
        B->n_bits.push_back(n);
        B->replicas.push_back(d->replicas[k]);
        B->entries.push_back(rp[n]);
        B->x_off.push_back(xs);
        B->e_off.push_back(B->total_replicas);
        xs += int64_t(n) * d->replicas[k];
        B->total_replicas += d->replicas[k];
        rows += n + 1;
        ents += rp[n];
    }
    if (B->n_total + rows > INT32_MAX || ents > INT32_MAX) return fail(QB_E_LIMIT, "the batch exceeds int32 indexing");
    // colour every instance (saturation order over its coupling graph) and lay it out
    std::vector<int32_t> row(1, 0), colp;
    std::vector<float> Jp, hp;
    std::vector<int32_t> bit;
    B->inst_pos.push_back(0);
    B->inst_cls.push_back(0);
    rows = 0;
    ents = 0;
    int64_t hoff = 0;
    for (int k = 0; k < K; ++k) {
        const int n = B->n_bits[k];
        const int32_t* rp = d->row_ptr + rows;
        const int32_t* cl = d->col + ents;
        const float* Jk = d->J + ents;
        std::vector<int32_t> pos_bit, cp;
        const int nc = colour_qubo(n, rp, cl, pos_bit, cp);
        std::vector<int32_t> pos_of(static_cast<size_t>(n));
        for (int p = 0; p < n; ++p) pos_of[size_t(pos_bit[size_t(p)])] = p;
        const int32_t base = B->inst_pos.back();
        for (int c = 0; c <= nc; ++c) B->cls_ptr.push_back(base + cp[size_t(c)]);
        B->inst_cls.push_back(int32_t(B->cls_ptr.size()));
        for (int p = 0; p < n; ++p) {
            const int i = pos_bit[size_t(p)];
            for (int e = rp[i]; e < rp[i + 1]; ++e) {
                colp.push_back(base + pos_of[size_t(cl[e])]);
                Jp.push_back(Jk[e]);
            }
            row.push_back(int32_t(colp.size()));
            hp.push_back(d->h[hoff + i]);
            bit.push_back(i);
            B->pos_bit.push_back(i);
        }
        B->inst_pos.push_back(base + n);
        B->seed.push_back(d->seed[k]);
        rows += n + 1;
        ents += rp[n];
        hoff += n;
    }
    B->n_total = B->inst_pos.back();
    // the rows sliced by 64 positions of each class, entry-major (padding repeats a valid entry)
    std::vector<int32_t> ctile, tptr, tcol;
    std::vector<float> tJ;
    for (size_t c = 0; c + 1 < B->cls_ptr.size(); ++c) {
        ctile.push_back(int32_t(tptr.size()));
        const int lo = B->cls_ptr[c], hi = B->cls_ptr[c + 1];
        if (hi < lo) continue;   // the boundary between two instances' class lists
        for (int t0 = lo; t0 < hi; t0 += 64) {
            const int t1 = std::min(hi, t0 + 64);
            int width = 0;
            for (int p = t0; p < t1; ++p) width = std::max(width, row[size_t(p) + 1] - row[size_t(p)]);
            if (int64_t(tcol.size()) + int64_t(width) * 64 > INT32_MAX)
                return fail(QB_E_LIMIT, "the batch's sliced rows exceed int32 indexing");
            tptr.push_back(int32_t(tcol.size()));
            const size_t base = tcol.size();
            tcol.resize(base + size_t(width) * 64, 0);
            tJ.resize(base + size_t(width) * 64, 0.0f);
            for (int p = t0; p < t1; ++p) {
                const int r0 = row[size_t(p)], deg = row[size_t(p) + 1] - r0;
                for (int u = 0; u < width; ++u) {
                    const int e = r0 + std::min(u, std::max(deg - 1, 0));
                    const size_t at = base + size_t(u) * 64 + size_t(p - t0);
                    tcol[at] = deg > 0 ? colp[size_t(e)] : p;
                    tJ[at] = deg > 0 ? Jp[size_t(e)] : 0.0f;
                }
            }
        }
    }
    ctile.push_back(int32_t(tptr.size()));
    QB_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    QB_HIP(hipEventCreate(&B->ev0));
    QB_HIP(hipEventCreate(&B->ev1));
    QB_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
    for (int i = 0; i < qb_qubo_batch::kBranches; ++i) {
        QB_HIP(hipStreamCreateWithFlags(&B->side[i], hipStreamNonBlocking));
        QB_HIP(hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming));
    }
    hipStream_t st = B->stream;
    int rc;
    if ((rc = B->b_inst_pos.upload(B->inst_pos, st)) || (rc = B->b_inst_cls.upload(B->inst_cls, st)) ||
        (rc = B->b_cls.upload(B->cls_ptr, st)) || (rc = B->b_row.upload(row, st)) || (rc = B->b_col.upload(colp, st)) ||
        (rc = B->b_J.upload(Jp, st)) || (rc = B->b_h.upload(hp, st)) || (rc = B->b_bit.upload(bit, st)) ||
        (rc = B->b_seed.upload(B->seed, st)) || (rc = B->b_ctile.upload(ctile, st)) ||
        (rc = B->b_tptr.upload(tptr, st)) || (rc = B->b_tcol.upload(tcol, st)) || (rc = B->b_tJ.upload(tJ, st)))
        return rc;
    QB_HIP(hipStreamSynchronize(st));
    *out = B.release();
    return QB_OK;
}

extern "C" int qb_qubo_batch_destroy(qb_qubo_batch* b) {
    delete b;
    return QB_OK;
}

// Jobs at replica block rb: instance k's replicas in blocks of rb, each block a job with
// its own spin words. Uploaded in instance order, and grouped by XCD queue: instances in
// decreasing cost (rows + entries, times replicas) each go to the least loaded of the
// eight queues, and a queue lists its instances in that order with each instance's jobs
// side by side.
static int qubo_jobs_for(qb_qubo_batch* B, int rb, const qb_qubo_batch::Jobs** out) {
    auto it = B->jobs.find(rb);
    if (it != B->jobs.end()) {
        *out = it->second.get();
        return QB_OK;
    }
    auto Jb = std::make_unique<qb_qubo_batch::Jobs>();
    int64_t words = 0;
    for (int k = 0; k < B->n_inst; ++k) {
        Jb->first.push_back(int32_t(Jb->host.size()));
        for (int r0 = 0; r0 < B->replicas[size_t(k)]; r0 += rb) {
            QuboJob j{};
            j.instance = k;
            j.replica0 = r0;
            j.replicas = std::min(rb, B->replicas[size_t(k)] - r0);
            j.mask_offset = int32_t(words);
            j.energy_offset = B->e_off[size_t(k)] + r0;
            words += B->n_bits[size_t(k)];
            Jb->host.push_back(j);
        }
    }
    Jb->first.push_back(int32_t(Jb->host.size()));
    if (words > INT32_MAX) return fail(QB_E_LIMIT, "the batch's spin words exceed int32 indexing");
    Jb->mask_words = words;
    auto cost = [&](int k) {
        return (int64_t(B->n_bits[size_t(k)]) + B->entries[size_t(k)]) * int64_t(B->replicas[size_t(k)]);
    };
    std::vector<int> order(static_cast<size_t>(B->n_inst));
    for (int k = 0; k < B->n_inst; ++k) order[size_t(k)] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    std::vector<std::vector<int>> queue_inst(kXcds);
    std::vector<int64_t> load(kXcds, 0);
    for (int k : order) {
        const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
        queue_inst[size_t(x)].push_back(k);
        load[size_t(x)] += cost(k);
    }
    std::vector<QuboJob> by_xcd;
    by_xcd.reserve(Jb->host.size());
    for (int x = 0; x < kXcds; ++x) {
        Jb->segs.begin[x] = int32_t(by_xcd.size());
        for (int k : queue_inst[size_t(x)])
            for (int j = Jb->first[size_t(k)]; j < Jb->first[size_t(k) + 1]; ++j) by_xcd.push_back(Jb->host[size_t(j)]);
    }
    Jb->segs.begin[kXcds] = int32_t(by_xcd.size());
    int rc;
    if ((rc = Jb->by_xcd.upload(by_xcd, B->stream)) || (rc = Jb->by_inst.upload(Jb->host, B->stream))) return rc;
    QB_HIP(hipStreamSynchronize(B->stream));
    *out = Jb.get();
    B->jobs[rb] = std::move(Jb);
    return QB_OK;
}

static DevQuboBatch qubo_dev(const qb_qubo_batch* B, const qb_qubo_batch::Jobs& J, bool by_xcd) {
    return DevQuboBatch{int(J.host.size()), by_xcd ? J.by_xcd.as<QuboJob>() : J.by_inst.as<QuboJob>(),
                        B->b_inst_pos.as<int32_t>(), B->b_inst_cls.as<int32_t>(), B->b_cls.as<int32_t>(),
                        B->b_row.as<int32_t>(), B->b_col.as<int32_t>(), B->b_J.as<float>(), B->b_h.as<float>(),
                        B->b_bit.as<int32_t>(), B->b_seed.as<uint64_t>(), B->b_ctile.as<int32_t>(),
                        B->b_tptr.as<int32_t>(), B->b_tcol.as<int32_t>(), B->b_tJ.as<float>()};
}

// One anneal of the whole batch at replica block rb: spin words cleared (all-zero x),
// then the persistent launch or one launch per instance, timed with events.
// The anneal's launches: the persistent launch, or one per instance. With `branches`
// (while a graph is recorded) the per-instance launches go round-robin to side streams
// forked from and joined back into the recorded stream, so the graph holds independent
// branches whose launches may overlap on the device.
static int qubo_launches(qb_qubo_batch* B, const qb_qubo_batch::Jobs& Jb, int rb, int block, int n_sweeps, int mode,
                         int workgroups, int* launches, bool branches = false) {
    hipStream_t st = B->stream;
    const size_t set = size_t(kXcds) * size_t(kQueueStride);
    *launches = 0;
    if (mode == 1 && branches) {
        QB_HIP(hipEventRecord(B->fork, st));
        for (int i = 0; i < qb_qubo_batch::kBranches; ++i) QB_HIP(hipStreamWaitEvent(B->side[i], B->fork, 0));
    }
    if (mode == 0) {
        const DevQuboBatch q = qubo_dev(B, Jb, true);
        QB_HIP(launch_qubo_batch(q, rb, block, B->w_mask.as<uint32_t>(), B->w_beta.as<double>(), n_sweeps,
                                 B->w_queue.as<unsigned>(), Jb.segs, std::min(workgroups, q.n_jobs), st));
        *launches = 1;
    } else {
        const DevQuboBatch q = qubo_dev(B, Jb, false);
        for (int k = 0; k < B->n_inst; ++k) {
            const int f = Jb.first[size_t(k)], cnt = Jb.first[size_t(k) + 1] - f;
            QuboQueues one{};   // the instance's jobs in queue 0, the others empty
            one.begin[0] = f;
            for (int x = 1; x <= kXcds; ++x) one.begin[x] = f + cnt;
            hipStream_t ls = branches ? B->side[k % qb_qubo_batch::kBranches] : st;
            QB_HIP(launch_qubo_batch(q, rb, block, B->w_mask.as<uint32_t>(), B->w_beta.as<double>(), n_sweeps,
                                     B->w_queue.as<unsigned>() + set * size_t(k), one, cnt, ls));
            ++*launches;
        }
        if (branches)
            for (int i = 0; i < qb_qubo_batch::kBranches; ++i) {
                QB_HIP(hipEventRecord(B->join[i], B->side[i]));
                QB_HIP(hipStreamWaitEvent(st, B->join[i], 0));
            }
    }
    return QB_OK;
}

// Resets the state (spin words, queue counters) and runs the launches; with `graph`, the
// reset (by fill kernels: a recorded graph holds no memset node) and the launches are
// recorded into a hipGraph on first use and replayed after.
static int qubo_run(qb_qubo_batch* B, const qb_qubo_batch::Jobs& Jb, int rb, int block, int n_sweeps, int mode,
                    int workgroups, bool graph, int* launches, float* ms, bool* replayed) {
    hipStream_t st = B->stream;
    // one set of XCD queue counters per launch (kXcds lines each)
    const size_t set = size_t(kXcds) * size_t(kQueueStride);
    const size_t queue_bytes = sizeof(unsigned) * set * size_t(mode == 0 ? 1 : B->n_inst);
    int rc;
    if ((rc = B->w_mask.reserve(sizeof(uint32_t) * size_t(Jb.mask_words))) ||
        (rc = B->w_queue.reserve(queue_bytes)))
        return rc;
    if (replayed) *replayed = false;
    if (!graph) {
        QB_HIP(hipMemsetAsync(B->w_mask.p, 0, sizeof(uint32_t) * size_t(Jb.mask_words), st));
        QB_HIP(hipMemsetAsync(B->w_queue.p, 0, queue_bytes, st));
        QB_HIP(hipEventRecord(B->ev0, st));
        if ((rc = qubo_launches(B, Jb, rb, block, n_sweeps, mode, workgroups, launches))) return rc;
    } else {
        qb_qubo_batch::Recorded* hit = nullptr;
        for (auto& r : B->recorded)
            if (r.rb == rb && r.block == block && r.n_sweeps == n_sweeps && r.mode == mode &&
                r.workgroups == workgroups && r.mask == B->w_mask.p && r.queue == B->w_queue.p && r.beta == B->w_beta.p)
                hit = &r;
        if (!hit) {
            qb_qubo_batch::Recorded r{rb, block, n_sweeps, mode, workgroups, B->w_mask.p, B->w_queue.p, B->w_beta.p};
            QB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int lrc = QB_OK;
            hipError_t e = launch_fill_u32(B->w_mask.as<uint32_t>(), size_t(Jb.mask_words), 0u, st);
            if (e == hipSuccess) e = launch_fill_u32(B->w_queue.as<uint32_t>(), queue_bytes / sizeof(uint32_t), 0u, st);
            if (e == hipSuccess) lrc = qubo_launches(B, Jb, rb, block, n_sweeps, mode, workgroups, launches, true);
            const hipError_t ee = hipStreamEndCapture(st, &r.graph);
            if (e != hipSuccess) return fail(QB_E_HIP, "recording the anneal's graph: %s", hipGetErrorString(e));
            if (lrc) {
                if (r.graph) (void)hipGraphDestroy(r.graph);
                return lrc;
            }
            if (ee != hipSuccess) return fail(QB_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ee));
            const hipError_t ie = hipGraphInstantiate(&r.exec, r.graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(r.graph);
                return fail(QB_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
            }
            B->recorded.push_back(r);
            hit = &B->recorded.back();
        } else {
            *launches = mode == 0 ? 1 : B->n_inst;
            if (replayed) *replayed = true;
        }
        QB_HIP(hipEventRecord(B->ev0, st));
        QB_HIP(hipGraphLaunch(hit->exec, st));
    }
    QB_HIP(hipEventRecord(B->ev1, st));
    QB_HIP(hipEventSynchronize(B->ev1));
    QB_HIP(hipGetLastError());
    QB_HIP(hipEventElapsedTime(ms, B->ev0, B->ev1));
    return QB_OK;
}

extern "C" int qb_qubo_batch_anneal(qb_qubo_batch* B, const qb_schedule* s, const qb_qubo_opts* opts, uint8_t* x,
                                    double* energy, qb_qubo_report* rep) {
    if (!B) return fail(QB_E_ARG, "qb_qubo_batch_anneal: null batch");
    int rc = check_schedule(s);
    if (rc) return rc;
    qb_qubo_opts o{0, 0, 0, 0, 0};
    if (opts) o = *opts;
    if (o.mode != 0 && o.mode != 1) return fail(QB_E_ARG, "mode must be 0 (persistent) or 1 (per instance)");
    if (o.replica_block != 0 && o.replica_block != 4 && o.replica_block != 8 && o.replica_block != 16 &&
        o.replica_block != 32)
        return fail(QB_E_ARG, "replica_block must be 0, 4, 8, 16 or 32 (got %d)", o.replica_block);
    if (o.workgroups < 0) return fail(QB_E_ARG, "negative workgroups");
    if (o.graph != 0 && o.graph != 1) return fail(QB_E_ARG, "graph must be 0 or 1 (got %d)", o.graph);
    if (o.block != 0 && o.block != 256 && o.block != 1024)
        return fail(QB_E_ARG, "block must be 0, 256 or 1024 (got %d)", o.block);
    if ((o.block == 0) != (o.replica_block == 0))
        return fail(QB_E_ARG, "block and replica_block are tuned together: give both or neither");
    hipDeviceProp_t prop;
    if ((rc = device_props(prop))) return rc;
    auto grid_for = [&](int rbv, int blk) {
        const int per_cu = std::max(1, qubo_blocks_per_cu(rbv, blk));
        return o.workgroups > 0 ? o.workgroups : per_cu * prop.multiProcessorCount;
    };
    const std::vector<double> beta = betas_of(*s);
    const int n_sweeps = int(beta.size());
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = B->w_beta.reserve(sizeof(double) * beta.size()))) return rc;
    QB_HIP(hipMemcpyAsync(B->w_beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice, B->stream));
    int rb = o.replica_block, block = o.block;
    if (rb == 0) {
        // the (threads per workgroup, replicas per job) shape that anneals a short prefix
        // of the schedule fastest
        static const int shapes[][2] = {{256, 8}, {256, 16}, {1024, 4}, {1024, 8}};
        float best = 0.0f;
        for (const auto& sh : shapes) {
            const qb_qubo_batch::Jobs* Jb = nullptr;
            if ((rc = qubo_jobs_for(B, sh[1], &Jb))) return rc;
            int l = 0;
            float ms = 0.0f;
            if ((rc = qubo_run(B, *Jb, sh[1], sh[0], std::min(n_sweeps, 8), o.mode, grid_for(sh[1], sh[0]), false, &l,
                               &ms, nullptr)))
                return rc;
            if (rb == 0 || ms < best) {
                rb = sh[1];
                block = sh[0];
                best = ms;
            }
        }
    }
    const qb_qubo_batch::Jobs* Jb = nullptr;
    if ((rc = qubo_jobs_for(B, rb, &Jb))) return rc;
    int launches = 0;
    float ms = 0.0f;
    bool replayed = false;
    if ((rc = qubo_run(B, *Jb, rb, block, n_sweeps, o.mode, grid_for(rb, block), o.graph != 0, &launches, &ms,
                       &replayed)))
        return rc;
    hipStream_t st = B->stream;
    if (energy) {
        if ((rc = B->w_energy.reserve(sizeof(double) * size_t(B->total_replicas)))) return rc;
        QB_HIP(launch_qubo_energy(qubo_dev(B, *Jb, false), B->w_mask.as<uint32_t>(), B->w_energy.as<double>(), st));
        QB_HIP(hipMemcpyAsync(energy, B->w_energy.p, sizeof(double) * size_t(B->total_replicas), hipMemcpyDeviceToHost, st));
    }
    if (x) {
        B->h_mask.resize(size_t(Jb->mask_words));
        QB_HIP(hipMemcpyAsync(B->h_mask.data(), B->w_mask.p, sizeof(uint32_t) * size_t(Jb->mask_words),
                              hipMemcpyDeviceToHost, st));
    }
    QB_HIP(hipStreamSynchronize(st));
    if (x) {
        for (const QuboJob& j : Jb->host) {
            const int n = B->n_bits[size_t(j.instance)];
            const int32_t base = B->inst_pos[size_t(j.instance)];
            uint8_t* xk = x + B->x_off[size_t(j.instance)];
            for (int p = 0; p < n; ++p) {
                const uint32_t w = B->h_mask[size_t(j.mask_offset) + size_t(p)];
                const int i = B->pos_bit[size_t(base + p)];
                for (int r = 0; r < j.replicas; ++r) xk[size_t(j.replica0 + r) * size_t(n) + size_t(i)] = uint8_t((w >> r) & 1u);
            }
        }
    }
    if (rep) {
        rep->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->kernel_ms = ms;
        rep->launches = launches;
        rep->replica_block = rb;
        rep->block = block;
        rep->jobs = int(Jb->host.size());
        rep->graph_replayed = replayed ? 1 : 0;
        double upd = 0.0;
        for (int k = 0; k < B->n_inst; ++k) upd += double(B->replicas[size_t(k)]) * double(B->n_bits[size_t(k)]);
        rep->spin_updates = upd * n_sweeps;
    }
    return QB_OK;
}

extern "C" int qb_qubo_batch_colouring(const qb_qubo_batch* B, int k, int32_t* order, int32_t* class_ptr,
                                       int* n_classes) {
    if (!B || !order || !class_ptr || !n_classes) return fail(QB_E_ARG, "qb_qubo_batch_colouring: null argument");
    if (k < 0 || k >= B->n_inst) return fail(QB_E_ARG, "instance %d of %d", k, B->n_inst);
    const int32_t base = B->inst_pos[size_t(k)];
    const int n = B->n_bits[size_t(k)];
    std::memcpy(order, B->pos_bit.data() + base, sizeof(int32_t) * size_t(n));
    const int c0 = B->inst_cls[size_t(k)], c1 = B->inst_cls[size_t(k) + 1];
    for (int c = c0; c < c1; ++c) class_ptr[c - c0] = B->cls_ptr[size_t(c)] - base;
    *n_classes = c1 - c0 - 1;
    return QB_OK;
}

// ==== dense QUBO field engine ==============================================================

struct qb_dense {
    hipStream_t stream = nullptr;
    int n = 0;
    DevBuf b_Q, b_h, w_beta, w_x, w_e, w_err, w_team, w_terr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~qb_dense() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" int qb_dense_create(int n, const float* Q, const float* h, qb_dense** out) {
    if (!out) return fail(QB_E_ARG, "qb_dense_create: null output");
    *out = nullptr;
    if (n <= 0 || !Q || !h) return fail(QB_E_ARG, "qb_dense_create: n must be positive and Q, h given");
    if (dense_lds_bytes(n, 0) > size_t(150 * 1024))
        return fail(QB_E_LIMIT, "n = %d: the FP64 fields, spins and window block of one replica must fit 150 KB of LDS", n);
    for (int i = 0; i < n; ++i) {
        if (Q[size_t(i) * size_t(n) + size_t(i)] != 0.0f) return fail(QB_E_ARG, "Q[%d][%d] is not zero", i, i);
        if (!std::isfinite(h[i])) return fail(QB_E_ARG, "h[%d] is not finite", i);
        for (int j = i + 1; j < n; ++j) {
            const float a = Q[size_t(i) * size_t(n) + size_t(j)], b = Q[size_t(j) * size_t(n) + size_t(i)];
            if (a != b || !std::isfinite(a)) return fail(QB_E_ARG, "Q is not symmetric and finite at (%d, %d)", i, j);
        }
    }
    auto D = std::make_unique<qb_dense>();
    D->n = n;
    QB_HIP(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
    QB_HIP(hipEventCreate(&D->ev0));
    QB_HIP(hipEventCreate(&D->ev1));
    int rc;
    if ((rc = D->b_Q.reserve(sizeof(float) * size_t(n) * size_t(n))) || (rc = D->b_h.reserve(sizeof(float) * size_t(n))))
        return rc;
    QB_HIP(hipMemcpyAsync(D->b_Q.p, Q, sizeof(float) * size_t(n) * size_t(n), hipMemcpyHostToDevice, D->stream));
    QB_HIP(hipMemcpyAsync(D->b_h.p, h, sizeof(float) * size_t(n), hipMemcpyHostToDevice, D->stream));
    QB_HIP(hipStreamSynchronize(D->stream));
    *out = D.release();
    return QB_OK;
}

extern "C" int qb_dense_destroy(qb_dense* d) {
    delete d;
    return QB_OK;
}

extern "C" int qb_dense_anneal(qb_dense* D, const qb_schedule* s, int R, uint64_t seed_offset, int precision,
                               int recompute_every, uint8_t* x, double* energy, double* field_err,
                               qb_dense_report* rep) {
    if (!D) return fail(QB_E_ARG, "qb_dense_anneal: null handle");
    int rc = check_schedule(s);
    if (rc) return rc;
    if (R <= 0) return fail(QB_E_ARG, "R must be positive");
    if (precision != QB_FP64 && precision != QB_FP32) return fail(QB_E_ARG, "precision must be QB_FP64 or QB_FP32");
    if (recompute_every < 0) return fail(QB_E_ARG, "negative recompute_every");
    const std::vector<double> beta = betas_of(*s);
    const int n = D->n;
    if ((rc = D->w_beta.reserve(sizeof(double) * beta.size())) || (rc = D->w_x.reserve(size_t(R) * size_t(n))) ||
        (rc = D->w_e.reserve(sizeof(double) * size_t(R))) || (rc = D->w_err.reserve(sizeof(double) * size_t(R))))
        return rc;
    hipStream_t st = D->stream;
    const auto t0 = std::chrono::steady_clock::now();
    QB_HIP(hipMemcpyAsync(D->w_beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice, st));
    const DevDense dd{n, D->b_Q.as<float>(), D->b_h.as<float>()};
    // up to 64 replicas: kDenseParts workgroups of one XCD per replica, the device's CUs shared
    // out; more replicas: one workgroup each
    const bool team = R <= kTeams * kDenseTeamSlots && n >= 64 * kDenseParts;
    if (team) {
        if ((rc = D->w_team.reserve(dense_team_work_bytes(n, R))) || (rc = D->w_terr.reserve(sizeof(int)))) return rc;
        QB_HIP(hipMemsetAsync(D->w_team.p, 0, dense_team_work_bytes(n, R), st));
        QB_HIP(hipMemsetAsync(D->w_terr.p, 0, sizeof(int), st));
    }
    const int k_rec = precision == QB_FP32 ? recompute_every : 0;
    QB_HIP(hipEventRecord(D->ev0, st));
    if (team)
        QB_HIP(launch_dense_team(dd, precision, R, D->w_beta.as<double>(), int(beta.size()), k_rec, seed_offset,
                                 D->w_x.as<uint8_t>(), D->w_e.as<double>(), D->w_err.as<double>(), D->w_team.p,
                                 D->w_terr.as<int>(), st));
    else
        QB_HIP(launch_dense(dd, precision, R, D->w_beta.as<double>(), int(beta.size()), k_rec, seed_offset,
                            D->w_x.as<uint8_t>(), D->w_e.as<double>(), D->w_err.as<double>(), st));
    QB_HIP(hipEventRecord(D->ev1, st));
    if (team) {
        int terr = 0;
        QB_HIP(hipMemcpyAsync(&terr, D->w_terr.p, sizeof(int), hipMemcpyDeviceToHost, st));
        QB_HIP(hipStreamSynchronize(st));
        if (terr) return fail(QB_E_TIMEOUT, "a dense replica's workgroups timed out waiting for each other");
    }
    if (x) QB_HIP(hipMemcpyAsync(x, D->w_x.p, size_t(R) * size_t(n), hipMemcpyDeviceToHost, st));
    if (energy) QB_HIP(hipMemcpyAsync(energy, D->w_e.p, sizeof(double) * size_t(R), hipMemcpyDeviceToHost, st));
    if (field_err) QB_HIP(hipMemcpyAsync(field_err, D->w_err.p, sizeof(double) * size_t(R), hipMemcpyDeviceToHost, st));
    QB_HIP(hipStreamSynchronize(st));
    if (rep) {
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, D->ev0, D->ev1));
        rep->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->kernel_ms = ms;
        rep->spin_updates = double(R) * double(beta.size()) * double(n);
    }
    return QB_OK;
}

// ---- exact region-of-interest posterior ------------------------------------------------

extern "C" int qb_roi_enumerate(int n, const double* h, const double* J, double beta, double* marginal,
                                uint8_t* map_state, qb_roi_result* out) {
    if (n < 1 || n > 36) return fail(QB_E_ARG, "qb_roi_enumerate: n = %d, must be 1..36", n);
    if (!h || !J || !out) return fail(QB_E_ARG, "qb_roi_enumerate: null argument");
    if (!(beta >= 0.0) || !std::isfinite(beta)) return fail(QB_E_ARG, "beta must be finite and >= 0");
    std::vector<double> Js(size_t(n) * size_t(n), 0.0);
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(h[i])) return fail(QB_E_ARG, "h[%d] is not finite", i);
        for (int j = i + 1; j < n; ++j) {
            const double v = J[size_t(i) * size_t(n) + size_t(j)];
            if (!std::isfinite(v)) return fail(QB_E_ARG, "J[%d][%d] is not finite", i, j);
            Js[size_t(i) * size_t(n) + size_t(j)] = Js[size_t(j) * size_t(n) + size_t(i)] = v;
        }
    }
    // 2^k states per thread (k <= 12 keeps the carried energy near a scratch sum), the
    // chunks split over launches of at most 2^22 threads
    const int k = std::max(0, std::min(n - 14, 12));
    const uint64_t chunks = uint64_t(1) << (n - k);
    const uint64_t per_launch = uint64_t(1) << 22;
    const int nmax = roi_nmax(n);
    const size_t max_blocks = size_t((std::min(chunks, per_launch) + 255) / 256);
    DevBuf d_h, d_J, d_e, d_s, d_m;
    int rc;
    if ((rc = d_h.reserve(sizeof(double) * size_t(n))) || (rc = d_J.reserve(sizeof(double) * Js.size())) ||
        (rc = d_e.reserve(sizeof(double) * max_blocks)) || (rc = d_s.reserve(sizeof(uint64_t) * max_blocks)) ||
        (rc = d_m.reserve(sizeof(double) * max_blocks * size_t(nmax + 1))))
        return rc;
    QB_HIP(hipMemcpy(d_h.p, h, sizeof(double) * size_t(n), hipMemcpyHostToDevice));
    QB_HIP(hipMemcpy(d_J.p, Js.data(), sizeof(double) * Js.size(), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    QB_HIP(hipEventCreate(&e0));
    QB_HIP(hipEventCreate(&e1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{e0, e1};
    double kernel_ms = 0.0;
    // pass 1: the minimum energy and its state (ties to the smallest index)
    double e_min = std::numeric_limits<double>::infinity();
    uint64_t s_min = 0;
    std::vector<double> be(max_blocks);
    std::vector<uint64_t> bs(max_blocks);
    for (uint64_t c0 = 0; c0 < chunks; c0 += per_launch) {
        const uint64_t cn = std::min(per_launch, chunks - c0);
        const size_t blocks = size_t((cn + 255) / 256);
        QB_HIP(hipEventRecord(e0, nullptr));
        QB_HIP(launch_roi_min(n, k, d_h.as<double>(), d_J.as<double>(), c0, cn, d_e.as<double>(), d_s.as<uint64_t>(),
                              nullptr));
        QB_HIP(hipEventRecord(e1, nullptr));
        QB_HIP(hipMemcpy(be.data(), d_e.p, sizeof(double) * blocks, hipMemcpyDeviceToHost));
        QB_HIP(hipMemcpy(bs.data(), d_s.p, sizeof(uint64_t) * blocks, hipMemcpyDeviceToHost));
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, e0, e1));
        kernel_ms += ms;
        for (size_t b = 0; b < blocks; ++b)
            if (be[b] < e_min || (be[b] == e_min && bs[b] < s_min)) {
                e_min = be[b];
                s_min = bs[b];
            }
    }
    // pass 2: Z and the weighted bit sums relative to the minimum, reduced in block order
    double z = 0.0;
    std::vector<double> msum(size_t(n), 0.0), bm(max_blocks * size_t(nmax + 1));
    for (uint64_t c0 = 0; c0 < chunks; c0 += per_launch) {
        const uint64_t cn = std::min(per_launch, chunks - c0);
        const size_t blocks = size_t((cn + 255) / 256);
        QB_HIP(hipEventRecord(e0, nullptr));
        QB_HIP(launch_roi_sum(n, k, d_h.as<double>(), d_J.as<double>(), beta, e_min, c0, cn, d_m.as<double>(), nullptr));
        QB_HIP(hipEventRecord(e1, nullptr));
        QB_HIP(hipMemcpy(bm.data(), d_m.p, sizeof(double) * blocks * size_t(nmax + 1), hipMemcpyDeviceToHost));
        float ms = 0.0f;
        QB_HIP(hipEventElapsedTime(&ms, e0, e1));
        kernel_ms += ms;
        for (size_t b = 0; b < blocks; ++b) {
            const double* row = bm.data() + b * size_t(nmax + 1);
            for (int i = 0; i < n; ++i) msum[size_t(i)] += row[i];
            z += row[nmax];
        }
    }
    if (marginal)
        for (int i = 0; i < n; ++i) marginal[i] = msum[size_t(i)] / z;
    if (map_state)
        for (int i = 0; i < n; ++i) map_state[i] = uint8_t((s_min >> i) & 1u);
    out->e_min = e_min;
    out->log_z = -beta * e_min + std::log(z);
    out->p_map = 1.0 / z;
    out->map_index = s_min;
    out->states = std::ldexp(1.0, n);
    out->kernel_ms = kernel_ms;
    return QB_OK;
}
