// argos/ising_recon.hpp -- binary branches on the MI355X engine, drop-in for QuBLAR's
// src/ising_recon.hpp (same names and signatures).
//
// anneal_branch and branch_fractions run on the GPU through libqublar_hip.so (FP64
// residuals, the reference's arithmetic); their trajectory visits bits by colour class
// with Philox draws keyed by (branch seed, sweep, bit) instead of index order with one
// splitmix64 stream (src/ising_recon.hpp:121-144). Branch b's seed is still
// b + 1 + seed_offset. Defining QUBLAR_REFERENCE_TRAJECTORY to 1 before including this
// header makes both follow the reference's own trajectory bit for bit instead
// (qb_anneal_branches_reference: one wave per branch, slow on large problems). The host helpers binary_energy, local_field_dE,
// freeze_provable_rock, roi_qubo and classify keep the reference's definitions, and
// muon_rows / build_binary_problem build the muon QUBO array for array as the
// reference's do, over the caller's medium and view types.
// Errors of the engine throw argos::EngineError.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../qublar.h"
#include "engine.hpp"

#ifndef QUBLAR_REFERENCE_TRAJECTORY
#define QUBLAR_REFERENCE_TRAJECTORY 0
#endif

namespace argos {

struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace engine_detail {
inline void check(int rc) {
    if (rc != QB_OK) throw EngineError(rc, std::string("qublar: ") + qb_last_error());
}
}  // namespace engine_detail

// Energy of a configuration from scratch (src/ising_recon.hpp:65-79).
inline double binary_energy(const BinaryProblem& p, const std::vector<uint8_t>& x) {
    std::vector<double> r(static_cast<size_t>(p.n_rays()));
    for (int b = 0; b < p.n_rays(); ++b) r[static_cast<size_t>(b)] = -p.d[static_cast<size_t>(b)];
    double prior = 0.0;
    for (int i = 0; i < p.n_vars(); ++i) {
        const size_t si = static_cast<size_t>(i);
        if (x[si])
            for (int k = p.row_ptr[si]; k < p.row_ptr[si + 1]; ++k)
                r[static_cast<size_t>(p.ray[static_cast<size_t>(k)])] += p.a[static_cast<size_t>(k)];
        prior += p.lin(i) * x[si];
        for (int k = p.nbr_ptr[si]; k < p.nbr_ptr[si + 1]; ++k) {
            const int j = p.nbr[static_cast<size_t>(k)];
            if (j > i && x[static_cast<size_t>(j)] != x[si]) prior += p.lambda;
        }
    }
    double data = 0.0;
    for (int b = 0; b < p.n_rays(); ++b) {
        const double rb = r[static_cast<size_t>(b)];
        data += 0.5 * p.w[static_cast<size_t>(b)] * rb * rb;
    }
    return data + prior;
}

// src/ising_recon.hpp:81-86
struct Schedule {
    double t_hot = 50.0;
    double t_cold = 1.0;
    int anneal_sweeps = 200;
    int hold_sweeps = 50;
};

namespace detail {
// SplitMix64 (src/ising_recon.hpp:89-95): drivers use it to build their own inputs.
inline uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform(uint64_t& s) { return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53; }
}  // namespace detail

// dE of flipping bit i given residuals r that match x (src/ising_recon.hpp:105-118).
inline double local_field_dE(const BinaryProblem& p, const std::vector<uint8_t>& x, const std::vector<double>& r, int i) {
    const size_t si = static_cast<size_t>(i);
    const double delta = x[si] ? -1.0 : 1.0;
    double dE = p.lin(i) * delta;
    for (int k = p.row_ptr[si]; k < p.row_ptr[si + 1]; ++k) {
        const size_t b = static_cast<size_t>(p.ray[static_cast<size_t>(k)]);
        const double a = static_cast<double>(p.a[static_cast<size_t>(k)]) * delta;
        dE += p.w[b] * (r[b] * a + 0.5 * a * a);
    }
    for (int k = p.nbr_ptr[si]; k < p.nbr_ptr[si + 1]; ++k)
        dE += p.lambda * (x[static_cast<size_t>(p.nbr[static_cast<size_t>(k)])] != x[si] ? -1.0 : 1.0);
    return dE;
}

// The problem resident on the device; reuse it to anneal one problem many times.
class DeviceProblem {
public:
    explicit DeviceProblem(const BinaryProblem& p) : n_(p.n_vars()) {
        qb_problem_desc d{};
        d.n_vars = p.n_vars();
        d.n_rays = p.n_rays();
        d.row_ptr = p.row_ptr.data();
        d.ray = p.ray.data();
        d.a = p.a.data();
        d.d = p.d.data();
        d.w = p.w.data();
        d.nbr_ptr = p.nbr_ptr.data();
        d.nbr = p.nbr.data();
        d.lambda = p.lambda;
        d.kappa = p.kappa;
        d.field = p.field.empty() ? nullptr : p.field.data();
        if (n_ > 0) engine_detail::check(qb_problem_create(&d, &h_));
    }
    ~DeviceProblem() { if (h_) qb_problem_destroy(h_); }
    DeviceProblem(const DeviceProblem&) = delete;
    DeviceProblem& operator=(const DeviceProblem&) = delete;

    int n_vars() const { return n_; }
    qb_problem* handle() const { return h_; }

    // R branches (seeds b + 1 + seed_offset). Fills whichever outputs are non-null.
    void anneal(const Schedule& s, int R, uint64_t seed_offset, std::vector<std::vector<uint8_t>>* branches,
                std::vector<float>* p_void, std::vector<uint8_t>* tri = nullptr, std::vector<float>* certainty = nullptr,
                std::vector<double>* energy = nullptr, qb_report* rep = nullptr,
                const qb_anneal_opts* opts = nullptr, bool reference_trajectory = false) const {
        if (R <= 0) throw EngineError(QB_E_ARG, "qublar: n_branches must be positive");
        const size_t n = static_cast<size_t>(n_);
        if (n_ == 0) {
            if (branches) branches->assign(static_cast<size_t>(R), {});
            if (p_void) p_void->clear();
            if (tri) tri->clear();
            if (certainty) certainty->clear();
            if (energy) energy->assign(static_cast<size_t>(R), 0.0);
            return;
        }
        const qb_schedule qs{s.t_hot, s.t_cold, s.anneal_sweeps, s.hold_sweeps};
        const qb_anneal_opts o = opts ? *opts : qb_anneal_opts{QB_FP64, 0, 0, 0, 0};
        std::vector<uint8_t> flat(branches ? static_cast<size_t>(R) * n : 0);
        if (p_void) p_void->assign(n, 0.0f);
        if (tri) tri->assign(n, 0);
        if (certainty) certainty->assign(n, 0.0f);
        if (energy) energy->assign(static_cast<size_t>(R), 0.0);
        if (reference_trajectory)
            engine_detail::check(qb_anneal_branches_reference(
                h_, &qs, R, seed_offset, branches ? flat.data() : nullptr, p_void ? p_void->data() : nullptr,
                tri ? tri->data() : nullptr, certainty ? certainty->data() : nullptr,
                energy ? energy->data() : nullptr, rep));
        else
            engine_detail::check(qb_anneal_branches(h_, &qs, R, seed_offset, &o, branches ? flat.data() : nullptr,
                                                    p_void ? p_void->data() : nullptr, tri ? tri->data() : nullptr,
                                                    certainty ? certainty->data() : nullptr,
                                                    energy ? energy->data() : nullptr, rep));
        if (branches) {
            branches->resize(static_cast<size_t>(R));
            for (size_t b = 0; b < static_cast<size_t>(R); ++b)
                (*branches)[b].assign(flat.begin() + static_cast<std::ptrdiff_t>(b * n),
                                      flat.begin() + static_cast<std::ptrdiff_t>((b + 1) * n));
        }
    }

private:
    int n_ = 0;
    qb_problem* h_ = nullptr;
};

namespace engine_detail {
// A 64-bit digest of every array and scalar of a problem (a multiply-xorshift over 8-byte
// words), so a problem passed again, at any address, is recognised by its content.
inline uint64_t digest(const BinaryProblem& p) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    auto mix = [&h](const void* data, size_t bytes) {
        const unsigned char* c = static_cast<const unsigned char*>(data);
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            h = (h ^ w) * 0xD6E8FEB86659FD93ull;
            h ^= h >> 32;
        }
        uint64_t tail = bytes;
        for (; i < bytes; ++i) tail = (tail << 8) | c[i];
        h = (h ^ tail) * 0xD6E8FEB86659FD93ull;
        h ^= h >> 29;
    };
    const int64_t sizes[6] = {p.n_vars(), p.n_rays(), int64_t(p.ray.size()), int64_t(p.nbr.size()),
                              int64_t(p.field.size()), int64_t(p.a.size())};
    mix(sizes, sizeof(sizes));
    mix(&p.lambda, sizeof(double));
    mix(&p.kappa, sizeof(double));
    mix(p.row_ptr.data(), 4 * p.row_ptr.size());
    mix(p.ray.data(), 4 * p.ray.size());
    mix(p.a.data(), 4 * p.a.size());
    mix(p.d.data(), 8 * p.d.size());
    mix(p.w.data(), 8 * p.w.size());
    mix(p.nbr_ptr.data(), 4 * p.nbr_ptr.size());
    mix(p.nbr.data(), 4 * p.nbr.size());
    mix(p.field.data(), 8 * p.field.size());
    return h;
}

// A problem resident on the device, with the lock its anneals take (one caller at a time
// per handle).
struct Resident {
    std::mutex busy;
    DeviceProblem dev;
    explicit Resident(const BinaryProblem& p) : dev(p) {}
};

// The last kResidentProblems problems annealed through anneal_branch / branch_fractions stay
// resident, so calling again on the same problem (check_ising.cu calls branch_fractions six
// times) skips the colouring and the upload. The cache is never destroyed: its handles are
// released with the process rather than in static destructors that run after the HIP
// runtime may already be gone.
constexpr size_t kResidentProblems = 4;

inline std::shared_ptr<Resident> resident(const BinaryProblem& p) {
    static std::mutex lock;
    static auto* cache = new std::vector<std::pair<uint64_t, std::shared_ptr<Resident>>>();
    const uint64_t key = digest(p);
    std::lock_guard<std::mutex> hold(lock);
    for (size_t k = 0; k < cache->size(); ++k)
        if ((*cache)[k].first == key) {
            auto hit = (*cache)[k];
            cache->erase(cache->begin() + static_cast<std::ptrdiff_t>(k));
            cache->push_back(hit);   // most recent last
            return hit.second;
        }
    auto made = std::make_shared<Resident>(p);
    if (cache->size() == kResidentProblems) cache->erase(cache->begin());
    cache->emplace_back(key, made);
    return made;
}
}  // namespace engine_detail

// One branch from all-zero x (src/ising_recon.hpp:121-144), on the GPU.
inline std::vector<uint8_t> anneal_branch(const BinaryProblem& p, const Schedule& s, uint64_t seed) {
    std::vector<std::vector<uint8_t>> br;
    const auto r = engine_detail::resident(p);
    std::lock_guard<std::mutex> hold(r->busy);
    // branch 0 has seed 0 + 1 + (seed - 1)
    r->dev.anneal(s, 1, seed - 1, &br, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                  QUBLAR_REFERENCE_TRAJECTORY != 0);
    return br.empty() ? std::vector<uint8_t>() : std::move(br[0]);
}

// R branches and the per-bit void fraction (src/ising_recon.hpp:148-166), on the GPU;
// n_threads is accepted for the reference's signature and has no effect.
inline std::vector<float> branch_fractions(const BinaryProblem& p, const Schedule& s, int n_branches,
                                           unsigned n_threads, std::vector<std::vector<uint8_t>>* keep = nullptr,
                                           uint64_t seed_offset = 0) {
    (void)n_threads;
    std::vector<float> frac;
    const auto r = engine_detail::resident(p);
    std::lock_guard<std::mutex> hold(r->busy);
    r->dev.anneal(s, n_branches, seed_offset, keep, &frac, nullptr, nullptr, nullptr, nullptr, nullptr,
                  QUBLAR_REFERENCE_TRAJECTORY != 0);
    if (p.n_vars() == 0) frac.clear();
    return frac;
}

// src/ising_recon.hpp:181-234: freeze every bit whose flip 0 -> 1 costs at least
// `margin` in every context (data bound sum w a (a/2 - d), prior bound kappa +
// lambda (frozen - free neighbours)), to a fixpoint; the reduced problem keeps the
// free bits with bonds to frozen rock folded into their field.
struct FreezeResult {
    BinaryProblem reduced;
    std::vector<int> kept;
    int rounds = 0;
};

inline FreezeResult freeze_provable_rock(const BinaryProblem& p, double margin) {
    const int n = p.n_vars();
    std::vector<double> data_bound(static_cast<size_t>(n), 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = p.row_ptr[static_cast<size_t>(i)]; k < p.row_ptr[static_cast<size_t>(i) + 1]; ++k) {
            const size_t b = static_cast<size_t>(p.ray[static_cast<size_t>(k)]);
            const double a = p.a[static_cast<size_t>(k)];
            data_bound[static_cast<size_t>(i)] += p.w[b] * a * (0.5 * a - p.d[b]);
        }
    std::vector<uint8_t> frozen(static_cast<size_t>(n), 0);
    FreezeResult out;
    bool changed = true;
    while (changed) {
        changed = false;
        for (int i = 0; i < n; ++i) {
            if (frozen[static_cast<size_t>(i)]) continue;
            int rock = 0, open = 0;
            for (int k = p.nbr_ptr[static_cast<size_t>(i)]; k < p.nbr_ptr[static_cast<size_t>(i) + 1]; ++k)
                (frozen[static_cast<size_t>(p.nbr[static_cast<size_t>(k)])] ? rock : open)++;
            if (data_bound[static_cast<size_t>(i)] + p.lin(i) + p.lambda * (rock - open) >= margin) {
                frozen[static_cast<size_t>(i)] = 1;
                changed = true;
            }
        }
        ++out.rounds;
    }
    std::vector<int> index(static_cast<size_t>(n), -1);
    for (int i = 0; i < n; ++i)
        if (!frozen[static_cast<size_t>(i)]) {
            index[static_cast<size_t>(i)] = static_cast<int>(out.kept.size());
            out.kept.push_back(i);
        }
    BinaryProblem& r = out.reduced;
    r.lambda = p.lambda;
    r.kappa = p.kappa;
    r.d = p.d;
    r.w = p.w;
    r.row_ptr.push_back(0);
    r.nbr_ptr.push_back(0);
    for (int i : out.kept) {
        const size_t si = static_cast<size_t>(i);
        r.var_voxel.push_back(p.var_voxel[si]);
        for (int k = p.row_ptr[si]; k < p.row_ptr[si + 1]; ++k) {
            r.ray.push_back(p.ray[static_cast<size_t>(k)]);
            r.a.push_back(p.a[static_cast<size_t>(k)]);
        }
        r.row_ptr.push_back(static_cast<int>(r.ray.size()));
        int rock = 0;
        for (int k = p.nbr_ptr[si]; k < p.nbr_ptr[si + 1]; ++k) {
            const int j = p.nbr[static_cast<size_t>(k)];
            if (frozen[static_cast<size_t>(j)]) ++rock;
            else r.nbr.push_back(index[static_cast<size_t>(j)]);
        }
        r.nbr_ptr.push_back(static_cast<int>(r.nbr.size()));
        r.field.push_back(p.lin(i) + p.lambda * rock);
    }
    r.voxel_var.assign(p.voxel_var.size(), -1);
    for (int j = 0; j < r.n_vars(); ++j) r.voxel_var[static_cast<size_t>(r.var_voxel[static_cast<size_t>(j)])] = j;
    return out;
}

// src/ising_recon.hpp:245-300: the exact conditional QUBO of a region R with every
// other bit held at x_bar: E(x_R) = const + sum h_i x_i + sum_{i<j} J_ij x_i x_j.
struct RoiQubo {
    std::vector<int> vars;
    std::vector<double> h;
    std::vector<double> J;   // n x n, upper triangle
};

inline RoiQubo roi_qubo(const BinaryProblem& p, const std::vector<int>& roi, const std::vector<uint8_t>& x_bar) {
    const int n = static_cast<int>(roi.size());
    RoiQubo q;
    q.vars = roi;
    q.h.assign(static_cast<size_t>(n), 0.0);
    q.J.assign(static_cast<size_t>(n) * static_cast<size_t>(n), 0.0);
    std::vector<int> pos(static_cast<size_t>(p.n_vars()), -1);
    for (int i = 0; i < n; ++i) pos[static_cast<size_t>(roi[static_cast<size_t>(i)])] = i;
    std::vector<double> r(static_cast<size_t>(p.n_rays()));
    for (int b = 0; b < p.n_rays(); ++b) r[static_cast<size_t>(b)] = -p.d[static_cast<size_t>(b)];
    for (int v = 0; v < p.n_vars(); ++v)
        if (x_bar[static_cast<size_t>(v)] && pos[static_cast<size_t>(v)] < 0)
            for (int k = p.row_ptr[static_cast<size_t>(v)]; k < p.row_ptr[static_cast<size_t>(v) + 1]; ++k)
                r[static_cast<size_t>(p.ray[static_cast<size_t>(k)])] += p.a[static_cast<size_t>(k)];
    std::vector<std::vector<std::pair<int, double>>> members(static_cast<size_t>(p.n_rays()));
    for (int i = 0; i < n; ++i) {
        const size_t vi = static_cast<size_t>(roi[static_cast<size_t>(i)]);
        for (int k = p.row_ptr[vi]; k < p.row_ptr[vi + 1]; ++k) {
            const size_t b = static_cast<size_t>(p.ray[static_cast<size_t>(k)]);
            const double a = p.a[static_cast<size_t>(k)];
            q.h[static_cast<size_t>(i)] += p.w[b] * (r[b] * a + 0.5 * a * a);
            members[b].push_back({i, a});
        }
    }
    for (size_t b = 0; b < members.size(); ++b)
        for (size_t u = 0; u < members[b].size(); ++u)
            for (size_t v = u + 1; v < members[b].size(); ++v) {
                int i = members[b][u].first, j = members[b][v].first;
                if (i > j) std::swap(i, j);
                q.J[static_cast<size_t>(i) * static_cast<size_t>(n) + static_cast<size_t>(j)] +=
                    p.w[b] * members[b][u].second * members[b][v].second;
            }
    for (int i = 0; i < n; ++i) {
        const int vi = roi[static_cast<size_t>(i)];
        q.h[static_cast<size_t>(i)] += p.lin(vi);
        for (int k = p.nbr_ptr[static_cast<size_t>(vi)]; k < p.nbr_ptr[static_cast<size_t>(vi) + 1]; ++k) {
            const int vj = p.nbr[static_cast<size_t>(k)];
            const int pj = pos[static_cast<size_t>(vj)];
            if (pj >= 0) {
                if (pj > i) {
                    q.h[static_cast<size_t>(i)] += p.lambda;
                    q.h[static_cast<size_t>(pj)] += p.lambda;
                    q.J[static_cast<size_t>(i) * static_cast<size_t>(n) + static_cast<size_t>(pj)] -= 2.0 * p.lambda;
                }
            } else {
                q.h[static_cast<size_t>(i)] += p.lambda * (x_bar[static_cast<size_t>(vj)] ? -1.0 : 1.0);
            }
        }
    }
    return q;
}

// ---- the muon sensor's rows and the QUBO builder -----------------------------------
// src/op_muon.hpp:22-50 and src/ising_recon.hpp:52-62. The medium and the views are
// the caller's types -- in the reference, VoxelMedium (src/muon.cuh:38-58) and MuonView /
// MuonBinnedData (src/muon_recon.hpp:44-70), defined in its CUDA headers -- so both are
// template parameters here. Members used: medium {lo.x, lo.y, lo.z, voxel, nx, ny, nz,
// mu_rock, mu[cell]}, view {chamber.x/y/z, data}, data {size(), open[b], det[b],
// bin_direction(b).x/y/z}.

namespace muon_detail {

// Amanatides & Woo (1987), "A fast voxel traversal algorithm for ray tracing", in float
// and in the order of operations of the reference's march_impl (src/muon.cuh:60-139), so
// the segments -- and the problem's entries -- come out bit for bit the same: the grid's
// slab entry and exit (components with |d| < 1e-12 never divided by), the start cell
// clamped into the grid, then one cell per step along the axis whose next boundary is
// nearest (ties to x, then y), visit(cell, segment) summed over every positive segment,
// at most 3 (nx + ny + nz) + 16 steps. As there, the boundaries are ray parameters from
// the origin p and the end is the exit parameter less the entry one.
template <typename Medium, typename Visit>
inline float march_cells(const Medium& m, const float p[3], const float d[3], Visit&& visit) {
    constexpr float kFlat = 1e-12f;
    const float lo[3] = {m.lo.x, m.lo.y, m.lo.z};
    const int n[3] = {m.nx, m.ny, m.nz};
    float enter = -1e30f, leave = 1e30f;
    for (int a = 0; a < 3; ++a) {
        if (std::fabs(d[a]) < kFlat) {
            if (p[a] < lo[a] || p[a] > lo[a] + n[a] * m.voxel) return 0.0f;
            continue;
        }
        const float t_lo = (lo[a] - p[a]) / d[a];
        const float t_hi = (lo[a] + n[a] * m.voxel - p[a]) / d[a];
        enter = std::max(enter, std::min(t_lo, t_hi));
        leave = std::min(leave, std::max(t_lo, t_hi));
    }
    if (leave <= std::max(enter, 0.0f)) return 0.0f;
    const float t_start = std::max(enter, 0.0f);
    const float t_end = leave - t_start;
    int cell[3];
    float next[3], step[3];
    for (int a = 0; a < 3; ++a) {
        const float q = p[a] + d[a] * t_start;
        cell[a] = std::min(std::max(static_cast<int>(std::floor((q - lo[a]) / m.voxel)), 0), n[a] - 1);
    }
    for (int a = 0; a < 3; ++a) {
        if (std::fabs(d[a]) < kFlat) {
            next[a] = 1e30f;
            step[a] = 0.0f;
            continue;
        }
        const float edge = static_cast<float>(d[a] > 0.0f ? cell[a] + 1 : cell[a]) * m.voxel + lo[a];
        next[a] = std::max((edge - p[a]) / d[a], 0.0f);
        step[a] = m.voxel / std::fabs(d[a]);
    }
    float total = 0.0f, done = 0.0f;
    const int max_steps = 3 * (m.nx + m.ny + m.nz) + 16;
    for (int it = 0; it < max_steps; ++it) {
        const int a = next[0] <= next[1] ? (next[0] <= next[2] ? 0 : 2) : (next[1] <= next[2] ? 1 : 2);
        const float t = std::min(next[a], t_end);
        const float seg = t - done;
        if (seg > 0.0f) total += visit((cell[2] * m.ny + cell[1]) * m.nx + cell[0], seg);
        if (t >= t_end) break;
        done = t;
        cell[a] += d[a] > 0.0f ? 1 : -1;
        next[a] += step[a];
        if (cell[a] < 0 || cell[a] >= n[a]) break;
    }
    return total;
}

}  // namespace muon_detail

// One row per sky bin of each view with at least min_open open-sky muons whose
// bin-centre ray crosses a bit: d = tau0 - t (the model's optical depth, marched in
// float, minus the measured -ln(det / open)), w = max(det, 1), and on every bit the ray
// crosses a = a_per_metre * segment (NaN: the medium's mu_rock, the void path).
template <typename Medium, typename View>
inline OperatorRows muon_rows(const Medium& m_model, const std::vector<View>& views, const BitField& bits,
                              double min_open = 30.0, float a_per_metre = std::numeric_limits<float>::quiet_NaN()) {
    OperatorRows op;
    const float a_scale = std::isnan(a_per_metre) ? m_model.mu_rock : a_per_metre;
    for (const View& view : views) {
        const auto& data = *view.data;
        const float p[3] = {view.chamber.x, view.chamber.y, view.chamber.z};
        for (int b = 0; b < data.size(); ++b) {
            if (static_cast<double>(data.open[b]) < min_open) continue;
            const auto dir = data.bin_direction(b);
            const float d[3] = {dir.x, dir.y, dir.z};
            const double tau0 = muon_detail::march_cells(m_model, p, d, [&](int c, float seg) {
                return m_model.mu[c] * seg;
            });
            const double det = std::max(static_cast<double>(data.det[b]), 1.0);
            const double t = -std::log(det / static_cast<double>(data.open[b]));
            const int row = op.n_rows();
            const size_t first = op.entries.size();
            muon_detail::march_cells(m_model, p, d, [&](int c, float seg) {
                const int bit = bits.site_bit[static_cast<size_t>(c)];
                if (bit >= 0) op.entries.push_back({bit, row, a_scale * seg});
                return 0.0f;
            });
            if (op.entries.size() == first) continue;
            op.d.push_back(tau0 - t);
            op.w.push_back(det);
        }
    }
    return op;
}

// The QUBO of binned muon data: a grid bit field over the domain, the muon rows, the
// Ising prior -- assembled into one BinaryProblem.
template <typename Medium, typename View>
inline BinaryProblem build_binary_problem(const Medium& m_model, const std::vector<View>& views,
                                          const std::vector<char>& domain, double lambda, double kappa,
                                          double min_open = 30.0,
                                          float a_per_metre = std::numeric_limits<float>::quiet_NaN()) {
    const BitField bits = BitField::grid(m_model.nx, m_model.ny, m_model.nz, domain);
    const OperatorRows muons = muon_rows(m_model, views, bits, min_open, a_per_metre);
    return assemble(bits, {&muons}, IsingPrior{lambda, kappa});
}

// src/ising_recon.hpp:302-309
enum class Bit : uint8_t { Exists = 0, NotThere = 1, Undecided = 2 };

inline Bit classify(float p_void, float lo = 0.1f, float hi = 0.9f) {
    if (p_void >= hi) return Bit::NotThere;
    if (p_void <= lo) return Bit::Exists;
    return Bit::Undecided;
}

}  // namespace argos
