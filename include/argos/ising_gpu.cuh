// argos/ising_gpu.cuh -- the integer path on the MI355X, drop-in for QuBLAR's
// src/ising_gpu.cuh (same names and signatures; plain host C++ over libqublar_hip.so,
// the file name kept so drivers include it unchanged).
//
// Plan/make_plan, scaled_budget and temperatures keep the reference's rules.
// DeviceGraph owns the graph on the device; run_batch anneals replicas with seeds
// seed_base .. seed_base + R - 1 in the reference's own trajectory (one wave per replica,
// its spins and fields in LDS, or in global memory for graphs past 32k vertices: vertices
// in index order, one splitmix64 stream drawn on uphill proposals), so replica r is
// wising::anneal(g, b, seed_base + r), as the reference's run_batch is; wising::anneal
// itself stays the host annealer, the independent side of that comparison.
// device_energies is the independent edge-list energy. Engine errors throw
// wising::EngineError (the reference printed them and called exit(3),
// src/ising_gpu.cuh:44-52).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../qublar.h"
#include "ising_weighted.hpp"

namespace argos {
namespace gising {

using wising::Budget;
using wising::WeightedGraph;

constexpr int kWarpsPerBlock = 4;          // kept for source compatibility
constexpr int kSharedStateLimit = 16384;   // the reference's shared-state rule (bytes per replica)

// src/ising_gpu.cuh:59-72 (+ force_global, remembered for DeviceGraph)
struct Plan {
    int n = 0;
    int64_t m = 0;
    int n_pad = 0;
    int64_t B = 0;
    int64_t max_w = 0;
    int field_bytes = 4;
    int weight_bytes = 4;
    bool dense = false;
    bool shared_state = false;
    double sigma = 0.0;
    int64_t w_min = 0;
    std::string refused;
    bool force_global = false;
};

// src/ising_gpu.cuh:74-101: n_pad, the field bound B = max_i sum_j |w_ij|, weight
// statistics, storage widths, dense/shared flags, sigma = sqrt(2 sum w^2 / n).
inline Plan make_plan(const WeightedGraph& g, bool force_global = false) {
    Plan p;
    p.n = g.n;
    p.m = g.m;
    p.n_pad = (g.n + 15) / 16 * 16;
    p.force_global = force_global;
    double sum_sq = 0.0;
    for (int32_t w : g.ew) {
        const int64_t a = w < 0 ? -int64_t(w) : int64_t(w);
        sum_sq += double(a) * double(a);
        p.max_w = std::max(p.max_w, a);
        if (a != 0 && (p.w_min == 0 || a < p.w_min)) p.w_min = a;
    }
    for (int i = 0; i < g.n; ++i) {
        int64_t bound = 0;
        for (int64_t k = g.row[static_cast<size_t>(i)]; k < g.row[static_cast<size_t>(i) + 1]; ++k) {
            const int32_t w = g.w[static_cast<size_t>(k)];
            bound += w < 0 ? -int64_t(w) : int64_t(w);
        }
        p.B = std::max(p.B, bound);
    }
    if (p.B > 2147483647LL) p.refused = "field bound exceeds int32";
    if (p.w_min == 0) p.refused = "all weights are zero";
    p.field_bytes = p.B <= 127 ? 1 : (p.B <= 32767 ? 2 : 4);
    p.weight_bytes = p.max_w <= 127 ? 1 : 4;
    p.dense = g.n <= 16384 && double(2 * g.m) / double(g.n) >= double(g.n) / 4.0;
    p.shared_state = !force_global && int64_t(g.n) * (p.field_bytes + 1) <= kSharedStateLimit;
    p.sigma = std::sqrt(2.0 * sum_sq / double(g.n));
    return p;
}

// T_hot = sigma, T_cold = 0.1 w_min (src/ising_gpu.cuh:104-111)
inline Budget scaled_budget(const Plan& p, int anneal, int hold) {
    Budget b;
    b.t_hot = p.sigma;
    b.t_cold = 0.1 * double(p.w_min);
    b.anneal_sweeps = anneal;
    b.hold_sweeps = hold;
    return b;
}

// The temperature of every sweep (src/ising_gpu.cuh:114-124).
inline std::vector<double> temperatures(const Budget& b) {
    const int total = b.anneal_sweeps + b.hold_sweeps;
    std::vector<double> t(static_cast<size_t>(std::max(total, 0)));
    for (int s = 0; s < total; ++s) {
        const double frac = std::min(1.0, double(s) / double(std::max(1, b.anneal_sweeps - 1)));
        t[static_cast<size_t>(s)] = s < b.anneal_sweeps ? b.t_hot * std::pow(b.t_cold / b.t_hot, frac) : b.t_cold;
    }
    return t;
}

// The graph on the device (src/ising_gpu.cuh:278-343).
class DeviceGraph {
public:
    Plan plan;
    DeviceGraph(const WeightedGraph& g, const Plan& p) : plan(p), h_(wising::upload_graph(g, p.force_global)) {}
    ~DeviceGraph() { if (h_) qb_graph_destroy(h_); }
    DeviceGraph(const DeviceGraph&) = delete;
    DeviceGraph& operator=(const DeviceGraph&) = delete;
    qb_graph* handle() const { return h_; }

private:
    qb_graph* h_ = nullptr;
};

// src/ising_gpu.cuh:345-359
struct BatchOut {
    int R = 0, n = 0, n_pad = 0;
    std::vector<int8_t> s;           // R x n_pad spins (+-1)
    std::vector<long long> energy;   // R tracked energies (-cut)
    double wall_s = 0.0;
    double kernel_ms = 0.0;
    int launches = 0;
    int sweeps_per_launch = 0;
    std::vector<uint8_t> x(int r) const {
        std::vector<uint8_t> out(static_cast<size_t>(n));
        const int8_t* row = s.data() + static_cast<size_t>(r) * static_cast<size_t>(n_pad);
        for (int i = 0; i < n; ++i) out[static_cast<size_t>(i)] = row[i] > 0 ? 1u : 0u;
        return out;
    }
};

// Replicas with seeds seed_base .. seed_base + R - 1 (src/ising_gpu.cuh:396-458);
// updates_per_launch bounds one launch's work and never changes results.
inline BatchOut run_batch(const DeviceGraph& dg, const Budget& b, uint64_t seed_base, int R,
                          double updates_per_launch = 1.0e9) {
    BatchOut out;
    out.R = R;
    out.n = dg.plan.n;
    out.n_pad = dg.plan.n_pad;
    out.s.resize(static_cast<size_t>(R) * static_cast<size_t>(out.n_pad));
    std::vector<int64_t> e(static_cast<size_t>(R));
    const qb_schedule qs{b.t_hot, b.t_cold, b.anneal_sweeps, b.hold_sweeps};
    qb_batch_report rep{};
    const int rc =
        qb_run_batch_reference(dg.handle(), &qs, seed_base, R, updates_per_launch, out.s.data(), e.data(), &rep);
    if (rc != QB_OK) throw wising::EngineError(rc, std::string("qublar: ") + qb_last_error());
    out.energy.assign(e.begin(), e.end());
    out.wall_s = rep.wall_s;
    out.kernel_ms = rep.kernel_ms;
    out.launches = rep.launches;
    out.sweeps_per_launch = rep.sweeps_per_launch;
    return out;
}

// Energies (-cut) of K states (K x n bytes) by the device edge-list kernel
// (src/ising_gpu.cuh:461-475).
inline std::vector<long long> device_energies(const DeviceGraph& dg, const std::vector<uint8_t>& X, int K) {
    std::vector<int64_t> e(static_cast<size_t>(K));
    const int rc = qb_device_energies(dg.handle(), X.data(), K, e.data());
    if (rc != QB_OK) throw wising::EngineError(rc, std::string("qublar: ") + qb_last_error());
    return std::vector<long long>(e.begin(), e.end());
}

}  // namespace gising
}  // namespace argos
