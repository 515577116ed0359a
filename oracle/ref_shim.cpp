// ref_shim.cpp -- C entry points over the reference's own integer path and model
// layer, compiled from the reference sources where they lie (see oracle/Makefile).
// TEST INFRASTRUCTURE ONLY: the output, oracle/_ref/libqb_ref.so, is a checker for
// the oracle; nothing in the product links it.
//
// Compiles src/ising_weighted.hpp (std only) and src/engine.hpp (std only).
// src/ising_recon.hpp is not buildable here: its include chain reaches
// cuda_runtime.h (ising_recon.hpp:28 -> muon_recon.hpp -> muon.cuh -> lidar.cuh
// -> trace.cuh), which this image does not have.
#include "engine.hpp"
#include "ising_weighted.hpp"

#include <cstring>
#include <vector>

namespace wi = argos::wising;

static wi::WeightedGraph make_graph(int n, int64_t m, const int32_t* eu, const int32_t* ev,
                                    const int32_t* ew) {
    // Route through the reference loader so the CSR is the reference's own.
    std::string t = std::to_string(n) + " " + std::to_string(m) + "\n";
    for (int64_t e = 0; e < m; ++e)
        t += std::to_string(eu[e] + 1) + " " + std::to_string(ev[e] + 1) + " " + std::to_string(ew[e]) + "\n";
    wi::WeightedGraph g;
    std::string err;
    wi::parse_rudy(t.data(), t.size(), g, err);
    return g;
}

extern "C" {

// wising::anneal (ising_weighted.hpp:250-267) for one seed.
int ref_anneal_int(int n, int64_t m, const int32_t* eu, const int32_t* ev, const int32_t* ew,
                   double t_hot, double t_cold, int anneal, int hold, uint64_t seed, uint8_t* x_out,
                   int64_t* energy) {
    const wi::WeightedGraph g = make_graph(n, m, eu, ev, ew);
    wi::Budget b;
    b.t_hot = t_hot;
    b.t_cold = t_cold;
    b.anneal_sweeps = anneal;
    b.hold_sweeps = hold;
    const wi::RunOut r = wi::anneal(g, b, seed);
    std::memcpy(x_out, r.x.data(), size_t(n));
    *energy = r.energy_tracked;
    return 0;
}

// parse_rudy (ising_weighted.hpp:83-149): 0 = accepted; fills n, m, sum_w.
int ref_parse_rudy(const char* text, size_t size, int* n, int64_t* m, int64_t* sum_w) {
    wi::WeightedGraph g;
    std::string err;
    if (!wi::parse_rudy(text, size, g, err)) return 1;
    *n = g.n;
    *m = g.m;
    *sum_w = g.sum_w;
    return 0;
}

// parse_rudy's whole result as one FNV-1a 64 digest over n, m, the edge lists, the CSR
// and the weight counts (0 when refused); returns 0 when accepted.
int ref_parse_rudy_digest(const char* text, size_t size, uint64_t* digest) {
    wi::WeightedGraph g;
    std::string err;
    *digest = 0;
    if (!wi::parse_rudy(text, size, g, err)) return 1;
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void* p, size_t bytes) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < bytes; ++i) h = (h ^ c[i]) * 1099511628211ull;
    };
    const int64_t head[6] = {g.n, g.m, g.sum_w, g.n_pos, g.n_neg, g.n_other};
    mix(head, sizeof(head));
    mix(g.eu.data(), 4 * g.eu.size());
    mix(g.ev.data(), 4 * g.ev.size());
    mix(g.ew.data(), 4 * g.ew.size());
    mix(g.row.data(), 8 * g.row.size());
    mix(g.col.data(), 4 * g.col.size());
    mix(g.w.data(), 4 * g.w.size());
    *digest = h;
    return 0;
}

int64_t ref_cut_of(int n, int64_t m, const int32_t* eu, const int32_t* ev, const int32_t* ew,
                   const uint8_t* x) {
    const wi::WeightedGraph g = make_graph(n, m, eu, ev, ew);
    return wi::cut_of(g, std::vector<uint8_t>(x, x + n));
}

void ref_tts99_bootstrap(const int* success, const double* time, size_t r, int resamples,
                         uint64_t seed, double* lo, double* hi) {
    const wi::TtsCi ci = wi::tts99_bootstrap(std::vector<int>(success, success + r),
                                            std::vector<double>(time, time + r), resamples, seed);
    *lo = ci.lo;
    *hi = ci.hi;
}

void ref_sha256_hex(const void* data, size_t n, char out[65]) {
    const std::string h = wi::Sha256::of(data, n);
    std::memcpy(out, h.c_str(), 65);
}

// BitField::grid + assemble (engine.hpp:61-149) over one set of explicit entries.
// Output sizes are queried with null pointers first.
int ref_assemble(int nx, int ny, int nz, const char* domain, int n_rows, const double* d,
                 const double* w, int n_entries, const int* e_bit, const int* e_row, const float* e_a,
                 int* n_vars_out, int* nnz_out, int* nnbr_out, int* row_ptr, int* ray, float* a,
                 int* nbr_ptr, int* nbr, int* var_voxel) {
    const std::vector<char> dom(domain, domain + size_t(nx) * ny * nz);
    const argos::BitField f = argos::BitField::grid(nx, ny, nz, dom);
    argos::OperatorRows op;
    op.d.assign(d, d + n_rows);
    op.w.assign(w, w + n_rows);
    for (int k = 0; k < n_entries; ++k) op.entries.push_back({e_bit[k], e_row[k], e_a[k]});
    const argos::BinaryProblem p = argos::assemble(f, {&op}, argos::IsingPrior{0.0, 0.0});
    *n_vars_out = p.n_vars();
    *nnz_out = int(p.ray.size());
    *nnbr_out = int(p.nbr.size());
    if (row_ptr) std::memcpy(row_ptr, p.row_ptr.data(), sizeof(int) * p.row_ptr.size());
    if (ray) std::memcpy(ray, p.ray.data(), sizeof(int) * p.ray.size());
    if (a) std::memcpy(a, p.a.data(), sizeof(float) * p.a.size());
    if (nbr_ptr) std::memcpy(nbr_ptr, p.nbr_ptr.data(), sizeof(int) * p.nbr_ptr.size());
    if (nbr) std::memcpy(nbr, p.nbr.data(), sizeof(int) * p.nbr.size());
    if (var_voxel) std::memcpy(var_voxel, p.var_voxel.data(), sizeof(int) * p.var_voxel.size());
    return 0;
}

// adjoint_mismatch (engine.hpp:156-174) of explicit entries.
double ref_adjoint_mismatch(int n_bits, int n_rows, int n_entries, const int* e_bit, const int* e_row,
                            const float* e_a, uint64_t seed) {
    argos::OperatorRows op;
    op.d.assign(size_t(n_rows), 0.0);
    op.w.assign(size_t(n_rows), 1.0);
    for (int k = 0; k < n_entries; ++k) op.entries.push_back({e_bit[k], e_row[k], e_a[k]});
    return argos::adjoint_mismatch(op, n_bits, seed);
}

}  // extern "C"
