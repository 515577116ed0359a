// argos/ising_weighted.hpp -- the integer path (weighted MaxCut / Ising with exact
// int32 weights and int64 energies), drop-in for QuBLAR's src/ising_weighted.hpp.
//
//   x in {0,1}^n, s = 2x - 1, cut(x) = sum_edges w_uv [x_u != x_v], E = -cut
//   f_i = sum_j w_ij s_j, dE_i = -s_i f_i
//
// anneal() is the host annealer, as the reference's (src/ising_weighted.hpp:250-267):
// the device engine is gising::run_batch (argos/ising_gpu.cuh), and the reference's
// harnesses check one against the other. The loader, exact energies, State, TTS99 and
// SHA-256 are host code with the reference's behaviour.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../qublar.h"

namespace argos {
namespace wising {

// src/ising_weighted.hpp:35-42
struct WeightedGraph {
    int n = 0;
    int64_t m = 0;
    std::vector<int32_t> eu, ev, ew;   // edges as read, 0-indexed
    std::vector<int64_t> row;          // CSR offsets (n + 1), both directions
    std::vector<int32_t> col, w;
    int64_t sum_w = 0, n_pos = 0, n_neg = 0, n_other = 0;
};

namespace rudy {
// The text is read line by line. A line's tokens are its maximal runs of characters
// other than ' ', '\t', '\r' (a line ends at '\n'); lines without tokens are skipped.
// The first tokenful line opens with "n m", and whatever follows m on that line is read
// as an edge line of its own; every later tokenful line is an edge "u v" or "u v w".
using Tokens = std::vector<std::pair<const char*, const char*>>;

inline bool is_separator(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Splits [b, e) into tokens; returns the number found.
inline size_t split(const char* b, const char* e, Tokens& out) {
    out.clear();
    const char* p = b;
    while (p != e) {
        if (is_separator(*p)) { ++p; continue; }
        const char* q = p;
        while (q != e && !is_separator(*q)) ++q;
        out.emplace_back(p, q);
        p = q;
    }
    return out.size();
}

// An optionally signed run of decimal digits whose magnitude is at most 2^40.
inline bool integer(const std::pair<const char*, const char*>& tok, int64_t& value) {
    const char* p = tok.first;
    int sign = 1;
    if (p != tok.second && (*p == '-' || *p == '+')) sign = (*p++ == '-') ? -1 : 1;
    if (p == tok.second) return false;
    constexpr uint64_t kCap = uint64_t(1) << 40;
    uint64_t mag = 0;
    for (; p != tok.second; ++p) {
        const unsigned digit = unsigned(static_cast<unsigned char>(*p)) - unsigned('0');
        if (digit > 9u) return false;
        mag = mag * 10u + digit;
        if (mag > kCap) return false;   // checked per digit, so mag never overflows
    }
    value = sign * int64_t(mag);
    return true;
}

inline std::string text_of(const std::pair<const char*, const char*>& tok) {
    return std::string(tok.first, tok.second);
}
}  // namespace rudy

// rudy / Gset text (src/ising_weighted.hpp:83-149): "n m", then m edges "u v [w]",
// 1-indexed, a missing w is +1. A file is refused (false, err says why) when a weight is
// not an integer or leaves int32, an edge line has one token or more than three, an edge
// is a self-loop or repeats another in either orientation, an index leaves [1, n], or the
// edge count differs from m.
inline bool parse_rudy(const char* data, size_t size, WeightedGraph& g, std::string& err) {
    g = WeightedGraph();
    const char* const end = data + size;
    rudy::Tokens tok;
    int64_t n = -1, m = -1;
    int64_t line_no = 0;
    auto refuse = [&](const std::string& why) {
        err = "rudy line " + std::to_string(line_no) + ": " + why;
        return false;
    };
    // one edge from tokens [first, tok.size())
    auto edge = [&](size_t first) {
        const size_t have = tok.size() - first;
        if (have < 2) return refuse("an edge needs two vertex indices");
        if (have > 3) return refuse("an edge line holds at most u, v and a weight");
        int64_t u = 0, v = 0, w = 1;
        if (!rudy::integer(tok[first], u) || !rudy::integer(tok[first + 1], v))
            return refuse("vertex indices must be integers");
        if (have == 3) {
            if (!rudy::integer(tok[first + 2], w))
                return refuse("weight '" + rudy::text_of(tok[first + 2]) + "' is not an integer");
            if (w < int64_t(std::numeric_limits<int32_t>::min()) || w > int64_t(std::numeric_limits<int32_t>::max()))
                return refuse("weight does not fit int32");
        }
        if (std::min(u, v) < 1 || std::max(u, v) > n) return refuse("vertex index outside 1.." + std::to_string(n));
        if (u == v) return refuse("edge joins a vertex to itself");
        g.eu.push_back(int32_t(u - 1));
        g.ev.push_back(int32_t(v - 1));
        g.ew.push_back(int32_t(w));
        return true;
    };
    for (const char* b = data; b < end;) {
        const char* e = static_cast<const char*>(std::memchr(b, '\n', size_t(end - b)));
        if (!e) e = end;
        ++line_no;
        if (rudy::split(b, e, tok) > 0) {
            if (n < 0) {
                if (!rudy::integer(tok[0], n) || n < 1 || n > (int64_t(1) << 30))
                    return refuse("the vertex count must be an integer in 1..2^30");
                if (tok.size() < 2 || !rudy::integer(tok[1], m) || m < 0)
                    return refuse("the edge count must follow the vertex count as an integer >= 0");
                if (tok.size() > 2 && !edge(2)) return false;
            } else if (!edge(0)) {
                return false;
            }
        }
        b = e + 1;
    }
    if (n < 0) {
        err = "rudy: no header line";
        return false;
    }
    const int64_t got = int64_t(g.eu.size());
    if (got != m) {
        err = "rudy: " + std::to_string(got) + " edges listed, the header announced " + std::to_string(m);
        return false;
    }
    g.n = int(n);
    g.m = m;
    // Both directions of every edge, rows filled in edge order.
    g.row.assign(size_t(n) + 1, 0);
    for (int64_t k = 0; k < m; ++k) {
        ++g.row[size_t(g.eu[size_t(k)]) + 1];
        ++g.row[size_t(g.ev[size_t(k)]) + 1];
    }
    for (size_t i = 1; i < g.row.size(); ++i) g.row[i] += g.row[i - 1];
    g.col.resize(size_t(2 * m));
    g.w.resize(size_t(2 * m));
    std::vector<int64_t> fill(g.row.begin(), g.row.end() - 1);
    for (int64_t k = 0; k < m; ++k) {
        const int32_t u = g.eu[size_t(k)], v = g.ev[size_t(k)], w = g.ew[size_t(k)];
        int64_t& fu = fill[size_t(u)];
        g.col[size_t(fu)] = v;
        g.w[size_t(fu++)] = w;
        int64_t& fv = fill[size_t(v)];
        g.col[size_t(fv)] = u;
        g.w[size_t(fv++)] = w;
    }
    // A repeated edge (either orientation) is a neighbour listed twice in a row: one
    // pass over the rows with a stamp per vertex.
    std::vector<int32_t> stamp(size_t(n), -1);
    for (int32_t i = 0; i < int32_t(n); ++i) {
        for (int64_t k = g.row[size_t(i)]; k < g.row[size_t(i) + 1]; ++k) {
            const int32_t j = g.col[size_t(k)];
            if (stamp[size_t(j)] == i) {
                g = WeightedGraph();
                err = "rudy: edge " + std::to_string(i + 1) + "-" + std::to_string(j + 1) + " is listed twice";
                return false;
            }
            stamp[size_t(j)] = i;
        }
    }
    for (int32_t w : g.ew) {
        g.sum_w += w;
        if (w == 1) ++g.n_pos;
        else if (w == -1) ++g.n_neg;
        else ++g.n_other;
    }
    return true;
}

// The whole file into buf (src/ising_weighted.hpp:151-168).
inline bool read_file(const std::string& path, std::vector<char>& buf) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    buf.clear();
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

// cut over the edge list (src/ising_weighted.hpp:173-178)
inline int64_t cut_of(const WeightedGraph& g, const std::vector<uint8_t>& x) {
    int64_t cut = 0;
    for (size_t e = 0; e < g.eu.size(); ++e)
        if (x[static_cast<size_t>(g.eu[e])] != x[static_cast<size_t>(g.ev[e])]) cut += g.ew[e];
    return cut;
}

// -cut over the CSR, each edge once (j > i) (src/ising_weighted.hpp:181-189)
inline int64_t energy_scratch(const WeightedGraph& g, const std::vector<uint8_t>& x) {
    int64_t cut = 0;
    for (int i = 0; i < g.n; ++i)
        for (int64_t k = g.row[static_cast<size_t>(i)]; k < g.row[static_cast<size_t>(i) + 1]; ++k) {
            const int j = g.col[static_cast<size_t>(k)];
            if (j > i && x[static_cast<size_t>(i)] != x[static_cast<size_t>(j)]) cut += g.w[static_cast<size_t>(k)];
        }
    return -cut;
}

// Spins with a maintained local field (src/ising_weighted.hpp:193-223).
struct State {
    const WeightedGraph* g = nullptr;
    std::vector<int8_t> s;     // +-1
    std::vector<int64_t> f;    // f_i = sum_j w_ij s_j
    int64_t energy = 0;        // -cut, tracked

    explicit State(const WeightedGraph& graph)
        : g(&graph), s(static_cast<size_t>(graph.n), int8_t(-1)), f(static_cast<size_t>(graph.n), 0) {
        for (int i = 0; i < graph.n; ++i)
            for (int64_t k = graph.row[static_cast<size_t>(i)]; k < graph.row[static_cast<size_t>(i) + 1]; ++k)
                f[static_cast<size_t>(i)] -= graph.w[static_cast<size_t>(k)];
    }
    int64_t dE(int i) const { return -int64_t(s[static_cast<size_t>(i)]) * f[static_cast<size_t>(i)]; }
    void flip(int i) {
        const int64_t change = dE(i);
        const int64_t si = s[static_cast<size_t>(i)];
        for (int64_t k = g->row[static_cast<size_t>(i)]; k < g->row[static_cast<size_t>(i) + 1]; ++k)
            f[static_cast<size_t>(g->col[static_cast<size_t>(k)])] -= 2 * int64_t(g->w[static_cast<size_t>(k)]) * si;
        s[static_cast<size_t>(i)] = int8_t(-si);
        energy += change;
    }
    std::vector<uint8_t> x() const {
        std::vector<uint8_t> out(s.size());
        for (size_t i = 0; i < s.size(); ++i) out[i] = s[i] > 0 ? 1u : 0u;
        return out;
    }
};

// src/ising_weighted.hpp:227-232
struct Budget {
    double t_hot = 5.0;
    double t_cold = 1e-3;
    int anneal_sweeps = 400;
    int hold_sweeps = 20;
};

namespace detail {
inline uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform(uint64_t& s) { return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53; }
}  // namespace detail

struct RunOut {
    std::vector<uint8_t> x;
    int64_t energy_tracked = 0;
};

struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

inline qb_graph* upload_graph(const WeightedGraph& g, bool force_global) {
    qb_graph_desc d{g.n, g.m, g.eu.data(), g.ev.data(), g.ew.data()};
    qb_graph* h = nullptr;
    const int rc = qb_graph_create(&d, force_global ? 1 : 0, &h);
    if (rc != QB_OK) throw EngineError(rc, std::string("qublar: ") + qb_last_error());
    return h;
}

// One replica with seed `seed` (src/ising_weighted.hpp:250-267), on the host, as the
// reference's: the CPU annealer is the independent trajectory the reference's harnesses
// compare the device engine with (gpu_anneal's run_batch == anneal check, runs_cpu.csv,
// agreement.csv), so it is not routed to the device. Vertices in index order; one
// splitmix64 stream seeded seed * 0x2545F4914F6CDD1D + 1 is drawn only for uphill
// proposals; a flip iff dE <= 0 or u < exp(-dE / T); T(s) is the schedule's temperature.
inline RunOut anneal(const WeightedGraph& g, const Budget& b, uint64_t seed) {
    const int sweeps = b.anneal_sweeps + b.hold_sweeps;
    std::vector<double> temperature(static_cast<size_t>(std::max(sweeps, 0)));
    const double span = static_cast<double>(std::max(1, b.anneal_sweeps - 1));
    for (int s = 0; s < sweeps; ++s)
        temperature[static_cast<size_t>(s)] =
            s < b.anneal_sweeps ? b.t_hot * std::pow(b.t_cold / b.t_hot, std::min(1.0, double(s) / span)) : b.t_cold;
    State state(g);
    uint64_t draws = seed * 0x2545F4914F6CDD1Dull + 1;
    for (const double t : temperature)
        for (int i = 0; i < g.n; ++i) {
            const int64_t d = state.dE(i);
            if (d > 0 && !(detail::uniform(draws) < std::exp(-static_cast<double>(d) / t))) continue;
            state.flip(i);
        }
    RunOut out;
    out.x = state.x();
    out.energy_tracked = state.energy;
    return out;
}

// Time to reach the target with 99% confidence (src/ising_weighted.hpp:271-275).
inline double tts99(double t_mean, double p) {
    if (p <= 0.0) return std::numeric_limits<double>::infinity();
    if (p >= 0.99) return t_mean;
    return t_mean * std::log(0.01) / std::log(1.0 - p);
}

struct TtsCi { double lo, hi; };

// Bootstrap 95% interval of TTS99: `resamples` resamplings of the runs, indices drawn
// by splitmix64 from `seed`, percentiles by linear interpolation, infinite when a
// bracketing resample never succeeds (src/ising_weighted.hpp:279-303).
inline TtsCi tts99_bootstrap(const std::vector<int>& success, const std::vector<double>& time, int resamples = 2000,
                             uint64_t seed = 20260928ull) {
    const size_t r = success.size();
    std::vector<double> est;
    est.reserve(static_cast<size_t>(resamples));
    uint64_t st = seed;
    for (int k = 0; k < resamples; ++k) {
        int hits = 0;
        double total = 0.0;
        for (size_t q = 0; q < r; ++q) {
            const size_t j = static_cast<size_t>(detail::splitmix(st) % static_cast<uint64_t>(r));
            hits += success[j];
            total += time[j];
        }
        est.push_back(tts99(total / static_cast<double>(r), static_cast<double>(hits) / static_cast<double>(r)));
    }
    std::sort(est.begin(), est.end());
    auto quantile = [&](double q) {
        const double at = q * static_cast<double>(est.size() - 1);
        const size_t lo = static_cast<size_t>(std::floor(at)), hi = static_cast<size_t>(std::ceil(at));
        if (std::isinf(est[lo]) || std::isinf(est[hi])) return std::numeric_limits<double>::infinity();
        return est[lo] + (est[hi] - est[lo]) * (at - static_cast<double>(lo));
    };
    return {quantile(0.025), quantile(0.975)};
}

// SHA-256 (FIPS 180-4) for assignment and file hashes (src/ising_weighted.hpp:307-379).
class Sha256 {
public:
    Sha256() { reset(); }
    void reset() {
        static const uint32_t iv[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                       0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
        std::memcpy(state_, iv, sizeof state_);
        total_ = 0;
        used_ = 0;
    }
    void update(const void* data, size_t n) {
        const uint8_t* p = static_cast<const uint8_t*>(data);
        total_ += n;
        while (n > 0) {
            const size_t take = std::min(n, sizeof block_ - used_);
            std::memcpy(block_ + used_, p, take);
            used_ += take;
            p += take;
            n -= take;
            if (used_ == sizeof block_) {
                compress(block_);
                used_ = 0;
            }
        }
    }
    std::string hex() {
        const uint64_t bits = total_ * 8u;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (used_ != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; ++i) len[i] = static_cast<uint8_t>(bits >> (56 - 8 * i));
        update(len, 8);
        static const char digits[] = "0123456789abcdef";
        std::string out;
        for (uint32_t word : state_)
            for (int sh = 28; sh >= 0; sh -= 4) out.push_back(digits[(word >> sh) & 15u]);
        return out;
    }
    static std::string of(const void* data, size_t n) {
        Sha256 h;
        h.update(data, n);
        return h.hex();
    }

private:
    static uint32_t rotr(uint32_t v, int k) { return (v >> k) | (v << (32 - k)); }
    void compress(const uint8_t* blk) {
        static const uint32_t K[64] = {
            0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
            0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
            0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
            0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
            0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
            0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
            0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
            0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = uint32_t(blk[4 * i]) << 24 | uint32_t(blk[4 * i + 1]) << 16 | uint32_t(blk[4 * i + 2]) << 8 |
                   uint32_t(blk[4 * i + 3]);
        for (int i = 16; i < 64; ++i)
            w[i] = w[i - 16] + (rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3)) + w[i - 7] +
                   (rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10));
        uint32_t v[8];
        std::memcpy(v, state_, sizeof v);
        for (int i = 0; i < 64; ++i) {
            const uint32_t t1 = v[7] + (rotr(v[4], 6) ^ rotr(v[4], 11) ^ rotr(v[4], 25)) +
                                ((v[4] & v[5]) ^ (~v[4] & v[6])) + K[i] + w[i];
            const uint32_t t2 = (rotr(v[0], 2) ^ rotr(v[0], 13) ^ rotr(v[0], 22)) +
                                ((v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]));
            std::memmove(v + 1, v, 7 * sizeof(uint32_t));
            v[4] += t1;
            v[0] = t1 + t2;
        }
        for (int i = 0; i < 8; ++i) state_[i] += v[i];
    }
    uint32_t state_[8];
    uint8_t block_[64];
    uint64_t total_ = 0;
    size_t used_ = 0;
};

// SHA-256 of the '0'/'1' string of x (src/ising_weighted.hpp:381-385).
inline std::string assignment_sha256(const std::vector<uint8_t>& x) {
    std::string bits(x.size(), '0');
    for (size_t i = 0; i < x.size(); ++i)
        if (x[i]) bits[i] = '1';
    return Sha256::of(bits.data(), bits.size());
}

}  // namespace wising
}  // namespace argos
