// argos/engine.hpp -- the model layer of the QUBO engine, drop-in for QuBLAR's
// src/engine.hpp (same names, members and behaviour), host C++ only.
//
//   E(x) = 1/2 sum_rows w (sum_bits a x - d)^2 + lambda sum_<uv>[x_u != x_v] + kappa sum x
//
// BitField (the bits, their sites and 6-neighbour adjacency), OperatorRows (one
// sensor's rows and each bit's share of them), IsingPrior, assemble() (rows of any
// number of sensors stacked over one field) and adjoint_mismatch().
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace argos {

// src/engine.hpp:30-49
struct BinaryProblem {
    std::vector<int> var_voxel;   // bit -> site
    std::vector<int> voxel_var;   // site -> bit, or -1
    std::vector<double> d, w;     // one entry per row (ray)
    std::vector<int> row_ptr, ray;  // bit -> (row, a), CSR
    std::vector<float> a;
    std::vector<int> nbr_ptr, nbr;  // bit adjacency, CSR
    double lambda = 0.0, kappa = 0.0;
    std::vector<double> field;    // optional per-bit linear term (replaces kappa)

    int n_vars() const { return static_cast<int>(var_voxel.size()); }
    double lin(int i) const { return field.empty() ? kappa : field[static_cast<size_t>(i)]; }
    int n_rays() const { return static_cast<int>(d.size()); }
};

// src/engine.hpp:52-86
struct BitField {
    std::vector<int> site;      // bit -> site
    std::vector<int> site_bit;  // site -> bit, or -1
    std::vector<int> nbr_ptr, nbr;

    int n_bits() const { return static_cast<int>(site.size()); }

    // Bits on the cells of an nx x ny x nz grid where domain != 0, cells in x-fastest
    // order; neighbours among bits in the order -x, +x, -y, +y, -z, +z.
    static BitField grid(int nx, int ny, int nz, const std::vector<char>& domain) {
        BitField f;
        const size_t cells = static_cast<size_t>(nx) * static_cast<size_t>(ny) * static_cast<size_t>(nz);
        f.site_bit.assign(cells, -1);
        for (size_t c = 0; c < cells; ++c) {
            if (!domain[c]) continue;
            f.site_bit[c] = static_cast<int>(f.site.size());
            f.site.push_back(static_cast<int>(c));
        }
        f.nbr_ptr.assign(static_cast<size_t>(f.n_bits()) + 1, 0);
        const int step[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int b = 0; b < f.n_bits(); ++b) {
            const int c = f.site[static_cast<size_t>(b)];
            const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
            for (const auto& s : step) {
                const int qx = x + s[0], qy = y + s[1], qz = z + s[2];
                if (qx < 0 || qy < 0 || qz < 0 || qx >= nx || qy >= ny || qz >= nz) continue;
                const int u = f.site_bit[(static_cast<size_t>(qz) * static_cast<size_t>(ny) + static_cast<size_t>(qy)) *
                                             static_cast<size_t>(nx) + static_cast<size_t>(qx)];
                if (u >= 0) f.nbr.push_back(u);
            }
            f.nbr_ptr[static_cast<size_t>(b) + 1] = static_cast<int>(f.nbr.size());
        }
        return f;
    }
};

// src/engine.hpp:90-109
struct OperatorRows {
    struct Entry { int bit, row; float a; };
    std::vector<double> d, w;
    std::vector<Entry> entries;

    int n_rows() const { return static_cast<int>(d.size()); }

    // y = A x
    std::vector<double> apply(const std::vector<double>& x) const {
        std::vector<double> y(static_cast<size_t>(n_rows()), 0.0);
        for (const Entry& e : entries) y[static_cast<size_t>(e.row)] += static_cast<double>(e.a) * x[static_cast<size_t>(e.bit)];
        return y;
    }
    // g = A^T r
    std::vector<double> adjoint(const std::vector<double>& r, int n_bits) const {
        std::vector<double> g(static_cast<size_t>(n_bits), 0.0);
        for (const Entry& e : entries) g[static_cast<size_t>(e.bit)] += static_cast<double>(e.a) * r[static_cast<size_t>(e.row)];
        return g;
    }
};

// src/engine.hpp:111-114
struct IsingPrior {
    double lambda = 0.0;  // cost of an unlike neighbour pair
    double kappa = 0.0;   // cost of a 1-bit
};

// src/engine.hpp:120-149: the sensors' rows stacked (row offsets in sensor order) and
// the entries grouped by bit with std::sort on the bit index, so a single-sensor
// problem is array-identical to the reference's assembly.
inline BinaryProblem assemble(const BitField& f, const std::vector<const OperatorRows*>& sensors,
                              const IsingPrior& prior) {
    BinaryProblem p;
    p.lambda = prior.lambda;
    p.kappa = prior.kappa;
    p.var_voxel = f.site;
    p.voxel_var = f.site_bit;
    struct Item { int bit, row; float a; };
    std::vector<Item> items;
    for (const OperatorRows* s : sensors) {
        const int base = p.n_rays();
        p.d.insert(p.d.end(), s->d.begin(), s->d.end());
        p.w.insert(p.w.end(), s->w.begin(), s->w.end());
        for (const auto& e : s->entries) items.push_back({e.bit, e.row + base, e.a});
    }
    std::sort(items.begin(), items.end(), [](const Item& u, const Item& v) { return u.bit < v.bit; });
    p.row_ptr.assign(static_cast<size_t>(p.n_vars()) + 1, 0);
    for (const Item& it : items) ++p.row_ptr[static_cast<size_t>(it.bit) + 1];
    for (int i = 0; i < p.n_vars(); ++i) p.row_ptr[static_cast<size_t>(i) + 1] += p.row_ptr[static_cast<size_t>(i)];
    p.ray.reserve(items.size());
    p.a.reserve(items.size());
    for (const Item& it : items) {
        p.ray.push_back(it.row);
        p.a.push_back(it.a);
    }
    p.nbr_ptr = f.nbr_ptr;
    p.nbr = f.nbr;
    return p;
}

// src/engine.hpp:156-174: relative mismatch of <A x, r> and <x, A^T r> for random x, r
// from the 64-bit LCG (6364136223846793005, 1442695040888963407), uniform in [-1/2, 1/2).
inline double adjoint_mismatch(const OperatorRows& op, int n_bits, uint64_t seed = 7) {
    uint64_t state = seed;
    auto draw = [&state] {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<double>(state >> 11) * 0x1.0p-53 - 0.5;
    };
    std::vector<double> x(static_cast<size_t>(n_bits)), r(static_cast<size_t>(op.n_rows()));
    for (double& v : x) v = draw();
    for (double& v : r) v = draw();
    const std::vector<double> ax = op.apply(x);
    const std::vector<double> atr = op.adjoint(r, n_bits);
    double lhs = 0.0, rhs = 0.0, scale = 0.0;
    for (size_t j = 0; j < r.size(); ++j) lhs += ax[j] * r[j];
    for (size_t i = 0; i < x.size(); ++i) {
        rhs += x[i] * atr[i];
        scale += std::fabs(x[i] * atr[i]);
    }
    return std::fabs(lhs - rhs) / std::max(scale, 1e-300);
}

}  // namespace argos
