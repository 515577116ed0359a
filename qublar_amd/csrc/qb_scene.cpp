// qb_scene.cpp -- host problem builders: the muon-tomography QUBO at real shape and
// the Erdos-Renyi max-cut graph.
//
// The muon QUBO follows the model of src/ising_recon.hpp:52-62 and
// src/op_muon.hpp:22-50: bits are the rock voxels of the surveyed (void-free) model
// outside 3 m detector rooms, in x-fastest site order with 6-neighbour adjacency
// (src/engine.hpp:61-85); one row per sky bin of each chamber whose bin-centre ray
// crosses at least one bit, with d = tau0 - t (model optical depth minus the
// measured -ln(det/open)), w = max(det, 1), and a = mu_rock * segment length on
// every bit the ray crosses. The geometry is that of src/muon_replica.hpp:23-35,
// 102-115 (pyramid, grid, chambers, 360 x 70 sky to 70 degrees); the counts are
// det ~ Poisson(N0 P(bin) exp(-tau_true)) with N0 muons per chamber under a
// cos^2(theta) sky, tau_true marched through the medium that holds the void.
#include "qublar.h"
#include "qb_host_problem.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>


namespace {

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform(uint64_t& s) { return double(splitmix(s) >> 11) * 0x1.0p-53; }

// Poisson variate: multiplication method below mean 10, else PTRS
// (Hormann 1993, "The transformed rejection method for generating Poisson random
// variables", Insurance: Mathematics and Economics 12).
int64_t poisson(double mean, uint64_t& s) {
    if (mean <= 0.0) return 0;
    if (mean < 10.0) {
        const double L = std::exp(-mean);
        int64_t k = 0;
        double p = 1.0;
        do { ++k; p *= uniform(s); } while (p > L);
        return k - 1;
    }
    const double slam = std::sqrt(mean), loglam = std::log(mean);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = uniform(s) - 0.5, V = uniform(s);
        const double us = 0.5 - std::fabs(U);
        const int64_t k = int64_t(std::floor((2.0 * a / us + b) * U + mean + 0.43));
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
            -mean + double(k) * loglam - std::lgamma(double(k) + 1.0))
            return k;
    }
}

struct Grid {
    int n;
    double h, lo[3];
};

// Amanatides-Woo traversal from p along unit d through the grid's box; visit(cell, seg)
// is called for every cell the ray crosses with a positive segment length.
template <typename Visit>
void march(const Grid& g, const double p[3], const double d[3], Visit visit) {
    double t0 = -1e300, t1 = 1e300;
    for (int a = 0; a < 3; ++a) {
        const double lo = g.lo[a], hi = g.lo[a] + g.n * g.h;
        if (std::fabs(d[a]) < 1e-12) {
            if (p[a] < lo || p[a] > hi) return;
        } else {
            const double ta = (lo - p[a]) / d[a], tb = (hi - p[a]) / d[a];
            t0 = std::max(t0, std::min(ta, tb));
            t1 = std::min(t1, std::max(ta, tb));
        }
    }
    const double ts = std::max(t0, 0.0);
    if (t1 <= ts) return;
    int vox[3], step[3];
    double tmax[3], tdelta[3];
    for (int a = 0; a < 3; ++a) {
        const double q = p[a] + d[a] * ts;
        vox[a] = std::min(g.n - 1, std::max(0, int(std::floor((q - g.lo[a]) / g.h))));
        step[a] = d[a] > 0.0 ? 1 : -1;
        if (std::fabs(d[a]) < 1e-12) {
            tmax[a] = 1e300;
            tdelta[a] = 0.0;
        } else {
            const double boundary = g.lo[a] + g.h * double(vox[a] + (d[a] > 0.0 ? 1 : 0));
            tmax[a] = std::max((boundary - p[a]) / d[a], ts);
            tdelta[a] = g.h / std::fabs(d[a]);
        }
    }
    double t = ts;
    for (int guard = 0; guard < 4 * g.n + 16; ++guard) {
        const int ax = tmax[0] <= tmax[1] ? (tmax[0] <= tmax[2] ? 0 : 2) : (tmax[1] <= tmax[2] ? 1 : 2);
        const double tn = std::min(tmax[ax], t1);
        if (tn > t) visit((size_t(vox[2]) * g.n + size_t(vox[1])) * g.n + size_t(vox[0]), tn - t);
        if (tn >= t1) break;
        t = tn;
        vox[ax] += step[ax];
        tmax[ax] += tdelta[ax];
        if (vox[ax] < 0 || vox[ax] >= g.n) break;
    }
}

}  // namespace

extern "C" int qb_muon_scene_defaults(qb_muon_scene* s) {
    if (!s) return QB_E_ARG;
    s->log2_muons = 27;
    s->rho = 2.0;
    s->mu_per_rho = 0.023;
    s->void_side = 30.0;
    s->void_z = 77.0;
    s->lambda = 2.0;
    s->kappa = std::log(999.0);
    s->min_open = 30.0;
    s->seed = 1;
    s->grid = 128;
    s->sky_az = 360;
    s->sky_th = 70;
    return QB_OK;
}

extern "C" int qb_scene_muon(const qb_muon_scene* s, qb_host_problem** out) {
    if (!s || !out) return QB_E_ARG;
    *out = nullptr;
    if (s->grid < 8 || s->grid > 512 || s->log2_muons < 4 || s->log2_muons > 40) return QB_E_ARG;
    if (s->sky_az < 1 || s->sky_th < 1 || s->sky_az > 4096 || s->sky_th > 4096) return QB_E_ARG;
    const float kHeight = 139.0f, kHalfBase = 115.0f;
    const int N = s->grid;
    const float voxel = 256.0f / float(N);
    const float lo[3] = {-128.0f, -128.0f, -10.0f};
    const float chambers[3][3] = {{0.0f, 0.0f, 40.0f}, {-40.0f, 0.0f, 25.0f}, {40.0f, 0.0f, 25.0f}};
    const double mu_rock = s->mu_per_rho * s->rho;
    const double mu_air = mu_rock * (1.2 / 2700.0);
    const size_t cells = size_t(N) * N * N;
    std::vector<uint8_t> rock(cells, 0), is_void(cells, 0);
    std::vector<int32_t> site_bit(cells, -1);
    auto H = new (std::nothrow) qb_host_problem();
    if (!H) return QB_E_NOMEM;
    std::vector<int32_t> site;
    const double hv = 0.5 * s->void_side;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t c = (size_t(k) * N + j) * N + i;
                const float x = lo[0] + (float(i) + 0.5f) * voxel;
                const float y = lo[1] + (float(j) + 0.5f) * voxel;
                const float z = lo[2] + (float(k) + 0.5f) * voxel;
                bool in_rock = false;
                if (z >= 0.0f && z <= kHeight) {
                    const float wz = kHalfBase * (1.0f - z / kHeight);
                    in_rock = std::fabs(x) <= wz && std::fabs(y) <= wz;
                }
                rock[c] = in_rock;
                is_void[c] = in_rock && std::fabs(x) <= hv && std::fabs(y) <= hv && std::fabs(z - s->void_z) <= hv;
                if (!in_rock) continue;
                bool room = false;
                for (const auto& ch : chambers) {
                    const float dx = x - ch[0], dy = y - ch[1], dz = z - ch[2];
                    room |= dx * dx + dy * dy + dz * dz < 3.0f * 3.0f;
                }
                if (!room) {
                    site_bit[c] = int32_t(site.size());
                    site.push_back(int32_t(c));
                }
            }
    const int n = int(site.size());
    H->n_vars = n;
    H->truth.resize(size_t(n));
    for (int b = 0; b < n; ++b) H->truth[size_t(b)] = is_void[size_t(site[size_t(b)])];
    // 6-neighbour adjacency, candidates in the order -x, +x, -y, +y, -z, +z
    H->nbr_ptr.assign(size_t(n) + 1, 0);
    for (int b = 0; b < n; ++b) {
        const int c = site[size_t(b)];
        const int x = c % N, y = (c / N) % N, z = c / (N * N);
        const int cand[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
        for (const auto& q : cand) {
            if (q[0] < 0 || q[0] >= N || q[1] < 0 || q[1] >= N || q[2] < 0 || q[2] >= N) continue;
            const int u = site_bit[(size_t(q[2]) * N + q[1]) * N + q[0]];
            if (u >= 0) H->nbr.push_back(u);
        }
        H->nbr_ptr[size_t(b) + 1] = int32_t(H->nbr.size());
    }
    // rays
    const Grid g{N, double(voxel), {double(lo[0]), double(lo[1]), double(lo[2])}};
    const int n_az = s->sky_az, n_th = s->sky_th;
    const double th_max = 70.0 * M_PI / 180.0;
    const double c_min = std::cos(th_max), norm = 1.0 - c_min * c_min * c_min;
    const double n0 = std::ldexp(1.0, s->log2_muons);
    const float a_scale = float(mu_rock);
    struct E { int32_t bit, row; float a; };
    std::vector<E> entries;
    for (int ch = 0; ch < 3; ++ch) {
        const double p[3] = {chambers[ch][0], chambers[ch][1], chambers[ch][2]};
        for (int bin = 0; bin < n_th * n_az; ++bin) {
            const int ith = bin / n_az, iaz = bin % n_az;
            const double thl = double(ith) / n_th * th_max, thh = double(ith + 1) / n_th * th_max;
            const double cl = std::cos(thh), chh = std::cos(thl);
            const double expected = n0 * (chh * chh * chh - cl * cl * cl) / norm / double(n_az);
            const double open = std::floor(expected + 0.5);
            if (open < s->min_open) continue;
            const double th = (double(ith) + 0.5) / n_th * th_max;
            const double az = (double(iaz) + 0.5) * 2.0 * M_PI / n_az;
            const double dir[3] = {std::sin(th) * std::cos(az), std::sin(th) * std::sin(az), std::cos(th)};
            double tau0 = 0.0, tau_true = 0.0;
            const int row = H->n_rays;
            const size_t first = entries.size();
            march(g, p, dir, [&](size_t c, double seg) {
                tau0 += (rock[c] ? mu_rock : mu_air) * seg;
                tau_true += (rock[c] && !is_void[c] ? mu_rock : mu_air) * seg;
                const int bit = site_bit[c];
                if (bit >= 0) entries.push_back(E{bit, row, float(double(a_scale) * seg)});
            });
            if (entries.size() == first) continue;
            uint64_t rs = s->seed * 0x9E3779B97F4A7C15ull + uint64_t(ch) * 0x100000001ull + uint64_t(bin) + 1;
            (void)splitmix(rs);
            const double det = std::max(double(poisson(open * std::exp(-tau_true), rs)), 1.0);
            const double t = -std::log(det / open);
            H->d.push_back(tau0 - t);
            H->w.push_back(det);
            ++H->n_rays;
        }
    }
    // assemble: entries grouped by bit, rows ascending inside a bit (src/engine.hpp:120-149)
    H->row_ptr.assign(size_t(n) + 1, 0);
    for (const E& e : entries) H->row_ptr[size_t(e.bit) + 1]++;
    for (int b = 0; b < n; ++b) H->row_ptr[size_t(b) + 1] += H->row_ptr[size_t(b)];
    H->ray.resize(entries.size());
    H->a.resize(entries.size());
    std::vector<int32_t> fill(H->row_ptr.begin(), H->row_ptr.end() - 1);
    for (const E& e : entries) {
        const int32_t k = fill[size_t(e.bit)]++;
        H->ray[size_t(k)] = e.row;
        H->a[size_t(k)] = e.a;
    }
    H->lambda = s->lambda;
    H->kappa = s->kappa;
    *out = H;
    return QB_OK;
}

// ---- gravimetry ---------------------------------------------------------------------
// The operator of src/op_gravity.hpp:17-74: a bit is a cell of volume h^3 carrying a
// density contrast drho, seen from a station as a point mass at the cell's centre, so
// a_sb = G drho h^3 (z_s - z_b) / r^3 in microGal. Here the rows are truncated at
// `radius` (the reference sums every cell). Cells lie below a flat surface z = 0, cell
// (i, j, k) centred at ((i + 1/2) h, (j + 1/2) h, -(k + 1/2) h), bits in x-fastest site
// order; one station at every column centre on the surface.

constexpr double kGravG = 6.67430e-11;    // m^3 kg^-1 s^-2 (CODATA 2018)
constexpr double kMicroGal = 1e-8;        // m / s^2

// Station offset (di, dj) columns away sees a cell of layer k iff r <= radius.
static bool sees(const qb_gravity_scene* s, int di, int dj, int k) {
    const double dx = di * s->h, dy = dj * s->h, dz = (k + 0.5) * s->h;
    return dx * dx + dy * dy + dz * dz <= s->radius * s->radius;
}

// Column reach of the rows: stations farther than the grid spans are never seen.
static int gravity_reach(const qb_gravity_scene* s) {
    return int(std::min(std::ceil(s->radius / s->h), double(std::max(s->nx, s->ny))));
}

// Lattice period of layer k: the smallest P >= 2 such that two cells of the layer whose
// columns differ by a nonzero multiple of P in x and in y see no common station, i.e.
// no lattice vector v = (a P, b P) is a difference p - q of two station offsets the
// layer sees. P = max(nx, ny) makes every cell of the layer a class of its own.
static int layer_period(const qb_gravity_scene* s, int k) {
    const int reach = gravity_reach(s);
    std::vector<std::pair<int, int>> off;
    for (int dj = -reach; dj <= reach; ++dj)
        for (int di = -reach; di <= reach; ++di)
            if (sees(s, di, dj, k)) off.emplace_back(di, dj);
    const int cap = std::max(2, std::max(s->nx, s->ny));
    for (int P = 2; P < cap; ++P) {
        bool ok = true;
        const int m = 2 * reach / P + 1;
        for (int b = -m; ok && b <= m; ++b)
            for (int a = -m; ok && a <= m; ++a) {
                if (a == 0 && b == 0) continue;
                for (const auto& p : off)
                    if (std::abs(p.first - a * P) <= reach && std::abs(p.second - b * P) <= reach &&
                        sees(s, p.first - a * P, p.second - b * P, k)) {
                        ok = false;
                        break;
                    }
            }
        if (ok) return P;
    }
    return cap;
}

namespace qb {
int gravity_finish(const qb_gravity_scene* s, qb_host_problem* H);
}

extern "C" int qb_gravity_scene_defaults(qb_gravity_scene* s) {
    if (!s) return QB_E_ARG;
    *s = qb_gravity_scene{};
    s->nx = 100;
    s->ny = 200;
    s->nz = 10;
    s->h = 5.0;
    s->radius = 50.0;
    s->drho = 1000.0;
    s->sigma = 10.0;
    const double lo[3] = {200.0, 400.0, -35.0}, hi[3] = {230.0, 440.0, -15.0};
    for (int a = 0; a < 3; ++a) {
        s->body_lo[a] = lo[a];
        s->body_hi[a] = hi[a];
    }
    s->lambda = 2.0;
    s->kappa = std::log(999.0);
    s->seed = 20261202;
    return QB_OK;
}

extern "C" int qb_scene_gravity(const qb_gravity_scene* s, qb_host_problem** out) {
    if (!s || !out) return QB_E_ARG;
    *out = nullptr;
    if (s->nx < 1 || s->ny < 1 || s->nz < 1 || !(s->h > 0.0) || !(s->radius >= 0.0) || !(s->sigma > 0.0))
        return QB_E_ARG;
    if (double(s->nx) * s->ny * s->nz > 2.0e7) return QB_E_LIMIT;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const int n = nx * ny * nz, n_st = nx * ny;
    auto H = std::unique_ptr<qb_host_problem>(new (std::nothrow) qb_host_problem());
    if (!H) return QB_E_NOMEM;
    H->n_vars = n;
    H->n_rays = n_st;
    H->truth.resize(size_t(n));
    const double m = s->drho * s->h * s->h * s->h;
    const int reach = gravity_reach(s);
    // per layer, the stations a cell sees as column offsets (row-major, so stations come
    // in ascending index) with their entry value
    std::vector<std::vector<std::pair<int, int>>> off(static_cast<size_t>(nz));
    std::vector<std::vector<float>> val(static_cast<size_t>(nz));
    for (int k = 0; k < nz; ++k)
        for (int dj = -reach; dj <= reach; ++dj)
            for (int di = -reach; di <= reach; ++di) {
                if (!sees(s, di, dj, k)) continue;
                const double dx = di * s->h, dy = dj * s->h, dz = (k + 0.5) * s->h;
                const double r2 = dx * dx + dy * dy + dz * dz;
                off[size_t(k)].emplace_back(di, dj);
                val[size_t(k)].push_back(float(kGravG * m * dz / (r2 * std::sqrt(r2)) / kMicroGal));
            }
    H->row_ptr.assign(size_t(n) + 1, 0);
    for (int b = 0; b < n; ++b) {
        const int i = b % nx, j = (b / nx) % ny, k = b / n_st;
        int cnt = 0;
        for (const auto& o : off[size_t(k)])
            cnt += i + o.first >= 0 && i + o.first < nx && j + o.second >= 0 && j + o.second < ny;
        if (int64_t(H->row_ptr[size_t(b)]) + cnt > INT32_MAX) return QB_E_LIMIT;
        H->row_ptr[size_t(b) + 1] = H->row_ptr[size_t(b)] + cnt;
    }
    H->ray.resize(size_t(H->row_ptr[size_t(n)]));
    H->a.resize(H->ray.size());
    for (int b = 0; b < n; ++b) {
        const int i = b % nx, j = (b / nx) % ny, k = b / n_st;
        int32_t e = H->row_ptr[size_t(b)];
        for (size_t q = 0; q < off[size_t(k)].size(); ++q) {
            const int si = i + off[size_t(k)][q].first, sj = j + off[size_t(k)][q].second;
            if (si < 0 || si >= nx || sj < 0 || sj >= ny) continue;
            H->ray[size_t(e)] = sj * nx + si;
            H->a[size_t(e)] = val[size_t(k)][q];
            ++e;
        }
    }
    if (int rc = qb::gravity_finish(s, H.get())) return rc;
    *out = H.release();
    return QB_OK;
}

namespace qb {

int gravity_reach_of(const qb_gravity_scene* s) { return gravity_reach(s); }

// The rest of the gravimetry problem once its rows (row_ptr, ray, a) are in place: the
// truth, d = the true body's anomaly (summed over the body's bits in bit and CSR order)
// plus Gaussian noise, w, the 6-neighbour lists, the layer-lattice colouring and the prior.
int gravity_finish(const qb_gravity_scene* s, qb_host_problem* H) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const int n = nx * ny * nz, n_st = nx * ny;
    H->truth.assign(size_t(n), 0);
    std::vector<double> model(size_t(n_st), 0.0);
    for (int b = 0; b < n; ++b) {
        const int i = b % nx, j = (b / nx) % ny, k = b / n_st;
        const double c[3] = {(i + 0.5) * s->h, (j + 0.5) * s->h, -(k + 0.5) * s->h};
        bool body = true;
        for (int a = 0; a < 3; ++a) body &= c[a] >= s->body_lo[a] && c[a] <= s->body_hi[a];
        H->truth[size_t(b)] = body;
        if (body)
            for (int32_t e = H->row_ptr[size_t(b)]; e < H->row_ptr[size_t(b) + 1]; ++e)
                model[size_t(H->ray[size_t(e)])] += double(H->a[size_t(e)]);
    }
    // d = the true body's anomaly plus Gaussian noise (Box-Muller over splitmix64)
    H->d.resize(size_t(n_st));
    H->w.assign(size_t(n_st), 1.0 / (s->sigma * s->sigma));
    for (int st = 0; st < n_st; ++st) {
        uint64_t rs = s->seed * 0x9E3779B97F4A7C15ull + uint64_t(st) + 1;
        const double u1 = 1.0 - uniform(rs), u2 = uniform(rs);
        H->d[size_t(st)] = model[size_t(st)] + s->sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    // 6-neighbour adjacency, candidates in the order -x, +x, -y, +y, -z, +z
    H->nbr.clear();
    H->nbr_ptr.assign(size_t(n) + 1, 0);
    for (int b = 0; b < n; ++b) {
        const int i = b % nx, j = (b / nx) % ny, k = b / n_st;
        const int cand[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (const auto& q : cand)
            if (q[0] >= 0 && q[0] < nx && q[1] >= 0 && q[1] < ny && q[2] >= 0 && q[2] < nz)
                H->nbr.push_back((q[2] * ny + q[1]) * nx + q[0]);
        H->nbr_ptr[size_t(b) + 1] = int32_t(H->nbr.size());
    }
    // colouring: layer k by its lattice of period P_k, layers in separate colour ranges
    H->colour.resize(size_t(n));
    int base = 0;
    for (int k = 0; k < nz; ++k) {
        const int P = layer_period(s, k);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) H->colour[size_t((k * ny + j) * nx + i)] = base + i % P + P * (j % P);
        base += P * P;
    }
    H->lambda = s->lambda;
    H->kappa = s->kappa;
    return QB_OK;
}

}  // namespace qb

// The reduced problem of freeze_provable_rock (src/ising_recon.hpp:206-233): the unfrozen
// bits in original order with their entries, every ray with its d and w, the neighbour
// lists between kept bits, and each kept bit's field lin + lambda * (its frozen
// neighbours), the bonds to frozen rock folded in.
extern "C" int qb_reduce_frozen(const qb_problem_desc* d, const uint8_t* frozen, qb_host_problem** out, int32_t* kept,
                                int* n_kept) {
    if (!d || !frozen || !out || !d->row_ptr || !d->nbr_ptr || (d->n_vars > 0 && (!d->ray || !d->a)) ||
        (d->n_rays > 0 && (!d->d || !d->w)))
        return QB_E_ARG;
    *out = nullptr;
    const int n = d->n_vars;
    auto H = std::unique_ptr<qb_host_problem>(new (std::nothrow) qb_host_problem());
    if (!H) return QB_E_NOMEM;
    std::vector<int32_t> new_index(size_t(std::max(n, 0)), -1);
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (!frozen[i]) new_index[size_t(i)] = m++;
    H->n_vars = m;
    H->n_rays = d->n_rays;
    H->d.assign(d->d, d->d + d->n_rays);
    H->w.assign(d->w, d->w + d->n_rays);
    H->row_ptr.reserve(size_t(m) + 1);
    H->nbr_ptr.reserve(size_t(m) + 1);
    H->row_ptr.push_back(0);
    H->nbr_ptr.push_back(0);
    std::vector<double> field;
    field.reserve(size_t(m));
    int k = 0;
    for (int i = 0; i < n; ++i) {
        if (frozen[i]) continue;
        if (kept) kept[k] = i;
        ++k;
        for (int e = d->row_ptr[i]; e < d->row_ptr[i + 1]; ++e) {
            H->ray.push_back(d->ray[e]);
            H->a.push_back(d->a[e]);
        }
        H->row_ptr.push_back(int32_t(H->ray.size()));
        int rock = 0;
        for (int e = d->nbr_ptr[i]; e < d->nbr_ptr[i + 1]; ++e) {
            if (frozen[d->nbr[e]]) ++rock;
            else H->nbr.push_back(new_index[size_t(d->nbr[e])]);
        }
        H->nbr_ptr.push_back(int32_t(H->nbr.size()));
        const double lin = d->field ? d->field[i] : d->kappa;
        field.push_back(lin + d->lambda * rock);
    }
    H->lambda = d->lambda;
    H->kappa = d->kappa;
    H->field = std::move(field);
    if (n_kept) *n_kept = m;
    *out = H.release();
    return QB_OK;
}

extern "C" int qb_host_problem_view(const qb_host_problem* h, qb_problem_desc* v) {
    if (!h || !v) return QB_E_ARG;
    v->n_vars = h->n_vars;
    v->n_rays = h->n_rays;
    v->row_ptr = h->row_ptr.data();
    v->ray = h->ray.data();
    v->a = h->a.data();
    v->d = h->d.data();
    v->w = h->w.data();
    v->nbr_ptr = h->nbr_ptr.data();
    v->nbr = h->nbr.data();
    v->lambda = h->lambda;
    v->kappa = h->kappa;
    v->field = h->field.empty() ? nullptr : h->field.data();
    v->draw_bit = nullptr;
    v->draw_key = nullptr;
    v->colour = h->colour.empty() ? nullptr : h->colour.data();
    return QB_OK;
}

extern "C" int qb_host_problem_truth(const qb_host_problem* h, uint8_t* truth) {
    if (!h || !truth) return QB_E_ARG;
    std::memcpy(truth, h->truth.data(), h->truth.size());
    return QB_OK;
}

extern "C" int qb_host_problem_colour(const qb_host_problem* h, int32_t* colour) {
    if (!h || !colour || h->colour.empty()) return QB_E_ARG;
    std::memcpy(colour, h->colour.data(), sizeof(int32_t) * h->colour.size());
    return QB_OK;
}

extern "C" int qb_host_problem_destroy(qb_host_problem* h) {
    delete h;
    return QB_OK;
}

extern "C" int qb_scene_er_graph(int n, double p, uint64_t seed, int32_t* eu, int32_t* ev, int32_t* ew, int64_t cap,
                                 int64_t* m_out) {
    if (n <= 0 || !m_out || cap < 0) return QB_E_ARG;
    uint64_t s = seed;
    int64_t m = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (uniform(s) < p) {
                const int32_t w = (splitmix(s) >> 63) == 0 ? 1 : -1;
                if (m < cap) {
                    eu[m] = i;
                    ev[m] = j;
                    ew[m] = w;
                }
                ++m;
            }
        }
    *m_out = m;
    return m <= cap ? QB_OK : QB_E_LIMIT;
}
