// check_dropin -- a reference-style driver built against the drop-in headers only
// (argos/engine.hpp, ising_recon.hpp, ising_weighted.hpp, ising_gpu.cuh), in the
// shape of QuBLAR's own checks (tools/gset_bench.cu --check, tools/gpu_anneal.cu
// --check, src/check_ising.cu A, src/check_engine.cu): one PASS/FAIL line per
// invariant, exit code 1 on any failure.
//   check_dropin            every check (needs the GPU)
//   check_dropin --host     the host-only checks
//   check_dropin --assemble / --muon ...   problem dumps for the CPU tests
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "argos/engine.hpp"
#include "argos/ising_gpu.cuh"
#include "argos/ising_recon.hpp"
#include "argos/ising_weighted.hpp"

using namespace argos;
namespace wi = argos::wising;
namespace gi = argos::gising;

static int g_fail = 0;

static void line(const char* what, bool ok, const std::string& detail) {
    std::printf("  %-62s %s  %s\n", what, ok ? "PASS" : "FAIL", detail.c_str());
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

static bool parse(const std::string& t, wi::WeightedGraph& g) {
    std::string err;
    return wi::parse_rudy(t.data(), t.size(), g, err);
}

// K_n with weights uniform in [-wmax, wmax] from splitmix64 (the self-check generator).
static wi::WeightedGraph random_complete(int n, int wmax, uint64_t seed) {
    std::string t = std::to_string(n) + " " + std::to_string(n * (n - 1) / 2) + "\n";
    uint64_t s = seed;
    for (int i = 1; i <= n; ++i)
        for (int j = i + 1; j <= n; ++j)
            t += std::to_string(i) + " " + std::to_string(j) + " " +
                 std::to_string(int(wi::detail::splitmix(s) % uint64_t(2 * wmax + 1)) - wmax) + "\n";
    wi::WeightedGraph g;
    parse(t, g);
    return g;
}

static wi::WeightedGraph random_sparse(int n, int m, uint64_t seed, int wmax) {
    uint64_t s = seed;
    std::string t = std::to_string(n) + " " + std::to_string(m) + "\n";
    std::vector<char> seen(size_t(n) * size_t(n), 0);
    for (int made = 0; made < m;) {
        const int u = int(wi::detail::splitmix(s) % uint64_t(n)), v = int(wi::detail::splitmix(s) % uint64_t(n));
        if (u == v) continue;
        const int a = std::min(u, v), b = std::max(u, v);
        if (seen[size_t(a) * size_t(n) + size_t(b)]) continue;
        seen[size_t(a) * size_t(n) + size_t(b)] = 1;
        t += std::to_string(a + 1) + " " + std::to_string(b + 1);
        if (wmax > 0) {
            int w = 0;
            while (w == 0) w = int(wi::detail::splitmix(s) % uint64_t(2 * wmax + 1)) - wmax;
            t += " " + std::to_string(w);
        }
        t += "\n";
        ++made;
    }
    wi::WeightedGraph g;
    parse(t, g);
    return g;
}

// The 16-variable problem of check_ising A (splitmix seed 7, 12 rays, 4x4 grid).
static BinaryProblem sixteen_var() {
    BinaryProblem p;
    uint64_t s = 7;
    p.lambda = 0.7;
    p.kappa = 0.4;
    for (int i = 0; i < 16; ++i) p.var_voxel.push_back(i);
    p.voxel_var = p.var_voxel;
    for (int b = 0; b < 12; ++b) {
        p.d.push_back(detail::uniform(s) * 0.3);
        p.w.push_back(20.0 + 80.0 * detail::uniform(s));
    }
    p.row_ptr.push_back(0);
    for (int i = 0; i < 16; ++i) {
        for (int b = 0; b < 12; ++b)
            if (detail::uniform(s) < 0.35) {
                p.ray.push_back(b);
                p.a.push_back(float(0.02 + 0.1 * detail::uniform(s)));
            }
        p.row_ptr.push_back(int(p.ray.size()));
    }
    p.nbr_ptr.push_back(0);
    for (int i = 0; i < 16; ++i) {
        const int x = i % 4, y = i / 4;
        if (x > 0) p.nbr.push_back(i - 1);
        if (x < 3) p.nbr.push_back(i + 1);
        if (y > 0) p.nbr.push_back(i - 4);
        if (y < 3) p.nbr.push_back(i + 4);
        p.nbr_ptr.push_back(int(p.nbr.size()));
    }
    return p;
}

static void host_checks() {
    // loader refusals (gset_bench.cu:313-334)
    const char* bad[] = {"3 1\n1 2 0.5\n", "3 1\n1 1 1\n", "3 2\n1 2 1\n2 1 1\n", "3 2\n1 2 1\n", "3 1\n1 4 1\n"};
    int refused = 0;
    for (const char* t : bad) {
        wi::WeightedGraph g;
        refused += parse(t, g) ? 0 : 1;
    }
    line("the loader refuses 5 malformed inputs", refused == 5, std::to_string(refused) + " of 5 refused");
    {
        wi::WeightedGraph g;
        const bool ok = parse("3 2\n1 2\n2 3 -7\n", g);
        line("a missing weight is +1, a signed weight is kept", ok && g.sum_w == -6 && g.m == 2,
             "sum of weights " + std::to_string(g.sum_w));
    }
    // gray-code walk: exact energies on every state of three integer K16 (gset_bench.cu:288-309)
    int64_t bad_states = 0, flips = 0;
    for (uint64_t seed = 1; seed <= 3; ++seed) {
        const wi::WeightedGraph g = random_complete(16, 1000, seed);
        wi::State st(g);
        std::vector<uint8_t> x(16, 0);
        for (uint64_t k = 1; k < (uint64_t(1) << 16); ++k) {
            int bit = 0;
            while (!((k >> bit) & 1)) ++bit;
            const int64_t before = st.energy, predicted = st.dE(bit);
            st.flip(bit);
            x[size_t(bit)] ^= 1u;
            ++flips;
            const int64_t cut = wi::cut_of(g, x);
            if (st.energy != -cut || st.energy != wi::energy_scratch(g, x) || st.energy - before != predicted)
                ++bad_states;
        }
    }
    line("integer energy exact on all 2^16 states, 3 K16", bad_states == 0,
         std::to_string(bad_states) + " mismatches over " + std::to_string(flips) + " flips");
    line("SHA-256 of 'abc' (FIPS 180-4)",
         wi::Sha256::of("abc", 3) == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad", "");
    // model layer: grid, assemble, adjoint (check_engine.cu:177-197)
    {
        const int nx = 6, ny = 5, nz = 4;
        std::vector<char> dom(size_t(nx * ny * nz), 0);
        for (size_t c = 0; c < dom.size(); ++c) dom[c] = (c * 7919u) % 5u != 0;
        const BitField f = BitField::grid(nx, ny, nz, dom);
        OperatorRows op;
        uint64_t s = 11;
        for (int r = 0; r < 40; ++r) {
            op.d.push_back(detail::uniform(s));
            op.w.push_back(1.0 + detail::uniform(s));
            for (int k = 0; k < 6; ++k)
                op.entries.push_back({int(wi::detail::splitmix(s) % uint64_t(f.n_bits())), r,
                                      float(detail::uniform(s))});
        }
        const BinaryProblem p = assemble(f, {&op, &op}, IsingPrior{1.0, 2.0});
        const double mis = adjoint_mismatch(op, f.n_bits());
        line("assemble stacks two sensors; adjoint mismatch < 1e-12",
             p.n_rays() == 80 && int(p.ray.size()) == 480 && p.row_ptr.back() == 480 && mis < 1e-12,
             "mismatch " + std::to_string(mis));
    }
    // ROI expansion and freeze bound (check_ising.cu:536-559)
    {
        const BinaryProblem p = sixteen_var();
        std::vector<uint8_t> xbar(16, 0);
        uint64_t s = 5;
        for (auto& v : xbar) v = uint8_t(wi::detail::splitmix(s) >> 63);
        const std::vector<int> roi = {1, 2, 5, 6, 9, 10, 13};
        const RoiQubo q = roi_qubo(p, roi, xbar);
        std::vector<uint8_t> base = xbar;
        for (int v : roi) base[size_t(v)] = 0;
        const double e0 = binary_energy(p, base);
        double worst = 0.0;
        for (int t = 0; t < 8; ++t) {
            std::vector<uint8_t> x = base, xr(roi.size());
            for (size_t i = 0; i < roi.size(); ++i) x[size_t(roi[i])] = xr[i] = uint8_t(wi::detail::splitmix(s) >> 63);
            double pred = 0.0;
            for (size_t i = 0; i < roi.size(); ++i) {
                pred += q.h[i] * xr[i];
                for (size_t j = i + 1; j < roi.size(); ++j) pred += q.J[i * roi.size() + j] * xr[i] * xr[j];
            }
            worst = std::max(worst, std::fabs(binary_energy(p, x) - e0 - pred));
        }
        line("ROI QUBO equals the full energy (8 random x)", worst < 1e-9, "worst abs err " + std::to_string(worst));
    }
}

static void gpu_checks() {
    // device energy oracle (gpu_anneal.cu:155-164)
    {
        int64_t bad = 0;
        for (uint64_t seed = 1; seed <= 3; ++seed) {
            const wi::WeightedGraph g = random_complete(16, 1000, seed);
            const gi::DeviceGraph dg(g, gi::make_plan(g));
            std::vector<uint8_t> X(size_t(1000) * 16);
            uint64_t s = 20260929ull + seed;
            for (auto& v : X) v = uint8_t(wi::detail::splitmix(s) >> 63);
            const std::vector<long long> e = gi::device_energies(dg, X, 1000);
            for (int k = 0; k < 1000; ++k) {
                const std::vector<uint8_t> x(X.begin() + k * 16, X.begin() + (k + 1) * 16);
                if (e[size_t(k)] != -wi::cut_of(g, x)) ++bad;
            }
        }
        line("device energy = host energy, 3 K16 x 1000 random states", bad == 0,
             std::to_string(bad) + " mismatches of 3000");
    }
    // run_batch = anneal seed for seed; energies consistent; launch split changes nothing
    auto compare = [](const wi::WeightedGraph& g, bool force_global, const char* name) {
        const gi::Plan p = gi::make_plan(g, force_global);
        const gi::DeviceGraph dg(g, p);
        const wi::Budget b = gi::scaled_budget(p, 400, 20);
        const gi::BatchOut out = gi::run_batch(dg, b, 1, 8);
        const gi::BatchOut split = gi::run_batch(dg, b, 1, 8, 8.0 * g.n * 13);
        int same = 0, consistent = 0;
        for (int r = 0; r < 8; ++r) {
            const wi::RunOut one = wi::anneal(g, b, uint64_t(r + 1));
            const std::vector<uint8_t> x = out.x(r);
            same += x == one.x && out.energy[size_t(r)] == one.energy_tracked && split.x(r) == x;
            const int64_t cut = wi::cut_of(g, x);
            consistent += cut == -out.energy[size_t(r)] && cut == -wi::energy_scratch(g, x);
        }
        line((std::string("run_batch = anneal = split launches, ") + name).c_str(), same == 8 && consistent == 8,
             std::to_string(same) + " of 8 identical, " + std::to_string(consistent) + " consistent, " +
                 std::to_string(split.launches) + " launches");
    };
    compare(random_complete(16, 1000, 1), false, "K16");
    compare(random_sparse(300, 3000, 42, 0), false, "sparse +1");
    compare(random_sparse(300, 3000, 42, 0), true, "sparse +1, global state");
    compare(random_sparse(300, 3000, 7, 1000), false, "signed sparse");
    // the 16-variable ground state (check_ising.cu:116-167)
    {
        const BinaryProblem p = sixteen_var();
        double exact = 1e300;
        std::vector<uint8_t> x(16);
        for (uint32_t c = 0; c < (1u << 16); ++c) {
            for (int i = 0; i < 16; ++i) x[size_t(i)] = (c >> i) & 1u;
            exact = std::min(exact, binary_energy(p, x));
        }
        Schedule map;
        map.t_hot = 5.0;
        map.t_cold = 1e-3;
        map.anneal_sweeps = 400;
        map.hold_sweeps = 20;
        std::vector<std::vector<uint8_t>> br;
        const std::vector<float> pv = branch_fractions(p, map, 8, 1, &br);
        double best = 1e300;
        for (const auto& b : br) best = std::min(best, binary_energy(p, b));
        const double one = binary_energy(p, anneal_branch(p, map, 3));
        line("annealer reaches the exact ground state (best of 8)", std::fabs(best - exact) < 1e-9 && pv.size() == 16,
             "annealed " + std::to_string(best) + ", exact " + std::to_string(exact));
        line("anneal_branch(seed 3) = branch 2 of branch_fractions", std::fabs(one - binary_energy(p, br[2])) == 0.0,
             "");
        // The problem stays resident between calls: a second call gives the same branches, and
        // a copy with one datum changed is a different problem, not the resident one.
        std::vector<std::vector<uint8_t>> again;
        branch_fractions(p, map, 8, 1, &again);
        BinaryProblem q = p;
        q.d[0] += 40.0;
        q.kappa = -3.0;
        std::vector<std::vector<uint8_t>> other;
        branch_fractions(q, map, 8, 1, &other);
        line("a resident problem anneals as a fresh one; a changed copy is not served stale", again == br && other != br,
             "");
    }
}

// --assemble nx ny nz seed rows: a random domain and sensor from splitmix64(seed),
// assembled; prints the arrays (floats as hex) for comparison with the reference.
static int assemble_dump(int nx, int ny, int nz, uint64_t seed, int rows) {
    uint64_t s = seed;
    std::vector<char> dom(size_t(nx) * size_t(ny) * size_t(nz));
    for (auto& c : dom) c = (wi::detail::splitmix(s) % 4u) != 0;
    const BitField f = BitField::grid(nx, ny, nz, dom);
    OperatorRows op;
    for (int r = 0; r < rows; ++r) {
        op.d.push_back(detail::uniform(s));
        op.w.push_back(1.0 + detail::uniform(s));
        const int k = 1 + int(wi::detail::splitmix(s) % 8u);
        for (int q = 0; q < k; ++q) {
            const int bit = int(wi::detail::splitmix(s) % uint64_t(f.n_bits()));
            op.entries.push_back({bit, r, float(detail::uniform(s))});
        }
    }
    const BinaryProblem p = assemble(f, {&op}, IsingPrior{0.0, 0.0});
    auto ints = [](const char* name, const std::vector<int>& v) {
        std::printf("%s", name);
        for (int x : v) std::printf(" %d", x);
        std::printf("\n");
    };
    ints("row_ptr", p.row_ptr);
    ints("ray", p.ray);
    std::printf("a");
    for (float x : p.a) std::printf(" %a", double(x));
    std::printf("\n");
    ints("nbr_ptr", p.nbr_ptr);
    ints("nbr", p.nbr);
    ints("var_voxel", p.var_voxel);
    std::printf("adjoint %a\n", adjoint_mismatch(op, f.n_bits()));
    return 0;
}

// --muon nx ny nz seed bins: build_binary_problem over a medium and views of this
// driver's own types with the members of the reference's VoxelMedium / MuonView /
// MuonBinnedData. The grid (voxel 2 m, random lo) holds random mu in a random domain; two
// views, one chamber inside the grid and one outside it, see `bins` directions each
// (random, plus axis-parallel ones) with random open and det counts. Prints the inputs
// and the problem (floats as hex) for a brute-force check.
struct Vec3 { float x, y, z; };
struct Medium {
    Vec3 lo;
    float voxel;
    int nx, ny, nz;
    float mu_rock, mu_air;
    const float* mu;
};
struct Binned {
    std::vector<Vec3> dirs;
    std::vector<unsigned long long> open, det;
    int size() const { return int(dirs.size()); }
    Vec3 bin_direction(int b) const { return dirs[size_t(b)]; }
};
struct View {
    Vec3 chamber;
    const Binned* data;
};

static int muon_dump(int nx, int ny, int nz, uint64_t seed, int bins) {
    uint64_t s = seed;
    auto u = [&] { return detail::uniform(s); };
    Medium m{};
    m.voxel = 2.0f;
    m.lo = Vec3{float(-nx + 4.0 * u()), float(-ny + 4.0 * u()), float(-3.0 * u())};
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    m.mu_rock = 0.046f;
    m.mu_air = 0.00002f;
    std::vector<float> mu(size_t(nx) * size_t(ny) * size_t(nz));
    std::vector<char> dom(mu.size());
    for (size_t c = 0; c < mu.size(); ++c) {
        dom[c] = u() < 0.7;
        mu[c] = dom[c] ? float(0.03 + 0.03 * u()) : m.mu_air;
    }
    m.mu = mu.data();
    std::vector<Binned> data(2);
    std::vector<View> views;
    // off every cell boundary (an origin on one makes the crossed cells ambiguous)
    const Vec3 chambers[2] = {Vec3{m.lo.x + float(nx) * 0.9f + 0.37f, m.lo.y + float(ny) * 1.1f + 0.21f, m.lo.z + 1.3f},
                              Vec3{m.lo.x - 7.5f, m.lo.y + float(ny) * 0.8f + 0.33f, m.lo.z - 2.25f}};
    for (int v = 0; v < 2; ++v) {
        Binned& D = data[size_t(v)];
        for (int b = 0; b < bins; ++b) {
            Vec3 d;
            if (b < 6) {   // axis-parallel directions: the flat-component guard
                d = Vec3{0.0f, 0.0f, 0.0f};
                (b % 2 ? d.x : (b % 3 ? d.y : d.z)) = b < 3 ? 1.0f : -1.0f;
            } else {
                const double th = std::acos(1.0 - 2.0 * u()), az = 6.283185307179586 * u();
                d = Vec3{float(std::sin(th) * std::cos(az)), float(std::sin(th) * std::sin(az)), float(std::cos(th))};
                if (v == 1) d.x = std::fabs(d.x);   // the outside chamber looks into the grid
            }
            D.dirs.push_back(d);
            const unsigned long long open = (unsigned long long)(5 + 200 * u());
            D.open.push_back(open);
            D.det.push_back((unsigned long long)(double(open) * u() * u()));
        }
        views.push_back(View{chambers[v], &D});
    }
    const float a_per_metre = seed % 2 ? std::numeric_limits<float>::quiet_NaN() : -0.021f;
    const BinaryProblem p = build_binary_problem(m, views, dom, 2.0, 6.9, 30.0, a_per_metre);
    std::printf("grid %d %d %d %a %a %a %a %a\n", nx, ny, nz, double(m.lo.x), double(m.lo.y), double(m.lo.z),
                double(m.voxel), double(std::isnan(a_per_metre) ? m.mu_rock : a_per_metre));
    std::printf("mu");
    for (float x : mu) std::printf(" %a", double(x));
    std::printf("\ndomain");
    for (char c : dom) std::printf(" %d", int(c));
    std::printf("\n");
    for (int v = 0; v < 2; ++v) {
        std::printf("view %a %a %a\n", double(chambers[v].x), double(chambers[v].y), double(chambers[v].z));
        std::printf("dirs");
        for (const Vec3& d : data[size_t(v)].dirs) std::printf(" %a %a %a", double(d.x), double(d.y), double(d.z));
        std::printf("\nopen");
        for (auto o : data[size_t(v)].open) std::printf(" %llu", o);
        std::printf("\ndet");
        for (auto o : data[size_t(v)].det) std::printf(" %llu", o);
        std::printf("\n");
    }
    auto ints = [](const char* name, const std::vector<int>& v) {
        std::printf("%s", name);
        for (int x : v) std::printf(" %d", x);
        std::printf("\n");
    };
    ints("row_ptr", p.row_ptr);
    ints("ray", p.ray);
    std::printf("a");
    for (float x : p.a) std::printf(" %a", double(x));
    std::printf("\nd");
    for (double x : p.d) std::printf(" %a", x);
    std::printf("\nw");
    for (double x : p.w) std::printf(" %a", x);
    std::printf("\n");
    ints("var_voxel", p.var_voxel);
    ints("nbr_ptr", p.nbr_ptr);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 8 && std::strcmp(argv[1], "--anneal-int") == 0) {
        // wising::anneal on a rudy file: "energy <tracked> x <0/1 string>"
        std::vector<char> buf;
        wi::WeightedGraph g;
        std::string err;
        if (!wi::read_file(argv[2], buf) || !wi::parse_rudy(buf.data(), buf.size(), g, err)) {
            std::fprintf(stderr, "cannot load %s: %s\n", argv[2], err.c_str());
            return 2;
        }
        wi::Budget b;
        b.t_hot = std::atof(argv[3]);
        b.t_cold = std::atof(argv[4]);
        b.anneal_sweeps = std::atoi(argv[5]);
        b.hold_sweeps = std::atoi(argv[6]);
        const wi::RunOut r = wi::anneal(g, b, std::strtoull(argv[7], nullptr, 10));
        std::printf("energy %lld x ", static_cast<long long>(r.energy_tracked));
        for (uint8_t v : r.x) std::putchar(v ? '1' : '0');
        std::printf("\n");
        return 0;
    }
    if (argc >= 2 && std::strcmp(argv[1], "--rudy-digest") == 0) {
        // one line per file: "refused" or "ok <FNV-1a 64 of the parsed graph>"
        for (int f = 2; f < argc; ++f) {
            std::vector<char> buf;
            wi::WeightedGraph g;
            std::string err;
            if (!wi::read_file(argv[f], buf) || !wi::parse_rudy(buf.data(), buf.size(), g, err)) {
                std::printf("refused\n");
                continue;
            }
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
            std::printf("ok %llu\n", static_cast<unsigned long long>(h));
        }
        return 0;
    }
    if (argc == 7 && std::strcmp(argv[1], "--muon") == 0)
        return muon_dump(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]),
                         std::strtoull(argv[5], nullptr, 10), std::atoi(argv[6]));
    if (argc == 7 && std::strcmp(argv[1], "--assemble") == 0)
        return assemble_dump(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]),
                             std::strtoull(argv[5], nullptr, 10), std::atoi(argv[6]));
    const bool host_only = argc > 1 && std::strcmp(argv[1], "--host") == 0;
    std::printf("qublar drop-in check (%s)\n", host_only ? "host" : "host + GPU");
    host_checks();
    if (!host_only) {
        try {
            gpu_checks();
        } catch (const std::exception& e) {
            line("GPU checks ran", false, e.what());
        }
    }
    std::printf(g_fail ? "\nFAILED: %d\n" : "\nall checks passed\n", g_fail);
    return g_fail ? 1 : 0;
}
