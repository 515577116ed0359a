/*
 * qb_oracle.c -- CPU oracle (test infrastructure only; see qb_oracle.h).
 *
 * Build with -ffp-contract=off and no -ffast-math: every floating-point
 * expression here is evaluated in the order written, so the MI355X engine,
 * compiled under the same rule, can reproduce each decision bit for bit.
 */
#define _GNU_SOURCE
#include "qb_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- splitmix64 (src/ising_recon.hpp:89-95; src/ising_weighted.hpp:236-242) ---- */
uint64_t orc_splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double orc_uniform(uint64_t* s) { return (double)(orc_splitmix(s) >> 11) * 0x1.0p-53; }

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"): 10 rounds of the 4x32 S-box; round i uses key + i*(W0, W1). ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The draw of bit `bit` in sweep `sweep` of the branch whose seed is `seed`:
   key = seed (low word, high word), counter = (bit, sweep, 0, 0); the first two
   output words form a 64-bit integer whose top 53 bits scale into [0, 1). */
double orc_philox_uniform(uint64_t seed, uint32_t sweep, uint32_t bit) {
    const uint32_t ctr[4] = {bit, sweep, 0u, 0u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const uint64_t z = ((uint64_t)o[0] << 32) | (uint64_t)o[1];
    return (double)(z >> 11) * 0x1.0p-53;
}

/* ---- qexp: the engine's exponential, defined by basic IEEE operations only ----
 * exp(x) = 2^k e^r, k = floor(x/ln2 + 1/2), r = (x - k ln2_hi) - k ln2_lo
 * (Cody & Waite reduction with fdlibm's split of ln2: k*ln2_hi is exact for
 * |k| < 2^11), e^r by its Taylor series to degree 13 in Horner form
 * (|r| <= 0.347, truncation < 5e-18 relative), then an exact power-of-two
 * scale. Accuracy is checked against the C library exp in the tests. */
static double pow2i(int k) {
    const uint64_t b = (uint64_t)(k + 1023) << 52;
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

double orc_qexp(double x) {
    if (!(x > -745.0)) return 0.0;
    if (x >= 709.0) return HUGE_VAL;
    const double kd = floor(x * 0x1.71547652b82fep+0 + 0.5);
    const int k = (int)kd;
    const double r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
    double q = 0x1.6124613a86d09p-33;          /* 1/13! */
    q = q * r + 0x1.1eed8eff8d898p-29;         /* 1/12! */
    q = q * r + 0x1.ae64567f544e4p-26;         /* 1/11! */
    q = q * r + 0x1.27e4fb7789f5cp-22;         /* 1/10! */
    q = q * r + 0x1.71de3a556c734p-19;         /* 1/9!  */
    q = q * r + 0x1.a01a01a01a01ap-16;         /* 1/8!  */
    q = q * r + 0x1.a01a01a01a01ap-13;         /* 1/7!  */
    q = q * r + 0x1.6c16c16c16c17p-10;         /* 1/6!  */
    q = q * r + 0x1.1111111111111p-7;          /* 1/5!  */
    q = q * r + 0x1.5555555555555p-5;          /* 1/4!  */
    q = q * r + 0x1.5555555555555p-3;          /* 1/3!  */
    q = q * r + 0.5;
    q = q * r + 1.0;
    q = q * r + 1.0;
    if (k >= -1021) return q * pow2i(k);
    return (q * pow2i(k + 600)) * pow2i(-600);
}

/* Schedule (src/ising_recon.hpp:129-131; src/ising_weighted.hpp:255-257). */
double orc_temperature(const orc_schedule* s, int sweep) {
    const int denom = s->anneal_sweeps - 1 > 1 ? s->anneal_sweeps - 1 : 1;
    double frac = (double)sweep / (double)denom;
    if (frac > 1.0) frac = 1.0;
    return sweep < s->anneal_sweeps ? s->t_hot * pow(s->t_cold / s->t_hot, frac) : s->t_cold;
}

/* ---- float path ---------------------------------------------------------------- */

static double lin_of(const orc_problem* p, int i) { return p->field ? p->field[i] : p->kappa; }

/* residuals from scratch in bit-major order (the loop of binary_energy, :66-71) */
static void residuals_scratch(const orc_problem* p, const uint8_t* x, double* r) {
    for (int b = 0; b < p->n_rays; ++b) r[b] = -p->d[b];
    for (int i = 0; i < p->n_vars; ++i)
        if (x[i])
            for (int k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k) r[p->ray[k]] += (double)p->a[k];
}

void orc_energy_parts(const orc_problem* p, const uint8_t* x, double* data_out, double* prior_out) {
    double* r = (double*)malloc(sizeof(double) * (size_t)(p->n_rays > 0 ? p->n_rays : 1));
    for (int b = 0; b < p->n_rays; ++b) r[b] = -p->d[b];
    double prior = 0.0;
    for (int i = 0; i < p->n_vars; ++i) {
        if (x[i])
            for (int k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k) r[p->ray[k]] += (double)p->a[k];
        prior += lin_of(p, i) * (double)x[i];
        for (int k = p->nbr_ptr[i]; k < p->nbr_ptr[i + 1]; ++k)
            if (p->nbr[k] > i && x[p->nbr[k]] != x[i]) prior += p->lambda;
    }
    double data = 0.0;
    for (int b = 0; b < p->n_rays; ++b) data += 0.5 * p->w[b] * r[b] * r[b];
    free(r);
    *data_out = data;
    *prior_out = prior;
}

double orc_binary_energy(const orc_problem* p, const uint8_t* x) {
    double data, prior;
    orc_energy_parts(p, x, &data, &prior);
    return data + prior;
}

/* local_field_dE (src/ising_recon.hpp:105-118), on FP64 residuals. */
double orc_local_field_dE(const orc_problem* p, const uint8_t* x, const double* r, int i) {
    const double delta = x[i] ? -1.0 : 1.0;
    double dE = lin_of(p, i) * delta;
    for (int k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k) {
        const double a = (double)p->a[k] * delta;
        dE += p->w[p->ray[k]] * (r[p->ray[k]] * a + 0.5 * a * a);
    }
    for (int k = p->nbr_ptr[i]; k < p->nbr_ptr[i + 1]; ++k) {
        const int before = x[p->nbr[k]] != x[i];
        dE += p->lambda * (before ? -1.0 : 1.0);
    }
    return dE;
}

/* The data term of entry k: w (r a + a^2 / 2) with a = a_k delta (ising_recon.hpp:110-111);
   r64 or r32 holds the carried residuals. */
static double term(const orc_problem* p, const double* r64, const float* r32, int k, double delta) {
    const double a = (double)p->a[k] * delta;
    const double rv = r32 ? (double)r32[p->ray[k]] : r64[p->ray[k]];
    return p->w[p->ray[k]] * (rv * a + 0.5 * a * a);
}

/* dE under a mode: the entries summed in one chain (segment == 0, the reference's
   order) or in chains of `segment` entries, the first seeded with lin * delta; the
   chain sums c_j are added in runs of ORC_CHAIN_GROUP, G_g = ((c_8g + c_8g+1) + ...),
   and the runs in order, dE = ((G_0 + G_1) + ...) -- with at most ORC_CHAIN_GROUP
   chains, the chain sums added in order. Then the neighbour terms in CSR order. */
#define ORC_CHAIN_GROUP 8
static double mode_dE(const orc_problem* p, const uint8_t* x, const double* r64, const float* r32, int i,
                      int segment) {
    const double delta = x[i] ? -1.0 : 1.0;
    const int k0 = p->row_ptr[i], k1 = p->row_ptr[i + 1];
    double dE = lin_of(p, i) * delta;
    if (segment <= 0 || k1 - k0 <= segment) {
        for (int k = k0; k < k1; ++k) dE += term(p, r64, r32, k, delta);
    } else {
        const int n_chain = (k1 - k0 + segment - 1) / segment;
        double g = 0.0;
        for (int j = 0; j < n_chain; ++j) {
            const int s0 = k0 + j * segment;
            const int s1 = s0 + segment < k1 ? s0 + segment : k1;
            double c;
            if (j == 0) {
                c = dE;
                for (int k = s0; k < s1; ++k) c += term(p, r64, r32, k, delta);
            } else {
                c = term(p, r64, r32, s0, delta);
                for (int k = s0 + 1; k < s1; ++k) c += term(p, r64, r32, k, delta);
            }
            g = j % ORC_CHAIN_GROUP == 0 ? c : g + c;
            if (j % ORC_CHAIN_GROUP == ORC_CHAIN_GROUP - 1 || j == n_chain - 1)
                dE = j < ORC_CHAIN_GROUP ? g : dE + g;
        }
    }
    for (int k = p->nbr_ptr[i]; k < p->nbr_ptr[i + 1]; ++k) {
        const int before = x[p->nbr[k]] != x[i];
        dE += p->lambda * (before ? -1.0 : 1.0);
    }
    return dE;
}

/* anneal_branch (src/ising_recon.hpp:121-144) under mode m. */
void orc_anneal_branch(const orc_problem* p, const orc_schedule* s, uint64_t seed,
                       const orc_mode* m, uint8_t* x, double* trace, double* r_final) {
    const int n = p->n_vars, nr = p->n_rays;
    memset(x, 0, (size_t)n);
    double* r64 = (double*)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
    float* r32 = m->precision == ORC_FP32 ? (float*)malloc(sizeof(float) * (size_t)(nr > 0 ? nr : 1)) : NULL;
    for (int b = 0; b < nr; ++b) r64[b] = -p->d[b];
    if (r32)
        for (int b = 0; b < nr; ++b) r32[b] = (float)r64[b];
    uint64_t state = seed * 0x2545F4914F6CDD1Dull + 1;
    const int total = s->anneal_sweeps + s->hold_sweeps;
    const int n_steps = m->order ? m->class_ptr[m->n_classes] : n;
    for (int sweep = 0; sweep < total; ++sweep) {
        const double temp = orc_temperature(s, sweep);
        const double beta = 1.0 / temp;
        for (int j = 0; j < n_steps; ++j) {
            const int i = m->order ? m->order[j] : j;
            const double delta = x[i] ? -1.0 : 1.0;
            const double dE = mode_dE(p, x, r64, r32, i, m->segment);
            int accept;
            if (m->rng == ORC_RNG_REFERENCE)
                accept = dE <= 0.0 || orc_uniform(&state) < exp(-dE / temp);
            else
                accept = dE <= 0.0 ||
                         orc_philox_uniform(seed, (uint32_t)sweep, (uint32_t)i) < orc_qexp(-dE * beta);
            if (accept) {
                x[i] ^= 1u;
                if (r32) {
                    for (int k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k) {
                        const float ad = delta > 0.0 ? p->a[k] : -p->a[k];
                        r32[p->ray[k]] = r32[p->ray[k]] + ad;
                    }
                } else {
                    for (int k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k)
                        r64[p->ray[k]] += (double)p->a[k] * delta;
                }
            }
        }
        if (m->recompute_every > 0 && (sweep + 1) % m->recompute_every == 0) {
            residuals_scratch(p, x, r64);
            if (r32)
                for (int b = 0; b < nr; ++b) r32[b] = (float)r64[b];
        }
        if (trace) trace[sweep] = orc_binary_energy(p, x);
    }
    if (r_final)
        for (int b = 0; b < nr; ++b) r_final[b] = r32 ? (double)r32[b] : r64[b];
    free(r64);
    free(r32);
}

typedef struct {
    const orc_problem* p;
    const orc_schedule* s;
    const orc_mode* m;
    int n_branches, n_threads, t;
    uint64_t seed_offset;
    uint8_t* branches;
} bf_job;

static void* bf_worker(void* arg) {
    bf_job* j = (bf_job*)arg;
    for (int b = j->t; b < j->n_branches; b += j->n_threads)
        orc_anneal_branch(j->p, j->s, (uint64_t)b + 1 + j->seed_offset, j->m,
                          j->branches + (size_t)b * (size_t)j->p->n_vars, NULL, NULL);
    return NULL;
}

void orc_branch_fractions(const orc_problem* p, const orc_schedule* s, int n_branches, int n_threads,
                          uint64_t seed_offset, const orc_mode* m, uint8_t* keep, float* p_out) {
    const size_t n = (size_t)p->n_vars;
    uint8_t* br = keep ? keep : (uint8_t*)malloc(n * (size_t)n_branches);
    if (n_threads < 1) n_threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)n_threads);
    bf_job* jobs = (bf_job*)malloc(sizeof(bf_job) * (size_t)n_threads);
    for (int t = 0; t < n_threads; ++t) {
        jobs[t] = (bf_job){p, s, m, n_branches, n_threads, t, seed_offset, br};
        pthread_create(&th[t], NULL, bf_worker, &jobs[t]);
    }
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    if (p_out) {
        for (size_t i = 0; i < n; ++i) p_out[i] = 0.0f;
        for (int b = 0; b < n_branches; ++b)
            for (size_t i = 0; i < n; ++i) p_out[i] += (float)br[(size_t)b * n + i];
        for (size_t i = 0; i < n; ++i) p_out[i] /= (float)n_branches;
    }
    if (!keep) free(br);
    free(th);
    free(jobs);
}

/* Bit::Exists = 0, NotThere = 1, Undecided = 2 (src/ising_recon.hpp:302-309). */
int orc_classify(float p, float lo, float hi) {
    if (p >= hi) return 1;
    if (p <= lo) return 0;
    return 2;
}

/* The data part of freeze_provable_rock's bound (src/ising_recon.hpp:189-192): each
   entry's w a (a / 2 - d), summed in CSR order. */
static void freeze_data_bounds(const orc_problem* p, double* lb) {
    for (int i = 0; i < p->n_vars; ++i) {
        lb[i] = 0.0;
        for (int k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k)
            lb[i] += p->w[p->ray[k]] * (double)p->a[k] * (0.5 * (double)p->a[k] - p->d[p->ray[k]]);
    }
}

/* Bit i's bound given the frozen set (src/ising_recon.hpp:198-202): data + lin +
   lambda (rock - free) over its neighbours. */
static double freeze_bound_of(const orc_problem* p, const double* lb, const uint8_t* frozen, int i) {
    int rock = 0, fr = 0;
    for (int k = p->nbr_ptr[i]; k < p->nbr_ptr[i + 1]; ++k) {
        if (frozen[p->nbr[k]]) ++rock; else ++fr;
    }
    return lb[i] + lin_of(p, i) + p->lambda * (double)(rock - fr);
}

/* freeze_provable_rock's fixpoint (src/ising_recon.hpp:187-205). */
int orc_freeze_mask(const orc_problem* p, double margin, uint8_t* frozen) {
    const int n = p->n_vars;
    double* lb = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    freeze_data_bounds(p, lb);
    memset(frozen, 0, (size_t)n);
    int rounds = 0;
    for (int changed = 1; changed; ++rounds) {
        changed = 0;
        for (int i = 0; i < n; ++i) {
            if (frozen[i]) continue;
            if (freeze_bound_of(p, lb, frozen, i) >= margin) { frozen[i] = 1; changed = 1; }
        }
    }
    free(lb);
    return rounds;
}

/* Every bit's bound with `frozen` held fixed (the quantity the rounds above compare with
   the margin), for checking it against a brute-force minimum of dE(0 -> 1). */
void orc_freeze_bounds(const orc_problem* p, const uint8_t* frozen, double* bound) {
    const int n = p->n_vars;
    double* lb = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    freeze_data_bounds(p, lb);
    for (int i = 0; i < n; ++i) bound[i] = freeze_bound_of(p, lb, frozen, i);
    free(lb);
}

/* roi_qubo (src/ising_recon.hpp:251-300). Pairs sharing a ray are visited ray by ray,
   each ray's members in the order R's variables list them. */
void orc_roi_qubo(const orc_problem* p, const int* roi, int n, const uint8_t* x_bar, double* h,
                  double* J) {
    int* pos = (int*)malloc(sizeof(int) * (size_t)(p->n_vars > 0 ? p->n_vars : 1));
    for (int v = 0; v < p->n_vars; ++v) pos[v] = -1;
    for (int i = 0; i < n; ++i) pos[roi[i]] = i;
    for (int i = 0; i < n; ++i) h[i] = 0.0;
    for (size_t q = 0; q < (size_t)n * (size_t)n; ++q) J[q] = 0.0;
    double* r = (double*)malloc(sizeof(double) * (size_t)(p->n_rays > 0 ? p->n_rays : 1));
    for (int b = 0; b < p->n_rays; ++b) r[b] = -p->d[b];
    for (int v = 0; v < p->n_vars; ++v)
        if (x_bar[v] && pos[v] < 0)
            for (int k = p->row_ptr[v]; k < p->row_ptr[v + 1]; ++k) r[p->ray[k]] += (double)p->a[k];
    /* per-ray member lists, in insertion order */
    int* cnt = (int*)calloc((size_t)(p->n_rays + 1), sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int k = p->row_ptr[roi[i]]; k < p->row_ptr[roi[i] + 1]; ++k) cnt[p->ray[k] + 1]++;
    for (int b = 0; b < p->n_rays; ++b) cnt[b + 1] += cnt[b];
    const int tot = cnt[p->n_rays];
    int* mem_i = (int*)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
    double* mem_a = (double*)malloc(sizeof(double) * (size_t)(tot > 0 ? tot : 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(p->n_rays > 0 ? p->n_rays : 1));
    for (int b = 0; b < p->n_rays; ++b) fill[b] = cnt[b];
    for (int i = 0; i < n; ++i)
        for (int k = p->row_ptr[roi[i]]; k < p->row_ptr[roi[i] + 1]; ++k) {
            const int b = p->ray[k];
            const double a = (double)p->a[k];
            h[i] += p->w[b] * (r[b] * a + 0.5 * a * a);
            mem_i[fill[b]] = i;
            mem_a[fill[b]] = a;
            fill[b]++;
        }
    for (int b = 0; b < p->n_rays; ++b)
        for (int u = cnt[b]; u < cnt[b + 1]; ++u)
            for (int v = u + 1; v < cnt[b + 1]; ++v) {
                int i = mem_i[u], j = mem_i[v];
                if (i > j) { const int t = i; i = j; j = t; }
                J[(size_t)i * (size_t)n + (size_t)j] += p->w[b] * mem_a[u] * mem_a[v];
            }
    for (int i = 0; i < n; ++i) {
        const int vi = roi[i];
        h[i] += lin_of(p, vi);
        for (int k = p->nbr_ptr[vi]; k < p->nbr_ptr[vi + 1]; ++k) {
            const int vj = p->nbr[k];
            if (pos[vj] >= 0) {
                if (pos[vj] > i) {
                    h[i] += p->lambda;
                    h[pos[vj]] += p->lambda;
                    J[(size_t)i * (size_t)n + (size_t)pos[vj]] -= 2.0 * p->lambda;
                }
            } else {
                h[i] += p->lambda * (x_bar[vj] ? -1.0 : 1.0);
            }
        }
    }
    free(pos); free(r); free(cnt); free(mem_i); free(mem_a); free(fill);
}

/* ---- integer path (src/ising_weighted.hpp) ---------------------------------------- */

void orc_graph_csr(int n, int64_t m, const int32_t* eu, const int32_t* ev, const int32_t* ew,
                   int64_t* row, int32_t* col, int32_t* w) {
    for (int i = 0; i <= n; ++i) row[i] = 0;
    for (int64_t e = 0; e < m; ++e) { row[eu[e] + 1]++; row[ev[e] + 1]++; }
    for (int i = 0; i < n; ++i) row[i + 1] += row[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) fill[i] = row[i];
    for (int64_t e = 0; e < m; ++e) {
        const int32_t a = eu[e], b = ev[e], wt = ew[e];
        col[fill[a]] = b; w[fill[a]++] = wt;
        col[fill[b]] = a; w[fill[b]++] = wt;
    }
    free(fill);
}

/* tokeniser of parse_rudy (ising_weighted.hpp:50-80) */
typedef struct { const char* p; const char* end; } cursor;
static int is_blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }
static void skip_blank(cursor* c) { while (c->p < c->end && is_blank(*c->p)) ++c->p; }
static void skip_ws(cursor* c) { while (c->p < c->end && (is_blank(*c->p) || *c->p == '\n')) ++c->p; }
static int at_eol(cursor* c) { skip_blank(c); return c->p >= c->end || *c->p == '\n'; }
static int token(cursor* c, const char** s, size_t* len) {
    skip_blank(c);
    if (c->p >= c->end || *c->p == '\n') return 0;
    const char* b = c->p;
    while (c->p < c->end && !is_blank(*c->p) && *c->p != '\n') ++c->p;
    *s = b;
    *len = (size_t)(c->p - b);
    return 1;
}
static int parse_int(const char* t, size_t len, int64_t* v) {
    if (len == 0) return 0;
    size_t i = 0;
    int neg = 0;
    if (t[0] == '-' || t[0] == '+') { neg = t[0] == '-'; i = 1; }
    if (i >= len) return 0;
    int64_t acc = 0;
    for (; i < len; ++i) {
        if (t[i] < '0' || t[i] > '9') return 0;
        acc = acc * 10 + (t[i] - '0');
        if (acc > ((int64_t)1 << 40)) return 0;
    }
    *v = neg ? -acc : acc;
    return 1;
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

#define FAIL(msg) do { snprintf(err, err_cap, "%s", msg); return 1; } while (0)

int orc_parse_rudy(const char* data, size_t size, int* n_out, int64_t* m_out, int32_t* eu, int32_t* ev,
                   int32_t* ew, int64_t cap, char* err, size_t err_cap) {
    cursor c = {data, data + size};
    const char* t;
    size_t tl;
    int64_t n = 0, m = 0;
    skip_ws(&c);
    if (!token(&c, &t, &tl) || !parse_int(t, tl, &n) || n <= 0 || n > ((int64_t)1 << 30)) FAIL("bad header: n");
    if (!token(&c, &t, &tl) || !parse_int(t, tl, &m) || m < 0) FAIL("bad header: m");
    *n_out = (int)n;
    *m_out = m;
    int64_t read = 0;
    for (;;) {
        skip_ws(&c);
        if (c.p >= c.end) break;
        int64_t u = 0, v = 0, w = 1;
        if (!token(&c, &t, &tl) || !parse_int(t, tl, &u)) FAIL("bad vertex token");
        if (!token(&c, &t, &tl) || !parse_int(t, tl, &v)) FAIL("bad vertex token");
        if (token(&c, &t, &tl)) {
            if (!parse_int(t, tl, &w)) FAIL("weight is not an integer; the float path is BinaryProblem");
            if (w < INT32_MIN || w > INT32_MAX) FAIL("weight out of int32 range");
            if (!at_eol(&c)) FAIL("trailing token on an edge line");
        }
        if (u < 1 || u > n || v < 1 || v > n) FAIL("vertex index out of range");
        if (u == v) FAIL("self-loop");
        if (read < cap) {
            eu[read] = (int32_t)(u - 1);
            ev[read] = (int32_t)(v - 1);
            ew[read] = (int32_t)w;
        }
        ++read;
    }
    if (read != m) FAIL("edge count differs from the header");
    if (m > cap) FAIL("caller buffer too small");
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(m > 0 ? m : 1));
    for (int64_t e = 0; e < m; ++e) {
        const uint64_t a = (uint64_t)(eu[e] < ev[e] ? eu[e] : ev[e]);
        const uint64_t b = (uint64_t)(eu[e] < ev[e] ? ev[e] : eu[e]);
        key[e] = (a << 32) | b;
    }
    qsort(key, (size_t)m, sizeof(uint64_t), cmp_u64);
    for (int64_t k = 1; k < m; ++k)
        if (key[k] == key[k - 1]) { free(key); FAIL("duplicate edge"); }
    free(key);
    return 0;
}

int64_t orc_cut_of(const orc_graph* g, const uint8_t* x) {
    int64_t cut = 0;
    for (int64_t e = 0; e < g->m; ++e)
        if (x[g->eu[e]] != x[g->ev[e]]) cut += g->ew[e];
    return cut;
}

int64_t orc_energy_scratch(const orc_graph* g, const uint8_t* x) {
    int64_t cut = 0;
    for (int i = 0; i < g->n; ++i)
        for (int64_t k = g->row[i]; k < g->row[i + 1]; ++k) {
            const int j = g->col[k];
            if (j > i && x[i] != x[j]) cut += g->w[k];
        }
    return -cut;
}

/* State + anneal (ising_weighted.hpp:193-267): s = +-1, f_i = sum_j w_ij s_j,
   dE_i = -s_i f_i, a flip scatters f_j -= 2 w_ij s_i. */
void orc_anneal_int(const orc_graph* g, const orc_schedule* b, uint64_t seed, const orc_mode* m,
                    uint8_t* x_out, int64_t* energy_tracked) {
    const int n = g->n;
    int8_t* s = (int8_t*)malloc((size_t)(n > 0 ? n : 1));
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        s[i] = -1;
        int64_t acc = 0;
        for (int64_t k = g->row[i]; k < g->row[i + 1]; ++k) acc -= g->w[k];
        f[i] = acc;
    }
    int64_t energy = 0;
    uint64_t state = seed * 0x2545F4914F6CDD1Dull + 1;
    const int total = b->anneal_sweeps + b->hold_sweeps;
    const int n_steps = m->order ? m->class_ptr[m->n_classes] : n;
    for (int sweep = 0; sweep < total; ++sweep) {
        const double temp = orc_temperature(b, sweep);
        const double beta = 1.0 / temp;
        for (int j = 0; j < n_steps; ++j) {
            const int i = m->order ? m->order[j] : j;
            const int64_t dEi = -(int64_t)s[i] * f[i];
            const double dE = (double)dEi;
            int accept;
            if (m->rng == ORC_RNG_REFERENCE)
                accept = dE <= 0.0 || orc_uniform(&state) < exp(-dE / temp);
            else
                accept = dE <= 0.0 ||
                         orc_philox_uniform(seed, (uint32_t)sweep, (uint32_t)i) < orc_qexp(-dE * beta);
            if (accept) {
                const int64_t si = s[i];
                for (int64_t k = g->row[i]; k < g->row[i + 1]; ++k) f[g->col[k]] -= 2 * (int64_t)g->w[k] * si;
                s[i] = (int8_t)(-si);
                energy += dEi;
            }
        }
    }
    for (int i = 0; i < n; ++i) x_out[i] = s[i] > 0 ? 1u : 0u;
    *energy_tracked = energy;
    free(s);
    free(f);
}

double orc_tts99(double t_mean, double p) {
    if (p <= 0.0) return HUGE_VAL;
    if (p >= 0.99) return t_mean;
    return t_mean * log(0.01) / log(1.0 - p);
}

static int cmp_double(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

void orc_tts99_bootstrap(const int* success, const double* time, size_t r, int resamples, uint64_t seed,
                         double* lo, double* hi) {
    double* v = (double*)malloc(sizeof(double) * (size_t)resamples);
    uint64_t st = seed;
    for (int b = 0; b < resamples; ++b) {
        int s = 0;
        double t = 0.0;
        for (size_t k = 0; k < r; ++k) {
            const size_t j = (size_t)(orc_splitmix(&st) % (uint64_t)r);
            s += success[j];
            t += time[j];
        }
        v[b] = orc_tts99(t / (double)r, (double)s / (double)r);
    }
    qsort(v, (size_t)resamples, sizeof(double), cmp_double);
    double out[2];
    const double qs[2] = {0.025, 0.975};
    for (int q = 0; q < 2; ++q) {
        const double pos = qs[q] * (double)(resamples - 1);
        const size_t a = (size_t)floor(pos), c = (size_t)ceil(pos);
        if (isinf(v[a]) || isinf(v[c])) out[q] = HUGE_VAL;
        else out[q] = v[a] + (v[c] - v[a]) * (pos - (double)a);
    }
    *lo = out[0];
    *hi = out[1];
    free(v);
}

/* ---- SHA-256 (FIPS 180-4) ---- */
static uint32_t rotr32(uint32_t x, int k) { return (x >> k) | (x << (32 - k)); }

static void sha_block(uint32_t h[8], const uint8_t* blk) {
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
        w[i] = (uint32_t)blk[4 * i] << 24 | (uint32_t)blk[4 * i + 1] << 16 | (uint32_t)blk[4 * i + 2] << 8 |
               (uint32_t)blk[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
        const uint32_t s0 = rotr32(w[i - 15], 7) ^ rotr32(w[i - 15], 18) ^ (w[i - 15] >> 3);
        const uint32_t s1 = rotr32(w[i - 2], 17) ^ rotr32(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
        const uint32_t t1 = hh + (rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25)) + ((e & f) ^ (~e & g)) + K[i] + w[i];
        const uint32_t t2 = (rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

void orc_sha256_hex(const void* data, size_t n, char out[65]) {
    uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                     0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    const uint8_t* p = (const uint8_t*)data;
    size_t full = n / 64;
    for (size_t b = 0; b < full; ++b) sha_block(h, p + 64 * b);
    uint8_t tail[128];
    const size_t rem = n - 64 * full;
    memset(tail, 0, sizeof tail);
    memcpy(tail, p + 64 * full, rem);
    tail[rem] = 0x80;
    const size_t tl = rem < 56 ? 64 : 128;
    const uint64_t bits = (uint64_t)n * 8u;
    for (int i = 0; i < 8; ++i) tail[tl - 1 - i] = (uint8_t)(bits >> (8 * i));
    sha_block(h, tail);
    if (tl == 128) sha_block(h, tail + 64);
    static const char* hexd = "0123456789abcdef";
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 8; ++k) out[8 * i + k] = hexd[(h[i] >> (28 - 4 * k)) & 15u];
    out[64] = 0;
}

/* ---- batched sparse QUBOs -------------------------------------------------------------
   One instance E(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j (J listed in both rows),
   replica r annealed from all-zero x, classes visited in the given order;
   dE = (1 - 2 x_i)(h_i + sum_j J_ij x_j) summed in FP64 over the row in CSR order; accept
   iff dE <= 0 or u < qexp(-dE * beta), u = the (r mod 2)-th uniform of the pair drawn by
   one Philox4x32-10 call with key = seed and counter (bit, sweep, r / 2, 0x51B0B47C):
   the top 53 bits of (o0 << 32 | o1), then of (o2 << 32 | o3). No reference counterpart
   beyond anneal_branch's test and schedule (src/ising_recon.hpp:121-144); the batch
   engine's trajectory is defined here. */
static double qubo_draw(uint64_t seed, uint32_t sweep, uint32_t bit, int r) {
    const uint32_t ctr[4] = {bit, sweep, (uint32_t)(r / 2), 0x51B0B47Cu};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const uint64_t z = r % 2 == 0 ? (((uint64_t)o[0] << 32) | (uint64_t)o[1]) : (((uint64_t)o[2] << 32) | (uint64_t)o[3]);
    return (double)(z >> 11) * 0x1.0p-53;
}

double orc_qubo_energy(int n, const int* row_ptr, const int* col, const float* J, const float* h,
                       const uint8_t* x) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!x[i]) continue;
        double pair = 0.0;
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
            if (x[col[k]]) pair += (double)J[k];
        e += (double)h[i] + 0.5 * pair;
    }
    return e;
}

void orc_qubo_anneal(int n, const int* row_ptr, const int* col, const float* J, const float* h, const int* order,
                     const orc_schedule* s, uint64_t seed, int R, uint8_t* x_out) {
    const int total = s->anneal_sweeps + s->hold_sweeps;
    for (int r = 0; r < R; ++r) {
        uint8_t* x = x_out + (size_t)r * (size_t)n;
        memset(x, 0, (size_t)n);
        for (int sweep = 0; sweep < total; ++sweep) {
            const double beta = 1.0 / orc_temperature(s, sweep);
            for (int t = 0; t < n; ++t) {
                const int i = order[t];
                double acc = (double)h[i];
                for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
                    if (x[col[k]]) acc += (double)J[k];
                const double dE = x[i] ? -acc : acc;
                if (dE <= 0.0 || qubo_draw(seed, (uint32_t)sweep, (uint32_t)i, r) < orc_qexp(-dE * beta))
                    x[i] ^= 1u;
            }
        }
    }
}

/* ---- dense QUBO field engine -------------------------------------------------------------
   qlog: natural log from IEEE operations only (the engine's qlog, restated): x = m 2^e with
   m in [sqrt(1/2), sqrt(2)), s = (m - 1)/(m + 1), ln m = 2 s P(s^2) with P the series
   1 + z/3 + ... + z^11/23 in Horner form, then e ln2_hi + (e ln2_lo + ln m). */
double orc_qlog(double x) {
    if (!(x > 0.0)) return -HUGE_VAL;
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    int e = (int)((b >> 52) & 0x7ffu) - 1023;
    const uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &mb, sizeof m);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    static const double inv[12] = {1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0,
                                   1.0 / 11.0, 1.0 / 9.0, 1.0 / 7.0, 1.0 / 5.0, 1.0 / 3.0, 1.0};
    double p = inv[0];
    for (int k = 1; k < 12; ++k) p = p * z + inv[k];
    const double lnm = 2.0 * s * p;
    const double ed = (double)e;
    return ed * 0x1.62e42fee00000p-1 + (ed * 0x1.a39ef35793c76p-33 + lnm);
}

/* E(x) = sum_i x_i (h_i + 1/2 sum_j Q_ij x_j), Q dense row-major. */
double orc_dense_energy(int n, const float* Q, const float* h, const uint8_t* x) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!x[i]) continue;
        double pair = 0.0;
        for (int j = 0; j < n; ++j)
            if (x[j]) pair += (double)Q[(size_t)i * (size_t)n + (size_t)j];
        e += (double)h[i] + 0.5 * pair;
    }
    return e;
}

/* One replica of the dense engine: bits in index order every sweep; fields f = h + Q x
   carried incrementally in FP64 (precision 0) or FP32 (precision 1: f_j = f_j + Q_ij * delta
   in float on every flip of bit i, and f_j = float(h_j + sum_k x_k Q_kj) summed in FP64,
   k ascending, after every recompute_every-th sweep). Accept iff dE = (1 - 2 x_i) f_i <= 0
   or dE * beta < -qlog(u), u = philox_uniform(seed, sweep, i). f_final (nullable) receives
   the carried fields at the end. */
void orc_dense_anneal(int n, const float* Q, const float* h, const orc_schedule* s, uint64_t seed, int precision,
                      int recompute_every, uint8_t* x, double* f_final) {
    double* f64 = (double*)malloc(sizeof(double) * (size_t)n);
    float* f32 = (float*)malloc(sizeof(float) * (size_t)n);
    memset(x, 0, (size_t)n);
    for (int i = 0; i < n; ++i) {
        f64[i] = (double)h[i];
        f32[i] = h[i];
    }
    const int total = s->anneal_sweeps + s->hold_sweeps;
    for (int sweep = 0; sweep < total; ++sweep) {
        const double beta = 1.0 / orc_temperature(s, sweep);
        for (int i = 0; i < n; ++i) {
            const double fi = precision ? (double)f32[i] : f64[i];
            const double dE = x[i] ? -fi : fi;
            const double L = -orc_qlog(orc_philox_uniform(seed, (uint32_t)sweep, (uint32_t)i));
            if (dE <= 0.0 || dE * beta < L) {
                const float dl = x[i] ? -1.0f : 1.0f;
                x[i] ^= 1u;
                const float* row = Q + (size_t)i * (size_t)n;
                if (precision)
                    for (int j = 0; j < n; ++j) f32[j] = f32[j] + row[j] * dl;
                else
                    for (int j = 0; j < n; ++j) f64[j] = f64[j] + (double)row[j] * (double)dl;
            }
        }
        if (precision && recompute_every > 0 && (sweep + 1) % recompute_every == 0)
            for (int i = 0; i < n; ++i) {
                double acc = (double)h[i];
                for (int k = 0; k < n; ++k)
                    if (x[k]) acc += (double)Q[(size_t)k * (size_t)n + (size_t)i];
                f32[i] = (float)acc;
            }
    }
    if (f_final)
        for (int i = 0; i < n; ++i) f_final[i] = precision ? (double)f32[i] : f64[i];
    free(f64);
    free(f32);
}

/* ---- exact ROI posterior ---------------------------------------------------------------
   tools/branches_blaze.py:45-53: every state's energy from scratch, p = exp(-(E - E_min))
   normalised (beta = 1 there), marginals sum_x p(x) x_i (:80-81). */
static double roi_energy(int n, const double* h, const double* J, uint64_t s) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!((s >> i) & 1u)) continue;
        e += h[i];
        for (int j = i + 1; j < n; ++j)
            if ((s >> j) & 1u) e += J[(size_t)i * (size_t)n + (size_t)j];
    }
    return e;
}

double orc_roi_enumerate(int n, const double* h, const double* J, double beta, double* marginal, double* log_z,
                         uint64_t* map_index) {
    const uint64_t states = (uint64_t)1 << n;
    double e_min = HUGE_VAL;
    uint64_t s_min = 0;
    for (uint64_t s = 0; s < states; ++s) {
        const double e = roi_energy(n, h, J, s);
        if (e < e_min) {
            e_min = e;
            s_min = s;
        }
    }
    /* Kahan-compensated sums (Kahan 1965, "Further remarks on reducing truncation
       errors"): over 2^n terms a plain running sum would lose up to 2^n ulp */
    double z = 0.0, zc = 0.0;
    double mc[64];
    for (int i = 0; i < n; ++i) marginal[i] = mc[i] = 0.0;
    for (uint64_t s = 0; s < states; ++s) {
        const double w = exp(-beta * (roi_energy(n, h, J, s) - e_min));
        const double y = w - zc, t = z + y;
        zc = (t - z) - y;
        z = t;
        for (int i = 0; i < n; ++i)
            if ((s >> i) & 1u) {
                const double yi = w - mc[i], ti = marginal[i] + yi;
                mc[i] = (ti - marginal[i]) - yi;
                marginal[i] = ti;
            }
    }
    for (int i = 0; i < n; ++i) marginal[i] /= z;
    *log_z = -beta * e_min + log(z);
    *map_index = s_min;
    return e_min;
}


/* ---- problem builders ------------------------------------------------------------------ */

uint32_t orc_hash_u32(uint32_t x) {   /* src/lidar.cuh:90-97 */
    x = (x ^ 61u) ^ (x >> 16);
    x *= 9u;
    x = x ^ (x >> 4);
    x *= 0x27d4eb2du;
    x = x ^ (x >> 15);
    return x;
}

float orc_hash_unit(uint32_t x) { return (float)(orc_hash_u32(x) >> 8) * (1.0f / 16777216.0f); }

static int medium_index(const orc_medium* m, int i, int j, int k) { return (k * m->ny + j) * m->nx + i; }

/* src/muon.cuh:66-134, every operation in float as written there */
float orc_march(const orc_medium* m, const float p[3], const float d[3], int* cells, float* segs, int cap,
                int* n_cells) {
    const float eps = 1e-12f;
    float t0 = -1e30f, t1 = 1e30f;
    const int nc[3] = {m->nx, m->ny, m->nz};
    int count = 0;
    *n_cells = 0;
    for (int a = 0; a < 3; ++a) {
        if (fabsf(d[a]) < eps) {
            if (p[a] < m->lo[a] || p[a] > m->lo[a] + (float)nc[a] * m->voxel) return 0.0f;
        } else {
            const float ta = (m->lo[a] - p[a]) / d[a];
            const float tb = (m->lo[a] + (float)nc[a] * m->voxel - p[a]) / d[a];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        }
    }
    if (t1 <= fmaxf(t0, 0.0f)) return 0.0f;
    const float t_start = fmaxf(t0, 0.0f);
    const float t_end = t1 - t_start;
    float q[3];
    for (int a = 0; a < 3; ++a) q[a] = p[a] + d[a] * t_start;
    int vox[3];
    float tmax[3], tdelta[3];
    for (int a = 0; a < 3; ++a) vox[a] = (int)floorf((q[a] - m->lo[a]) / m->voxel);
    for (int a = 0; a < 3; ++a) {
        if (vox[a] < 0) vox[a] = 0;
        if (vox[a] >= nc[a]) vox[a] = nc[a] - 1;
        if (fabsf(d[a]) < eps) {
            tmax[a] = 1e30f;
            tdelta[a] = 0.0f;
        } else {
            const float boundary = (d[a] > 0.0f ? (float)(vox[a] + 1) : (float)vox[a]) * m->voxel + m->lo[a];
            tmax[a] = (boundary - p[a]) / d[a];
            tdelta[a] = m->voxel / fabsf(d[a]);
            if (tmax[a] < 0.0f) tmax[a] = 0.0f;
        }
    }
    float tau = 0.0f, t_done = 0.0f;
    for (int steps = 0; steps < 3 * (m->nx + m->ny + m->nz) + 16; ++steps) {
        const int axis = (tmax[0] <= tmax[1]) ? ((tmax[0] <= tmax[2]) ? 0 : 2) : ((tmax[1] <= tmax[2]) ? 1 : 2);
        const float t_next = fminf(tmax[axis], t_end);
        const float seg = t_next - t_done;
        if (seg > 0.0f) {
            const int idx = medium_index(m, vox[0], vox[1], vox[2]);
            if (m->mu) tau += m->mu[idx] * seg;
            if (count < cap) {
                cells[count] = idx;
                segs[count] = seg;
            }
            ++count;
        }
        if (t_next >= t_end) break;
        t_done = t_next;
        vox[axis] += (d[axis] > 0.0f) ? 1 : -1;
        tmax[axis] += tdelta[axis];
        if (vox[axis] < 0 || vox[axis] >= nc[axis]) break;
    }
    *n_cells = count;
    return tau;
}

/* src/muon.cuh:160-224 (sample_sky, muon_expose_binned), float throughout */
void orc_muon_expose(const orc_medium* m, const float c[3], float c_min, float th_max, int n_az, int n_th,
                     int64_t n_cand, uint32_t seed, uint64_t* open, uint64_t* det) {
    for (int64_t i = 0; i < n_cand; ++i) {
        const uint32_t key = orc_hash_u32((uint32_t)i * 0x9E3779B9u);
        const float u1 = orc_hash_unit(key);
        const float u2 = orc_hash_unit(key ^ 0x68BC21EBu);
        const float c3 = c_min * c_min * c_min;
        const float cz = cbrtf(c3 + u1 * (1.0f - c3));
        const float sn = sqrtf(fmaxf(0.0f, 1.0f - cz * cz));
        const float az = 6.28318530718f * u2;
        const float d[3] = {sn * cosf(az), sn * sinf(az), cz};
        const float th = acosf(fminf(1.0f, fmaxf(-1.0f, cz)));
        int ith = (int)(th / th_max * (float)n_th);
        if (ith < 0) ith = 0;
        if (ith >= n_th) ith = n_th - 1;
        int iaz = (int)(az * 0.15915494f * (float)n_az);
        if (iaz < 0) iaz = 0;
        if (iaz >= n_az) iaz = n_az - 1;
        const int bin = ith * n_az + iaz;
        open[bin] += 1;
        int dummy = 0;
        const float tau = orc_march(m, c, d, NULL, NULL, 0, &dummy);
        const float u3 = orc_hash_unit(key ^ 0x2545F491u ^ seed);
        if (u3 < expf(-tau)) det[bin] += 1;
    }
}

/* src/op_muon.hpp:22-50 for one view */
void orc_muon_rows(const orc_medium* model, const float c[3], const float* dirs, int n_bins, const uint64_t* open,
                   const uint64_t* det, const int32_t* site_bit, double min_open, float a_scale, int row0,
                   double* d, double* w, int cap_rows, int32_t* ent_bit, int32_t* ent_row, float* ent_a, int cap_ent,
                   int* n_rows, int* n_ent) {
    const int cap_cells = 3 * (model->nx + model->ny + model->nz) + 16;
    int* cells = (int*)malloc(sizeof(int) * (size_t)cap_cells);
    float* segs = (float*)malloc(sizeof(float) * (size_t)cap_cells);
    int rows = 0, ents = 0;
    for (int b = 0; b < n_bins; ++b) {
        if ((double)open[b] < min_open) continue;
        int nc = 0;
        const float tau0 = orc_march(model, c, dirs + 3 * b, cells, segs, cap_cells, &nc);
        const double dt = (double)det[b] > 1.0 ? (double)det[b] : 1.0;
        const double t = -log(dt / (double)open[b]);
        int touches = 0;
        for (int k = 0; k < nc && k < cap_cells; ++k) {
            const int bit = site_bit[cells[k]];
            if (bit < 0) continue;
            if (ents < cap_ent) {
                ent_bit[ents] = bit;
                ent_row[ents] = row0 + rows;
                ent_a[ents] = a_scale * segs[k];
            }
            ++ents;
            touches = 1;
        }
        if (!touches) continue;
        if (rows < cap_rows) {
            d[rows] = (double)tau0 - t;
            w[rows] = dt;
        }
        ++rows;
    }
    free(cells);
    free(segs);
    *n_rows = rows;
    *n_ent = ents;
}

void orc_assemble_stable(int n_bits, int n_ent, const int32_t* ent_bit, const int32_t* ent_row, const float* ent_a,
                         int32_t* row_ptr, int32_t* ray, float* a) {
    for (int i = 0; i <= n_bits; ++i) row_ptr[i] = 0;
    for (int k = 0; k < n_ent; ++k) row_ptr[ent_bit[k] + 1]++;
    for (int i = 0; i < n_bits; ++i) row_ptr[i + 1] += row_ptr[i];
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_bits > 0 ? n_bits : 1));
    for (int i = 0; i < n_bits; ++i) fill[i] = row_ptr[i];
    for (int k = 0; k < n_ent; ++k) {
        const int q = fill[ent_bit[k]]++;
        ray[q] = ent_row[k];
        a[q] = ent_a[k];
    }
    free(fill);
}
