/*
 * qb_oracle.h -- CPU oracle for the simulated-annealing QUBO engine.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libqublar_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of QuBLAR's annealing path:
 *   src/ising_recon.hpp:65-166   binary_energy, local_field_dE, anneal_branch,
 *                                branch_fractions, Schedule, splitmix/uniform
 *   src/ising_recon.hpp:187-309  freeze_provable_rock, roi_qubo, classify
 *   src/ising_weighted.hpp:83-303 parse_rudy, cut_of, energy_scratch, State,
 *                                anneal, tts99, tts99_bootstrap
 *   src/ising_weighted.hpp:307-385 SHA-256 of the '0'/'1' assignment string
 *   src/muon.cuh:66-224, op_muon.hpp:22-50, engine.hpp:120-149  the muon exposure,
 *                                rows and assembly (the device builders' checker)
 *
 * Two trajectory definitions share one code path:
 *   ORC_RNG_REFERENCE  visit order = bit index, one splitmix64 stream per
 *                      branch drawn only when dE > 0, accept iff
 *                      u < exp(-dE / T) with the C library exp. This is the
 *                      reference's own trajectory.
 *   ORC_RNG_PHILOX     visit order = colour classes (given by the caller),
 *                      one Philox4x32-10 draw keyed by (branch seed, sweep,
 *                      bit), accept iff u < qexp(-dE * (1/T)). This is the
 *                      trajectory the MI355X engine reproduces.
 * and two residual carries for the float path: FP64 (the reference's) and
 * FP32 with an optional FP64 from-scratch recompute every k sweeps.
 *
 * Pinning status of each function is listed in oracle/README.md.
 */
#ifndef QB_ORACLE_H
#define QB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the float problem (src/engine.hpp:30-49 BinaryProblem) ---------------- */
typedef struct {
    int n_vars, n_rays;
    const int* row_ptr;   /* n_vars + 1: bit -> (ray, a) CSR */
    const int* ray;
    const float* a;
    const double* d;      /* n_rays */
    const double* w;      /* n_rays */
    const int* nbr_ptr;   /* n_vars + 1 */
    const int* nbr;
    double lambda, kappa;
    const double* field;  /* n_vars or NULL (then lin(i) = kappa) */
} orc_problem;

typedef struct {          /* src/ising_recon.hpp:81-86, src/ising_weighted.hpp:227-232 */
    double t_hot, t_cold;
    int anneal_sweeps, hold_sweeps;
} orc_schedule;

enum { ORC_RNG_REFERENCE = 0, ORC_RNG_PHILOX = 1 };
enum { ORC_FP64 = 0, ORC_FP32 = 1 };

typedef struct {
    int rng;               /* ORC_RNG_* */
    const int* order;      /* colour order: bits grouped by class; NULL = index order */
    const int* class_ptr;  /* n_classes + 1 offsets into order */
    int n_classes;
    int precision;         /* ORC_FP64 / ORC_FP32 residual carry */
    int recompute_every;   /* >0: FP64 from-scratch residuals after every k-th sweep */
    int segment;           /* 0: dE sums a bit's entries in one sequential chain (the
                              reference); S > 0: in chains of S entries, the first seeded
                              with lin*delta, the chain sums added in runs of 8 and the
                              runs in order */
} orc_mode;

/* ---- primitives ---- */
uint64_t orc_splitmix(uint64_t* s);                       /* ising_recon.hpp:89-94 */
double orc_uniform(uint64_t* s);                          /* ising_recon.hpp:95 */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_philox_uniform(uint64_t seed, uint32_t sweep, uint32_t bit);
double orc_qexp(double x);
double orc_temperature(const orc_schedule* s, int sweep);  /* ising_recon.hpp:129-131 */

/* ---- float path ---- */
double orc_binary_energy(const orc_problem* p, const uint8_t* x);      /* ising_recon.hpp:65-79 */
void orc_energy_parts(const orc_problem* p, const uint8_t* x, double* data, double* prior);
double orc_local_field_dE(const orc_problem* p, const uint8_t* x, const double* r, int i);
/* One branch from all-zero x. trace (nullable) receives binary_energy after each sweep;
   r_final (nullable, n_rays doubles) the carried residuals at the end. */
void orc_anneal_branch(const orc_problem* p, const orc_schedule* s, uint64_t seed,
                       const orc_mode* m, uint8_t* x_out, double* trace, double* r_final);
/* R branches with seeds b + 1 + seed_offset on n_threads pthreads (ising_recon.hpp:148-166).
   keep (nullable) receives R x n_vars bytes; p_out n_vars floats. */
void orc_branch_fractions(const orc_problem* p, const orc_schedule* s, int n_branches,
                          int n_threads, uint64_t seed_offset, const orc_mode* m,
                          uint8_t* keep, float* p_out);
int orc_classify(float p, float lo, float hi);             /* ising_recon.hpp:302-309 */
/* freeze_provable_rock (ising_recon.hpp:187-234): frozen[i] = 1 for frozen bits; returns rounds. */
int orc_freeze_mask(const orc_problem* p, double margin, uint8_t* frozen);
/* Each bit's freeze bound with the frozen set held fixed (ising_recon.hpp:198-202). */
void orc_freeze_bounds(const orc_problem* p, const uint8_t* frozen, double* bound);
/* roi_qubo (ising_recon.hpp:251-300): h[n], J[n*n] (upper triangle). */
void orc_roi_qubo(const orc_problem* p, const int* roi, int n, const uint8_t* x_bar,
                  double* h, double* J);

/* ---- integer path (src/ising_weighted.hpp) ---- */
typedef struct {
    int n;
    int64_t m;
    const int32_t *eu, *ev, *ew;  /* edge list, 0-indexed */
    const int64_t* row;           /* n + 1, symmetric CSR as parse_rudy builds it */
    const int32_t *col, *w;
} orc_graph;

/* CSR of parse_rudy (ising_weighted.hpp:134-147): row n+1, col/w 2m. */
void orc_graph_csr(int n, int64_t m, const int32_t* eu, const int32_t* ev, const int32_t* ew,
                   int64_t* row, int32_t* col, int32_t* w);
/* parse_rudy (ising_weighted.hpp:83-149). Returns 0 on success and fills n, m and, when the
   caller's buffers are large enough (cap edges), eu/ev/ew; otherwise 1 and err. */
int orc_parse_rudy(const char* data, size_t size, int* n, int64_t* m, int32_t* eu, int32_t* ev,
                   int32_t* ew, int64_t cap, char* err, size_t err_cap);
int64_t orc_cut_of(const orc_graph* g, const uint8_t* x);         /* :173-178 */
int64_t orc_energy_scratch(const orc_graph* g, const uint8_t* x); /* :181-189 */
/* anneal (ising_weighted.hpp:250-267) under either trajectory definition. */
void orc_anneal_int(const orc_graph* g, const orc_schedule* b, uint64_t seed, const orc_mode* m,
                    uint8_t* x_out, int64_t* energy_tracked);
double orc_tts99(double t_mean, double p);                        /* :271-275 */
void orc_tts99_bootstrap(const int* success, const double* time, size_t r, int resamples,
                         uint64_t seed, double* lo, double* hi);  /* :279-303 */
void orc_sha256_hex(const void* data, size_t n, char out[65]);   /* :307-379 */

/* ---- batched sparse QUBOs (the batch engine's trajectory) ---- */
double orc_qubo_energy(int n, const int* row_ptr, const int* col, const float* J, const float* h,
                       const uint8_t* x);
/* order: the instance's bits grouped by colour class (a proper colouring of its
   coupling graph, so bits of a class are independent and any order inside it gives
   the same result). x_out: R x n. */
void orc_qubo_anneal(int n, const int* row_ptr, const int* col, const float* J, const float* h, const int* order,
                     const orc_schedule* s, uint64_t seed, int R, uint8_t* x_out);

/* ---- dense QUBO field engine (the precision study's dense instance) ---- */
double orc_qlog(double x);
double orc_dense_energy(int n, const float* Q, const float* h, const uint8_t* x);
void orc_dense_anneal(int n, const float* Q, const float* h, const orc_schedule* s, uint64_t seed, int precision,
                      int recompute_every, uint8_t* x, double* f_final);

/* The exact posterior of a small QUBO over all 2^n states, by brute force in state index
   order (tools/branches_blaze.py:45-53 exact_posterior, 80-81 marginals; the QUBO is
   roi_qubo's, src/ising_recon.hpp:245-300): E(x) = sum h_i x_i + sum_{i<j} J_ij x_i x_j
   (J n x n, upper triangle), p ~ exp(-beta (E - E_min)). Returns E_min; marginal[i],
   *log_z, *map_index (first state of minimum energy, bit i = x_i). n <= 30. */
double orc_roi_enumerate(int n, const double* h, const double* J, double beta, double* marginal, double* log_z,
                         uint64_t* map_index);

/* ---- problem builders (src/muon.cuh:66-224, src/op_muon.hpp:22-50, src/engine.hpp:120-149) ----
   A voxel medium: cells (i, j, k), x fastest, lower corner lo, edge voxel, attenuation mu. */
typedef struct {
    float lo[3];
    float voxel;
    int nx, ny, nz;
    const float* mu;
} orc_medium;

/* The Wang hash and its unit float (src/lidar.cuh:90-101). */
uint32_t orc_hash_u32(uint32_t x);
float orc_hash_unit(uint32_t x);
/* The Amanatides-Woo march of src/muon.cuh:66-134 in float: optical depth sum mu * segment
   from p along unit d; cells[k], segs[k] receive the crossed cells with a positive segment
   (cap entries at most; *n_cells the count, which may exceed cap). */
float orc_march(const orc_medium* m, const float p[3], const float d[3], int* cells, float* segs, int cap,
                int* n_cells);
/* muon_expose_binned (src/muon.cuh:195-224): candidates 0 .. n_cand - 1 from chamber c, each
   a hash-sampled cos^2 sky direction binned into n_th x n_az bins up to th_max, counted open,
   and detected when hash_unit(key ^ 0x2545F491 ^ seed) < exp(-tau). Adds to open and det. */
void orc_muon_expose(const orc_medium* m, const float c[3], float c_min, float th_max, int n_az, int n_th,
                     int64_t n_cand, uint32_t seed, uint64_t* open, uint64_t* det);
/* muon_rows (src/op_muon.hpp:22-50) for one view, bin directions given (dirs: 3 floats per
   bin): for each bin with open >= min_open whose march of the model medium crosses a bit,
   one row: d = tau0 - t with t = -log(max(det, 1) / open), w = max(det, 1), and an entry
   (bit, row, a_scale * seg) per crossed bit in march order. Appends at most cap_rows rows
   and cap_ent entries; returns the counts through *n_rows, *n_ent. */
void orc_muon_rows(const orc_medium* model, const float c[3], const float* dirs, int n_bins, const uint64_t* open,
                   const uint64_t* det, const int32_t* site_bit, double min_open, float a_scale, int row0,
                   double* d, double* w, int cap_rows, int32_t* ent_bit, int32_t* ent_row, float* ent_a, int cap_ent,
                   int* n_rows, int* n_ent);
/* assemble's CSR (src/engine.hpp:120-149) with entries grouped by bit in their given order
   inside a bit (a stable sort by bit). */
void orc_assemble_stable(int n_bits, int n_ent, const int32_t* ent_bit, const int32_t* ent_row, const float* ent_a,
                         int32_t* row_ptr, int32_t* ray, float* a);

#ifdef __cplusplus
}
#endif
#endif
