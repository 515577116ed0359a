/*
 * qublar.h -- C ABI of libqublar_hip.so, the MI355X simulated-annealing QUBO engine.
 *
 * Plain pointers and sizes; every function returns a status code (QB_OK = 0) and
 * leaves a message for qb_last_error() on failure. Nothing calls exit(). Input
 * arrays are copied; output buffers belong to the caller. A handle owns one HIP
 * stream and is thread-compatible (one caller at a time per handle).
 *
 * Trajectory: sweeps visit bits by colour class (bits in one class share no ray
 * and are not neighbours, so their updates commute), and every Metropolis draw is
 * Philox4x32-10 keyed by (branch seed, sweep, bit). Branch b's seed is
 * b + 1 + seed_offset (float path) or seed_base + r (integer path), as in the
 * reference. Results are bit-identical to the CPU oracle run with the same
 * colouring and keys.
 *
 * Reference interfaces each entry replaces are cited per declaration
 * (paths relative to the QuBLAR tree).
 */
#ifndef QUBLAR_H
#define QUBLAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QB_OK 0
#define QB_E_ARG 1      /* invalid argument */
#define QB_E_HIP 2      /* HIP runtime error (replaces GISING_CUDA's exit(3), ising_gpu.cuh:44-52) */
#define QB_E_REFUSED 3  /* the plan refuses the problem (Plan::refused, ising_gpu.cuh:93-94) */
#define QB_E_NOMEM 4    /* device or host allocation failed */
#define QB_E_LIMIT 5    /* a stated size limit is exceeded */
#define QB_E_TIMEOUT 6  /* a device-side barrier wait exceeded its spin limit */

#define QB_FP64 0       /* residuals carried in FP64 (the reference's arithmetic) */
#define QB_FP32 1       /* residuals carried in FP32, FP64 recompute every k sweeps */

/* Message of the last failure on the calling thread ("" if none). */
const char* qb_last_error(void);
/* Library ABI version (major * 10000 + minor * 100 + patch). */
int qb_version(void);

typedef struct {
    char name[64];
    char arch[32];
    int compute_units;
    int xcds;
    int64_t hbm_bytes;
    int lds_bytes_per_cu;
} qb_device_info;
int qb_get_device_info(qb_device_info* out);

/* No counterpart in the reference. A plain device-to-device streaming copy of `bytes`
   (a positive multiple of 16) between two buffers the call allocates, repeated `reps`
   times; *ms_per_copy is the mean duration from HIP events. It calibrates HBM byte
   counters against a kernel of known traffic (2 x bytes) and gives the bandwidth a pure
   read + write stream sustains. */
int qb_copy_probe(int64_t bytes, int reps, double* ms_per_copy);

/* No counterpart in the reference. A check, on the device, of what the sweep's team barrier
   relies on: that a team's plain stores, drained before the barrier, are seen by the
   L2-hitting loads of every member after it. For `steps` steps every thread of every
   workgroup stores the step number, the team passes the barrier, and every thread reads its
   word of each member; *stale receives the reads that saw an older step (0 on a device the
   sweep is correct on), *us_per_step (nullable) the mean step time. workgroups as for
   qb_anneal_opts. */
int qb_team_barrier_check(int steps, int workgroups, uint64_t* stale, double* us_per_step);

/* ==== float path: the ray x voxel QUBO ================================================ */

/* BinaryProblem (src/engine.hpp:30-49): bit -> (ray, a) CSR, per-ray d and w, the
   neighbour CSR, lambda, kappa and an optional per-bit linear field. */
typedef struct {
    int n_vars, n_rays;
    const int32_t* row_ptr;   /* n_vars + 1 */
    const int32_t* ray;       /* row_ptr[n_vars] */
    const float* a;
    const double* d;          /* n_rays */
    const double* w;          /* n_rays */
    const int32_t* nbr_ptr;   /* n_vars + 1 */
    const int32_t* nbr;
    double lambda, kappa;
    const double* field;      /* n_vars, or NULL: lin(i) = kappa */
    /* No counterpart in the reference (both may be NULL). The draw of bit i in sweep s
       for branch b is keyed by seed b + 1 + seed_offset + draw_key[i] and counter
       (draw_bit[i], s); NULL means draw_key = 0 and draw_bit[i] = i. A problem
       assembled from independent requests (block-diagonal, each request's bits keeping
       its own local index and key) thereby anneals every request exactly as alone. */
    const int32_t* draw_bit;  /* n_vars */
    const uint64_t* draw_key; /* n_vars */
    /* No counterpart in the reference (may be NULL). A colouring to use instead of the
       engine's own: colour[i] >= 0, no two bits that share a ray or are neighbours
       with the same colour (checked; QB_E_ARG otherwise). Classes are visited in
       increasing colour, unused colours skipped. The union of independently coloured
       requests is coloured by their own colourings side by side. */
    const int32_t* colour;    /* n_vars */
} qb_problem_desc;

typedef struct qb_problem qb_problem;

/* Copies the problem to the device in colour-class layout. Replaces constructing a
   BinaryProblem for branch_fractions (src/ising_recon.hpp:148-151). */
int qb_problem_create(const qb_problem_desc* desc, qb_problem** out);
int qb_problem_destroy(qb_problem* p);

typedef struct {
    int n_vars, n_rays;
    int64_t nnz, n_nbr;
    int n_classes;            /* colour classes = dependent steps per sweep */
    int max_class, min_class;
    int max_entries_per_bit;
    int max_bits_per_ray;     /* lower bound on n_classes */
    int64_t device_bytes;     /* structure bytes resident in HBM */
} qb_problem_info;
int qb_problem_get_info(const qb_problem* p, qb_problem_info* out);
/* The colouring: order[n_vars] (bits grouped by class, ascending inside a class)
   and class_ptr[n_classes + 1]. */
int qb_problem_colouring(const qb_problem* p, int32_t* order, int32_t* class_ptr);
/* The same colouring computed on the host without a device (class_ptr needs room for
   n_vars + 1 entries). The visit order of the sweep is new; the reference visits bits
   in index order (src/ising_recon.hpp:132). */
int qb_colour_problem(const qb_problem_desc* desc, int32_t* order, int32_t* class_ptr, int* n_classes);

/* Schedule (src/ising_recon.hpp:81-86) and Budget (src/ising_weighted.hpp:227-232). */
typedef struct {
    double t_hot, t_cold;
    int anneal_sweeps, hold_sweeps;
} qb_schedule;

typedef struct {
    int precision;            /* QB_FP64 or QB_FP32 */
    int recompute_every;      /* > 0: residuals recomputed from scratch in FP64 after every k-th sweep */
    int workgroups;           /* persistent workgroups, a multiple of 8 (one team per XCD); 0 = one per CU */
    int sweeps_per_launch;    /* 0 = all sweeps in one launch; results do not depend on it */
    int group_width;          /* branches per group: 2, 4, 8 or 16; 0 = chosen from n_branches */
} qb_anneal_opts;

typedef struct {
    double wall_s;            /* state init to outputs on the host (excludes the problem upload) */
    double kernel_ms;         /* sweep kernels only, HIP events on the handle's stream */
    int launches;
    double spin_updates;      /* branches x sweeps x n_vars */
    double fp32_residual_err; /* max |r_carried - r_scratch_fp64| over rays and branches (FP32 carry) */
} qb_report;

/* n_branches branches with seeds b + 1 + seed_offset annealed from all-zero x:
   branch_fractions (src/ising_recon.hpp:148-166) and anneal_branch (:121-144).
   Outputs (each nullable): branches n_branches x n_vars bytes (0/1), p_void n_vars
   floats (sum of branches / n_branches), tri n_vars bytes (classify, :302-309:
   0 Exists, 1 NotThere, 2 Undecided), certainty n_vars floats (max(p, 1 - p)),
   energy n_branches doubles (binary_energy of each final branch, FP64, :65-79). */
int qb_anneal_branches(qb_problem* p, const qb_schedule* s, int n_branches, uint64_t seed_offset,
                       const qb_anneal_opts* opts, uint8_t* branches, float* p_void, uint8_t* tri,
                       float* certainty, double* energy, qb_report* rep);

/* branch_fractions in the reference's own trajectory (src/ising_recon.hpp:119-166): branch b
   anneals from x = 0, r = -d with seed b + 1 + seed_offset, bits in index order, dE by
   local_field_dE's arithmetic and order (:105-118), one splitmix64 stream drawn only for
   uphill proposals, a flip iff dE <= 0 or u < exp(-dE / T). Each branch is anneal_branch's,
   except where a draw falls within an ulp or two of exp(-dE / T) (the device's exponential
   against the host's libm). One wave per branch: for many-branch or small problems; the
   colour-class engine above is the fast path. draw_bit and draw_key do not apply. Outputs as
   qb_anneal_branches (energies FP64 from scratch); rep.fp32_residual_err is 0. */
int qb_anneal_branches_reference(qb_problem* p, const qb_schedule* s, int n_branches, uint64_t seed_offset,
                                 uint8_t* branches, float* p_void, uint8_t* tri, float* certainty, double* energy,
                                 qb_report* rep);
/* The same with the launch split stated: each launch runs at most sweeps_per_launch sweeps
   (0: the library bounds a launch at about 2^25 proposals), and the next continues from the
   spins, residuals and splitmix64 state the last one left, so the split never changes
   results. rep.launches counts the launches. No counterpart in the reference, whose CPU
   loop has no launches (src/ising_recon.hpp:121-144). */
int qb_anneal_branches_reference_split(qb_problem* p, const qb_schedule* s, int n_branches, uint64_t seed_offset,
                                       int sweeps_per_launch, uint8_t* branches, float* p_void, uint8_t* tri,
                                       float* certainty, double* energy, qb_report* rep);

/* The frozen set of freeze_provable_rock (src/ising_recon.hpp:187-205) on the device:
   frozen[i] = 1 (original bit order) where bit i is provably rock at `margin` nats. The
   bounds are the reference's, each bit's data bound summed in CSR order; the rounds test
   every bit against the previous round's frozen set, which for lambda >= 0 reaches the same
   fixpoint as the reference's in-order rounds (QB_E_ARG for lambda < 0). *rounds (nullable):
   the rounds run, the last one freezing nothing. The reduced problem is built on the host
   (argos::freeze_provable_rock). */
int qb_freeze(qb_problem* p, double margin, uint8_t* frozen, int* rounds);

/* binary_energy (src/ising_recon.hpp:65-79) of K states (K x n_vars bytes, 0/1),
   computed on the device in FP64; data and prior parts separately (each nullable). */
int qb_energy(qb_problem* p, const uint8_t* X, int K, double* total, double* data, double* prior);

/* No counterpart in the reference. qb_energy for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
   same kernels and results, bit for bit, with every device buffer of the call (the staged
   states, their packed masks, the partial sums and the three energy vectors) carved from a
   static device scratch of 32 MiB that each device holds until the process ends. The call
   makes no hipMalloc or hipFree. It needs about K * n_vars * 1.125 bytes plus the partial
   sums; a call that needs more returns QB_E_LIMIT and does nothing. Every call on one device
   shares that device's scratch, so calls that overlap on one device (from several threads
   or on several problems) race: serialise them. */
int qb_energy_noalloc(qb_problem* p, const uint8_t* X, int K, double* total, double* data, double* prior);

/* ---- staged solves (continuous batching) ----------------------------------------------
   No counterpart in the reference, which anneals one problem per branch_fractions call
   (src/ising_recon.hpp:148-166). A stage keeps the branch state (spins and residuals) of a
   problem -- in practice a union of independent requests built with draw_bit / draw_key /
   colour -- on the device between calls, so the union can be annealed a few sweeps at a time
   while requests join and leave. Every bit carries a clock: its own sweep is the stage's sweep
   count plus its span's offset, its temperature and Philox counter are those of its own sweep,
   and once its own sweep reaches the schedule's end it stays put; a span may also be limited to
   its request's first n_branches branches. A request that joins a running union with its state
   imported therefore follows, spin for spin, the trajectory of qb_anneal_branches on it alone
   (with seed_offset = its key), whatever it shares the union with and whenever it joins. With
   QB_FP32 and a recompute period k, spans are set and moved only at multiples of k sweeps, so the
   FP64 recompute falls on every request's own k-th sweeps. Calls on one stage are serialised by
   its problem's stream; the problem must outlive its stages. */
typedef struct qb_stage qb_stage;
/* A stage of n_branches branches (seed offset 0: keys come from draw_key), every bit at own
   sweep 0 with all branches, from x = 0, r = -d. opts: precision, recompute_every, workgroups,
   group_width as for qb_anneal_branches (sweeps_per_launch unused). */
int qb_stage_create(qb_problem* p, const qb_schedule* s, int n_branches, const qb_anneal_opts* opts, qb_stage** out);
int qb_stage_destroy(qb_stage* st);
/* Before the stage's first run: bits [bit0, bit0 + n_bits) (original index) start at own sweep
   sweeps_done, annealing their first n_branches branches (0: all; at most 255). */
int qb_stage_set_span(qb_stage* st, int bit0, int n_bits, int sweeps_done, int n_branches);
/* Before dst's first run: the state of src bits [src_bit0, + n_bits) and rays [src_ray0, + n_rays)
   -- spins, clocks (own sweeps kept) and residuals of the groups both stages hold -- into dst
   bits [dst_bit0, ...) and rays [dst_ray0, ...). The stages must agree in precision, group width,
   schedule length and recompute period, and src must be at a multiple of that period. */
int qb_stage_import(qb_stage* dst, const qb_stage* src, int src_bit0, int dst_bit0, int n_bits, int src_ray0,
                    int dst_ray0, int n_rays);
/* Runs the stage's next n_sweeps sweeps (one launch). rep.spin_updates counts every proposal
   slot of the launch, live or not. */
int qb_stage_run(qb_stage* st, int n_sweeps, qb_report* rep);
/* The outputs of bits [bit0, + n_bits) over their first n_branches branches, as
   qb_anneal_branches gives them for a problem of those bits alone (branch_fractions' p_v and
   classify, src/ising_recon.hpp:160-163, 302-309): branches n_branches x n_bits (nullable),
   p_void, tri, certainty (n_bits each, nullable), in the span's bit order. */
int qb_stage_outputs(qb_stage* st, int bit0, int n_bits, int n_branches, uint8_t* branches, float* p_void,
                     uint8_t* tri, float* certainty);
/* Sweeps the stage has run. */
int qb_stage_sweeps(const qb_stage* st, int* sweeps);

/* ==== integer path: weighted MaxCut / Ising ========================================== */

/* WeightedGraph edge list (src/ising_weighted.hpp:35-42), 0-indexed. */
typedef struct {
    int n;
    int64_t m;
    const int32_t* eu;
    const int32_t* ev;
    const int32_t* ew;
} qb_graph_desc;

typedef struct qb_graph qb_graph;

/* Plan (src/ising_gpu.cuh:59-101): the same bound and statistics fields. */
typedef struct {
    int n;
    int64_t m;
    int n_pad;
    int64_t B;                /* max_i sum_j |w_ij| */
    int64_t max_w;
    int field_bytes, weight_bytes;
    int dense, shared_state;
    double sigma;             /* sqrt(2 sum w^2 / n) */
    int64_t w_min;
    char refused[64];         /* "" when accepted */
    int n_classes;
} qb_plan;

/* make_plan + DeviceGraph (ising_gpu.cuh:74-101, 278-343). */
int qb_graph_create(const qb_graph_desc* desc, int force_global, qb_graph** out);
int qb_graph_destroy(qb_graph* g);
int qb_graph_get_plan(const qb_graph* g, qb_plan* out);
int qb_graph_colouring(const qb_graph* g, int32_t* order, int32_t* class_ptr);
/* The graph colouring on the host without a device (class_ptr: room for n + 1);
   the reference visits vertices in index order (src/ising_weighted.hpp:258). */
int qb_colour_graph(const qb_graph_desc* desc, int32_t* order, int32_t* class_ptr, int* n_classes);

typedef struct {
    double wall_s;
    double kernel_ms;
    int launches;
    int sweeps_per_launch;
} qb_batch_report;

/* run_batch (ising_gpu.cuh:396-458): replicas with seeds seed_base .. seed_base + R - 1.
   s_out: R x n_pad spins (+1 / -1, original bit order); energy_out: R tracked
   energies (-cut). updates_per_launch bounds one launch's work and never changes results. */
int qb_run_batch(qb_graph* g, const qb_schedule* b, uint64_t seed_base, int R, double updates_per_launch,
                 int8_t* s_out, int64_t* energy_out, qb_batch_report* rep);

/* run_batch in the reference's own trajectory (src/ising_gpu.cuh:396-458 with the semantics it
   shares with wising::anneal, src/ising_weighted.hpp:250-267): replica r anneals from s = -1
   with seed seed_base + r, bits in index order, one splitmix64 stream per replica drawn only
   for uphill proposals, a flip iff dE <= 0 or u < exp(-dE / T). Each replica's assignment and
   tracked energy are those of wising::anneal(g, b, seed_base + r), except where a draw falls
   within an ulp or two of exp(-dE / T) (the device's exponential against the host's libm).
   Same outputs and launch split as qb_run_batch. A replica's state (n_pad * 5 bytes) is held
   in LDS up to 160 KB and in global memory beyond; the trajectory is the same either way. */
int qb_run_batch_reference(qb_graph* g, const qb_schedule* b, uint64_t seed_base, int R, double updates_per_launch,
                           int8_t* s_out, int64_t* energy_out, qb_batch_report* rep);

/* device_energies (ising_gpu.cuh:461-475): -cut of K states (K x n bytes, 0/1). */
int qb_device_energies(qb_graph* g, const uint8_t* X, int K, int64_t* energy_out);

/* ---- exact region-of-interest posterior ------------------------------------------------
   The conditional QUBO of a region (roi_qubo, src/ising_recon.hpp:245-300: h[n], J[n x n]
   row-major with the upper triangle used) enumerated over all 2^n states on the device, as
   tools/branches_blaze.py:45-53, 80-81 does on the host: p(x) = exp(-beta E(x)) / Z with
   E(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j. marginal[i] = p(x_i = 1) (may be NULL);
   map_state[i] = bit i of the minimum-energy state, ties to the smallest state index
   (bit i of the index = x_i; may be NULL). 1 <= n <= 36, beta >= 0 finite. */
typedef struct {
    double e_min;           /* minimum energy */
    double log_z;           /* log sum_x exp(-beta E(x)) */
    double p_map;           /* probability of the minimum-energy state */
    uint64_t map_index;     /* that state, bit i = x_i */
    double states;          /* 2^n */
    double kernel_ms;       /* both enumeration passes, device time */
} qb_roi_result;

int qb_roi_enumerate(int n, const double* h, const double* J, double beta, double* marginal, uint8_t* map_state,
                     qb_roi_result* out);

/* ==== problem builders (host) ======================================================== */

typedef struct qb_host_problem qb_host_problem;

/* The muon-tomography scene: a square pyramid (height 139 m, half-base 115 m) in a
   128^3 grid of 2 m voxels, three point chambers, 360 x 70 sky bins to 70 degrees,
   counts det ~ Poisson(N0 cos^2(theta) exp(-mu L)) per bin, a cubic void, and the
   QUBO of src/ising_recon.hpp:52-62 over the rock voxels outside 3 m rooms. */
typedef struct {
    int log2_muons;           /* muons per chamber = 2^log2_muons (27) */
    double rho;               /* rock density, g/cm^3 (2.0) */
    double mu_per_rho;        /* attenuation per density, 1/m per g/cm^3 (0.023: mu = 0.046 /m) */
    double void_side;         /* cube edge, m (30) */
    double void_z;            /* cube centre height, m (77) */
    double lambda, kappa;     /* prior (2, ln 999) */
    double min_open;          /* bins with fewer open-sky muons are dropped (30) */
    uint64_t seed;            /* Poisson draws */
    int grid;                 /* voxels per axis (128); smaller grids scale the scene */
    int sky_az, sky_th;       /* sky bins per chamber: azimuth x zenith (360 x 70) */
} qb_muon_scene;

int qb_muon_scene_defaults(qb_muon_scene* s);
int qb_scene_muon(const qb_muon_scene* s, qb_host_problem** out);
/* Pointers into the host problem, valid until it is destroyed. */
int qb_host_problem_view(const qb_host_problem* h, qb_problem_desc* view);
/* truth[n_vars]: 1 where the bit's voxel is inside the void. */
int qb_host_problem_truth(const qb_host_problem* h, uint8_t* truth);
int qb_host_problem_destroy(qb_host_problem* h);

/* The reduced problem of freeze_provable_rock (src/ising_recon.hpp:206-233), on the host, from
   the problem and a frozen set (qb_freeze's): the unfrozen bits in original order with their
   entries, every ray with its d and w, the neighbour lists between kept bits, and each kept
   bit's field lin + lambda * (its frozen neighbours). kept[n_kept] (nullable; room for n_vars)
   maps reduced bits to original ones. The result is a host problem (qb_host_problem_view;
   its view carries the field). */
int qb_reduce_frozen(const qb_problem_desc* desc, const uint8_t* frozen, qb_host_problem** out, int32_t* kept,
                     int* n_kept);

/* The gravimetry scene: the operator of src/op_gravity.hpp:17-74 (a cell of volume h^3
   and contrast drho is a point mass at its centre; a_sb = G drho h^3 (z_s - z_b) / r^3 in
   microGal, one row per station, w = 1 / sigma^2) with every row truncated at `radius`
   (the reference sums all cells, op_gravity.hpp:59-74). nx x ny x nz cells of edge h lie
   below the surface z = 0, bits in x-fastest site order, one station at each column
   centre on the surface; d = the anomaly of the true body (cells whose centre lies in
   the box body_lo..body_hi) plus N(0, sigma^2) noise; the kappa / lambda prior over
   6-neighbour cells. The builder also colours the problem: layer k by a lattice of
   period P_k columns, the smallest P_k >= 2 for which two cells of a class see no
   common station (qb_host_problem_colour; the view passes it as desc.colour). */
typedef struct {
    int nx, ny, nz;             /* cells (100 x 200 x 10) */
    double h;                   /* cell edge, m (5) */
    double radius;              /* kernel cut-off, m (50) */
    double drho;                /* density contrast, kg/m^3 (1000) */
    double sigma;               /* station noise, microGal (10) */
    double body_lo[3], body_hi[3];   /* the true body, m (200..230, 400..440, -35..-15) */
    double lambda, kappa;       /* prior (2, ln 999) */
    uint64_t seed;              /* noise draws */
} qb_gravity_scene;

int qb_gravity_scene_defaults(qb_gravity_scene* s);
int qb_scene_gravity(const qb_gravity_scene* s, qb_host_problem** out);
/* colour[n_vars]: the builder's colouring; QB_E_ARG when the builder gave none. */
int qb_host_problem_colour(const qb_host_problem* h, int32_t* colour);

/* ---- problem builders on the device --------------------------------------------------
   A voxel medium (src/muon.cuh:38-58): cells (i, j, k), x fastest, lower corner lo, edge
   voxel, attenuation mu per cell (host array, copied). A binned sky (src/muon.cuh:177-186). */
typedef struct {
    float lo[3];
    float voxel;
    int nx, ny, nz;
    const float* mu;
} qb_voxel_medium;

typedef struct {
    float c_min;              /* cos of the zenith cap (0.34202: 70 degrees) */
    float th_max;             /* the cap, radians (1.22173) */
    int n_az, n_th;           /* azimuth x zenith bins; bin = ith * n_az + iaz */
} qb_muon_sky;

/* muon_expose_binned (src/muon.cuh:195-224) on the device: candidates 0 .. n_candidates - 1
   from the chamber, each a hash-sampled cos^2 sky direction binned into the sky, marched
   through the medium (src/muon.cuh:66-134) and detected when its hash draw falls below
   exp(-tau). Adds the open and detected counts per bin to open[] and det[] (n_az x n_th).
   The direction, binning and detection use the device's float cbrt, sin, cos, acos and exp,
   as the reference's kernel does, so a candidate near a bin edge or a detection threshold
   may land differently from a host's libm. */
int qb_muon_expose(const qb_voxel_medium* m, const float chamber[3], const qb_muon_sky* sky, int64_t n_candidates,
                   uint32_t seed, uint64_t* open, uint64_t* det);
/* muon_rows (src/op_muon.hpp:22-50) plus assemble (src/engine.hpp:120-149) and the
   6-neighbour lists (BitField::grid, src/engine.hpp:61-85) on the device, for n_views views:
   chambers 3 floats per view, dirs 3 floats per (view, bin) (the bin directions, as
   MuonBinnedData::bin_direction gives them), open / det counts per (view, bin), site_bit per
   cell of the model medium (-1 outside the domain; the domain cells numbered 0, 1, ... in
   cell order). A (view, bin) with open >= min_open whose march of the model medium crosses a
   bit is a row, in (view, bin) order: d = tau0 - t, t = -log(max(det, 1) / open), w =
   max(det, 1), entries a = a_scale * segment. Entries are grouped by bit in row order (the
   reference's std::sort leaves the order inside a bit to its library). The result is a host
   problem (qb_host_problem_view) with lambda and kappa. */
int qb_muon_rows_device(const qb_voxel_medium* model, int n_views, const float* chambers, const float* dirs,
                        int n_bins, const uint64_t* open, const uint64_t* det, const int32_t* site_bit,
                        double min_open, float a_scale, double lambda, double kappa, qb_host_problem** out);
/* The reference's muon pipeline for the scene of qb_muon_scene, built on the device: the
   pyramid's true medium (with the void) exposed from each chamber by 2^log2_muons candidate
   muons, rows of the model medium (without the void), assembly and neighbour lists. Its
   counts are the reference's candidate Monte Carlo, not qb_scene_muon's per-bin Poisson
   draws, so the two problems share their bits, rows and entries but not d and w. *build_ms
   (nullable): host wall time of the build. */
int qb_scene_muon_device(const qb_muon_scene* s, qb_host_problem** out, double* build_ms);
/* qb_scene_gravity with its rows (row_ptr, ray, a) built on the device: the same problem,
   array for array. */
int qb_scene_gravity_device(const qb_gravity_scene* s, qb_host_problem** out, double* build_ms);

/* ---- batched sparse QUBOs ---------------------------------------------------------
   No single counterpart in the reference: it solves one QUBO per call (anneal_branch /
   branch_fractions, src/ising_recon.hpp:121-166, with a QUBO's pair terms encoded as
   two-bit rays, src/check_ree_qubo.cu:172-192). Here many independent QUBOs
       E_k(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j,   x in {0,1}^n_k,
   each with its own replicas and seed, are annealed in one persistent launch. Replica r
   of instance k anneals from all-zero x, visiting the instance's colour classes in
   order; a flip is accepted iff dE <= 0 or u < qexp(-dE / T), with dE summed in FP64
   over the bit's CSR row in order, and u the (r mod 2)-th 53-bit uniform of one
   Philox4x32-10 call keyed by seed_k with counter (bit, sweep, r / 2, 0x51B0B47C).
   Results are identical in every launch shape and to the CPU oracle. */
typedef struct {
    int n_instances;
    const int32_t* n_bits;      /* per instance */
    const int32_t* row_ptr;     /* per instance n_k + 1 offsets into its own entries, concatenated
                                   (sum (n_k + 1) values; each instance's list starts at 0) */
    const int32_t* col;         /* entries of all instances, concatenated; instance-local bit index */
    const float* J;             /* coupling per entry; J_ij listed in row i and in row j, equal */
    const float* h;             /* per bit, concatenated */
    const int32_t* replicas;    /* per instance, 1..32 */
    const uint64_t* seed;       /* per instance */
} qb_qubo_batch_desc;

typedef struct qb_qubo_batch qb_qubo_batch;

/* Validates (symmetric rows, no self-coupling, indices in range), colours each instance
   and uploads the batch. */
int qb_qubo_batch_create(const qb_qubo_batch_desc* desc, qb_qubo_batch** out);
int qb_qubo_batch_destroy(qb_qubo_batch* b);

typedef struct {
    int mode;                 /* 0: one persistent launch for the whole batch; 1: one launch per instance */
    int replica_block;        /* replicas per job: 4, 8, 16 or 32 (0 with block 0: tuned) */
    int workgroups;           /* persistent workgroups; 0 = as many as fit the device */
    int block;                /* threads per workgroup: 256 or 1024; 0 with replica_block 0: the
                                 (block, replica block) shape that runs a short prefix fastest */
    int graph;                /* 1: the anneal's launch schedule (state reset, the persistent launch or
                                 the per-instance launches, these as 4 independent branches that may
                                 overlap on the device) is recorded once into a hipGraph per (shape,
                                 mode, sweeps, workgroups) and replayed by later calls, which then make
                                 one graph launch; results are those of graph = 0. No counterpart in
                                 the reference, whose chunked relaunch loop is plain
                                 (src/ising_gpu.cuh:424-445) */
} qb_qubo_opts;

typedef struct {
    double wall_s;            /* state init to outputs on the host */
    double kernel_ms;         /* annealing kernels only, HIP events */
    int launches;
    int replica_block;        /* the replica block the run used */
    int jobs;
    int block;                /* the threads per workgroup it used */
    double spin_updates;      /* sum over instances of replicas x sweeps x n_k */
    int graph_replayed;       /* 1 when the anneal replayed a graph recorded by an earlier call */
} qb_qubo_report;

/* Anneals every replica of every instance. x: sum_k R_k n_k bytes, instance-major then
   replica-major (nullable); energy: sum_k R_k doubles (nullable), FP64 from scratch. */
int qb_qubo_batch_anneal(qb_qubo_batch* b, const qb_schedule* s, const qb_qubo_opts* opts, uint8_t* x,
                         double* energy, qb_qubo_report* rep);
/* The colouring qb_qubo_batch_create gives an instance, on the host without a device
   (rows checked as there; class_ptr needs room for n + 1). */
int qb_colour_qubo(int n, const int32_t* row_ptr, const int32_t* col, const float* J, int32_t* order,
                   int32_t* class_ptr, int* n_classes);
/* Instance k's colouring: order[n_k] (bits grouped by class) and class_ptr[n_classes + 1];
   *n_classes receives the count (class_ptr needs room for n_k + 1). */
int qb_qubo_batch_colouring(const qb_qubo_batch* b, int instance, int32_t* order, int32_t* class_ptr,
                            int* n_classes);

/* ---- dense QUBO field engine (precision study) ------------------------------------
   No counterpart in the reference. One dense QUBO E(x) = sum_i h_i x_i + sum_{i<j} Q_ij
   x_i x_j (Q symmetric, zero diagonal), R replicas (seed b + 1 + seed_offset) annealed
   from all-zero x, bits in index order every sweep, local fields f = h + Q x carried
   incrementally: FP64, or FP32 (one float add of +-Q_ij per flip) with an FP64
   recompute every recompute_every sweeps. Accept iff dE = (1 - 2 x_i) f_i <= 0 or
   dE / T < -qlog(u), u = Philox4x32-10 keyed by the replica seed, counter (bit, sweep). */
typedef struct qb_dense qb_dense;
int qb_dense_create(int n, const float* Q, const float* h, qb_dense** out);
int qb_dense_destroy(qb_dense* d);

typedef struct {
    double wall_s;            /* state init to outputs */
    double kernel_ms;         /* the annealing kernel, HIP events */
    double spin_updates;      /* R x sweeps x n */
} qb_dense_report;

/* x: R x n (nullable); energy: R FP64 energies from scratch (nullable); field_err: R values
   of max_i |f_carried - f_scratch(FP64)| at the end (nullable). */
int qb_dense_anneal(qb_dense* d, const qb_schedule* s, int R, uint64_t seed_offset, int precision,
                    int recompute_every, uint8_t* x, double* energy, double* field_err, qb_dense_report* rep);

/* The Erdos-Renyi max-cut workload: splitmix64 from `seed` over pairs i < j in order;
   a pair is an edge when u < p, and its weight is +1 when the next draw's top bit is
   0, else -1. Fills up to cap edges; *m_out = edge count. */
int qb_scene_er_graph(int n, double p, uint64_t seed, int32_t* eu, int32_t* ev, int32_t* ew, int64_t cap,
                      int64_t* m_out);

#ifdef __cplusplus
}
#endif
#endif
