// qb_internal.h -- device layouts shared by the host orchestration (qb_host.cpp)
// and the gfx950 kernels (qb_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace qb {

// Branches are annealed in groups of 16: a group's spins for one bit form a 16-bit
// mask, and its residuals for one ray are 16 consecutive values, so the 16 lanes
// that evaluate one bit for the 16 branches of a group read one contiguous line.
constexpr int kGroup = 16;

// A bit's dE sums its CSR entries in chains of kSeg entries, the first seeded with
// lin * delta; the chain sums are added in runs of kChainRun (each run in order) and
// the run sums in order -- with at most kChainRun chains, the chain sums in order.
// Bits with more than kSeg entries come first inside their class.
constexpr int kSeg = 16;
constexpr int kChainRun = 8;

// One CSC entry of a bit in colour-class position order: the ray, the bit's share
// a of it, and the ray's weight w (copied per entry so the sweep streams it).
struct RayEnt {
    int32_t ray;
    float a;
    double w;
};

// One entry of a ray's bit list, in increasing original bit index (the order the
// from-scratch residual sum of binary_energy visits them).
struct BitEnt {
    int32_t pos;
    float a;
};

// The float problem on the device, in colour-class position order: position p
// holds original bit pos_bit[p]; class c is positions [cls_ptr[c], cls_ptr[c+1]).
struct DevProblem {
    int n, n_rays, n_classes;
    int64_t nnz, n_nbr;
    const int32_t* cls_ptr;
    const int32_t* pos_bit;
    const int32_t* bit_pos;     // original bit -> position
    const int32_t* col_ptr;
    const RayEnt* ent;
    const int32_t* nb_ptr;
    const int32_t* nb;      // neighbour positions, reference CSR order
    const int32_t* nb_bit;  // neighbour original indices (for the energy's j > i rule)
    const double* lin;
    const double* d;
    const double* w;
    const int32_t* rt_ptr;
    const BitEnt* rt;
    const uint8_t* seq_update;  // 1 where a ray repeats in the column: scatter in CSR order
    const int32_t* draw_bit;    // the counter index of each position's draws
    const uint64_t* draw_key;   // added to the branch seed of each position's draws
    // Light bits (no entry, at most kLightNb neighbours) close every class, from
    // light_begin[c]; light_nb[p * kLightNb + k] is light position p's k-th neighbour in
    // CSR order (-1 past its last).
    const int32_t* light_begin;
    const int32_t* light_nb;
    double lambda;
};
constexpr int kLightNb = 8;

// Work batches of one sweep step at one group width: a workgroup of BLOCK threads
// has BLOCK / BW slots of BW lanes, and a bit with k entries needs
// max(1, ceil(k / kSeg)) slots (one per chain). A batch packs whole bits into one
// workgroup's slots; a bit needing more slots than a workgroup has is a batch of its
// own (count -1) and is evaluated in several passes.
//
// Each slot of a packed batch has a descriptor holding everything static the slot
// needs, so the sweep can fetch a step's structure before the step's state is ready:
//   pos   the bit's position (-1: the slot is empty)
//   ka    the first CSR entry of the slot's chain
//   nb0   the bit's first neighbour entry
//   info  len (bits 0-4, entries in the chain, 0..kSeg) | j (5-12, the chain's index
//         in its bit; the bit's first slot is the slot's index minus j) | nch (13-20,
//         the bit's chains) | n_nb (21-30, the bit's neighbours) | seq (31, a ray
//         repeats in the column)
struct SlotDesc {
    int32_t pos, ka, nb0;
    uint32_t info;
};
constexpr int kMaxPackedNeighbours = 1023;

// The static operands of a packed batch as one block, copied whole into LDS a step
// ahead: its n_used slots (a prefix of the batch) each hold one chain of kSeg entries
// (past the chain's length: ray -1, a 0, w 0), entry u of slot s at 16-byte granule
// s * kSeg + (u ^ (s % kSeg)) -- the swizzle keeps the slots' reads of one entry on
// different LDS banks -- followed by the n_used slots' heads.
constexpr int kNb = 6;   // neighbour positions carried per bit; more are read from nb
struct SlotHead {
    int32_t pos;
    uint32_t info;
    int32_t nb0;
    int32_t bit;          // draw counter index
    int32_t q[kNb];       // first neighbour positions (the last repeated past the bit's count)
    double lin;
    uint64_t key;         // draw key offset
    int32_t pad[2];
};
static_assert(sizeof(SlotHead) == 64, "a head is four 16-byte granules");
static_assert(sizeof(RayEnt) == 16, "an entry is one 16-byte granule");
constexpr int kHeadGranules = int(sizeof(SlotHead) / 16);
__host__ __device__ inline int blk_entry(int slot, int u) { return slot * kSeg + (u ^ (slot & (kSeg - 1))); }

struct DevBatches {
    const int32_t* cls_batch;   // n_classes + 1: class c's batches, its multi-pass batches first
    const int32_t* cls_heavy;   // n_classes: class c's multi-pass batches
    const int32_t* first;       // per batch: first item
    const int32_t* count;       // per batch: items, or -1 for one bit in several passes
    const int32_t* pos;         // per item: the bit's position
    const int32_t* slot0;       // per item: its first slot in the batch
    const SlotDesc* desc;       // per batch: BLOCK / BW slot descriptors (all empty for count -1)
    const uint4* blk;           // the packed batches' blocks, in 16-byte granules
    const int2* blk_meta;       // per batch: {first granule of its block, slots used (0: multi-pass)}
    int n_batches;
    int heavy_lds;              // 1 when some batch is multi-pass: the sweep's dynamic LDS then
                                // holds the staging area after the class tables
};

// The sweep keeps its class tables (cls_batch, cls_heavy, light_begin, cls_ptr end) in LDS.
constexpr int kSweepTableLimit = 96 * 1024;
__host__ __device__ inline size_t sweep_table_bytes(int n_classes) { return sizeof(int32_t) * size_t(4 * n_classes + 2); }

// Workgroups of a team reserved for a class's multi-pass batches, the packed batches
// going round-robin over the others; with none reserved every task of the class,
// multi-pass first, goes round-robin over the whole team. Reserving pays only where
// a class has more tasks than the team has workgroups and a multi-pass batch costs
// about one packed batch per reserved workgroup's share -- not the case for the sweep
// as measured -- so none are reserved: a class that fits the team gets one task per
// workgroup either way (its packing leaves room for the multi-pass batches).
__host__ __device__ inline int heavy_ranks(int heavy_tasks, int team_size) {
    (void)heavy_tasks;
    (void)team_size;
    return 0;
}

// Mutable state of one anneal: n_groups groups of BW branches (BW in {2, 4, 8, 16}).
template <typename T>
struct DevState {
    int n_groups, R;
    uint64_t seed_offset;
    uint16_t* x;            // [g][n] spin masks, bit b = branch g*BW + b
    T* r;                   // [g][n_rays][BW] residuals
    const double* beta;     // 1 / T per sweep
    unsigned* bar;          // [0]: grid barrier; [64 + 64 t + r]: team t member r's arrival word (r < 64)
    unsigned* tickets;      // [64 t]: team t's member count; [64 t + 32]: its light-task counter
    int* err;               // 1: a barrier wait timed out; 2: a team came up short
    // A staged solve (null clock otherwise): per position, the bit's own sweep is the
    // launch's sweep plus clock & kClockOffsetMask, and clock >> 24 is the number of
    // branches its request anneals (0: all). A bit whose own sweep has reached
    // total_sweeps, or a branch past its request's, stays put (no draw, no flip); beta
    // is then indexed by the own sweep.
    const uint32_t* clock;
    int total_sweeps;
};
constexpr uint32_t kClockOffsetMask = 0xFFFFFFu;
constexpr int kClockMaxBranches = 255;

// Unsigned words of the barrier array and the ticket array (zeroed before a launch).
constexpr int kBarWords = 64 * 9;
constexpr int kTicketWords = 64 * 8;

// The integer graph on the device, in colour-class position order.
struct DevGraph {
    int n, n_pad, n_classes;
    int64_t m;
    const int32_t* cls_ptr;
    const int32_t* pos_bit;
    const int32_t* nb_ptr;
    const int32_t* nb;
    const int32_t* nw;
    const int32_t* f0;      // fields of the all-zero assignment
    const int32_t* bit_pos; // original bit -> position
    const int32_t* eu;      // edge list, original indices
    const int32_t* ev;
    const int32_t* ew;
};

// ---- launchers (qb_kernels.hip) ----
// Writes every packed batch's block (DevBatches::blk) from its slot descriptors.
hipError_t launch_batch_blocks(const DevProblem& p, const SlotDesc* desc, const int2* meta, int n_batches, int slots,
                               uint4* blk, hipStream_t st);
hipError_t launch_ray_init(const DevProblem& p, int precision, int bw, void* x, void* r, int n_groups,
                           hipStream_t st);
// The persistent sweep on team_size * 8 workgroups (team_size per XCD).
// With a non-null `clock` the staged form runs (DevState::clock): sweep_begin .. sweep_end
// count the stage's own sweeps and every bit keeps its own.
hipError_t launch_ray_sweeps(const DevProblem& p, const DevBatches& bt, int precision, int bw, void* x, void* r,
                             int n_groups, int R,
                             uint64_t seed_offset, const double* beta, int sweep_begin, int sweep_end,
                             int recompute_every, unsigned* bar, unsigned* tickets, int* err, int team_size,
                             hipStream_t st, const uint32_t* clock = nullptr, int total_sweeps = 0);
// Co-resident sweep workgroups per CU for this precision and group width, with
// `table_bytes` of dynamic LDS (0 when they do not fit), plain or staged form.
int ray_sweep_blocks_per_cu(int precision, int bw, size_t table_bytes, bool staged = false);
constexpr int kSweepBlock = 256;   // threads per sweep workgroup
constexpr int kTeams = 8;          // XCD teams of the sweep (group g belongs to team g mod 8)
constexpr int kMaxTeamSize = 64;   // members of a team (one arrival word each in its barrier area)
// A multi-pass bit whose chains fit kHeavyPasses rounds of a workgroup's slots is staged
// whole in LDS. The staging area follows the class tables in the sweep's dynamic LDS
// (16-byte aligned), reserved only when the problem has a multi-pass bit at the group
// width, so problems without one keep that LDS for their tables.
constexpr int kHeavyPasses = 3;
// The sweep's dynamic LDS: the class tables, two block buffers (the step's and the next
// one's), and a work area X. A packed step transposes its residuals in X; a multi-pass
// step stages its column's first kSlots chains in its own (empty) block buffer and the
// rest, with its chain sums, in X.
__host__ __device__ inline size_t sweep_tables_bytes_aligned(int n_classes) {
    return (sweep_table_bytes(n_classes) + 15) & ~size_t(15);
}
__host__ __device__ inline size_t sweep_block_bytes(int bw) {
    return size_t(kSweepBlock / bw) * (size_t(kSeg) * sizeof(RayEnt) + sizeof(SlotHead));
}
__host__ __device__ inline size_t sweep_trow_bytes(int precision) {
    return size_t(kSweepBlock) * size_t(kSeg) * (precision == 1 ? sizeof(float) : sizeof(double));
}
__host__ __device__ inline size_t heavy_extra_entries(int bw) {
    return size_t(kHeavyPasses - 1) * size_t(kSweepBlock / bw) * size_t(kSeg);
}
__host__ __device__ inline size_t heavy_extra_bytes(int bw) {
    return heavy_extra_entries(bw) * sizeof(RayEnt) + size_t(kHeavyPasses) * size_t(kSweepBlock / bw) * size_t(bw) *
                                                          sizeof(double);
}
__host__ __device__ inline size_t sweep_x_bytes(int bw, int precision, bool heavy) {
    const size_t t = sweep_trow_bytes(precision);
    return heavy && heavy_extra_bytes(bw) > t ? heavy_extra_bytes(bw) : t;
}
inline size_t sweep_dynamic_lds(int n_classes, int bw, bool heavy, int precision) {
    return sweep_tables_bytes_aligned(n_classes) + 2 * sweep_block_bytes(bw) + sweep_x_bytes(bw, precision, heavy);
}
// Entries of a bit with more chains than a sweep workgroup has slots at group width 2
// (a multi-pass bit there).
constexpr int kHeavyEntries = kSeg * (kSweepBlock / 2);
hipError_t launch_ray_outputs(const DevProblem& p, const uint16_t* x, int n_groups, int bw, int R, float* p_void,
                              uint8_t* tri, float* certainty, uint8_t* branches, hipStream_t st);
// FP64 energy of every branch: data and prior partial sums per block, reduced in a
// fixed order; resid_err receives max |r_carried - r_scratch| when r_carried != null;
// total (nullable) receives data + prior.
hipError_t launch_ray_energy(const DevProblem& p, const uint16_t* x, int n_groups, int bw, int R,
                             const void* r_carried, int precision, double* part, int n_part_blocks,
                             double* data, double* prior, float* resid_err, hipStream_t st,
                             double* total = nullptr);
int ray_energy_blocks(const DevProblem& p, int bw);
hipError_t launch_pack_states(const uint8_t* X, int K, int n, const int32_t* pos_bit, uint16_t* x,
                              hipStream_t st);

hipError_t launch_int_sweeps(const DevGraph& g, int R, uint64_t seed_base, const double* beta,
                             int sweep_begin, int sweep_end, int init, int8_t* s_state, int32_t* f_state,
                             long long* energy, bool lds_state, hipStream_t st);
// freeze_provable_rock on the device: per-position data bounds, then one fixpoint round
// (frozen -> next, *changed counts the bits newly frozen).
// The team barrier's visibility check: team_size * 8 workgroups, words holding 256
// unsigned per workgroup, *stale counting the reads that saw an older step.
hipError_t launch_team_barrier_probe(unsigned* words, unsigned* bar, unsigned* tickets, int* err, int steps,
                                     int team_size, unsigned long long* stale, hipStream_t st);
// Staged solves: a span's clock word; a span's state (spins, clocks, residuals of groups
// [0, groups)) from one stage into another; a span's branch agreement over rq branches.
hipError_t launch_stage_clock(const DevProblem& p, int bit0, int n_bits, uint32_t word, uint32_t* clock,
                              hipStream_t st);
hipError_t launch_stage_import(const DevProblem& src, const void* sx, const void* sr, const uint32_t* sclock,
                               int src_sweeps, const DevProblem& dst, void* dx, void* dr, uint32_t* dclock,
                               int precision, int bw, int sb0, int db0, int n_bits, int sr0, int dr0, int n_rays,
                               int groups, hipStream_t st);
hipError_t launch_stage_outputs(const DevProblem& p, const uint16_t* x, int bw, int bit0, int n_bits, int rq,
                                float* p_void, uint8_t* tri, float* certainty, uint8_t* branches, hipStream_t st);
hipError_t launch_freeze_bound(const DevProblem& p, double* lb, hipStream_t st);
hipError_t launch_freeze_round(const DevProblem& p, const double* lb, double margin, const uint8_t* frozen,
                               uint8_t* next, unsigned* changed, hipStream_t st);
// The reference's own float trajectory (anneal_branch): branch b's spins xs[b][bit order]
// (bytes), residuals rs[b][ray] (FP64) and splitmix64 state sm[b] in global memory; one
// launch runs sweeps [sweep_begin, sweep_end), the first starting from x = 0.
hipError_t launch_ray_ref(const DevProblem& p, int R, uint64_t seed_offset, const double* temp, int sweep_begin,
                          int sweep_end, uint8_t* xs, double* rs, uint64_t* sm, hipStream_t st);
// The reference's own integer trajectory (wising::anneal), one wave per replica with its
// state in LDS (n_pad * 5 bytes; four replicas per workgroup while that fits), or in global
// memory when one replica's state exceeds kRefLdsLimit: temp[sweep] is the temperature,
// rng[r] the replica's splitmix64 state between launches.
constexpr size_t kRefLdsLimit = 160 * 1024;
size_t int_ref_lds_bytes(const DevGraph& g);
bool int_ref_in_lds(const DevGraph& g);
hipError_t launch_int_ref(const DevGraph& g, int R, uint64_t seed_base, const double* temp, int sweep_begin,
                          int sweep_end, int init, int8_t* s_state, int32_t* f_state, uint64_t* rng,
                          long long* energy, hipStream_t st);
hipError_t launch_int_unpermute(const DevGraph& g, int R, const int8_t* s_state, int8_t* s_out,
                                hipStream_t st);
hipError_t launch_int_energies(const DevGraph& g, const uint8_t* X, int K, long long* out, hipStream_t st);
int int_lds_capacity_spins();

hipError_t launch_copy_probe(const void* src, void* dst, size_t n16, hipStream_t st);

// ---- batched sparse QUBOs (qb_qubo.hip) ----
// Positions: instances one after another, each in colour-major order (instance k's bits
// are positions [inst_pos[k], inst_pos[k+1]), its classes cls_ptr[inst_cls[k] ..
// inst_cls[k+1] - 1] as global positions). Rows keep each bit's CSR order.
struct QuboJob {
    int32_t instance, replica0, replicas;   // replicas [replica0, replica0 + replicas) of the instance
    int32_t mask_offset;                    // the job's spin words: masks[mask_offset + local position]
    int32_t energy_offset;                  // energy[energy_offset + r]
    int32_t pad[3];
};
struct DevQuboBatch {
    int n_jobs;
    const QuboJob* jobs;         // grouped by XCD queue (or in instance order)
    const int32_t* inst_pos;     // n_instances + 1
    const int32_t* inst_cls;     // n_instances + 1 offsets into cls_ptr
    const int32_t* cls_ptr;      // per instance: its n_classes + 1 class bounds (global positions)
    const int32_t* row_ptr;      // total bits + 1
    const int32_t* col;          // neighbour position
    const float* J;              // coupling per entry
    const float* h;              // field per position
    const int32_t* bit_id;       // per position: the bit's index inside its instance (draw counter)
    const uint64_t* seed;        // per instance
    // The rows again, sliced by 64 positions of a class (a wave's bits in one pass) and
    // stored entry-major: entry u of tile t's lane l at tile_ptr[t] + 64 u + l, so one load
    // of a wave reads 64 consecutive words. Tile t of class c is class_tile[c] + (pos - lo) / 64.
    const int32_t* class_tile;   // per class (global class index)
    const int32_t* tile_ptr;     // per tile
    const int32_t* tcol;         // neighbour position
    const float* tJ;             // coupling
};
constexpr int kQuboBlock = 256;      // threads of the energy kernel
constexpr int kQuboLdsBytes = 65536;  // LDS spin words of one annealing job
// block: 256 or 1024 threads per workgroup (one bit per thread in a class pass).
// The persistent kernel's job queues, one per XCD: jobs [begin[x], begin[x + 1]) of
// q.jobs are XCD x's, counted by queue[kQueueStride * x]. A workgroup takes from its own
// XCD's queue, then from the others in turn.
constexpr int kXcds = 8;
constexpr int kQueueStride = 16;   // one 64-byte line per counter
struct QuboQueues {
    int32_t begin[kXcds + 1];
};
hipError_t launch_qubo_batch(const DevQuboBatch& q, int rb, int block, uint32_t* masks, const double* beta,
                             int n_sweeps, unsigned* queue, const QuboQueues& segs, int workgroups, hipStream_t st);
int qubo_blocks_per_cu(int rb, int block);
hipError_t launch_qubo_energy(const DevQuboBatch& q, const uint32_t* masks, double* energy, hipStream_t st);
// p[0 .. n) = v: a kernel, so a recorded graph holds no memset node.
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st);

// ---- dense QUBO field engine (qb_dense.hip) ----
struct DevDense {
    int n;
    const float* Q;   // n x n, symmetric, zero diagonal
    const float* h;
};
constexpr int kDenseBlock = 1024;
size_t dense_lds_bytes(int n, int precision);
// One workgroup per replica (seed b + 1 + seed_offset); x_out R x n, energy and the final
// max |f_carried - f_scratch| per replica.
hipError_t launch_dense(const DevDense& d, int precision, int R, const double* beta, int n_sweeps,
                        int recompute_every, uint64_t seed_offset, uint8_t* x_out, double* energy, double* field_err,
                        hipStream_t st);
// The same anneal with kDenseParts workgroups of one XCD per replica (R <= 8 * kDenseTeamSlots,
// n >= 64 * kDenseParts): bit for bit the same spins and fields. `work` holds
// dense_team_work_bytes(n, R) zeroed bytes; *err (zeroed) is set when a wait times out.
constexpr int kDenseParts = 4;
constexpr int kDenseTeamSlots = 8;   // replicas per XCD (32 CUs / kDenseParts)
size_t dense_team_work_bytes(int n, int R);
size_t dense_team_lds_bytes(int n, int precision);
hipError_t launch_dense_team(const DevDense& d, int precision, int R, const double* beta, int n_sweeps,
                             int recompute_every, uint64_t seed_offset, uint8_t* x_out, double* energy,
                             double* field_err, void* work, int* err, hipStream_t st);


#if defined(__HIPCC__)
// The XCD (accelerator complex die) the calling wave runs on: HW_REG_XCC_ID bits 0-3.
__device__ __forceinline__ unsigned xcc_id() { return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)); }
// A load that skips the CU's L1 and reads the XCD's L2 (sc1): within one XCD, the point of
// coherence for stores that have completed.
template <typename V>
__device__ __forceinline__ V ld_l2(const V* p) {
    return __hip_atomic_load(const_cast<V*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
#endif

// ---- exact region-of-interest posterior (qb_roi.hip) -----------------------------------
// Couplings symmetric with a zero diagonal (Js, n x n), chunks of 2^k states per thread.
int roi_nmax(int n);
hipError_t launch_roi_min(int n, int k, const double* h, const double* Js, uint64_t chunk0, uint64_t chunks,
                          double* blk_e, uint64_t* blk_s, hipStream_t st);
// blk: per block NMAX weighted bit sums and Z (roi_nmax(n) + 1 doubles).
hipError_t launch_roi_sum(int n, int k, const double* h, const double* Js, double beta, double e_ref, uint64_t chunk0,
                          uint64_t chunks, double* blk, hipStream_t st);

}  // namespace qb
