// qb_qubo.hip -- gfx950 kernels of the batched sparse-QUBO engine.
//
// A batch holds many independent QUBOs, E(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j
// over x in {0,1}^n, each with its own replicas. The work unit is a job: one instance
// and a block of RB of its replicas. Jobs share nothing, so the persistent kernel takes
// them from queues with no inter-workgroup synchronisation, and a workgroup anneals its
// job alone. Each instance belongs to one XCD's queue (instances spread over the eight
// by decreasing cost), its jobs side by side, so the jobs of an instance run together on
// one XCD and read its rows through that XCD's L2; a workgroup whose queue is empty takes
// from the others. Inside a job: every sweep visits the instance's colour classes in
// order, a class is a set of independent flips, and the workgroup syncs between classes.
//
// Inside a class each lane takes one bit for all the job's replicas: it walks the bit's
// CSR row once, kRowChunk neighbours at a time (columns and couplings, then the
// neighbours' spin words, all in flight together), and adds each coupling into the
// accumulator of every replica whose neighbour spin is set -- in row order, per replica.
// The spin words of a job (bit r of word i = replica r0 + r's x_i) live in the
// workgroup's LDS while it anneals (global memory only for instances too large for it),
// so the random reads of neighbours' spins never leave the CU.
//
// dE of flipping bit i = (1 - 2 x_i)(h_i + sum_j J_ij x_j), summed in FP64 in CSR order,
// and a flip is accepted iff dE <= 0 or u < qexp(-dE * beta) -- the ray x voxel engine's
// test, decided through a float exponential with a guard band (uphill_accept). Replicas 2m and 2m + 1 share one Philox4x32-10 call keyed by the instance seed with
// counter (bit, sweep, m, tag), one 53-bit uniform each (draw2_u01).
#include <algorithm>

#include "qb_internal.h"
#include "qb_rng.h"

namespace qb {

constexpr int kRowChunk = 16;

// Spin words of a job in LDS: one byte per bit up to 8 replicas per job, 16 or 32 bits
// beyond; kQuboLdsBytes of them hold instances up to kQuboLdsBytes / sizeof(word) bits.
template <int RB>
struct LdsWord {
    using type = uint32_t;
};
template <>
struct LdsWord<4> {
    using type = uint8_t;
};
template <>
struct LdsWord<8> {
    using type = uint8_t;
};
template <>
struct LdsWord<16> {
    using type = uint16_t;
};

// One job's sweeps with its spin words at `mk` (LDS or global), local position index.
template <int RB, int BLOCK, typename W>
__device__ __forceinline__ void qubo_job(const DevQuboBatch& q, const QuboJob& jb, W* mk, const double* beta,
                                         int n_sweeps) {
    const int k = jb.instance, rb = jb.replicas;
    const int bit0 = q.inst_pos[k];
    const int c0 = q.inst_cls[k], c1 = q.inst_cls[k + 1];
    const uint64_t key = q.seed[k];
    const uint32_t pair0 = uint32_t(jb.replica0) / 2u;   // replica0 is a multiple of RB
    for (int sweep = 0; sweep < n_sweeps; ++sweep) {
        const double b = beta[sweep];
        for (int c = c0; c + 1 < c1; ++c) {
            const int lo = q.cls_ptr[c], hi = q.cls_ptr[c + 1];
            for (int pos = lo + int(threadIdx.x); pos < hi; pos += BLOCK) {
                const uint32_t mi = mk[pos - bit0];
                const double hp = double(q.h[pos]);
                double acc[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) acc[r] = hp;
                const int deg = q.row_ptr[pos + 1] - q.row_ptr[pos];
                const int* tc = q.tcol + q.tile_ptr[q.class_tile[c] + (pos - lo) / 64] + ((pos - lo) & 63);
                const float* tw = q.tJ + (tc - q.tcol);
                for (int ea = 0; ea < deg; ea += kRowChunk) {
                    int cj[kRowChunk];
                    float wj[kRowChunk];
#pragma unroll
                    for (int u = 0; u < kRowChunk; ++u) {
                        const int e = min(ea + u, deg - 1);
                        cj[u] = tc[64 * e];
                        wj[u] = tw[64 * e];
                    }
                    uint32_t mj[kRowChunk];
#pragma unroll
                    for (int u = 0; u < kRowChunk; ++u) mj[u] = ea + u < deg ? uint32_t(mk[cj[u] - bit0]) : 0u;
                    // acc += w where the neighbour's spin is set, as fma(w, x_j, acc): w * 1 and
                    // w * 0 are exact, so each addition rounds once as acc + w does, and an
                    // unset neighbour leaves acc (up to the sign of an exact zero, which no
                    // decision sees) -- three instructions per (entry, replica), no select
#pragma unroll
                    for (int u = 0; u < kRowChunk; ++u) {
                        const double w = double(wj[u]);
#pragma unroll
                        for (int r = 0; r < RB; ++r) acc[r] = fma(w, double((mj[u] >> r) & 1u), acc[r]);
                    }
                }
                const uint32_t bit = uint32_t(q.bit_id[pos]);
                uint32_t flips = 0;
#pragma unroll
                for (int r = 0; r < RB; r += 2) {
                    if (r < rb) {
                        double u[2];
                        draw2_u01(key, uint32_t(sweep), bit, pair0 + uint32_t(r / 2), u[0], u[1]);
#pragma unroll
                        for (int v = 0; v < 2; ++v) {
                            const double dE = ((mi >> (r + v)) & 1u) ? -acc[r + v] : acc[r + v];
                            if (r + v < rb && (dE <= 0.0 || uphill_accept(-dE * b, u[v]))) flips |= 1u << (r + v);
                        }
                    }
                }
                if (flips) mk[pos - bit0] = W(mi ^ flips);
            }
            __syncthreads();
        }
    }
}

// The next job for this workgroup: from its own XCD's queue, else from the others in turn
// (a queue already read as exhausted costs no atomic); -1 when every queue is empty.
__device__ __forceinline__ int take_job(unsigned* queue, const QuboQueues& segs, int home) {
    for (int t = 0; t < kXcds; ++t) {
        const int x = (home + t) % kXcds;
        const int len = segs.begin[x + 1] - segs.begin[x];
        unsigned* ctr = queue + kQueueStride * x;
        if (len <= 0 || __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= unsigned(len)) continue;
        const unsigned k = atomicAdd(ctr, 1u);
        if (k < unsigned(len)) return segs.begin[x] + int(k);
    }
    return -1;
}

template <int RB, int BLOCK>
__global__ void __launch_bounds__(BLOCK) qubo_batch_kernel(DevQuboBatch q, uint32_t* masks, const double* beta,
                                                            int n_sweeps, unsigned* queue, QuboQueues segs) {
    static_assert(RB % 2 == 0, "replicas are drawn in pairs");
    using W = typename LdsWord<RB>::type;
    constexpr int kLdsBits = kQuboLdsBytes / int(sizeof(W));
    __shared__ W lmask[kLdsBits];
    __shared__ int s_job;
    const int home = int(xcc_id() % kXcds);
    for (;;) {
        if (threadIdx.x == 0) s_job = take_job(queue, segs, home);
        __syncthreads();
        const int jn = s_job;
        __syncthreads();   // s_job is written again by the next take
        if (jn < 0) return;
        const QuboJob jb = q.jobs[jn];
        const int n = q.inst_pos[jb.instance + 1] - q.inst_pos[jb.instance];
        uint32_t* gm = masks + jb.mask_offset;
        if (n <= kLdsBits) {
            // the job's spins live in LDS from all-zero x; written out once at the end
            for (int i = int(threadIdx.x); i < n; i += BLOCK) lmask[i] = W(0);
            __syncthreads();
            qubo_job<RB, BLOCK, W>(q, jb, lmask, beta, n_sweeps);
            for (int i = int(threadIdx.x); i < n; i += BLOCK) gm[i] = uint32_t(lmask[i]);
            __syncthreads();
        } else {
            qubo_job<RB, BLOCK, uint32_t>(q, jb, gm, beta, n_sweeps);
        }
    }
}

// FP64 energy of every replica of every job: one workgroup per job, thread t sums
// positions t, t + kQuboBlock, ... for each of the job's replicas, then a tree
// reduction per replica.
__global__ void __launch_bounds__(kQuboBlock) qubo_energy_kernel(DevQuboBatch q, const uint32_t* masks,
                                                                 double* energy) {
    __shared__ double part[kQuboBlock];
    const QuboJob jb = q.jobs[blockIdx.x];
    const int k = jb.instance;
    const uint32_t* mk = masks + jb.mask_offset;
    const int bit0 = q.inst_pos[k], bit1 = q.inst_pos[k + 1];
    for (int r = 0; r < jb.replicas; ++r) {
        double e = 0.0;
        for (int pos = bit0 + int(threadIdx.x); pos < bit1; pos += kQuboBlock) {
            if (!((mk[pos - bit0] >> r) & 1u)) continue;
            double f = double(q.h[pos]);
            double pair = 0.0;
            for (int i = q.row_ptr[pos]; i < q.row_ptr[pos + 1]; ++i)
                if ((mk[q.col[i] - bit0] >> r) & 1u) pair += double(q.J[i]);
            e += f + 0.5 * pair;
        }
        part[threadIdx.x] = e;
        __syncthreads();
        for (int w = kQuboBlock / 2; w > 0; w /= 2) {
            if (int(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) energy[jb.energy_offset + r] = part[0];
        __syncthreads();
    }
}

template <int BLOCK>
static hipError_t launch_qubo_block(const DevQuboBatch& q, int rb, uint32_t* masks, const double* beta, int n_sweeps,
                                    unsigned* queue, const QuboQueues& segs, int workgroups, hipStream_t st) {
    switch (rb) {
        case 4: qubo_batch_kernel<4, BLOCK><<<workgroups, BLOCK, 0, st>>>(q, masks, beta, n_sweeps, queue, segs); break;
        case 8: qubo_batch_kernel<8, BLOCK><<<workgroups, BLOCK, 0, st>>>(q, masks, beta, n_sweeps, queue, segs); break;
        case 16: qubo_batch_kernel<16, BLOCK><<<workgroups, BLOCK, 0, st>>>(q, masks, beta, n_sweeps, queue, segs); break;
        case 32: qubo_batch_kernel<32, BLOCK><<<workgroups, BLOCK, 0, st>>>(q, masks, beta, n_sweeps, queue, segs); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_qubo_batch(const DevQuboBatch& q, int rb, int block, uint32_t* masks, const double* beta,
                             int n_sweeps, unsigned* queue, const QuboQueues& segs, int workgroups, hipStream_t st) {
    for (int x = 0; x < kXcds; ++x)
        if (segs.begin[x] < 0 || segs.begin[x + 1] < segs.begin[x] || segs.begin[x + 1] > q.n_jobs)
            return hipErrorInvalidValue;
    if (block == 256) return launch_qubo_block<256>(q, rb, masks, beta, n_sweeps, queue, segs, workgroups, st);
    if (block == 1024) return launch_qubo_block<1024>(q, rb, masks, beta, n_sweeps, queue, segs, workgroups, st);
    return hipErrorInvalidValue;
}

int qubo_blocks_per_cu(int rb, int block) {
    int per_cu = 0;
    const void* k = nullptr;
    if (block == 256)
        k = rb == 4 ? (const void*)&qubo_batch_kernel<4, 256> : rb == 8 ? (const void*)&qubo_batch_kernel<8, 256>
          : rb == 16 ? (const void*)&qubo_batch_kernel<16, 256> : (const void*)&qubo_batch_kernel<32, 256>;
    else
        k = rb == 4 ? (const void*)&qubo_batch_kernel<4, 1024> : rb == 8 ? (const void*)&qubo_batch_kernel<8, 1024>
          : rb == 16 ? (const void*)&qubo_batch_kernel<16, 1024> : (const void*)&qubo_batch_kernel<32, 1024>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0) != hipSuccess) return 0;
    return per_cu;
}

// A fill kernel in place of a memset where the launch schedule is recorded into a graph.
__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    fill_u32_kernel<<<unsigned(blocks), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_qubo_energy(const DevQuboBatch& q, const uint32_t* masks, double* energy, hipStream_t st) {
    qubo_energy_kernel<<<q.n_jobs, kQuboBlock, 0, st>>>(q, masks, energy);
    return hipGetLastError();
}

}  // namespace qb
