// qb_qubo.hip -- gfx950 kernels of the batched sparse-QUBO engine.
//
// A batch holds many independent QUBOs, E(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j
// over x in {0,1}^n, each with its own replicas. The work unit is a job: one instance
// and a block of RB of its replicas. Jobs share nothing, so the persistent kernel takes
// them from a queue (longest first) with no inter-workgroup synchronisation, and a
// workgroup anneals its job alone: every sweep visits the instance's colour classes in
// order, a class is a set of independent flips, and the workgroup syncs between classes.
//
// Inside a class each bit is RB lanes, one per replica: the lanes walk the bit's CSR row
// together, so a row's columns and couplings are read once for all RB replicas, and each
// lane picks its replica's bit out of the neighbour's spin word. The spin words of a job
// (bit r of word i = replica r0 + r's x_i) live in global memory and stay in L2.
//
// dE of flipping bit i = (1 - 2 x_i)(h_i + sum_j J_ij x_j), summed in FP64 in CSR order;
// the draw is Philox4x32-10 keyed by the replica seed (instance seed + replica + 1) with
// counter (bit, sweep), and a flip is accepted iff dE <= 0 or u < qexp(-dE * beta) --
// the same test and draws as the ray x voxel engine.
#include "qb_internal.h"
#include "qb_rng.h"

namespace qb {

template <int RB>
__global__ void __launch_bounds__(kQuboBlock) qubo_batch_kernel(DevQuboBatch q, uint32_t* masks, const double* beta,
                                                                 int n_sweeps, unsigned* queue, int job_first,
                                                                 int job_count) {
    constexpr int kBitsPerPass = kQuboBlock / RB;
    __shared__ int s_job;
    const int lane_r = int(threadIdx.x) % RB;
    const int slot = int(threadIdx.x) / RB;
    const unsigned group = unsigned(threadIdx.x & 63u) & ~unsigned(RB - 1);
    for (;;) {
        if (threadIdx.x == 0) s_job = int(atomicAdd(queue, 1u));
        __syncthreads();
        const int jn = s_job;
        __syncthreads();   // s_job is written again by the next take
        if (jn >= job_count) return;
        const QuboJob jb = q.jobs[job_first + jn];
        const int k = jb.instance, rb = jb.replicas;
        uint32_t* mk = masks + jb.mask_offset;
        const int bit0 = q.inst_pos[k];
        const int c0 = q.inst_cls[k], c1 = q.inst_cls[k + 1];
        const uint64_t key = q.seed[k] + uint64_t(jb.replica0) + uint64_t(lane_r) + 1u;
        const bool live = lane_r < rb;
        for (int sweep = 0; sweep < n_sweeps; ++sweep) {
            const double b = beta[sweep];
            for (int c = c0; c + 1 < c1; ++c) {
                const int lo = q.cls_ptr[c], hi = q.cls_ptr[c + 1];
                for (int pos = lo + slot; pos < hi; pos += kBitsPerPass) {
                    const uint32_t mi = mk[pos - bit0];
                    double acc = double(q.h[pos]);
                    const int e0 = q.row_ptr[pos], e1 = q.row_ptr[pos + 1];
                    for (int e = e0; e < e1; ++e) {
                        const uint32_t mj = mk[q.col[e] - bit0];
                        const double nxt = acc + double(q.J[e]);
                        acc = ((mj >> lane_r) & 1u) ? nxt : acc;
                    }
                    const double dE = ((mi >> lane_r) & 1u) ? -acc : acc;
                    const double u = draw_u01(key, uint32_t(sweep), uint32_t(q.bit_id[pos]));
                    const bool flip = live && (dE <= 0.0 || u < qexp(-dE * b));
                    const uint64_t bal = __ballot(flip);
                    const uint32_t flips = uint32_t(bal >> group) & (RB == 32 ? 0xffffffffu : ((1u << RB) - 1u));
                    if (lane_r == 0 && flips) mk[pos - bit0] = mi ^ flips;
                }
                __syncthreads();
            }
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

hipError_t launch_qubo_batch(const DevQuboBatch& q, int rb, uint32_t* masks, const double* beta, int n_sweeps,
                             unsigned* queue, int job_first, int job_count, int workgroups, hipStream_t st) {
    switch (rb) {
        case 4: qubo_batch_kernel<4><<<workgroups, kQuboBlock, 0, st>>>(q, masks, beta, n_sweeps, queue, job_first, job_count); break;
        case 8: qubo_batch_kernel<8><<<workgroups, kQuboBlock, 0, st>>>(q, masks, beta, n_sweeps, queue, job_first, job_count); break;
        case 16: qubo_batch_kernel<16><<<workgroups, kQuboBlock, 0, st>>>(q, masks, beta, n_sweeps, queue, job_first, job_count); break;
        case 32: qubo_batch_kernel<32><<<workgroups, kQuboBlock, 0, st>>>(q, masks, beta, n_sweeps, queue, job_first, job_count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_qubo_energy(const DevQuboBatch& q, const uint32_t* masks, double* energy, hipStream_t st) {
    qubo_energy_kernel<<<q.n_jobs, kQuboBlock, 0, st>>>(q, masks, energy);
    return hipGetLastError();
}

}  // namespace qb
