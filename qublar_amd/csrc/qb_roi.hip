// qb_roi.hip -- gfx950 kernels of the exact region-of-interest posterior.
//
// A small QUBO E(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j over n <= 36 bits (the
// conditional QUBO of roi_qubo, src/ising_recon.hpp:245-300) is enumerated over all 2^n
// states, as tools/branches_blaze.py:45-53 does on the host: the minimum energy and its
// state, then Z = sum exp(-beta (E - E_min)) and every bit's weight sum_x w(x) x_i.
//
// One thread owns a chunk of 2^k states: state index s = (c << k) | g(t), with c the
// thread's chunk (the high bits, fixed) and g(t) = t ^ (t >> 1) the Gray code of the step t,
// so step t flips bit ctz(t) -- the same bit in every lane of a wave, whose coupling row
// is then one LDS broadcast. The thread carries the fields f_j = h_j + sum_i J_ji x_i (FP64,
// from scratch at the chunk's start): flipping bit b changes E by (1 - 2 x_b) f_b and every
// f_j by J_bj (x_b' - x_b). Chunks are at most 2^12 states, so the carried energy stays
// within ~1e-13 of a scratch sum. Per-block results go to global memory; the host reduces
// them in block order, so a result does not depend on anything but n, k and the launch
// split (both fixed by n).
//
// The step loop is unrolled by 2^kRoiUnrollLog: within a group, step i > 0 flips bit
// ctz(i), known at compile time, so only a group's first step selects its field and
// coupling row by a runtime index. Pass 2 sums a group's weights first and adds that sum
// to the bits the group does not change (chunks shorter than a group run the plain loop).
#include "qb_internal.h"

namespace qb {

constexpr int kRoiBlock = 256;

// Symmetric couplings (zero diagonal) and h, NMAX-strided, into LDS.
template <int NMAX>
__device__ __forceinline__ void roi_stage(int n, const double* h, const double* Js, double* sJ, double* sh) {
    for (int i = int(threadIdx.x); i < NMAX * NMAX; i += kRoiBlock) {
        const int r = i / NMAX, c = i % NMAX;
        sJ[i] = r < n && c < n ? Js[size_t(r) * size_t(n) + size_t(c)] : 0.0;
    }
    for (int i = int(threadIdx.x); i < NMAX; i += kRoiBlock) sh[i] = i < n ? h[i] : 0.0;
    __syncthreads();
}

// The chunk's first state x = c << k: its fields and energy from scratch.
template <int NMAX>
__device__ __forceinline__ double roi_start(int n, uint64_t x, const double* sJ, const double* sh, double (&f)[NMAX]) {
    double e = 0.0;
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
        double fj = sh[j];
        double below = 0.0;   // sum over i < j of J_ji x_i
        for (int i = 0; i < n; ++i) {
            const double v = ((x >> i) & 1u) ? sJ[j * NMAX + i] : 0.0;
            fj += v;
            below += i < j ? v : 0.0;
        }
        f[j] = fj;
        if (j < n && ((x >> j) & 1u)) e += sh[j] + below;
    }
    return e;
}

// One Gray step from step t - 1 to t: flips b = ctz(t); returns the new energy.
template <int NMAX>
__device__ __forceinline__ double roi_step(uint64_t& x, int b, double e, const double* sJ, double (&f)[NMAX]) {
    double fb = 0.0;
#pragma unroll
    for (int j = 0; j < NMAX; ++j) fb = j == b ? f[j] : fb;
    const bool on = (x >> b) & 1u;
    e += on ? -fb : fb;
    x ^= uint64_t(1) << b;
    const double d = on ? -1.0 : 1.0;
    const double* row = sJ + b * NMAX;
    // row[j] * d is exact (d = +-1), so the fused form rounds as f + row * d does
#pragma unroll
    for (int j = 0; j < NMAX; ++j) f[j] = fma(row[j], d, f[j]);
    return e;
}

constexpr int kRoiUnrollLog = 3;   // unrolled steps per group: 2^kRoiUnrollLog
constexpr int kRoiUnroll = 1 << kRoiUnrollLog;

template <int NMAX>
__global__ void __launch_bounds__(kRoiBlock) roi_min_kernel(int n, int k, const double* h, const double* Js,
                                                             uint64_t chunk0, uint64_t chunks, double* blk_e,
                                                             uint64_t* blk_s) {
    __shared__ double sJ[NMAX * NMAX];
    __shared__ double sh[NMAX];
    __shared__ double re[kRoiBlock];
    __shared__ uint64_t rs[kRoiBlock];
    roi_stage<NMAX>(n, h, Js, sJ, sh);
    const uint64_t local = uint64_t(blockIdx.x) * kRoiBlock + threadIdx.x;
    const bool live = local < chunks;
    uint64_t x = (chunk0 + (live ? local : 0)) << k;
    double f[NMAX];
    double e = roi_start<NMAX>(n, x, sJ, sh, f);
    double best = e;
    uint64_t best_s = x;
    const uint64_t steps = uint64_t(1) << k;
    auto keep = [&]() {
        if (e < best || (e == best && x < best_s)) {
            best = e;
            best_s = x;
        }
    };
    if (k >= kRoiUnrollLog) {
        for (uint64_t tb = 0; tb < (steps >> kRoiUnrollLog); ++tb) {
            if (tb > 0) {
                e = roi_step<NMAX>(x, kRoiUnrollLog + __builtin_ctzll(tb), e, sJ, f);
                keep();
            }
#pragma unroll
            for (int i = 1; i < kRoiUnroll; ++i) {
                e = roi_step<NMAX>(x, __builtin_ctz(unsigned(i)), e, sJ, f);
                keep();
            }
        }
    } else {
        for (uint64_t t = 1; t < steps; ++t) {
            e = roi_step<NMAX>(x, __builtin_ctzll(t), e, sJ, f);
            keep();
        }
    }
    re[threadIdx.x] = live ? best : __builtin_huge_val();
    rs[threadIdx.x] = live ? best_s : ~uint64_t(0);
    __syncthreads();
    for (int w = kRoiBlock / 2; w > 0; w /= 2) {
        if (int(threadIdx.x) < w) {
            const double eo = re[threadIdx.x + w];
            const uint64_t so = rs[threadIdx.x + w];
            if (eo < re[threadIdx.x] || (eo == re[threadIdx.x] && so < rs[threadIdx.x])) {
                re[threadIdx.x] = eo;
                rs[threadIdx.x] = so;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        blk_e[blockIdx.x] = re[0];
        blk_s[blockIdx.x] = rs[0];
    }
}

template <int NMAX>
__global__ void __launch_bounds__(kRoiBlock) roi_sum_kernel(int n, int k, const double* h, const double* Js,
                                                             double beta, double e_ref, uint64_t chunk0,
                                                             uint64_t chunks, double* blk /*[blocks][NMAX + 1]*/) {
    __shared__ double sJ[NMAX * NMAX];
    __shared__ double sh[NMAX];
    __shared__ double red[kRoiBlock];
    roi_stage<NMAX>(n, h, Js, sJ, sh);
    const uint64_t local = uint64_t(blockIdx.x) * kRoiBlock + threadIdx.x;
    const bool live = local < chunks;
    uint64_t x = (chunk0 + (live ? local : 0)) << k;
    double f[NMAX];
    double e = roi_start<NMAX>(n, x, sJ, sh, f);
    double z = 0.0;
    double m[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) m[j] = 0.0;
    auto add = [&]() {
        const double w = exp(-beta * (e - e_ref));
        z += w;
#pragma unroll
        for (int j = 0; j < NMAX; ++j) m[j] += ((x >> j) & 1u) ? w : 0.0;
    };
    const uint64_t steps = uint64_t(1) << k;
    if (k >= kRoiUnrollLog) {
        // Bits kRoiUnrollLog.. stay fixed over a group (its first step flips the last of
        // them to change): the group's weight sum goes to them once; only the low bits are
        // summed state by state.
        for (uint64_t tb = 0; tb < (steps >> kRoiUnrollLog); ++tb) {
            if (tb > 0) e = roi_step<NMAX>(x, kRoiUnrollLog + __builtin_ctzll(tb), e, sJ, f);
            double wg = 0.0, lo[kRoiUnrollLog];
#pragma unroll
            for (int l = 0; l < kRoiUnrollLog; ++l) lo[l] = 0.0;
#pragma unroll
            for (int i = 0; i < kRoiUnroll; ++i) {
                if (i > 0) e = roi_step<NMAX>(x, __builtin_ctz(unsigned(i)), e, sJ, f);
                const double w = exp(-beta * (e - e_ref));
                wg += w;
#pragma unroll
                for (int l = 0; l < kRoiUnrollLog; ++l) lo[l] += ((x >> l) & 1u) ? w : 0.0;
            }
            z += wg;
#pragma unroll
            for (int j = 0; j < NMAX; ++j) m[j] += j < kRoiUnrollLog ? lo[j < kRoiUnrollLog ? j : 0]
                                                                    : (((x >> j) & 1u) ? wg : 0.0);
        }
    } else {
        for (uint64_t t = 0; t < steps; ++t) {
            if (t > 0) e = roi_step<NMAX>(x, __builtin_ctzll(t), e, sJ, f);
            add();
        }
    }
    double* out = blk + size_t(blockIdx.x) * (NMAX + 1);
    for (int q = 0; q <= NMAX; ++q) {
        double v = q == NMAX ? z : 0.0;
#pragma unroll
        for (int j = 0; j < NMAX; ++j) v = j == q ? m[j] : v;
        red[threadIdx.x] = live ? v : 0.0;
        __syncthreads();
        for (int w = kRoiBlock / 2; w > 0; w /= 2) {
            if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = red[0];
        __syncthreads();
    }
}

int roi_nmax(int n) { return n <= 16 ? 16 : (n <= 24 ? 24 : (n <= 32 ? 32 : 36)); }

hipError_t launch_roi_min(int n, int k, const double* h, const double* Js, uint64_t chunk0, uint64_t chunks,
                          double* blk_e, uint64_t* blk_s, hipStream_t st) {
    const unsigned blocks = unsigned((chunks + kRoiBlock - 1) / kRoiBlock);
    switch (roi_nmax(n)) {
        case 16: roi_min_kernel<16><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, chunk0, chunks, blk_e, blk_s); break;
        case 24: roi_min_kernel<24><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, chunk0, chunks, blk_e, blk_s); break;
        case 32: roi_min_kernel<32><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, chunk0, chunks, blk_e, blk_s); break;
        default: roi_min_kernel<36><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, chunk0, chunks, blk_e, blk_s); break;
    }
    return hipGetLastError();
}

hipError_t launch_roi_sum(int n, int k, const double* h, const double* Js, double beta, double e_ref, uint64_t chunk0,
                          uint64_t chunks, double* blk, hipStream_t st) {
    const unsigned blocks = unsigned((chunks + kRoiBlock - 1) / kRoiBlock);
    switch (roi_nmax(n)) {
        case 16: roi_sum_kernel<16><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, beta, e_ref, chunk0, chunks, blk); break;
        case 24: roi_sum_kernel<24><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, beta, e_ref, chunk0, chunks, blk); break;
        case 32: roi_sum_kernel<32><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, beta, e_ref, chunk0, chunks, blk); break;
        default: roi_sum_kernel<36><<<blocks, kRoiBlock, 0, st>>>(n, k, h, Js, beta, e_ref, chunk0, chunks, blk); break;
    }
    return hipGetLastError();
}

}  // namespace qb
