// qb_dense.hip -- gfx950 kernel of the dense-QUBO field engine (the precision study's
// dense instance).
//
// E(x) = sum_i h_i x_i + sum_{i<j} Q_ij x_i x_j with Q dense, symmetric, zero diagonal.
// Every bit couples to every other, so a sweep is sequential: bits in index order. The
// engine keeps the local fields f_i = h_i + sum_j Q_ij x_j and updates them on every
// flip: f += delta_t Q_t (delta_t = +1 for 0 -> 1, -1 for 1 -> 0). One workgroup anneals
// one replica with x and f in LDS:
//
//   - a window of 64 bits is decided by one wave, one bit per lane: the lanes prefetch the
//     window's 64 x 64 block of Q into registers and their draws, then step t = 0..63 lets
//     lane t decide (dE = (1 - 2 x_t) f_t) and, on a flip, every lane adds Q_{t,lane}
//     delta_t to its field -- so later bits of the window see earlier flips;
//   - the whole workgroup then adds the window's flipped rows, in flip order, to every field
//     outside the window (a rank-k update, k = flips in the window).
//
// Every field therefore receives the same additions in the same order as the sequential
// sweep. Fields are FP64, or FP32 carried incrementally (one float add of +-Q per flip) with
// an FP64 recompute f = h + Q x every k sweeps. A flip is accepted iff dE <= 0 or
// dE * beta < -qlog(u), u from Philox keyed by (replica seed, sweep, bit): the Metropolis
// test written with the log, so the sequential steps need no exponential.
#include "qb_internal.h"
#include "qb_rng.h"

namespace qb {

constexpr int kDenseWin = 64;
constexpr int kRankChunk = 8;
constexpr int kFieldGroup = 4;   // fields a thread updates together (i, i + block, ...)

// Fields i = g0 + q kDenseBlock (q < kFieldGroup, i < n) from scratch:
// h_i + sum_j x_j Q_ji in FP64, j ascending (Q_ji = Q_ij), additions of unset bits
// skipped; every field's loads of a 16-column chunk in flight together.
__device__ __forceinline__ void scratch_fields(const DevDense& d, const uint8_t* x, int g0, double (&s)[kFieldGroup]) {
    constexpr int kChunk = 16;
    const int n = d.n;
    int col[kFieldGroup];
#pragma unroll
    for (int q = 0; q < kFieldGroup; ++q) {
        col[q] = min(g0 + q * kDenseBlock, n - 1);
        s[q] = double(d.h[col[q]]);
    }
    for (int j0 = 0; j0 < n; j0 += kChunk) {
        float v[kFieldGroup][kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
#pragma unroll
            for (int q = 0; q < kFieldGroup; ++q) v[q][u] = d.Q[size_t(min(j0 + u, n - 1)) * size_t(n) + size_t(col[q])];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            const bool set = j0 + u < n && x[min(j0 + u, n - 1)];
#pragma unroll
            for (int q = 0; q < kFieldGroup; ++q) {
                const double nxt = s[q] + double(v[q][u]);
                s[q] = set ? nxt : s[q];
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kDenseBlock) dense_anneal_kernel(DevDense d, const double* beta, int n_sweeps,
                                                                    int recompute_every, uint64_t seed_offset,
                                                                    uint8_t* x_out, double* energy, double* field_err) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = d.n;
    T* f = reinterpret_cast<T*>(smem);
    uint8_t* x = reinterpret_cast<uint8_t*>(f + n);
    __shared__ int s_k;
    __shared__ int s_t[kDenseWin];
    __shared__ float s_delta[kDenseWin];
    __shared__ double s_red[kDenseBlock / 64][2];
    const int tid = int(threadIdx.x);
    const uint64_t key = seed_offset + uint64_t(blockIdx.x) + 1u;
    for (int i = tid; i < n; i += kDenseBlock) {
        x[i] = 0;
        f[i] = T(d.h[i]);
    }
    __syncthreads();
    for (int sweep = 0; sweep < n_sweeps; ++sweep) {
        const double b = beta[sweep];
        for (int i0 = 0; i0 < n; i0 += kDenseWin) {
            const int w = min(kDenseWin, n - i0);
            if (tid < 64) {
                const int lane = tid;
                const bool live = lane < w;
                const int i = i0 + (live ? lane : 0);
                float qb[kDenseWin];
#pragma unroll
                for (int t = 0; t < kDenseWin; ++t) qb[t] = d.Q[size_t(i0 + min(t, w - 1)) * size_t(n) + size_t(i)];
                const double L = -qlog(draw_u01(key, uint32_t(sweep), uint32_t(i)));
                T fl = f[i];
                int xl = x[i];
                int k = 0;
#pragma unroll
                for (int t = 0; t < kDenseWin; ++t) {
                    if (t < w) {
                        const double dE = xl ? -double(fl) : double(fl);
                        const bool flip = lane == t && (dE <= 0.0 || dE * b < L);
                        if (__ballot(flip)) {
                            const float dl = __shfl(xl ? -1.0f : 1.0f, t);
                            fl = fl + T(qb[t]) * T(dl);
                            if (lane == t) xl ^= 1;
                            if (lane == 0) {
                                s_t[k] = t;
                                s_delta[k] = dl;
                            }
                            ++k;
                        }
                    }
                }
                if (live) {
                    f[i] = fl;
                    x[i] = uint8_t(xl);
                }
                if (lane == 0) s_k = k;
            }
            __syncthreads();
            const int k = s_k;
            if (k > 0) {
                // kRankChunk flipped rows at a time: their loads in flight together, the
                // additions in flip order
                for (int g0 = tid; g0 < n; g0 += kFieldGroup * kDenseBlock) {
                    int col[kFieldGroup];
                    bool upd[kFieldGroup];
                    T acc[kFieldGroup];
#pragma unroll
                    for (int q = 0; q < kFieldGroup; ++q) {
                        const int i = g0 + q * kDenseBlock;
                        upd[q] = i < n && !(i >= i0 && i < i0 + w);
                        col[q] = min(i, n - 1);
                        acc[q] = f[col[q]];
                    }
                    for (int m0 = 0; m0 < k; m0 += kRankChunk) {
                        float v[kFieldGroup][kRankChunk];
#pragma unroll
                        for (int u = 0; u < kRankChunk; ++u) {
                            const size_t row = size_t(i0 + s_t[min(m0 + u, k - 1)]) * size_t(n);
#pragma unroll
                            for (int q = 0; q < kFieldGroup; ++q) v[q][u] = d.Q[row + size_t(col[q])];
                        }
#pragma unroll
                        for (int u = 0; u < kRankChunk; ++u) {
                            const T dl = T(s_delta[min(m0 + u, k - 1)]);
#pragma unroll
                            for (int q = 0; q < kFieldGroup; ++q) {
                                const T nxt = acc[q] + T(v[q][u]) * dl;
                                acc[q] = m0 + u < k ? nxt : acc[q];
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < kFieldGroup; ++q)
                        if (upd[q]) f[g0 + q * kDenseBlock] = acc[q];
                }
            }
            __syncthreads();
        }
        if (recompute_every > 0 && (sweep + 1) % recompute_every == 0) {
            // f = h + Q x from scratch in FP64, j ascending (column i of Q = row i)
            for (int g0 = tid; g0 < n; g0 += kFieldGroup * kDenseBlock) {
                double sc[kFieldGroup];
                scratch_fields(d, x, g0, sc);   // reads x and Q only, so fields may change meanwhile
#pragma unroll
                for (int q = 0; q < kFieldGroup; ++q)
                    if (g0 + q * kDenseBlock < n) f[g0 + q * kDenseBlock] = T(sc[q]);
            }
            __syncthreads();
        }
    }
    // outputs: spins, the FP64 energy and max |f_carried - f_scratch|
    double e = 0.0, err = 0.0;
    for (int g0 = tid; g0 < n; g0 += kFieldGroup * kDenseBlock) {
        double sc[kFieldGroup];
        scratch_fields(d, x, g0, sc);
#pragma unroll
        for (int q = 0; q < kFieldGroup; ++q) {
            const int i = g0 + q * kDenseBlock;
            if (i >= n) continue;
            err = fmax(err, fabs(double(f[i]) - sc[q]));
            if (x[i]) e += double(d.h[i]) + 0.5 * (sc[q] - double(d.h[i]));
            x_out[size_t(blockIdx.x) * size_t(n) + size_t(i)] = x[i];
        }
    }
    for (int o = 32; o > 0; o /= 2) {
        e += __shfl_down(e, o);
        err = fmax(err, __shfl_down(err, o));
    }
    if ((tid & 63) == 0) {
        s_red[tid / 64][0] = e;
        s_red[tid / 64][1] = err;
    }
    __syncthreads();
    if (tid == 0) {
        double es = 0.0, em = 0.0;
        for (int wv = 0; wv < kDenseBlock / 64; ++wv) {
            es += s_red[wv][0];
            em = fmax(em, s_red[wv][1]);
        }
        energy[blockIdx.x] = es;
        field_err[blockIdx.x] = em;
    }
}

size_t dense_lds_bytes(int n, int precision) {
    return size_t(n) * (precision == 1 ? sizeof(float) : sizeof(double)) + size_t(n);
}

hipError_t launch_dense(const DevDense& d, int precision, int R, const double* beta, int n_sweeps,
                        int recompute_every, uint64_t seed_offset, uint8_t* x_out, double* energy, double* field_err,
                        hipStream_t st) {
    const size_t lds = dense_lds_bytes(d.n, precision);
    if (precision == 1) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dense_anneal_kernel<float>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        dense_anneal_kernel<float><<<R, kDenseBlock, lds, st>>>(d, beta, n_sweeps, recompute_every, seed_offset, x_out,
                                                                energy, field_err);
    } else {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dense_anneal_kernel<double>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        dense_anneal_kernel<double><<<R, kDenseBlock, lds, st>>>(d, beta, n_sweeps, recompute_every, seed_offset, x_out,
                                                                 energy, field_err);
    }
    return hipGetLastError();
}

}  // namespace qb
