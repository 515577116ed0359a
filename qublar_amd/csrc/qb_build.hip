// qb_build.hip -- problem builders on the device (gfx950): the muon exposure, the muon
// rows, their assembly into the bit-major CSR, the 6-neighbour lists, and the gravimetry
// rows. The host keeps only the O(rays) and O(stations) scalar work (bin directions,
// the noise of the gravimetry data) and the colouring.
//
//   muon_expose_kernel  src/muon.cuh:195-224: one hash-sampled candidate muon per thread,
//                       binned, marched through the true medium, detected with
//                       Beer-Lambert; open / detected counts per sky bin.
//   rows_count_kernel,  src/op_muon.hpp:22-50: per (view, sky bin) ray of the model
//   rows_write_kernel   medium, the optical depth and the crossed bits (a = a_scale * seg);
//                       rows that cross no bit are dropped; entries in row order.
//   assembly            src/engine.hpp:120-149: entries grouped by bit with a stable radix
//                       sort, so inside a bit they stay in row order.
//   grid_*_kernel       src/engine.hpp:61-85: bits over the domain cells (x fastest) and
//                       their 6-neighbour lists in the order -x, +x, -y, +y, -z, +z.
//   gravity_*_kernel    src/op_gravity.hpp:59-74 with rows truncated at a radius.
//
// The march is the reference's Amanatides-Woo stepping (src/muon.cuh:66-134) in float,
// with every operation as written there; with -ffp-contract=off it gives the host's
// float results bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "qb_host_problem.h"
#include "qublar.h"

namespace qb {
int build_fail(int code, const char* msg);
}

namespace {

struct Medium {
    float lo[3];
    float voxel;
    int nx, ny, nz;
    const float* mu;   // null: the march records cells only
};

// src/muon.cuh:66-134 in float. visit(cell, seg) per crossed cell with seg > 0.
template <typename Visit>
__host__ __device__ inline float march(const Medium& m, const float p[3], const float d[3], Visit&& visit) {
    const float eps = 1e-12f;
    float t0 = -1e30f, t1 = 1e30f;
    const int nc[3] = {m.nx, m.ny, m.nz};
    for (int a = 0; a < 3; ++a) {
        if (fabsf(d[a]) < eps) {
            if (p[a] < m.lo[a] || p[a] > m.lo[a] + float(nc[a]) * m.voxel) return 0.0f;
        } else {
            const float ta = (m.lo[a] - p[a]) / d[a];
            const float tb = (m.lo[a] + float(nc[a]) * m.voxel - p[a]) / d[a];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        }
    }
    if (t1 <= fmaxf(t0, 0.0f)) return 0.0f;
    const float t_start = fmaxf(t0, 0.0f);
    const float t_end = t1 - t_start;
    int vox[3];
    float tmax[3], tdelta[3];
    for (int a = 0; a < 3; ++a) vox[a] = int(floorf((p[a] + d[a] * t_start - m.lo[a]) / m.voxel));
    for (int a = 0; a < 3; ++a) {
        if (vox[a] < 0) vox[a] = 0;
        if (vox[a] >= nc[a]) vox[a] = nc[a] - 1;
        if (fabsf(d[a]) < eps) {
            tmax[a] = 1e30f;
            tdelta[a] = 0.0f;
        } else {
            const float boundary = (d[a] > 0.0f ? float(vox[a] + 1) : float(vox[a])) * m.voxel + m.lo[a];
            tmax[a] = (boundary - p[a]) / d[a];
            tdelta[a] = m.voxel / fabsf(d[a]);
            if (tmax[a] < 0.0f) tmax[a] = 0.0f;
        }
    }
    float tau = 0.0f, t_done = 0.0f;
    for (int steps = 0; steps < 3 * (m.nx + m.ny + m.nz) + 16; ++steps) {
        const int axis = (tmax[0] <= tmax[1]) ? ((tmax[0] <= tmax[2]) ? 0 : 2) : ((tmax[1] <= tmax[2]) ? 1 : 2);
        const float t_next = fminf(tmax[axis], t_end);
        const float seg = t_next - t_done;
        if (seg > 0.0f) {
            const int idx = (vox[2] * m.ny + vox[1]) * m.nx + vox[0];
            if (m.mu) tau += m.mu[idx] * seg;
            visit(idx, seg);
        }
        if (t_next >= t_end) break;
        t_done = t_next;
        vox[axis] += (d[axis] > 0.0f) ? 1 : -1;
        tmax[axis] += tdelta[axis];
        if (vox[axis] < 0 || vox[axis] >= nc[axis]) break;
    }
    return tau;
}

// src/lidar.cuh:90-101
__device__ inline uint32_t hash_u32(uint32_t x) {
    x = (x ^ 61u) ^ (x >> 16);
    x *= 9u;
    x = x ^ (x >> 4);
    x *= 0x27d4eb2du;
    x = x ^ (x >> 15);
    return x;
}
__device__ inline float hash_unit(uint32_t x) { return float(hash_u32(x) >> 8) * (1.0f / 16777216.0f); }

__global__ void muon_expose_kernel(Medium m, float cx, float cy, float cz, float c_min, float th_max, int n_az,
                                   int n_th, int64_t n_cand, uint32_t seed, unsigned long long* open,
                                   unsigned long long* det) {
    const float c[3] = {cx, cy, cz};
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n_cand;
         i += int64_t(gridDim.x) * blockDim.x) {
        const uint32_t key = hash_u32(uint32_t(i) * 0x9E3779B9u);
        const float u1 = hash_unit(key);
        const float u2 = hash_unit(key ^ 0x68BC21EBu);
        const float c3 = c_min * c_min * c_min;
        const float z = cbrtf(c3 + u1 * (1.0f - c3));
        const float s = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float az = 6.28318530718f * u2;
        const float d[3] = {s * cosf(az), s * sinf(az), z};
        const float th = acosf(fminf(1.0f, fmaxf(-1.0f, z)));
        int ith = int(th / th_max * float(n_th));
        ith = ith < 0 ? 0 : (ith >= n_th ? n_th - 1 : ith);
        int iaz = int(az * 0.15915494f * float(n_az));
        iaz = iaz < 0 ? 0 : (iaz >= n_az ? n_az - 1 : iaz);
        const int bin = ith * n_az + iaz;
        atomicAdd(open + bin, 1ull);
        const float tau = march(m, c, d, [](int, float) {});
        const float u3 = hash_unit(key ^ 0x2545F491u ^ seed);
        if (u3 < expf(-tau)) atomicAdd(det + bin, 1ull);
    }
}

// One candidate ray per (view, bin): its bit entries (0 when open < min_open) and the
// model medium's optical depth along it.
__global__ void rows_count_kernel(Medium model, const float* chambers, const float* dirs, int n_bins, int n_rays,
                                  const unsigned long long* open, double min_open, const int32_t* site_bit, int* cnt,
                                  float* tau0) {
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= n_rays) return;
    const int v = r / n_bins;
    int k = 0;
    float t = 0.0f;
    if (double(open[r]) >= min_open)
        t = march(model, chambers + 3 * v, dirs + 3 * size_t(r), [&](int idx, float) { k += site_bit[idx] >= 0; });
    cnt[r] = k;
    tau0[r] = t;
}

// The kept rays' entries (bit, row, a) at their offsets, and the rows' d and w.
__global__ void rows_write_kernel(Medium model, const float* chambers, const float* dirs, int n_bins, int n_rays,
                                  const unsigned long long* open, const unsigned long long* det, const int32_t* site_bit,
                                  const int* cnt, const float* tau0, const int* ent_off, const int* row_of, float a_scale,
                                  int32_t* ent_bit, int32_t* ent_row, float* ent_a, double* d, double* w) {
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= n_rays || cnt[r] == 0) return;
    const int v = r / n_bins;
    const int row = row_of[r];
    int e = ent_off[r];
    Medium cells = model;
    cells.mu = nullptr;
    march(cells, chambers + 3 * v, dirs + 3 * size_t(r), [&](int idx, float seg) {
        const int bit = site_bit[idx];
        if (bit >= 0) {
            ent_bit[e] = bit;
            ent_row[e] = row;
            ent_a[e] = a_scale * seg;
            ++e;
        }
    });
    const double dt = double(det[r]) > 1.0 ? double(det[r]) : 1.0;
    d[row] = double(tau0[r]) + log(dt / double(open[r]));
    w[row] = dt;
}

__global__ void keep_kernel(const int* cnt, int n, int* keep) {
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < n) keep[r] = cnt[r] > 0 ? 1 : 0;
}

__global__ void iota_kernel(int32_t* v, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) v[i] = i;
}

__global__ void gather_kernel(const int32_t* perm, const int32_t* ent_row, const float* ent_a, int n, int32_t* ray,
                              float* a) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    ray[k] = ent_row[perm[k]];
    a[k] = ent_a[perm[k]];
}

__global__ void bit_count_kernel(const int32_t* ent_bit, int n, int* count) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < n) atomicAdd(count + ent_bit[k], 1);
}

// BitField::grid (src/engine.hpp:61-85) over a domain flag per cell.
__global__ void grid_count_kernel(const int32_t* site, int n_bits, const int32_t* site_bit, int nx, int ny, int nz,
                                  int* cnt) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= n_bits) return;
    const int c = site[b];
    const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    const int cand[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
    int k = 0;
    for (int q = 0; q < 6; ++q) {
        const int* p = cand[q];
        if (p[0] < 0 || p[0] >= nx || p[1] < 0 || p[1] >= ny || p[2] < 0 || p[2] >= nz) continue;
        k += site_bit[(p[2] * ny + p[1]) * nx + p[0]] >= 0;
    }
    cnt[b] = k;
}

__global__ void grid_write_kernel(const int32_t* site, int n_bits, const int32_t* site_bit, int nx, int ny, int nz,
                                  const int* off, int32_t* nbr) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= n_bits) return;
    const int c = site[b];
    const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    const int cand[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
    int e = off[b];
    for (int q = 0; q < 6; ++q) {
        const int* p = cand[q];
        if (p[0] < 0 || p[0] >= nx || p[1] < 0 || p[1] >= ny || p[2] < 0 || p[2] >= nz) continue;
        const int u = site_bit[(p[2] * ny + p[1]) * nx + p[0]];
        if (u >= 0) nbr[e++] = u;
    }
}

// The muon scene's cells (the pyramid of the host builder): the true and the model
// attenuation, and the domain flag (rock outside 3 m rooms around the chambers).
__global__ void muon_cells_kernel(int N, float voxel, float lo0, float lo1, float lo2, float height, float half_base,
                                  float hv, float void_z, float mu_rock, float mu_air, const float* chambers,
                                  float* mu_true, float* mu_model, int* domain, uint8_t* is_void) {
    const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= N * N * N) return;
    const int i = c % N, j = (c / N) % N, k = c / (N * N);
    const float x = lo0 + (float(i) + 0.5f) * voxel;
    const float y = lo1 + (float(j) + 0.5f) * voxel;
    const float z = lo2 + (float(k) + 0.5f) * voxel;
    bool rock = false;
    if (z >= 0.0f && z <= height) {
        const float wz = half_base * (1.0f - z / height);
        rock = fabsf(x) <= wz && fabsf(y) <= wz;
    }
    const bool vd = rock && fabsf(x) <= hv && fabsf(y) <= hv && fabsf(z - void_z) <= hv;
    bool room = false;
    for (int q = 0; q < 3; ++q) {
        const float dx = x - chambers[3 * q], dy = y - chambers[3 * q + 1], dz = z - chambers[3 * q + 2];
        room |= dx * dx + dy * dy + dz * dz < 3.0f * 3.0f;
    }
    mu_model[c] = rock ? mu_rock : mu_air;
    mu_true[c] = rock && !vd ? mu_rock : mu_air;
    domain[c] = rock && !room ? 1 : 0;
    is_void[c] = vd ? 1 : 0;
}

__global__ void site_kernel(const int* domain, const int* pos, int cells, int32_t* site, int32_t* site_bit) {
    const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= cells) return;
    if (domain[c]) {
        site[pos[c]] = c;
        site_bit[c] = pos[c];
    } else {
        site_bit[c] = -1;
    }
}

// Gravimetry rows (op_gravity.hpp:59-74, truncated): per bit, the stations within the
// radius in row-major offset order, a = G drho h^3 dz / r^3 in microGal.
struct GravityGrid {
    int nx, ny, nz, reach;
    double h, radius2, mass;
};

__device__ inline bool grav_sees(const GravityGrid& g, int di, int dj, int k) {
    const double dx = di * g.h, dy = dj * g.h, dz = (k + 0.5) * g.h;
    return dx * dx + dy * dy + dz * dz <= g.radius2;
}

__global__ void gravity_count_kernel(GravityGrid g, int* cnt) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= g.nx * g.ny * g.nz) return;
    const int i = b % g.nx, j = (b / g.nx) % g.ny, k = b / (g.nx * g.ny);
    int c = 0;
    for (int dj = -g.reach; dj <= g.reach; ++dj)
        for (int di = -g.reach; di <= g.reach; ++di)
            c += grav_sees(g, di, dj, k) && i + di >= 0 && i + di < g.nx && j + dj >= 0 && j + dj < g.ny;
    cnt[b] = c;
}

__global__ void gravity_write_kernel(GravityGrid g, const int* off, int32_t* ray, float* a) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= g.nx * g.ny * g.nz) return;
    const int i = b % g.nx, j = (b / g.nx) % g.ny, k = b / (g.nx * g.ny);
    constexpr double kG = 6.67430e-11, kMicroGal = 1e-8;
    int e = off[b];
    for (int dj = -g.reach; dj <= g.reach; ++dj)
        for (int di = -g.reach; di <= g.reach; ++di) {
            if (!grav_sees(g, di, dj, k)) continue;
            const int si = i + di, sj = j + dj;
            if (si < 0 || si >= g.nx || sj < 0 || sj >= g.ny) continue;
            const double dx = di * g.h, dy = dj * g.h, dz = (k + 0.5) * g.h;
            const double r2 = dx * dx + dy * dy + dz * dz;
            ray[e] = sj * g.nx + si;
            a[e] = float(kG * g.mass * dz / (r2 * sqrt(r2)) / kMicroGal);
            ++e;
        }
}

// ---- host side ------------------------------------------------------------------------

#define BD_HIP(call)                                                                         \
    do {                                                                                     \
        const hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return qb::build_fail(QB_E_HIP, hipGetErrorString(e_));        \
    } while (0)

// Device buffers of one build, freed together.
struct Arena {
    std::vector<void*> ptrs;
    ~Arena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int get(T** out, size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess)
            return qb::build_fail(QB_E_NOMEM, "hipMalloc failed in a device builder");
        ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return QB_OK;
    }
};

unsigned blocks_for(size_t n, unsigned bs = 256) { return unsigned((n + bs - 1) / bs); }

// Exclusive prefix sum of n ints into out (length n + 1: out[n] = the total).
int exclusive_sum(Arena& A, const int* in, int* out, int n, hipStream_t st) {
    size_t tmp = 0;
    BD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n + 1, st));
    void* t = nullptr;
    if (int rc = A.get(reinterpret_cast<unsigned char**>(&t), tmp)) return rc;
    BD_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, in, out, n + 1, st));
    return QB_OK;
}

// A copy of `in` with one extra zero at the end (the exclusive sum's total slot).
int padded(Arena& A, const int* in, int n, int** out, hipStream_t st) {
    if (int rc = A.get(out, size_t(n) + 1)) return rc;
    BD_HIP(hipMemsetAsync(*out, 0, sizeof(int) * (size_t(n) + 1), st));
    if (n > 0) BD_HIP(hipMemcpyAsync(*out, in, sizeof(int) * size_t(n), hipMemcpyDeviceToDevice, st));
    return QB_OK;
}

template <typename T>
int to_host(std::vector<T>& v, const T* d, size_t n, hipStream_t st) {
    v.resize(n);
    if (n) BD_HIP(hipMemcpyAsync(v.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
    return QB_OK;
}

// The rows of n_views views (bins x views candidate rays) and their assembly into the
// bit-major CSR; neighbour lists from the domain's sites. Fills H's row_ptr, ray, a, d, w,
// nbr_ptr, nbr (n_vars and site arrays given on the device).
int rows_and_assembly(Arena& A, const Medium& model, const float* d_ch, const float* d_dirs, int n_bins, int n_views,
                      const unsigned long long* d_open, const unsigned long long* d_det, const int32_t* d_site_bit,
                      const int32_t* d_site, int n_bits, double min_open, float a_scale, qb_host_problem* H,
                      hipStream_t st) {
    const int n_cand = n_bins * n_views;
    // entries are counted and offset in int32: refuse a build that could pass 2^31
    if (int64_t(n_cand) * (3 * (int64_t(model.nx) + model.ny + model.nz) + 16) > int64_t(INT32_MAX))
        return qb::build_fail(QB_E_LIMIT, "muon rows: too many rays for int32 entry offsets");
    int *cnt, *keep, *cnt1, *keep1, *eoff, *rid;
    float* tau0;
    int rc;
    if ((rc = A.get(&cnt, n_cand)) || (rc = A.get(&keep, n_cand)) || (rc = A.get(&tau0, n_cand)) ||
        (rc = A.get(&eoff, size_t(n_cand) + 1)) || (rc = A.get(&rid, size_t(n_cand) + 1)))
        return rc;
    rows_count_kernel<<<blocks_for(n_cand), 256, 0, st>>>(model, d_ch, d_dirs, n_bins, n_cand, d_open, min_open,
                                                          d_site_bit, cnt, tau0);
    BD_HIP(hipGetLastError());
    keep_kernel<<<blocks_for(n_cand), 256, 0, st>>>(cnt, n_cand, keep);
    BD_HIP(hipGetLastError());
    if ((rc = padded(A, cnt, n_cand, &cnt1, st)) || (rc = padded(A, keep, n_cand, &keep1, st)) ||
        (rc = exclusive_sum(A, cnt1, eoff, n_cand, st)) || (rc = exclusive_sum(A, keep1, rid, n_cand, st)))
        return rc;
    int totals[2] = {0, 0};
    BD_HIP(hipMemcpyAsync(totals, eoff + n_cand, sizeof(int), hipMemcpyDeviceToHost, st));
    BD_HIP(hipMemcpyAsync(totals + 1, rid + n_cand, sizeof(int), hipMemcpyDeviceToHost, st));
    BD_HIP(hipStreamSynchronize(st));
    const int nnz = totals[0], n_rows = totals[1];
    int32_t *ent_bit, *ent_row, *perm, *perm_sorted, *bit_sorted, *ray;
    float *ent_a, *a;
    double *d, *w;
    if ((rc = A.get(&ent_bit, nnz)) || (rc = A.get(&ent_row, nnz)) || (rc = A.get(&ent_a, nnz)) ||
        (rc = A.get(&perm, nnz)) || (rc = A.get(&perm_sorted, nnz)) || (rc = A.get(&bit_sorted, nnz)) ||
        (rc = A.get(&ray, nnz)) || (rc = A.get(&a, nnz)) || (rc = A.get(&d, n_rows)) || (rc = A.get(&w, n_rows)))
        return rc;
    rows_write_kernel<<<blocks_for(n_cand), 256, 0, st>>>(model, d_ch, d_dirs, n_bins, n_cand, d_open, d_det,
                                                          d_site_bit, cnt, tau0, eoff, rid, a_scale, ent_bit, ent_row,
                                                          ent_a, d, w);
    BD_HIP(hipGetLastError());
    // assemble: a stable radix sort of the entries by bit keeps each bit's entries in row order
    iota_kernel<<<blocks_for(nnz), 256, 0, st>>>(perm, nnz);
    BD_HIP(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 31 && (1 << end_bit) < n_bits) ++end_bit;
    size_t tmp = 0;
    BD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ent_bit, bit_sorted, perm, perm_sorted, nnz, 0, end_bit, st));
    unsigned char* t = nullptr;
    if ((rc = A.get(&t, tmp))) return rc;
    BD_HIP(hipcub::DeviceRadixSort::SortPairs(t, tmp, ent_bit, bit_sorted, perm, perm_sorted, nnz, 0, end_bit, st));
    gather_kernel<<<blocks_for(nnz), 256, 0, st>>>(perm_sorted, ent_row, ent_a, nnz, ray, a);
    BD_HIP(hipGetLastError());
    int *bcnt, *rowp;
    if ((rc = A.get(&bcnt, size_t(n_bits) + 1)) || (rc = A.get(&rowp, size_t(n_bits) + 1))) return rc;
    BD_HIP(hipMemsetAsync(bcnt, 0, sizeof(int) * (size_t(n_bits) + 1), st));
    bit_count_kernel<<<blocks_for(nnz), 256, 0, st>>>(ent_bit, nnz, bcnt);
    BD_HIP(hipGetLastError());
    if ((rc = exclusive_sum(A, bcnt, rowp, n_bits, st))) return rc;
    // neighbours
    int *ncnt, *ncnt1, *noff;
    if ((rc = A.get(&ncnt, n_bits)) || (rc = A.get(&noff, size_t(n_bits) + 1))) return rc;
    grid_count_kernel<<<blocks_for(n_bits), 256, 0, st>>>(d_site, n_bits, d_site_bit, model.nx, model.ny, model.nz,
                                                          ncnt);
    BD_HIP(hipGetLastError());
    if ((rc = padded(A, ncnt, n_bits, &ncnt1, st)) || (rc = exclusive_sum(A, ncnt1, noff, n_bits, st))) return rc;
    int n_nbr = 0;
    BD_HIP(hipMemcpyAsync(&n_nbr, noff + n_bits, sizeof(int), hipMemcpyDeviceToHost, st));
    BD_HIP(hipStreamSynchronize(st));
    int32_t* nbr;
    if ((rc = A.get(&nbr, n_nbr))) return rc;
    grid_write_kernel<<<blocks_for(n_bits), 256, 0, st>>>(d_site, n_bits, d_site_bit, model.nx, model.ny, model.nz,
                                                          noff, nbr);
    BD_HIP(hipGetLastError());
    H->n_vars = n_bits;
    H->n_rays = n_rows;
    if ((rc = to_host(H->row_ptr, rowp, size_t(n_bits) + 1, st)) || (rc = to_host(H->ray, ray, size_t(nnz), st)) ||
        (rc = to_host(H->a, a, size_t(nnz), st)) || (rc = to_host(H->d, d, size_t(n_rows), st)) ||
        (rc = to_host(H->w, w, size_t(n_rows), st)) || (rc = to_host(H->nbr_ptr, noff, size_t(n_bits) + 1, st)) ||
        (rc = to_host(H->nbr, nbr, size_t(n_nbr), st)))
        return rc;
    BD_HIP(hipStreamSynchronize(st));
    return QB_OK;
}

}  // namespace

// ---- C ABI ----------------------------------------------------------------------------

extern "C" int qb_muon_expose(const qb_voxel_medium* m, const float chamber[3], const qb_muon_sky* sky,
                              int64_t n_candidates, uint32_t seed, uint64_t* open, uint64_t* det) {
    if (!m || !m->mu || !chamber || !sky || !open || !det) return qb::build_fail(QB_E_ARG, "qb_muon_expose: null argument");
    if (m->nx <= 0 || m->ny <= 0 || m->nz <= 0 || !(m->voxel > 0.0f) || sky->n_az <= 0 || sky->n_th <= 0 ||
        n_candidates < 0 || int64_t(m->nx) * m->ny * m->nz > (int64_t(1) << 31))
        return qb::build_fail(QB_E_ARG, "qb_muon_expose: bad medium, sky or candidate count");
    const size_t cells = size_t(m->nx) * m->ny * m->nz, bins = size_t(sky->n_az) * sky->n_th;
    hipStream_t st = nullptr;
    BD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    Arena A;
    float* mu;
    unsigned long long *o, *d;
    int rc;
    if ((rc = A.get(&mu, cells)) || (rc = A.get(&o, bins)) || (rc = A.get(&d, bins))) return rc;
    BD_HIP(hipMemcpyAsync(mu, m->mu, sizeof(float) * cells, hipMemcpyHostToDevice, st));
    BD_HIP(hipMemsetAsync(o, 0, sizeof(unsigned long long) * bins, st));
    BD_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long) * bins, st));
    const Medium md{{m->lo[0], m->lo[1], m->lo[2]}, m->voxel, m->nx, m->ny, m->nz, mu};
    const unsigned grid = unsigned(std::min<int64_t>(std::max<int64_t>((n_candidates + 255) / 256, 1), 65536));
    muon_expose_kernel<<<grid, 256, 0, st>>>(md, chamber[0], chamber[1], chamber[2], sky->c_min, sky->th_max,
                                              sky->n_az, sky->n_th, n_candidates, seed, o, d);
    BD_HIP(hipGetLastError());
    std::vector<unsigned long long> ho(bins), hd(bins);
    BD_HIP(hipMemcpyAsync(ho.data(), o, sizeof(unsigned long long) * bins, hipMemcpyDeviceToHost, st));
    BD_HIP(hipMemcpyAsync(hd.data(), d, sizeof(unsigned long long) * bins, hipMemcpyDeviceToHost, st));
    BD_HIP(hipStreamSynchronize(st));
    for (size_t b = 0; b < bins; ++b) {
        open[b] += ho[b];
        det[b] += hd[b];
    }
    return QB_OK;
}

extern "C" int qb_muon_rows_device(const qb_voxel_medium* model, int n_views, const float* chambers,
                                   const float* dirs, int n_bins, const uint64_t* open, const uint64_t* det,
                                   const int32_t* site_bit, double min_open, float a_scale, double lambda,
                                   double kappa, qb_host_problem** out) {
    if (!model || !model->mu || !chambers || !dirs || !open || !det || !site_bit || !out || n_views <= 0 ||
        n_bins <= 0)
        return qb::build_fail(QB_E_ARG, "qb_muon_rows_device: null or empty argument");
    *out = nullptr;
    const int64_t cells64 = int64_t(model->nx) * model->ny * model->nz;
    if (model->nx <= 0 || model->ny <= 0 || model->nz <= 0 || cells64 > (int64_t(1) << 30) ||
        int64_t(n_bins) * n_views > (int64_t(1) << 30))
        return qb::build_fail(QB_E_ARG, "qb_muon_rows_device: bad medium or too many rays");
    const size_t cells = size_t(cells64), n_cand = size_t(n_bins) * n_views;
    std::vector<int32_t> site;
    for (size_t c = 0; c < cells; ++c) {
        if (site_bit[c] >= 0 && site_bit[c] != int32_t(site.size()))
            return qb::build_fail(QB_E_ARG, "qb_muon_rows_device: site_bit must number the domain cells 0, 1, ... in "
                                            "cell order (BitField::grid)");
        if (site_bit[c] >= 0) site.push_back(int32_t(c));
    }
    auto H = std::unique_ptr<qb_host_problem>(new (std::nothrow) qb_host_problem());
    if (!H) return QB_E_NOMEM;
    hipStream_t st = nullptr;
    BD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    Arena A;
    float *mu, *ch, *dr;
    unsigned long long *o, *d;
    int32_t *sb, *sites;
    int rc;
    if ((rc = A.get(&mu, cells)) || (rc = A.get(&ch, 3 * size_t(n_views))) || (rc = A.get(&dr, 3 * n_cand)) ||
        (rc = A.get(&o, n_cand)) || (rc = A.get(&d, n_cand)) || (rc = A.get(&sb, cells)) ||
        (rc = A.get(&sites, site.size())))
        return rc;
    BD_HIP(hipMemcpyAsync(mu, model->mu, sizeof(float) * cells, hipMemcpyHostToDevice, st));
    BD_HIP(hipMemcpyAsync(ch, chambers, sizeof(float) * 3 * size_t(n_views), hipMemcpyHostToDevice, st));
    BD_HIP(hipMemcpyAsync(dr, dirs, sizeof(float) * 3 * n_cand, hipMemcpyHostToDevice, st));
    BD_HIP(hipMemcpyAsync(o, open, sizeof(uint64_t) * n_cand, hipMemcpyHostToDevice, st));
    BD_HIP(hipMemcpyAsync(d, det, sizeof(uint64_t) * n_cand, hipMemcpyHostToDevice, st));
    BD_HIP(hipMemcpyAsync(sb, site_bit, sizeof(int32_t) * cells, hipMemcpyHostToDevice, st));
    if (!site.empty())
        BD_HIP(hipMemcpyAsync(sites, site.data(), sizeof(int32_t) * site.size(), hipMemcpyHostToDevice, st));
    const Medium md{{model->lo[0], model->lo[1], model->lo[2]}, model->voxel, model->nx, model->ny, model->nz, mu};
    if ((rc = rows_and_assembly(A, md, ch, dr, n_bins, n_views, o, d, sb, sites, int(site.size()), min_open, a_scale,
                                H.get(), st)))
        return rc;
    H->lambda = lambda;
    H->kappa = kappa;
    *out = H.release();
    return QB_OK;
}

// The reference's muon pipeline on the device for the scene of qb_muon_scene (pyramid,
// void, chambers, sky): exposure of each chamber through the true medium by candidate
// muons (2^log2_muons per chamber), rows of the model medium (no void), assembly and
// neighbour lists. Bin directions are computed on the host in float as the reference's
// MuonBinnedData::bin_direction does (src/muon_recon.hpp:52-58).
extern "C" int qb_scene_muon_device(const qb_muon_scene* s, qb_host_problem** out, double* build_ms) {
    if (!s || !out) return qb::build_fail(QB_E_ARG, "qb_scene_muon_device: null argument");
    *out = nullptr;
    if (s->grid < 8 || s->grid > 512 || s->log2_muons < 4 || s->log2_muons > 36 || s->sky_az < 1 || s->sky_th < 1 ||
        s->sky_az > 4096 || s->sky_th > 4096)
        return qb::build_fail(QB_E_ARG, "qb_scene_muon_device: scene parameters out of range");
    const auto t0 = std::chrono::steady_clock::now();
    const int N = s->grid;
    const size_t cells = size_t(N) * N * N;
    const float voxel = 256.0f / float(N);
    const float lo[3] = {-128.0f, -128.0f, -10.0f};
    const float chambers[9] = {0.0f, 0.0f, 40.0f, -40.0f, 0.0f, 25.0f, 40.0f, 0.0f, 25.0f};
    const float mu_rock = float(s->mu_per_rho * s->rho);
    const float mu_air = mu_rock * (1.2f / 2700.0f);
    const int n_az = s->sky_az, n_th = s->sky_th, n_bins = n_az * n_th;
    const float th_max = 1.22173f, c_min = 0.34202f;   // 70 degrees (src/muon.cuh:177-181)
    std::vector<float> dirs(size_t(3) * 3 * n_bins);
    for (int v = 0; v < 3; ++v)
        for (int b = 0; b < n_bins; ++b) {
            const int ith = b / n_az, iaz = b % n_az;
            const float th = (float(ith) + 0.5f) / float(n_th) * th_max;
            const float az = (float(iaz) + 0.5f) * 6.28318530718f / float(n_az);
            const float sn = std::sin(th);
            float* q = dirs.data() + 3 * (size_t(v) * n_bins + b);
            q[0] = sn * std::cos(az);
            q[1] = sn * std::sin(az);
            q[2] = std::cos(th);
        }
    auto H = std::unique_ptr<qb_host_problem>(new (std::nothrow) qb_host_problem());
    if (!H) return QB_E_NOMEM;
    hipStream_t st = nullptr;
    BD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    Arena A;
    float *mu_true, *mu_model, *ch, *dr;
    int *domain, *domain1, *pos;
    uint8_t* is_void;
    int32_t *site, *site_bit;
    unsigned long long *open, *det;
    int rc;
    if ((rc = A.get(&mu_true, cells)) || (rc = A.get(&mu_model, cells)) || (rc = A.get(&ch, 9)) ||
        (rc = A.get(&dr, dirs.size())) || (rc = A.get(&domain, cells)) || (rc = A.get(&pos, cells + 1)) ||
        (rc = A.get(&is_void, cells)) || (rc = A.get(&site_bit, cells)) || (rc = A.get(&open, 3 * size_t(n_bins))) ||
        (rc = A.get(&det, 3 * size_t(n_bins))))
        return rc;
    BD_HIP(hipMemcpyAsync(ch, chambers, sizeof(chambers), hipMemcpyHostToDevice, st));
    BD_HIP(hipMemcpyAsync(dr, dirs.data(), sizeof(float) * dirs.size(), hipMemcpyHostToDevice, st));
    muon_cells_kernel<<<blocks_for(cells), 256, 0, st>>>(N, voxel, lo[0], lo[1], lo[2], 139.0f, 115.0f,
                                                         float(0.5 * s->void_side), float(s->void_z), mu_rock, mu_air,
                                                         ch, mu_true, mu_model, domain, is_void);
    BD_HIP(hipGetLastError());
    if ((rc = padded(A, domain, int(cells), &domain1, st)) || (rc = exclusive_sum(A, domain1, pos, int(cells), st)))
        return rc;
    int n_bits = 0;
    BD_HIP(hipMemcpyAsync(&n_bits, pos + cells, sizeof(int), hipMemcpyDeviceToHost, st));
    BD_HIP(hipStreamSynchronize(st));
    if ((rc = A.get(&site, n_bits))) return rc;
    site_kernel<<<blocks_for(cells), 256, 0, st>>>(domain, pos, int(cells), site, site_bit);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemsetAsync(open, 0, sizeof(unsigned long long) * 3 * size_t(n_bins), st));
    BD_HIP(hipMemsetAsync(det, 0, sizeof(unsigned long long) * 3 * size_t(n_bins), st));
    const Medium truth{{lo[0], lo[1], lo[2]}, voxel, N, N, N, mu_true};
    const int64_t n_cand = int64_t(1) << s->log2_muons;
    const unsigned grid = unsigned(std::min<int64_t>((n_cand + 255) / 256, 65536));
    for (int v = 0; v < 3; ++v) {
        muon_expose_kernel<<<grid, 256, 0, st>>>(truth, chambers[3 * v], chambers[3 * v + 1], chambers[3 * v + 2],
                                                  c_min, th_max, n_az, n_th, n_cand, uint32_t(s->seed),
                                                  open + size_t(v) * n_bins, det + size_t(v) * n_bins);
        BD_HIP(hipGetLastError());
    }
    const Medium model{{lo[0], lo[1], lo[2]}, voxel, N, N, N, mu_model};
    if ((rc = rows_and_assembly(A, model, ch, dr, n_bins, 3, open, det, site_bit, site, n_bits, s->min_open, mu_rock,
                                H.get(), st)))
        return rc;
    std::vector<int32_t> hsite;
    std::vector<uint8_t> hvoid;
    if ((rc = to_host(hsite, site, size_t(n_bits), st)) || (rc = to_host(hvoid, is_void, cells, st))) return rc;
    BD_HIP(hipStreamSynchronize(st));
    H->truth.resize(size_t(n_bits));
    for (int b = 0; b < n_bits; ++b) H->truth[size_t(b)] = hvoid[size_t(hsite[size_t(b)])];
    H->lambda = s->lambda;
    H->kappa = s->kappa;
    if (build_ms) *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = H.release();
    return QB_OK;
}

// The gravimetry rows of qb_scene_gravity built on the device: the same row_ptr, ray and a
// arrays (each entry computed per (station, cell) pair in the builder's arithmetic); d (the
// body's anomaly plus noise), the neighbour lists and the colouring are completed on the
// host from them exactly as qb_scene_gravity does.
namespace qb {
int gravity_finish(const qb_gravity_scene* s, qb_host_problem* H);
int gravity_reach_of(const qb_gravity_scene* s);
}

extern "C" int qb_scene_gravity_device(const qb_gravity_scene* s, qb_host_problem** out, double* build_ms) {
    if (!s || !out) return qb::build_fail(QB_E_ARG, "qb_scene_gravity_device: null argument");
    *out = nullptr;
    if (s->nx < 1 || s->ny < 1 || s->nz < 1 || !(s->h > 0.0) || !(s->radius >= 0.0) || !(s->sigma > 0.0))
        return qb::build_fail(QB_E_ARG, "qb_scene_gravity_device: scene parameters out of range");
    if (double(s->nx) * s->ny * s->nz > 2.0e7) return qb::build_fail(QB_E_LIMIT, "qb_scene_gravity_device: too many cells");
    const auto t0 = std::chrono::steady_clock::now();
    const int n = s->nx * s->ny * s->nz;
    auto H = std::unique_ptr<qb_host_problem>(new (std::nothrow) qb_host_problem());
    if (!H) return QB_E_NOMEM;
    hipStream_t st = nullptr;
    BD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    Arena A;
    const GravityGrid g{s->nx, s->ny, s->nz, qb::gravity_reach_of(s), s->h, s->radius * s->radius,
                        s->drho * s->h * s->h * s->h};
    // entries are counted and offset in int32: refuse a grid whose rows could pass 2^31
    if (int64_t(n) * (2 * int64_t(g.reach) + 1) * (2 * int64_t(g.reach) + 1) > int64_t(INT32_MAX))
        return qb::build_fail(QB_E_LIMIT, "qb_scene_gravity_device: the rows could pass 2^31 entries");
    int *cnt, *cnt1, *off;
    int rc;
    if ((rc = A.get(&cnt, n)) || (rc = A.get(&off, size_t(n) + 1))) return rc;
    gravity_count_kernel<<<blocks_for(n), 256, 0, st>>>(g, cnt);
    BD_HIP(hipGetLastError());
    if ((rc = padded(A, cnt, n, &cnt1, st)) || (rc = exclusive_sum(A, cnt1, off, n, st))) return rc;
    int nnz = 0;
    BD_HIP(hipMemcpyAsync(&nnz, off + n, sizeof(int), hipMemcpyDeviceToHost, st));
    BD_HIP(hipStreamSynchronize(st));
    int32_t* ray;
    float* a;
    if ((rc = A.get(&ray, nnz)) || (rc = A.get(&a, nnz))) return rc;
    gravity_write_kernel<<<blocks_for(n), 256, 0, st>>>(g, off, ray, a);
    BD_HIP(hipGetLastError());
    H->n_vars = n;
    H->n_rays = s->nx * s->ny;
    if ((rc = to_host(H->row_ptr, off, size_t(n) + 1, st)) || (rc = to_host(H->ray, ray, size_t(nnz), st)) ||
        (rc = to_host(H->a, a, size_t(nnz), st)))
        return rc;
    BD_HIP(hipStreamSynchronize(st));
    if ((rc = qb::gravity_finish(s, H.get()))) return rc;
    if (build_ms) *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = H.release();
    return QB_OK;
}
