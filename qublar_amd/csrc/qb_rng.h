// qb_rng.h -- the engine's random draws and acceptance exponential, compiled for
// both the host and gfx950 from one source.
//
// Draws are counter-based: the draw of bit `bit` in sweep `sweep` of the branch
// whose seed is `seed` is Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) with
// key = (seed low word, seed high word) and counter = (bit, sweep, 0, 0); the top
// 53 bits of (o0 << 32 | o1) scale into [0, 1). A draw therefore depends only on
// (branch, sweep, bit) -- never on the launch, the slot or the visit order.
//
// qexp is the exponential of the Metropolis test u < exp(-dE / T). It is made of
// correctly rounded IEEE-754 operations only (no fused multiply-add: every file
// that includes this header is compiled with -ffp-contract=off), so the host and
// the device compute the same bits for every input.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define QB_HD __host__ __device__ __forceinline__
#else
#define QB_HD inline
#endif

namespace qb {

struct u32x4 { uint32_t x, y, z, w; };

QB_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
        const u32x4 n{uint32_t(p1 >> 32) ^ c.y ^ k0, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k1,
                      uint32_t(p0)};
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Uniform in [0, 1) with 53 random bits.
QB_HD double draw_u01(uint64_t seed, uint32_t sweep, uint32_t bit) {
    const u32x4 o = philox4x32_10(u32x4{bit, sweep, 0u, 0u}, uint32_t(seed), uint32_t(seed >> 32));
    const uint64_t z = (uint64_t(o.x) << 32) | uint64_t(o.y);
    return double(z >> 11) * 0x1.0p-53;
}

// Two uniforms from one Philox call, for the batched-QUBO engine's replica pairs: key =
// the instance seed, counter = (bit, sweep, pair, kPairTag); the first is the top 53 bits
// of (o0 << 32 | o1), the second of (o2 << 32 | o3). The tag keeps these counters apart
// from the single draws' (bit, sweep, 0, 0).
constexpr uint32_t kPairTag = 0x51B0B47Cu;
QB_HD void draw2_u01(uint64_t seed, uint32_t sweep, uint32_t bit, uint32_t pair, double& u0, double& u1) {
    const u32x4 o = philox4x32_10(u32x4{bit, sweep, pair, kPairTag}, uint32_t(seed), uint32_t(seed >> 32));
    const uint64_t z0 = (uint64_t(o.x) << 32) | uint64_t(o.y);
    const uint64_t z1 = (uint64_t(o.z) << 32) | uint64_t(o.w);
    u0 = double(z0 >> 11) * 0x1.0p-53;
    u1 = double(z1 >> 11) * 0x1.0p-53;
}

QB_HD double pow2i(int k) {
    const uint64_t b = uint64_t(k + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &b, sizeof d);
    return d;
}

// exp(x) = 2^k e^r with k = floor(x / ln 2 + 1/2) and r = (x - k ln2_hi) - k ln2_lo
// (Cody-Waite reduction; ln2_hi carries 32 significant bits, so k * ln2_hi is exact
// for |k| < 2^11). e^r is the degree-13 Taylor polynomial in Horner form: for
// |r| <= 0.347 its truncation error is below 5e-18 relative.
QB_HD double qexp(double x) {
    if (!(x > -745.0)) return 0.0;
    if (x >= 709.0) return __builtin_huge_val();
    const double kd = __builtin_floor(x * 0x1.71547652b82fep+0 + 0.5);
    const int k = int(kd);
    const double r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
    double q = 0x1.6124613a86d09p-33;
    q = q * r + 0x1.1eed8eff8d898p-29;
    q = q * r + 0x1.ae64567f544e4p-26;
    q = q * r + 0x1.27e4fb7789f5cp-22;
    q = q * r + 0x1.71de3a556c734p-19;
    q = q * r + 0x1.a01a01a01a01ap-16;
    q = q * r + 0x1.a01a01a01a01ap-13;
    q = q * r + 0x1.6c16c16c16c17p-10;
    q = q * r + 0x1.1111111111111p-7;
    q = q * r + 0x1.5555555555555p-5;
    q = q * r + 0x1.5555555555555p-3;
    q = q * r + 0.5;
    q = q * r + 1.0;
    q = q * r + 1.0;
    if (k >= -1021) return q * pow2i(k);
    return (q * pow2i(k + 600)) * pow2i(-600);
}

// Natural log of a positive normal double (0 gives -inf), from IEEE operations only:
// x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m - 1) / (m + 1) (|s| <= 0.1716, m - 1
// exact), ln m = 2 s (1 + s^2/3 + ... + s^22/23) in Horner form (truncation below 3e-19
// relative), then e ln2_hi + (e ln2_lo + ln m) with ln2_hi's 32 significant bits making
// e ln2_hi exact. The dense engine's acceptance test uses -qlog(u).
QB_HD double qlog(double x) {
    if (!(x > 0.0)) return -__builtin_huge_val();
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof b);
    int e = int((b >> 52) & 0x7ffu) - 1023;
    const uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    __builtin_memcpy(&m, &mb, sizeof m);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    const double lnm = 2.0 * s * p;
    const double ed = double(e);
    return ed * 0x1.62e42fee00000p-1 + (ed * 0x1.a39ef35793c76p-33 + lnm);
}

// The Metropolis test for an uphill move (dE > 0) at inverse temperature beta.
QB_HD bool metropolis_uphill(double dE, double beta, uint64_t seed, uint32_t sweep, uint32_t bit) {
    return draw_u01(seed, sweep, bit) < qexp(-dE * beta);
}

#if defined(__HIPCC__)
// The Metropolis decision u < qexp(-dE * beta) for dE > 0, decided by the hardware float
// exponential where that is unambiguous. For -80 < x = -dE beta <= 0, e_f = 2^(float(x)
// log2e) (v_exp_f32) is within 1.5e-5 of e^x (float(x): 4.8e-6 absolute; the product with
// log2e: 2.1e-5 in the exponent; v_exp_f32: a few ulp), so u < e_f (1 - 2^-14) implies
// u < qexp(x) and u > e_f (1 + 2^-14) implies u > qexp(x); a draw inside the band takes
// qexp itself. For x <= -80, qexp(x) < 2e-35 lies below every nonzero 53-bit draw, so
// only u = 0 can pass. Every decision is therefore qexp's, bit for bit.
__device__ __forceinline__ bool uphill_accept(double x, double u) {
    if (!(x > -80.0)) [[unlikely]] return u == 0.0 && qexp(x) > 0.0;
    constexpr double kBand = 0x1.0p-14;
    const double e = double(__builtin_amdgcn_exp2f(float(x) * 1.44269504f));
    const bool below = u < e * (1.0 - kBand), above = u > e * (1.0 + kBand);
    if (!below && !above) [[unlikely]] return u < qexp(x);   // inside the band
    return below;
}
#endif

}  // namespace qb
