// Writes tests/golden/philox_rocrand.txt: Philox4x32-10 outputs computed by rocRAND's
// header implementation (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), an
// implementation independent of the oracle's. rocRAND's engine takes key = seed
// (low, high word), counter.zw = subsequence and counter.xy = offset / 4.
// Build and run:  hipcc -O2 -o /tmp/mkphilox tests/golden/make_philox_vectors.cpp && /tmp/mkphilox > tests/golden/philox_rocrand.txt
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdint>
#include <cstdio>

static uint64_t sm(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main() {
    uint64_t s = 20261202ull;
    std::printf("# c0 c1 c2 c3 k0 k1 -> o0 o1 o2 o3 (hex), from rocRAND philox4x32_10_engine\n");
    for (int t = 0; t < 96; ++t) {
        const uint64_t seed = sm(s);
        // counters as the engine uses them: bit < 2^24, sweep < 2^20, upper words small
        const uint32_t c0 = uint32_t(sm(s) & 0xFFFFFFu), c1 = uint32_t(sm(s) & 0xFFFFFu);
        const uint32_t c2 = t % 3 == 0 ? uint32_t(sm(s) & 0xFFu) : 0u, c3 = t % 5 == 0 ? 1u : 0u;
        rocrand_device::philox4x32_10_engine e(seed, (uint64_t(c3) << 32) | c2, 0);
        e.discard(4ull * ((uint64_t(c1) << 32) | c0));
        const unsigned o0 = e(), o1 = e(), o2 = e(), o3 = e();
        std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c0, c1, c2, c3,
                    uint32_t(seed), uint32_t(seed >> 32), o0, o1, o2, o3);
    }
    return 0;
}
