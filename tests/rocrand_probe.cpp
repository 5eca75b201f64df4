// Host-side probe of rocRAND's Philox4x32-10 (rocrand/rocrand_philox4x32_10.h), used by
// tests/test_oracle_rng.py to cross-check the oracle's Philox.  Reads lines
// "c0 c1 c2 c3 k0 k1" (hex) from stdin and prints the four output words in hex.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct probe : rocrand_device::philox4x32_10_engine {
    static uint4 run(uint4 c, uint2 k)
    {
        probe p;
        return p.ten_rounds(c, k);
    }
};

int main()
{
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        const uint4 r = probe::run(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
