// Prints rocRAND's philox4x32-10 block function (its open-source header) on given
// counters/keys, to cross-check the library's own Philox (mrpnn_amd/csrc/philox.h).
// Host code only.  Input lines: c0 c1 c2 c3 k0 k1 (decimal); output: 4 words per line.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Exposed : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    Exposed e;
    while (scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = e.block(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
