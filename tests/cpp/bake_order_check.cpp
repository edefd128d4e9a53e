// Checks mrpnn_amd/csrc/bake_order.h on the host: for every grid shape up to 70 x 70 tiles and
// every centre tile, and for the march's tile grids of large frames with a few centres, ring_tile is a bijection from ranks onto tiles, its Chebyshev distance from
// the centre never decreases with the rank, and out-of-range ranks give -1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bake_order.h"

int main() {
    long long grids = 0;
    for (int ntu = 1; ntu <= 70; ntu += (ntu < 12 ? 1 : 7))
        for (int ntv = 1; ntv <= 70; ntv += (ntv < 12 ? 1 : 5))
            for (int ciu = 0; ciu < ntu; ++ciu)
                for (int civ = 0; civ < ntv; ++civ) {
                    const int n = ntu * ntv;
                    std::vector<int> seen(n, 0);
                    int last_d = 0;
                    for (int k = 0; k < n; ++k) {
                        const int t = mrp::ring_tile(k, ntu, ntv, ciu, civ);
                        if (t < 0 || t >= n || seen[t]) {
                            std::printf("FAIL grid %dx%d centre (%d,%d) rank %d -> %d\n", ntu, ntv, ciu, civ, k, t);
                            return 1;
                        }
                        seen[t] = 1;
                        const int du = std::abs(t % ntu - ciu), dv = std::abs(t / ntu - civ);
                        const int d = du > dv ? du : dv;
                        if (d < last_d) {
                            std::printf("FAIL order grid %dx%d centre (%d,%d) rank %d\n", ntu, ntv, ciu, civ, k);
                            return 1;
                        }
                        last_d = d;
                    }
                    if (mrp::ring_tile(n, ntu, ntv, ciu, civ) != -1 || mrp::ring_tile(-1, ntu, ntv, ciu, civ) != -1) {
                        std::printf("FAIL range grid %dx%d\n", ntu, ntv);
                        return 1;
                    }
                    ++grids;
                }
    // the march's tile grids (1080p and larger frames, every tile shape) with a few centres each
    const int big[][2] = {{240, 135}, {480, 68}, {120, 270}, {960, 34}, {1920, 17}, {256, 256}, {64, 1080}};
    for (const auto& g : big) {
        const int ntu = g[0], ntv = g[1], n = ntu * ntv;
        const int cs[][2] = {{ntu / 2, ntv / 2}, {0, 0}, {ntu - 1, ntv - 1}, {ntu / 3, ntv - 1}};
        for (const auto& c : cs) {
            std::vector<int> seen(n, 0);
            int last_d = 0;
            for (int k = 0; k < n; ++k) {
                const int t = mrp::ring_tile(k, ntu, ntv, c[0], c[1]);
                if (t < 0 || t >= n || seen[t]) {
                    std::printf("FAIL grid %dx%d centre (%d,%d) rank %d -> %d\n", ntu, ntv, c[0], c[1], k, t);
                    return 1;
                }
                seen[t] = 1;
                const int du = std::abs(t % ntu - c[0]), dv = std::abs(t / ntu - c[1]);
                const int d = du > dv ? du : dv;
                if (d < last_d) {
                    std::printf("FAIL order grid %dx%d rank %d\n", ntu, ntv, k);
                    return 1;
                }
                last_d = d;
            }
            ++grids;
        }
    }
    std::printf("ok %lld\n", grids);
    return 0;
}
