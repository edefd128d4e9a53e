// Dispatch order of tiles by rings around a centre tile.  Host and device.
//
// The bake's tiles are light-sheared texel columns of very different sizes: a tile near the
// centre of the sheared (u', v') domain holds up to 16 texels in each of the 128 layers, a tile
// at the domain's edge a handful.  Workgroups take tiles in blockIdx order, so dispatching the
// heavy tiles first and the light ones last (longest-processing-time-first list scheduling)
// keeps the last wave of workgroups short.  The tile weight falls off with the distance from
// the projection of the cube's centre, so the order visits the tile grid in square rings of
// growing Chebyshev distance around that centre tile, each ring row by row, clipped to the grid
// (no table and no sort: rank -> tile in closed form).
#pragma once
#include "philox.h"  // MRP_HD

namespace mrp {

MRP_HD int ring_imin(int a, int b) { return a < b ? a : b; }
MRP_HD int ring_imax(int a, int b) { return a > b ? a : b; }

// Tiles of the grid within Chebyshev distance d of the centre tile: the square [-d, d]^2 of
// offsets clipped to the grid's offset range [a0, a1] x [b0, b1].
MRP_HD int ring_count_within(int d, int a0, int a1, int b0, int b1) {
    if (d < 0) return 0;
    const int nu = ring_imin(d, a1) - ring_imax(-d, a0) + 1, nv = ring_imin(d, b1) - ring_imax(-d, b0) + 1;
    return nu > 0 && nv > 0 ? nu * nv : 0;
}

// Tile index tu + ntu * tv of rank k (0 <= k < ntu * ntv) in ring order around the centre tile
// (ciu, civ) (0 <= ciu < ntu, 0 <= civ < ntv).  Within ring d: the row dv = -d (if in the grid),
// then the rows -d < dv < d with their cells du = -d and du = +d (those in the grid), then the
// row dv = +d.  Returns -1 for k out of range.
MRP_HD int ring_tile(int k, int ntu, int ntv, int ciu, int civ) {
    if (k < 0 || k >= ntu * ntv) return -1;
    const int a0 = -ciu, a1 = ntu - 1 - ciu, b0 = -civ, b1 = ntv - 1 - civ;
    // the ring of rank k: the smallest d with more than k tiles within distance d (binary search;
    // every tile lies within distance dmax)
    int lo = 0, hi = ring_imax(ring_imax(-a0, a1), ring_imax(-b0, b1));
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ring_count_within(mid, a0, a1, b0, b1) > k) hi = mid;
        else lo = mid + 1;
    }
    const int d = lo;
    int j = k - ring_count_within(d - 1, a0, a1, b0, b1);
    const int rlo = ring_imax(-d, a0), rhi = ring_imin(d, a1);  // clipped extent of a full row
    const int nrow = rhi - rlo + 1;
    if (-d >= b0) {  // first row of the ring
        if (j < nrow) return (ciu + rlo + j) + ntu * (civ - d);
        j -= nrow;
    }
    if (d > 0) {  // rows strictly inside: at most the two side cells each
        const bool left = -d >= a0, right = d <= a1;
        const int per = (left ? 1 : 0) + (right ? 1 : 0);
        const int mlo = ring_imax(-d + 1, b0), mhi = ring_imin(d - 1, b1);
        const int m = mhi >= mlo ? mhi - mlo + 1 : 0;
        if (j < m * per) {
            const int dv = mlo + j / per;
            const int du = per == 2 ? (j % per == 0 ? -d : d) : (left ? -d : d);
            return (ciu + du) + ntu * (civ + dv);
        }
        j -= m * per;
        if (d <= b1 && j < nrow) return (ciu + rlo + j) + ntu * (civ + d);  // last row
    }
    return -1;
}

}  // namespace mrp
