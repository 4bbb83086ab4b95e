// yk_sfc.hpp -- space-filling-curve keys of the particle re-sort.
//
// Slots are kept in the order of a Hilbert curve through a 2^g x 2^g x 2^g
// cell grid.  Consecutive Hilbert cells share a face, so 16 or 64
// consecutive slots form a compact group; a Morton (Z) order jumps between
// distant cells inside such groups and makes their bounding boxes longer
// (with uniformly random ions at about one per cell, 16-slot groups of Morton
// order have mean box edges of about 1.6, 1.3 and 1.0 cell-group widths where
// Hilbert groups are cubes, and the cluster pair lists built on them are
// about 13 % longer at a cutoff four group widths long).
#pragma once

namespace yk {

// Hilbert index of cell (x, y, z), each coordinate in [0, 2^g), g in [1, 10]:
// Skilling's transform of the coordinates to the transposed index
// (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707 (2004)
// 381-387), whose bits are then interleaved, the x transform's bit the most
// significant of each triple.  Indices run over [0, 8^g) once each.
__host__ __device__ inline unsigned hilbert_key3(unsigned x, unsigned y, unsigned z, int g)
{
    const unsigned M = 1u << (g - 1);
    for (unsigned Q = M; Q > 1; Q >>= 1) {   // undo the excess work, top bit down
        const unsigned P = Q - 1;
        if (x & Q) x ^= P;                    // axis 0: invert (the exchange is empty)
        if (y & Q) {                          // axis 1: invert or exchange with axis 0
            x ^= P;
        } else {
            const unsigned t = (x ^ y) & P;
            x ^= t;
            y ^= t;
        }
        if (z & Q) {                          // axis 2
            x ^= P;
        } else {
            const unsigned t = (x ^ z) & P;
            x ^= t;
            z ^= t;
        }
    }
    y ^= x;   // Gray encode
    z ^= y;
    unsigned t = 0;
    for (unsigned Q = M; Q > 1; Q >>= 1)
        if (z & Q) t ^= Q - 1;
    x ^= t;
    y ^= t;
    z ^= t;
    unsigned key = 0;
    for (int b = g - 1; b >= 0; --b)
        key = (key << 3) | (((x >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((z >> b) & 1u);
    return key;
}

}  // namespace yk
