// Two time steps per pass over the lattice (host side): the launch list of
// one blocked sweep.
//
// Step t of a chunk of K z planes is written to a ring of plane slots small
// enough for the 256 MB last-level cache to keep; step t+1 of the planes one
// behind reads that ring and writes the output lattice.  HBM then carries
// one lattice read and one lattice write per two steps, plus what the cache
// spills.  Per chunk i (planes [iK, iK + nz)):
//   A_i  step t   of planes [iK, iK + nz)                   lattice -> ring
//   B_i  step t+1 of planes [max(iK - 1, 1), iK + nz - 1)   ring -> lattice
// The launches run in the order A_0, B_0, A_1, B_1, ...  B_i reads step t of
// planes iK - 2 .. iK + nz - 1, the last K + 2 planes written, so a ring of
// K + 2 slots (plane p in slot p mod (K + 2)) has overwritten none of them.
// Planes lz - 1 and 0 of step t+1 need step t of planes lz - 2, lz - 1, 0
// and 1, which exist together only at the end: the A launches also copy
// planes 0 and 1 into a head buffer, and one last launch computes those two
// planes.  Periodic wraps in z are the plane indices taken modulo lz; in y
// and x the addressing is the single-step kernel's.  (Running A_i and
// B_{i-1} as one launch over a ring of 2K + 2 slots measured slower.)
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "kernels.h"

namespace d3q19 {

// Ring slots and buffer sizes in elements.
inline uint32_t block_ring_slots(uint32_t K) { return K + 2; }
inline uint64_t block_ring_elems(const Geom &g, uint32_t K) { return 19ull * block_ring_slots(K) * g.plane; }
inline uint64_t block_head_elems(const Geom &g) { return 19ull * 2 * g.plane; }

inline std::vector<BlockArgs> block_schedule(const Geom &g, const RatesD &r, uint32_t K, void *lat_in,
                                             void *lat_out, void *ring, void *head)
{
    const uint32_t lz = g.lz, slots = block_ring_slots(K);
    const PlaneStore full_in = {lat_in, g.N, lz, nullptr, 0, 0, 0};
    const PlaneStore full_out = {lat_out, g.N, lz, nullptr, 0, 0, 0};
    const PlaneStore ring_only = {ring, (uint64_t)slots * g.plane, slots, nullptr, 0, 0, 0};
    PlaneStore ring_head = ring_only;
    ring_head.head = head;
    ring_head.head_stride = 2ull * g.plane;
    ring_head.head_lo = 0;
    ring_head.head_n = std::min<uint32_t>(2, lz);

    std::vector<BlockArgs> out;
    auto add = [&](int32_t z0, uint32_t nz, const PlaneStore &src, const PlaneStore &dst, int nt) {
        BlockArgs a;
        a.g = g;
        a.dplane = make_fastdiv((uint32_t)g.plane);
        a.r = r;
        a.z0 = z0;
        a.nz = nz;
        a.src = src;
        a.dst = dst;
        a.nt = nt;
        out.push_back(a);
    };
    for (uint32_t z0 = 0; z0 < lz; z0 += K) {
        const uint32_t nz = std::min(K, lz - z0);
        add((int32_t)z0, nz, full_in, ring_head, 0);
        const uint32_t lo = std::max<uint32_t>(z0 == 0 ? 0 : z0 - 1, 1), hi = z0 + nz - 1;
        if (hi > lo)
            add((int32_t)lo, hi - lo, ring_only, full_out, 1);
    }
    // planes lz - 1 and 0 (one plane when lz == 1)
    add((int32_t)lz - 1, std::min<uint32_t>(2, lz), ring_head, full_out, 1);
    return out;
}

} // namespace d3q19
