// Two time steps per pass over the lattice (host side): the launch list of
// one blocked sweep.
//
// The lattice is swept slab by slab along z or y (kernels.h: PlaneStore).
// Step t of a chunk of K slabs is written to a ring small enough for the
// 256 MB last-level cache to keep; step t+1 of the slabs one behind reads
// that ring and writes the output lattice.  HBM then carries one lattice
// read and one lattice write per two steps, plus what the cache spills.  Per
// chunk i (slabs [iK, iK + n)):
//   A_i  step t   of slabs [iK, iK + n)                    lattice -> ring
//   B_i  step t+1 of slabs [max(iK - 1, 1), iK + n - 1)    ring -> lattice
// The launches run in the order A_0, B_0, A_1, B_1, ...  B_i reads step t of
// slabs iK - 2 .. iK + n - 1, the last K + 2 slabs written, so a ring of
// K + 2 slots (slab s in slot s mod (K + 2)) has overwritten none of them.
// Slabs ns - 1 and 0 of step t+1 need slabs of step t from both ends, which
// exist together only at the end: the A launches also copy slabs 0 and 1
// into a head buffer, and one last launch computes those two slabs.  The
// periodic wraps along the sweep axis are the slab indices taken modulo ns;
// along the rows they wrap within a slab.  (Running A_i and B_{i-1} as one
// launch over a ring of 2K + 2 slots measured slower.)
#pragma once

#include <string.h>

#include <stdint.h>

#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "kernels.h"

namespace d3q19 {

// Slab geometry of a sweep along axis 2 (z) or 1 (y).
struct BlockGeom {
    uint32_t ns, nr;                      // slabs, rows per slab
    uint64_t lat_slab_stride, lat_row_stride; // in the lattice (x fastest, then y, then z)
};

inline BlockGeom block_geom(const Geom &g, int axis)
{
    if (axis == 2)
        return BlockGeom{g.lz, g.ly, (uint64_t)g.lx * g.ly, g.lx};
    return BlockGeom{g.ly, g.lz, g.lx, (uint64_t)g.lx * g.ly};
}

// Ring slots and buffer sizes in elements (slabs of lx * nr nodes).
inline uint32_t block_ring_slots(uint32_t K) { return K + 2; }
inline uint64_t block_ring_elems(const Geom &g, int axis, uint32_t K)
{
    return 19ull * block_ring_slots(K) * g.lx * block_geom(g, axis).nr;
}
inline uint64_t block_head_elems(const Geom &g, int axis) { return 19ull * 2 * g.lx * block_geom(g, axis).nr; }

// The force of one of the two steps of a pair: rates with the uniform force
// (r.fx, r.fy, r.fz), or per_node with the collision's force fc and the one
// the fused fields were formed with, fm (BlockArgs::per_node).
struct StepForce {
    RatesD r;
    ForceSpec fc, fm;
    int per_node;
};

// Launch list of one pair: step t with force f0, step t+1 with force f1.
inline std::vector<BlockArgs> block_schedule(const Geom &g, int axis, const StepForce &f0, const StepForce &f1,
                                             uint32_t K, void *lat_in, void *lat_out, void *ring, void *head)
{
    const BlockGeom bg = block_geom(g, axis);
    const uint32_t ns = bg.ns, slots = block_ring_slots(K);
    const uint64_t slab = (uint64_t)g.lx * bg.nr;
    const PlaneStore full_in = {lat_in, g.N, ns, make_fastdiv(ns), bg.lat_slab_stride, bg.lat_row_stride, nullptr, 0,
                                0, 0};
    const PlaneStore full_out = {lat_out, g.N, ns, make_fastdiv(ns), bg.lat_slab_stride, bg.lat_row_stride, nullptr, 0,
                                 0, 0};
    const PlaneStore ring_only = {ring, slots * slab, slots, make_fastdiv(slots), slab, g.lx, nullptr, 0, 0, 0};
    PlaneStore ring_head = ring_only;
    ring_head.head = head;
    ring_head.head_stride = 2 * slab;
    ring_head.head_lo = 0;
    ring_head.head_n = std::min<uint32_t>(2, ns);

    std::vector<BlockArgs> out;
    auto add = [&](int32_t s0, uint32_t count, const PlaneStore &src, const PlaneStore &dst, int nt,
                   const StepForce &f) {
        BlockArgs a;
        a.g = g;
        a.ns = ns;
        a.nr = bg.nr;
        a.slab = (uint32_t)slab;
        a.slabp = slab_padded((uint32_t)slab);
        a.dslabp = make_fastdiv(a.slabp);
        a.r = f.r;
        a.fc = f.fc;
        a.fm = f.fm;
        a.per_node = f.per_node;
        a.s0 = s0;
        a.count = count;
        a.src = src;
        a.dst = dst;
        a.nt = nt;
        out.push_back(a);
    };
    for (uint32_t s0 = 0; s0 < ns; s0 += K) {
        const uint32_t n = std::min(K, ns - s0);
        add((int32_t)s0, n, full_in, ring_head, 0, f0);
        const uint32_t lo = std::max<uint32_t>(s0 == 0 ? 0 : s0 - 1, 1), hi = s0 + n - 1;
        if (hi > lo)
            add((int32_t)lo, hi - lo, ring_only, full_out, 1, f1);
    }
    // slabs ns - 1 and 0 (one slab when ns == 1)
    add((int32_t)ns - 1, std::min<uint32_t>(2, ns), ring_head, full_out, 1, f1);
    return out;
}

// Ring of the two-step blocked sweep: about 225 MB of the 256 MB last-level
// cache.  The step-t launches read the lattice with non-temporal loads, so
// that stream does not evict the ring; on the 192x768x384 channel rings of
// 202-269 MB then measured best in FP64 and FP32 (18-24 slots of 11.2 MB), and
// 291 MB lost 12 %.
constexpr double BLOCK_RING_BYTES = 225e6;
constexpr long BLOCK_MIN_SLABS = 4;
constexpr long BLOCK_MAX_SLABS = 32;

// The sweep lbm_create chooses for a lattice of elem-byte values: axis 2 (z)
// or 1 (y) and K slabs per chunk (0: single steps only).  Along each axis K is
// the number of slabs whose ring of K + 2 fits ring_bytes, at most
// BLOCK_MAX_SLABS and the number of slabs.  The axis whose chunk holds more
// nodes is taken (the larger slabs on a tie): the two extra ring slots are
// then the smallest share of the ring.  A chunk of fewer than min_slabs
// turns the sweep off.  axis_env / planes_env (may be null) are the values of
// LBM_BLOCK_AXIS ("z" or "y") and LBM_BLOCK_PLANES (K, 0 turns the sweep
// off), which override the choice; any other value of either is reported in
// `bad`.  blocked = false (2-byte storage, whose step is bound by
// arithmetic) makes the default K zero.
struct BlockChoice {
    int axis;
    uint32_t K;
    const char *bad; // null, or the variable whose value is invalid (the caller fails)
};

inline BlockChoice block_choice(const Geom &g, uint32_t elem, bool blocked, double ring_bytes, long min_slabs,
                                const char *axis_env, const char *planes_env)
{
    auto chunk = [&](int axis) -> long {
        const BlockGeom bg = block_geom(g, axis);
        const double slab_bytes = (double)g.lx * bg.nr * 19.0 * (double)elem;
        const long k = (long)(ring_bytes / slab_bytes) - 2;
        return std::max<long>(0, std::min<long>(std::min<long>(k, BLOCK_MAX_SLABS), (long)bg.ns));
    };
    auto chunk_nodes = [&](int axis) -> uint64_t {
        return (uint64_t)chunk(axis) * g.lx * block_geom(g, axis).nr;
    };
    // z slabs (lx * ly nodes) are the larger ones when ly >= lz
    const int big = g.ly >= g.lz ? 2 : 1, small = 3 - big;
    int axis = chunk_nodes(small) > chunk_nodes(big) ? small : big;
    if (axis_env) {
        if (strcmp(axis_env, "y") == 0)
            axis = 1;
        else if (strcmp(axis_env, "z") == 0)
            axis = 2;
        else
            return BlockChoice{axis, 0, "LBM_BLOCK_AXIS (y or z)"};
    }
    long kb = blocked ? chunk(axis) : 0;
    if (kb < min_slabs)
        kb = 0;
    if (planes_env) {
        char *end = nullptr;
        const long v = strtol(planes_env, &end, 10);
        if (end == planes_env || *end != 0 || v < 0)
            return BlockChoice{axis, 0, "LBM_BLOCK_PLANES (an integer >= 0)"};
        kb = v;
    }
    const BlockGeom bg = block_geom(g, axis);
    kb = std::min<long>(std::min<long>(kb, BLOCK_MAX_SLABS), (long)bg.ns);
    if (kb < 1 || (uint64_t)slab_padded(g.lx * bg.nr) * (uint64_t)kb >= (1ull << 31))
        kb = 0;
    return BlockChoice{axis, (uint32_t)kb, nullptr};
}

// A pair under one uniform force (r.fx, r.fy, r.fz).
inline std::vector<BlockArgs> block_schedule(const Geom &g, int axis, const RatesD &r, uint32_t K, void *lat_in,
                                             void *lat_out, void *ring, void *head)
{
    const ForceSpec u = {r.fx, r.fy, r.fz, nullptr, 0.0, 0.0, 0.0};
    const StepForce f = {r, u, u, 0};
    return block_schedule(g, axis, f, f, K, lat_in, lat_out, ring, head);
}

} // namespace d3q19
