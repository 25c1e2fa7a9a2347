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
// launch over a ring of 2K + 2 slots measured slower.)  When whole slabs are
// too large for a useful chunk, the slabs are cut into row windows and the
// whole list above runs once per window (block_windowed).
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

// Row windows.  W = 0 (or W >= nr) sweeps whole slabs.  Otherwise the rows
// are taken W at a time (the last window may be narrower), and each window is
// swept on its own: step t of rows r0 - 1 .. r0 + W (wrapping along the
// rows) goes to ring slabs of W + 2 rows, step t+1 of rows r0 .. r0 + W - 1
// reads them.  The two extra rows are computed twice, by this window and the
// neighbouring one; the values are the same, since each is a function of the
// step-t lattice alone.  Windows keep the ring's slabs small when the
// lattice's are too large for a useful chunk (the 2048x1024x512 channel).
inline bool block_windowed(const Geom &g, int axis, uint32_t W) { return W > 0 && W < block_geom(g, axis).nr; }
inline uint32_t block_ring_rows(const Geom &g, int axis, uint32_t W)
{
    return block_windowed(g, axis, W) ? W + 2 : block_geom(g, axis).nr;
}

// Ring slots and buffer sizes in elements (slabs of lx * ring rows nodes).
inline uint32_t block_ring_slots(uint32_t K) { return K + 2; }
inline uint64_t block_ring_elems(const Geom &g, int axis, uint32_t K, uint32_t W = 0)
{
    return 19ull * block_ring_slots(K) * g.lx * block_ring_rows(g, axis, W);
}
inline uint64_t block_head_elems(const Geom &g, int axis, uint32_t W = 0)
{
    return 19ull * 2 * g.lx * block_ring_rows(g, axis, W);
}

// The force of one of the two steps of a pair: rates with the uniform force
// (r.fx, r.fy, r.fz), or per_node with the collision's force fc and the one
// the fused fields were formed with, fm (BlockArgs::per_node).
struct StepForce {
    RatesD r;
    ForceSpec fc, fm;
    int per_node;
};

// Launch list of one pair: step t with force f0, step t+1 with force f1,
// whole slabs (W = 0) or row windows of W rows.
inline std::vector<BlockArgs> block_schedule(const Geom &g, int axis, const StepForce &f0, const StepForce &f1,
                                             uint32_t K, uint32_t W, void *lat_in, void *lat_out, void *ring,
                                             void *head)
{
    const BlockGeom bg = block_geom(g, axis);
    const uint32_t ns = bg.ns, nr = bg.nr, slots = block_ring_slots(K);
    const bool win = block_windowed(g, axis, W);
    const uint32_t rr = block_ring_rows(g, axis, W);
    const uint64_t rslab = (uint64_t)g.lx * rr; // nodes of a ring slab
    const PlaneStore full_in = {lat_in, g.N, ns, make_fastdiv(ns), bg.lat_slab_stride, bg.lat_row_stride, nullptr, 0,
                                0, 0};
    const PlaneStore full_out = {lat_out, g.N, ns, make_fastdiv(ns), bg.lat_slab_stride, bg.lat_row_stride, nullptr, 0,
                                 0, 0};
    const PlaneStore ring_only = {ring, slots * rslab, slots, make_fastdiv(slots), rslab, g.lx, nullptr, 0, 0, 0};
    PlaneStore ring_head = ring_only;
    ring_head.head = head;
    ring_head.head_stride = 2 * rslab;
    ring_head.head_lo = 0;
    ring_head.head_n = std::min<uint32_t>(2, ns);

    // local-row frames of the two launch kinds (BlockArgs): step t reads the
    // lattice and writes the ring, step t+1 reads the ring and writes the
    // lattice
    struct Rows {
        uint32_t nrw, src_r0, src_nr, dst_r0, node_r0;
    };
    std::vector<BlockArgs> out;
    auto add = [&](int32_t s0, uint32_t count, const PlaneStore &src, const PlaneStore &dst, int nt,
                   const StepForce &f, const Rows &rw) {
        BlockArgs a;
        a.g = g;
        a.ns = ns;
        a.nr = nr;
        a.nrw = rw.nrw;
        a.src_r0 = rw.src_r0;
        a.src_nr = rw.src_nr;
        a.dst_r0 = rw.dst_r0;
        a.node_r0 = rw.node_r0;
        a.slab = g.lx * rw.nrw;
        a.slabp = slab_padded(a.slab);
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
    const uint32_t step = win ? W : nr;
    for (uint32_t r0 = 0; r0 < nr; r0 += step) {
        const uint32_t w = std::min(step, nr - r0);
        Rows ra, rb;
        if (win) {
            const uint32_t lo = r0 == 0 ? nr - 1 : r0 - 1; // lattice row of ring row 0
            ra = Rows{w + 2, lo, nr, 0, lo};
            rb = Rows{w, 1, w + 2, r0, r0};
        } else {
            ra = Rows{nr, 0, nr, 0, 0};
            rb = ra;
        }
        for (uint32_t s0 = 0; s0 < ns; s0 += K) {
            const uint32_t n = std::min(K, ns - s0);
            add((int32_t)s0, n, full_in, ring_head, 0, f0, ra);
            const uint32_t lo = std::max<uint32_t>(s0 == 0 ? 0 : s0 - 1, 1), hi = s0 + n - 1;
            if (hi > lo)
                add((int32_t)lo, hi - lo, ring_only, full_out, 1, f1, rb);
        }
        // slabs ns - 1 and 0 (one slab when ns == 1)
        add((int32_t)ns - 1, std::min<uint32_t>(2, ns), ring_head, full_out, 1, f1, rb);
    }
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
// or 1 (y), K slabs per chunk (0: single steps only) and W rows per window
// (0: whole slabs).  Along each axis K is the number of slabs whose ring of
// K + 2 fits ring_bytes, at most BLOCK_MAX_SLABS and the number of slabs.
// The axis whose chunk holds more nodes is taken (the larger slabs on a
// tie): the two extra ring slots are then the smallest share of the ring.
// When that chunk has fewer than min_slabs slabs, the slabs are cut into
// row windows instead: as few windows, of equal height, as keep a window's
// slab within ring_bytes / (BLOCK_WINDOW_K + 2), swept along the axis with
// more slabs; with no such window the sweep is off.  axis_env / planes_env /
// rows_env (may be null) are the values of LBM_BLOCK_AXIS ("z" or "y"),
// LBM_BLOCK_PLANES (K, 0 turns the sweep off) and LBM_BLOCK_ROWS (W, 0 for
// whole slabs), which override the choice (K then follows from W's slabs
// unless LBM_BLOCK_PLANES is set); any other value is reported in `bad`.
// blocked = false (2-byte storage, whose step is bound by arithmetic) makes
// the default K zero.
constexpr long BLOCK_WINDOW_K = 18;

struct BlockChoice {
    int axis;
    uint32_t K;
    uint32_t W;
    const char *bad; // null, or the variable whose value is invalid (the caller fails)
};

inline BlockChoice block_choice(const Geom &g, uint32_t elem, bool blocked, double ring_bytes, long min_slabs,
                                const char *axis_env, const char *planes_env, const char *rows_env = nullptr)
{
    // slabs of `rows` rows per chunk along `axis`
    auto chunk_rows = [&](int axis, uint32_t rows) -> long {
        const double slab_bytes = (double)g.lx * rows * 19.0 * (double)elem;
        const long k = (long)(ring_bytes / slab_bytes) - 2;
        return std::max<long>(0, std::min<long>(std::min<long>(k, BLOCK_MAX_SLABS), (long)block_geom(g, axis).ns));
    };
    auto chunk = [&](int axis) -> long { return chunk_rows(axis, block_geom(g, axis).nr); };
    auto chunk_nodes = [&](int axis) -> uint64_t {
        return (uint64_t)chunk(axis) * g.lx * block_geom(g, axis).nr;
    };
    auto parse = [](const char *v, long &out) -> bool {
        char *end = nullptr;
        out = strtol(v, &end, 10);
        return end != v && *end == 0 && out >= 0;
    };
    // z slabs (lx * ly nodes) are the larger ones when ly >= lz
    const int big = g.ly >= g.lz ? 2 : 1, small = 3 - big;
    int axis = chunk_nodes(small) > chunk_nodes(big) ? small : big;
    if (blocked && chunk(axis) < min_slabs)
        axis = g.ly >= g.lz ? 1 : 2; // row windows, along the axis with more slabs
    if (axis_env) {
        if (strcmp(axis_env, "y") == 0)
            axis = 1;
        else if (strcmp(axis_env, "z") == 0)
            axis = 2;
        else
            return BlockChoice{axis, 0, 0, "LBM_BLOCK_AXIS (y or z)"};
    }
    const BlockGeom bg = block_geom(g, axis);
    long kb = blocked ? chunk(axis) : 0;
    long W = 0;
    if (blocked && kb < min_slabs) {
        const double row_bytes = (double)g.lx * 19.0 * (double)elem;
        const long wmax = (long)(ring_bytes / (double)(BLOCK_WINDOW_K + 2) / row_bytes) - 2;
        kb = 0;
        if (wmax >= 1 && wmax < (long)bg.nr) {
            const long nwin = ((long)bg.nr + wmax - 1) / wmax;
            W = ((long)bg.nr + nwin - 1) / nwin;
            kb = chunk_rows(axis, (uint32_t)W + 2);
            if (kb < min_slabs)
                kb = 0, W = 0;
        }
    }
    if (rows_env) {
        if (!parse(rows_env, W))
            return BlockChoice{axis, 0, 0, "LBM_BLOCK_ROWS (an integer >= 0)"};
        if (W >= (long)bg.nr)
            W = 0;
        if (W > 0) {
            kb = chunk_rows(axis, (uint32_t)W + 2);
        } else {
            kb = blocked ? chunk(axis) : 0;
            if (kb < min_slabs)
                kb = 0;
        }
    }
    if (planes_env) {
        long v;
        if (!parse(planes_env, v))
            return BlockChoice{axis, 0, 0, "LBM_BLOCK_PLANES (an integer >= 0)"};
        kb = v;
    }
    kb = std::min<long>(std::min<long>(kb, BLOCK_MAX_SLABS), (long)bg.ns);
    const uint32_t rows = W > 0 ? (uint32_t)W + 2 : bg.nr;
    if (kb < 1 || (uint64_t)slab_padded(g.lx * rows) * (uint64_t)kb >= (1ull << 31))
        kb = 0;
    return BlockChoice{axis, (uint32_t)kb, kb > 0 ? (uint32_t)W : 0u, nullptr};
}

// A pair under one uniform force (r.fx, r.fy, r.fz).
inline std::vector<BlockArgs> block_schedule(const Geom &g, int axis, const RatesD &r, uint32_t K, uint32_t W,
                                             void *lat_in, void *lat_out, void *ring, void *head)
{
    const ForceSpec u = {r.fx, r.fy, r.fz, nullptr, 0.0, 0.0, 0.0};
    const StepForce f = {r, u, u, 0};
    return block_schedule(g, axis, f, f, K, W, lat_in, lat_out, ring, head);
}

} // namespace d3q19
