// Host-side check of the two-step blocked sweep (csrc/blocked.h): the launch
// list is emulated on the CPU with the kernels' own addressing helpers
// (kernels.h), every stored value being a tag (state, node, population).
// Checked: every index lies inside its buffer; within one launch no location
// is both read and written (its threads run concurrently); every gathered
// value is the tag the single-step pull would read (P_s(x - c_i)_i, or
// P_s(x)_opp(i) at a wall), so no ring slot is read after being overwritten;
// every node and population of the output lattice holds its step-t+2 tag and
// is written exactly once.  Whole slabs and row windows (ragged last window,
// windows of one row, windows one row short of the slab).  Also: the default
// sweep (block_choice) for the tuned workloads and under the environment
// overrides.  Launches nothing on a GPU.
#include <stdio.h>
#include <stdlib.h>

#include <set>
#include <utility>
#include <vector>

#include "blocked.h"

using namespace d3q19;

struct Buf {
    std::vector<uint64_t> v;
};

static Geom geom(uint32_t lx, uint32_t ly, uint32_t lz)
{
    Geom g;
    g.lx = lx;
    g.ly = ly;
    g.lz = lz;
    g.plane = lx * ly;
    g.N = (uint64_t)lx * ly * lz;
    g.dlx = make_fastdiv(lx);
    g.dly = make_fastdiv(ly);
    g.remap = 0;
    return g;
}

struct Ctx {
    Geom g;
    Buf *bufs[4];
    bool ok = true;
    std::set<std::pair<const void *, uint64_t>> reads, writes; // locations touched by the current launch
    std::vector<std::pair<uint64_t *, uint64_t>> pending;       // stores of the current launch
    uint64_t tag(uint64_t s, uint64_t n, int i) const { return (s * g.N + n) * 19 + (uint64_t)i + 1; }
    Buf *find(void *p)
    {
        for (Buf *b : bufs)
            if ((void *)b->v.data() == p)
                return b;
        return nullptr;
    }
};

// Node index of (ix, r, slab s) for a sweep along axis 2 (z) or 1 (y).
static uint64_t node(const Geom &g, int axis, uint32_t ix, uint32_t r, uint32_t sl)
{
    const uint32_t iy = axis == 2 ? r : sl, iz = axis == 2 ? sl : r;
    return ix + (uint64_t)g.lx * (iy + (uint64_t)g.ly * iz);
}

template <int AXIS, int I>
static void gather_one(Ctx &c, const PlaneRef (&pl)[3], const BlockLane &l, uint32_t ix, uint32_t iy, uint32_t z,
                       int s)
{
    // (iy, z) here are (row, slab) of the sweep
    void *buf;
    uint64_t idx;
    block_src<AXIS, I>(pl, l, buf, idx);
    Buf *B = c.find(buf);
    if (!B || idx >= B->v.size()) {
        fprintf(stderr, "dir %d node (%u,%u,%u): index %llu outside its buffer\n", I, ix, iy, z,
                (unsigned long long)idx);
        c.ok = false;
        return;
    }
    const Geom &g = c.g;
    const uint32_t nr = AXIS == 2 ? g.ly : g.lz, ns = AXIS == 2 ? g.lz : g.ly;
    const int cr = AXIS == 2 ? CY[I] : CZ[I], cs = AXIS == 2 ? CZ[I] : CY[I];
    uint64_t want;
    const bool walled = (CX[I] == 1 && ix == 0) || (CX[I] == -1 && ix + 1 == g.lx);
    if (walled) {
        want = c.tag(s, node(g, AXIS, ix, iy, z), OPP[I]);
    } else {
        const uint32_t sx = ix - CX[I];
        const uint32_t sr = (iy + nr - cr) % nr, ss = (z + ns - cs) % ns;
        want = c.tag(s, node(g, AXIS, sx, sr, ss), I);
    }
    c.reads.insert({buf, idx});
    if (B->v[idx] != want) {
        fprintf(stderr, "dir %d node (%u,%u,%u) state %d: read tag %llu, want %llu\n", I, ix, iy, z, s,
                (unsigned long long)B->v[idx], (unsigned long long)want);
        c.ok = false;
    }
}

template <int AXIS, int... Is>
static void gather_all(Ctx &c, const PlaneRef (&pl)[3], const BlockLane &l, uint32_t ix, uint32_t iy, uint32_t z,
                       int s, std::integer_sequence<int, Is...>)
{
    (gather_one<AXIS, Is>(c, pl, l, ix, iy, z, s), ...);
}

// the kernel's store addressing: element (uint32) r * row_stride + ix past the
// slab offset (r: the local row plus the launch's dst_r0), population stride
// d.stride
static void store(Ctx &c, const PlaneRef &d, uint32_t ix, uint32_t r, uint64_t n, int s)
{
    Buf *B = c.find(d.buf);
    const uint32_t e = r * (uint32_t)d.row_stride + ix;
    for (int i = 0; i < 19; i++) {
        const uint64_t idx = d.off + (uint64_t)i * d.stride + e;
        if (!B || idx >= B->v.size()) {
            fprintf(stderr, "store outside its buffer\n");
            c.ok = false;
            return;
        }
        c.writes.insert({d.buf, idx});
        c.pending.push_back({&B->v[idx], c.tag(s, n, i)});
    }
}

static bool check(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t K, int axis, uint32_t W = 0)
{
    Ctx c;
    c.g = geom(lx, ly, lz);
    const Geom &g = c.g;
    Buf in, out, ring, head;
    in.v.resize(19 * g.N);
    out.v.assign(19 * g.N, 0);
    ring.v.assign(block_ring_elems(g, axis, K, W), 0);
    head.v.assign(block_head_elems(g, axis, W), 0);
    c.bufs[0] = &in;
    c.bufs[1] = &out;
    c.bufs[2] = &ring;
    c.bufs[3] = &head;
    for (uint64_t n = 0; n < g.N; n++)
        for (int i = 0; i < 19; i++)
            in.v[(uint64_t)i * g.N + n] = c.tag(0, n, i);
    RatesD r = {};
    const auto launches =
        block_schedule(g, axis, r, K, W, in.v.data(), out.v.data(), ring.v.data(), head.v.data());
    uint64_t out_writes = 0;
    for (const BlockArgs &a : launches) {
      c.reads.clear();
      c.writes.clear();
      c.pending.clear();
      {
        const int s = (a.src.ring == (void *)in.v.data()) ? 0 : 1; // A reads step t, B step t+1
        if (a.nt != (a.dst.ring == (void *)out.v.data() ? 1 : 0)) {
            fprintf(stderr, "non-temporal stores must go to the output lattice only\n");
            return false;
        }
        if (a.s0 < -(int32_t)a.ns || a.s0 >= (int32_t)a.ns || a.count > a.ns) {
            fprintf(stderr, "launch s0 %d count %u outside the range block_plane wraps once\n", a.s0, a.count);
            return false;
        }
        if (a.slab != g.lx * a.nrw || a.src_r0 >= a.src_nr || a.node_r0 >= a.nr ||
            (uint64_t)a.nrw + a.src_r0 > 2ull * a.src_nr || (uint64_t)a.nrw + a.node_r0 > 2ull * a.nr) {
            fprintf(stderr, "launch rows: nrw %u src_r0 %u src_nr %u node_r0 %u nr %u outside one wrap\n", a.nrw,
                    a.src_r0, a.src_nr, a.node_r0, a.nr);
            return false;
        }
        if (a.slabp % WAVE_LANES != 0 || a.slabp < a.slab || a.slabp - a.slab >= WAVE_LANES) {
            fprintf(stderr, "padded slab %u of slab %u is not the next multiple of a wave\n", a.slabp, a.slab);
            return false;
        }
        // whole blocks of the launch (256 or 512 threads; the larger covers both)
        const uint32_t nthreads = (a.slabp * a.count + 511u) / 512u * 512u;
        for (uint32_t t = 0; t < nthreads; t++) {
            // the kernel takes the slab of the wave's first lane
            const uint32_t sl = fdiv(t / WAVE_LANES * WAVE_LANES, a.dslabp);
            if (sl != fdiv(t, a.dslabp)) {
                fprintf(stderr, "thread %u: its wave spans two slabs\n", t);
                return false;
            }
            const uint32_t w = t - sl * a.slabp;
            if (sl >= a.count || w >= a.slab)
                continue;
            // the kernel's row mapping (block_row) and gather addressing
            const BlockRow br = block_row(a, w);
            const uint32_t ix = br.ix, rr = br.rg;
            const uint32_t sb = block_plane(a, sl);
            PlaneRef pl[3];
            block_planes(a, sb, pl);
            const BlockLane l = block_lane(ix, br.rs, a.src_nr, (uint32_t)a.src.row_stride, g.lx);
            if (axis == 2)
                gather_all<2>(c, pl, l, ix, rr, sb, s, std::make_integer_sequence<int, 19>{});
            else
                gather_all<1>(c, pl, l, ix, rr, sb, s, std::make_integer_sequence<int, 19>{});
            if (!c.ok)
                return false;
            const uint64_t n = node(g, axis, ix, rr, sb);
            store(c, plane_write(a.dst, sb), ix, br.rl + a.dst_r0, n, s + 1);
            if (a.dst.ring == (void *)out.v.data())
                out_writes += 19;
            const PlaneRef hd = plane_write_head(a.dst, sb);
            if (hd.buf != nullptr)
                store(c, hd, ix, br.rl + a.dst_r0, n, s + 1);
            if (!c.ok)
                return false;
        }
      }
      for (const auto &w : c.writes)
          if (c.reads.count(w)) {
              fprintf(stderr, "%ux%ux%u K=%u: a launch reads and writes the same location\n", lx, ly, lz, K);
              return false;
          }
      if (c.writes.size() != c.pending.size()) {
          fprintf(stderr, "%ux%ux%u K=%u: a launch writes a location twice\n", lx, ly, lz, K);
          return false;
      }
      for (const auto &w : c.pending)
          *w.first = w.second;
    }
    if (out_writes != 19 * g.N) {
        fprintf(stderr, "%ux%ux%u K=%u W=%u axis %d: %llu output stores for %llu values\n", lx, ly, lz, K, W, axis,
                (unsigned long long)out_writes, (unsigned long long)(19 * g.N));
        return false;
    }
    for (uint64_t n = 0; n < g.N; n++)
        for (int i = 0; i < 19; i++)
            if (out.v[(uint64_t)i * g.N + n] != c.tag(2, n, i)) {
                fprintf(stderr, "%ux%ux%u K=%u axis %d: output node %llu population %d not step t+2\n", lx, ly,
                        lz, K, axis, (unsigned long long)n, i);
                return false;
            }
    printf("ok %ux%ux%u K=%u W=%u axis %c (%zu launches)\n", lx, ly, lz, K, W, axis == 2 ? 'z' : 'y',
           launches.size());
    return true;
}

// lbm_create's default sweep (block_choice) for the workloads it was tuned on.
static bool check_choice(uint32_t lx, uint32_t ly, uint32_t lz, uint32_t elem, bool blocked, const char *axis_env,
                         const char *planes_env, int want_axis, uint32_t want_K, uint32_t want_W = 0,
                         const char *rows_env = nullptr)
{
    const Geom g = geom(lx, ly, lz);
    const BlockChoice c =
        block_choice(g, elem, blocked, BLOCK_RING_BYTES, BLOCK_MIN_SLABS, axis_env, planes_env, rows_env);
    if (c.bad || c.axis != want_axis || c.K != want_K || c.W != want_W) {
        fprintf(stderr, "%ux%ux%u elem %u: sweep axis %d K %u W %u, want axis %d K %u W %u\n", lx, ly, lz, elem,
                c.axis, c.K, c.W, want_axis, want_K, want_W);
        return false;
    }
    return true;
}

int main()
{
    // 192x768x384: 11.2 MB slabs, 18 to a chunk (224 MB rings): y slabs in
    // FP64, z slabs in FP32; FP16 storage single steps; the environment
    // overrides (z slabs of 22.4 MB in FP64: 8 to a chunk)
    bool choice = check_choice(192, 768, 384, 8, true, nullptr, nullptr, 1, 18) &&
                  check_choice(192, 768, 384, 4, true, nullptr, nullptr, 2, 18) &&
                  check_choice(192, 768, 384, 2, false, nullptr, nullptr, 2, 0) &&
                  check_choice(192, 768, 384, 8, true, "z", nullptr, 2, 8) &&
                  check_choice(192, 768, 384, 8, true, nullptr, "0", 1, 0) &&
                  check_choice(192, 768, 384, 2, false, "y", "12", 1, 12) &&
                  check_choice(192, 768, 384, 8, true, nullptr, "40", 1, 32) &&
                  // 2048x1024x512 in FP32: 80 MB z and 40 MB y slabs (three
                  // to a ring): y slabs in four windows of 256 rows (10 MB
                  // window slabs, 20 to a chunk); FP16 storage single steps;
                  // the overrides
                  check_choice(512, 2048, 1024, 4, true, nullptr, nullptr, 1, 20, 256) &&
                  check_choice(512, 2048, 1024, 2, false, nullptr, nullptr, 1, 0, 0) &&
                  check_choice(512, 2048, 1024, 4, true, "z", nullptr, 2, 20, 256) &&
                  check_choice(512, 2048, 1024, 4, true, nullptr, nullptr, 1, 0, 0, "0") &&
                  check_choice(512, 2048, 1024, 4, true, nullptr, "8", 1, 8, 128, "128") &&
                  check_choice(512, 2048, 1024, 4, true, nullptr, "0", 1, 0, 0) &&
                  // the 2048x1024x512 FP64 grid of half size (256x1024x512):
                  // 19.9 MB y slabs, 9 to a chunk, whole slabs
                  check_choice(256, 1024, 512, 8, true, nullptr, nullptr, 1, 9) &&
                  // windows on a small grid by the override only
                  check_choice(12, 10, 6, 8, true, nullptr, "3", 2, 3, 4, "4") &&
                  check_choice(12, 10, 6, 8, true, nullptr, nullptr, 2, 6, 0, "10") &&
                  // 96x384x192 in FP32 (2.8 MB z slabs, 1.4 MB y slabs): 32
                  // z slabs hold more nodes than 32 y slabs
                  check_choice(96, 384, 192, 4, true, nullptr, nullptr, 2, 32) &&
                  // small grids: at most 32 slabs and at most the slab count;
                  // equal chunks take the larger slabs
                  check_choice(32, 64, 32, 8, true, nullptr, nullptr, 2, 32) &&
                  check_choice(12, 10, 6, 8, true, nullptr, nullptr, 2, 6) &&
                  check_choice(12, 6, 10, 8, true, nullptr, nullptr, 1, 6) &&
                  check_choice(8, 80, 50, 8, true, nullptr, nullptr, 2, 32) &&
                  check_choice(12, 6, 40, 8, true, nullptr, nullptr, 1, 6);
    if (!choice)
        return 1;
    // invalid environment values are reported, not taken as a default
    const Geom g0 = geom(192, 768, 384);
    for (const char *bad_axis : {"Y", "x", "", "yz"})
        if (!block_choice(g0, 8, true, BLOCK_RING_BYTES, BLOCK_MIN_SLABS, bad_axis, nullptr).bad) {
            fprintf(stderr, "LBM_BLOCK_AXIS=\"%s\" accepted\n", bad_axis);
            return 1;
        }
    for (const char *bad_k : {"", "abc", "4x", "-1", " "}) {
        if (!block_choice(g0, 8, true, BLOCK_RING_BYTES, BLOCK_MIN_SLABS, nullptr, bad_k).bad) {
            fprintf(stderr, "LBM_BLOCK_PLANES=\"%s\" accepted\n", bad_k);
            return 1;
        }
        if (!block_choice(g0, 8, true, BLOCK_RING_BYTES, BLOCK_MIN_SLABS, nullptr, nullptr, bad_k).bad) {
            fprintf(stderr, "LBM_BLOCK_ROWS=\"%s\" accepted\n", bad_k);
            return 1;
        }
    }
    printf("ok sweep choice\n");

    const uint32_t cases[][4] = {{4, 3, 1, 2},  {4, 3, 2, 2},  {5, 4, 3, 2},  {6, 5, 7, 2},   {5, 3, 9, 4},
                                 {4, 4, 8, 4},  {3, 5, 10, 3}, {8, 6, 16, 1}, {8, 6, 17, 16}, {7, 5, 33, 8},
                                 {2, 1, 5, 2},  {6, 2, 4, 1},  {9, 7, 12, 5}, {4, 3, 3, 16},
                                 {4, 37, 41, 18}, {3, 70, 4, 32}, {2, 5, 75, 32}};
    bool ok = true;
    for (auto &k : cases)
        for (int axis : {2, 1})
            ok = check(k[0], k[1], k[2], k[3], axis) && ok;
    if (!ok)
        return 1;
    printf("ok blocked sweep\n");
    // row windows: {lx, ly, lz, K, W} (W rows of the slab's nr: ragged last
    // windows, windows of one row, W = nr - 1, one and two slabs)
    const uint32_t wcases[][5] = {{4, 7, 9, 3, 2},  {5, 8, 6, 2, 3},  {3, 10, 4, 4, 1},  {6, 5, 7, 2, 4},
                                  {4, 6, 6, 5, 5},  {3, 9, 2, 2, 2},  {4, 12, 5, 1, 4},  {2, 3, 1, 2, 2},
                                  {3, 4, 11, 16, 3}, {5, 13, 8, 6, 7}, {4, 2, 9, 3, 1},  {3, 11, 3, 32, 5}};
    for (auto &k : wcases)
        for (int axis : {2, 1}) {
            const uint32_t nr = axis == 2 ? k[1] : k[2];
            if (k[4] < nr)
                ok = check(k[0], k[1], k[2], k[3], axis, k[4]) && ok;
        }
    if (!ok)
        return 1;
    printf("ok row windows\n");
    return 0;
}
