// Host-side check of the lattice index arithmetic the kernels use
// (d3q19_common.h): for every node and direction the pull source lies inside
// the lattice, the gather (pull with bounce-back) is a bijection onto the
// 19*N post-collision slots, the fast division is exact, and the XCD block
// remap is a permutation.  Built with hipcc and run on the CPU; launches
// nothing on a GPU.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "d3q19_common.h"

using namespace d3q19;

template <int I>
static bool visit(const Nbr &b, uint64_t N, std::vector<uint8_t> &seen)
{
    uint64_t slot;
    if (walled<I>(b))
        slot = (uint64_t)OPP[I] * N + b.n;
    else {
        uint32_t s = src_of<I>(b);
        if (s >= N) {
            fprintf(stderr, "direction %d node %u: source %u out of range\n", I, b.n, s);
            return false;
        }
        slot = (uint64_t)I * N + s;
    }
    if (seen[slot]) {
        fprintf(stderr, "slot %llu gathered twice\n", (unsigned long long)slot);
        return false;
    }
    seen[slot] = 1;
    return true;
}

template <int... Is>
static bool visit_all(const Nbr &b, uint64_t N, std::vector<uint8_t> &seen, std::integer_sequence<int, Is...>)
{
    return (visit<Is>(b, N, seen) && ...);
}

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
    g.remap = 1;
    return g;
}

static bool check_full(uint32_t lx, uint32_t ly, uint32_t lz)
{
    Geom g = geom(lx, ly, lz);
    std::vector<uint8_t> seen((size_t)g.N * 19, 0);
    for (uint64_t n = 0; n < g.N; n++) {
        Nbr b = make_nbr((uint32_t)n, g);
        uint32_t ix = (uint32_t)(n % lx), iy = (uint32_t)((n / lx) % ly), iz = (uint32_t)(n / ((uint64_t)lx * ly));
        if (b.wlo != (ix == 0) || b.whi != (ix == lx - 1)) {
            fprintf(stderr, "wall flags wrong at node %llu\n", (unsigned long long)n);
            return false;
        }
        // the +y neighbour offset must land on (ix, iy+1 mod ly, iz)
        uint64_t yp = ix + (uint64_t)lx * (((iy + 1) % ly) + (uint64_t)ly * iz);
        if ((uint32_t)(n + b.yp) != yp) {
            fprintf(stderr, "y+ offset wrong at node %llu\n", (unsigned long long)n);
            return false;
        }
        uint64_t zm = ix + (uint64_t)lx * (iy + (uint64_t)ly * ((iz + lz - 1) % lz));
        if ((uint32_t)(n + b.zm) != zm) {
            fprintf(stderr, "z- offset wrong at node %llu\n", (unsigned long long)n);
            return false;
        }
        if (!visit_all(b, g.N, seen, std::make_integer_sequence<int, 19>{}))
            return false;
    }
    for (size_t k = 0; k < seen.size(); k++)
        if (!seen[k]) {
            fprintf(stderr, "slot %zu never gathered\n", k);
            return false;
        }
    return true;
}

template <int... Is>
static bool in_range_all(const Nbr &b, uint64_t N, std::integer_sequence<int, Is...>)
{
    return ((walled<Is>(b) || src_of<Is>(b) < N) && ...);
}

// Large grids: exact division on every plane boundary and a sample of nodes,
// sources in range.
static bool check_sampled(uint32_t lx, uint32_t ly, uint32_t lz)
{
    Geom g = geom(lx, ly, lz);
    uint64_t x = 88172645463325252ull;
    for (uint64_t k = 0; k < 20000000ull; k++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        uint64_t n = (k < 4096) ? (k < 2048 ? k : g.N - 1 - (k - 2048)) : x % g.N;
        if (fdiv((uint32_t)n, g.dlx) != n / lx || fdiv((uint32_t)(n / lx), g.dly) != (n / lx) / ly) {
            fprintf(stderr, "fast division wrong at %llu\n", (unsigned long long)n);
            return false;
        }
        Nbr b = make_nbr((uint32_t)n, g);
        if (!in_range_all(b, g.N, std::make_integer_sequence<int, 19>{})) {
            fprintf(stderr, "source out of range at %llu\n", (unsigned long long)n);
            return false;
        }
    }
    return true;
}

// The vector-run pull of k_step_run (VW nodes per thread, 256-thread blocks,
// 64-lane waves), emulated with slot indices as values: every load is in
// bounds and aligned, every shuffle reads the lane the kernel reads, and each
// node receives exactly the slot the one-node gather pulls.
struct RunCtx {
    Geom g;
    uint64_t nruns;
};

template <int VW, int I>
static int64_t run_start(const RunCtx &c, uint64_t t)
{
    if (t >= c.nruns)
        return -1; // inactive lane: loads nothing
    const Nbr b = make_nbr((uint32_t)(t * VW), c.g);
    return (int64_t)((uint64_t)I * c.g.N + src_yz<I>(b));
}

template <int VW, int I>
static bool check_run_dir(const RunCtx &c, uint64_t t)
{
    const uint64_t N = c.g.N;
    const bool active = t < c.nruns;
    const uint32_t lane = (uint32_t)(t & 63);
    const uint32_t n0 = active ? (uint32_t)(t * VW) : 0u;
    const Nbr b = make_nbr(n0, c.g);
    const uint32_t ix0 = n0 - fdiv(n0, c.g.dlx) * c.g.lx;
    const bool whi_last = (ix0 + VW == c.g.lx);
    const bool first = (lane == 0), last = (lane == 63) || (t + 1 >= c.nruns);
    const int64_t s0 = run_start<VW, I>(c, t);
    if (active && (s0 < 0 || (uint64_t)s0 + VW > 19 * N || s0 % VW != 0)) {
        fprintf(stderr, "run %llu dir %d: vector load at %lld out of range or unaligned\n", (unsigned long long)t, I,
                (long long)s0);
        return false;
    }
    int64_t F[VW];
    if constexpr (CX[I] == 0) {
        for (int k = 0; k < VW; k++)
            F[k] = s0 + k;
    } else if constexpr (CX[I] == 1) {
        // __shfl_up: lane - 1 of the same wave, own value on lane 0
        const int64_t src = lane > 0 ? run_start<VW, I>(c, t - 1) : s0;
        int64_t prev = src < 0 ? -1 : src + VW - 1;
        if (active && first && !b.wlo) {
            if (s0 < 1) {
                fprintf(stderr, "run %llu dir %d: edge load below 0\n", (unsigned long long)t, I);
                return false;
            }
            prev = s0 - 1;
        }
        if (active && b.wlo)
            prev = (int64_t)((uint64_t)OPP[I] * N + b.n);
        F[0] = prev;
        for (int k = 1; k < VW; k++)
            F[k] = s0 + k - 1;
    } else {
        // __shfl_down: lane + 1 of the same wave, own value on lane 63
        const int64_t src = lane < 63 ? run_start<VW, I>(c, t + 1) : s0;
        int64_t next = src < 0 ? -1 : src;
        if (active && last && !whi_last) {
            if ((uint64_t)s0 + VW >= 19 * N) {
                fprintf(stderr, "run %llu dir %d: edge load past the end\n", (unsigned long long)t, I);
                return false;
            }
            next = s0 + VW;
        }
        if (active && whi_last)
            next = (int64_t)((uint64_t)OPP[I] * N + b.n + (VW - 1));
        for (int k = 0; k < VW - 1; k++)
            F[k] = s0 + k + 1;
        F[VW - 1] = next;
    }
    if (!active)
        return true;
    for (int k = 0; k < VW; k++) {
        const Nbr bn = make_nbr(n0 + k, c.g);
        const int64_t want = walled<I>(bn) ? (int64_t)((uint64_t)OPP[I] * N + bn.n)
                                           : (int64_t)((uint64_t)I * N + src_of<I>(bn));
        if (F[k] != want) {
            fprintf(stderr, "run %llu node %u dir %d: pulled slot %lld, gather wants %lld\n", (unsigned long long)t,
                    n0 + k, I, (long long)F[k], (long long)want);
            return false;
        }
    }
    return true;
}

template <int VW, int... Is>
static bool check_run_all(const RunCtx &c, uint64_t t, std::integer_sequence<int, Is...>)
{
    return (check_run_dir<VW, Is>(c, t) && ...);
}

template <int VW>
static bool check_runs(uint32_t lx, uint32_t ly, uint32_t lz)
{
    RunCtx c{geom(lx, ly, lz), 0};
    if (lx % VW != 0)
        return true; // the launcher keeps one node per thread for this shape
    c.nruns = c.g.N / VW;
    const uint64_t threads = (c.nruns + 255) / 256 * 256;
    for (uint64_t t = 0; t < threads; t++)
        if (!check_run_all<VW>(c, t, std::make_integer_sequence<int, 19>{}))
            return false;
    return true;
}

static bool check_remap(uint32_t nb)
{
    std::vector<uint8_t> seen(nb, 0);
    for (uint32_t b = 0; b < nb; b++) {
        uint32_t l = xcd_remap(b, nb);
        if (l >= nb || seen[l]) {
            fprintf(stderr, "xcd_remap not a permutation for nb=%u\n", nb);
            return false;
        }
        seen[l] = 1;
    }
    return true;
}

int main()
{
    const uint32_t shapes[][3] = {{2, 1, 1},   {3, 1, 4},    {5, 7, 3},    {64, 3, 2},   {33, 17, 9},
                                  {130, 4, 3}, {12, 10, 6},  {32, 64, 32}, {16, 12, 8},  {48, 40, 24},
                                  {96, 384, 192}, {192, 768, 384}};
    for (auto &s : shapes) {
        if (!check_full(s[0], s[1], s[2])) {
            fprintf(stderr, "FAIL full %u %u %u\n", s[0], s[1], s[2]);
            return 1;
        }
        printf("ok full %u x %u x %u\n", s[0], s[1], s[2]);
        if (!check_runs<2>(s[0], s[1], s[2]) || !check_runs<4>(s[0], s[1], s[2])) {
            fprintf(stderr, "FAIL runs %u %u %u\n", s[0], s[1], s[2]);
            return 1;
        }
        printf("ok runs %u x %u x %u\n", s[0], s[1], s[2]);
    }
    const uint32_t big[][3] = {{512, 2048, 1024}, {256, 1024, 512}, {1023, 1021, 1025}};
    for (auto &s : big) {
        if (!check_sampled(s[0], s[1], s[2])) {
            fprintf(stderr, "FAIL sampled %u %u %u\n", s[0], s[1], s[2]);
            return 1;
        }
        printf("ok sampled %u x %u x %u\n", s[0], s[1], s[2]);
    }
    for (uint32_t nb : {1u, 7u, 8u, 9u, 15u, 16u, 17u, 1000u, 221184u, 4194304u})
        if (!check_remap(nb))
            return 1;
    printf("ok remap\n");
    return 0;
}
