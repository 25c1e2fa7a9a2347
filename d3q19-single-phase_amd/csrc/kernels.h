// Kernel argument blocks and per-precision launchers (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19_common.h"

namespace d3q19 {

// Relaxation rates and force as passed by the driver (FP64), with their FP32
// roundings for the kernels that compute in FP32 (set by with_f32 on the
// host, so no kernel converts them per thread).
struct RatesD {
    double s1, s2, s4, s9, s10, s13, s16;
    double omegepsl, omegepslj, omegxx;
    double fx, fy, fz;
    Rates<float> f32;
};

inline RatesD with_f32(RatesD r)
{
    r.f32 = Rates<float>{(float)r.s1,       (float)r.s2,        (float)r.s4,     (float)r.s9, (float)r.s10,
                         (float)r.s13,      (float)r.s16,       (float)r.omegepsl, (float)r.omegepslj,
                         (float)r.omegxx,   (float)r.fx,        (float)r.fy,     (float)r.fz};
    return r;
}

// The rates in the kernel's arithmetic type.
template <typename T>
__host__ __device__ __forceinline__ Rates<T> rates_as(const RatesD &r)
{
    if constexpr (std::is_same<T, float>::value) {
        return r.f32;
    } else {
        return Rates<T>{r.s1, r.s2, r.s4, r.s9, r.s10, r.s13, r.s16, r.omegepsl, r.omegepslj, r.omegxx,
                        r.fx, r.fy, r.fz};
    }
}

// Body force at the nodes: the force_realx/y/z arrays of
// Channel-Flow/var_inc.f90 as FORCING (collision.f90:515-527) and FORCINGP
// (collision.f90:529-602) set them.  tab == nullptr: (fx, fy, fz) at every
// node.  Otherwise the FORCINGP field: (fx, fy, fz) = (0, force_in_y, 0)
// outside its two bands, and tab holds the factors that depend on one
// coordinate only, computed once per grid:
//   tab[3*ix + 0..2]             band (0 none, 1, 2), 1 - cos(x9), sin(x9)
//   tab[3*lx + 4*iy + 0..3]      cos(3 y9), sin(3 y9) of band 1, then of band 2
//   tab[3*lx + 4*ly + 2*iz + 0..1] cos(2 z9), sin(2 z9)
// a, b, c are band 1's products of the amplitude with the constants in front
// of those factors (band 2 has the opposite signs).
struct ForceSpec {
    double fx, fy, fz;
    const double *tab;
    double a, b, c;
};

struct StepArgs {
    Geom g;
    const void *src; // P_k
    void *dst;       // P_{k+1}
    const double *rho, *ux, *uy, *uz;
    const int *solid; // node flags (>= 0: inside a particle) for the particle modes
    double rho_off;
    RatesD r;
    const float *mlp;  // learned closure parameters (mode bit 8), see mlp.h
    float inv_ystar;   // u*/nu: y+ per lattice unit of wall distance
    float yscale;      // y+ feature divisor
    int runs;          // 1: modes 0 and 1 of FP64/FP32 take the vector-run kernel when lx allows
    int w32;           // 1: 32-bit pull offsets when N allows (step mode + 32, or + 64 between adjacent
                       // rows above 2^29 nodes in 2-byte storage); 2: + 64 wherever N <= 2^30
    ForceSpec fc;      // mode bit 16: collision force (r.fx, r.fy, r.fz are then unused)
    ForceSpec fm;      // mode bit 16: force that the fused fields u = j + F/2 are formed with
};

struct MacroArgs {
    Geom g;
    const void *L;
    double *rho, *ux, *uy, *uz;
    ForceSpec f;
    double rho_off;
    unsigned long long *errbits;
};

struct CopyArgs {
    Geom g;
    void *L;
    double *aos; // 19 doubles per node, nodes [n0, n0+count)
    uint64_t n0, count;
};

struct InitArgs {
    Geom g;
    void *L;
    double *rho;
    const double *ux, *uy, *uz;
};

struct SumArgs {
    Geom g;
    const void *L;
    const double *rho;
    const int *solid; // nullable: when set, only fluid nodes (solid < 0) count
    double *partials; // per block: sum, then count when solid is set
};

struct PlaneArgs {
    Geom g;
    const void *L;
    const double *rho, *ux, *uy, *uz;
    const int *solid; // nullable: when set, only fluid nodes count
    ForceSpec f;
    double rho_off;
    double *partials;
};

// Plane sums per wall-normal index: 11 moments and the fluid-node count.
constexpr int NPLANE = 12;

// Two-step blocked sweep: the lattice is swept slab by slab along one
// periodic axis, z (AXIS 2) or y (AXIS 1).  Slab s holds the nodes with that
// coordinate equal to s; within it a node is (ix, r), r being the other
// periodic coordinate.
//
// Where the slabs of one lattice state live: slab s in slot s % ring_slots
// of `ring`, its element (I, r, ix) at I * ring_stride + slot * slab_stride +
// r * row_stride + ix; except slabs head_lo .. head_lo + head_n - 1, which a
// reader takes from `head` (same slab and row strides, population stride
// head_stride) and a writer also stores there.  A whole lattice is the store
// with ring_slots = the number of slabs and the lattice's own strides.
struct PlaneStore {
    void *ring;
    uint64_t ring_stride;
    uint32_t ring_slots;
    FastDiv dslots; // division by ring_slots
    uint64_t slab_stride, row_stride;
    void *head; // nullable
    uint64_t head_stride;
    int32_t head_lo;
    uint32_t head_n;
};

// Slab s of a store: buffer, population stride, element offset of the slab
// and stride between its rows.
struct PlaneRef {
    void *buf;
    uint64_t stride;
    uint64_t off;
    uint64_t row_stride;
};

// Ring slot of slab s (s < 2^31), by multiply-high rather than a division.
__host__ __device__ inline uint32_t ring_slot(const PlaneStore &st, uint32_t s)
{
    return s - fdiv(s, st.dslots) * st.ring_slots;
}

// Slab s as a reader sees it.
__host__ __device__ inline PlaneRef plane_read(const PlaneStore &st, uint32_t s)
{
    const uint32_t h = s - (uint32_t)st.head_lo; // wraps for s < head_lo
    if (st.head != nullptr && h < st.head_n)
        return PlaneRef{st.head, st.head_stride, (uint64_t)h * st.slab_stride, st.row_stride};
    return PlaneRef{st.ring, st.ring_stride, (uint64_t)ring_slot(st, s) * st.slab_stride, st.row_stride};
}

// Slab s as a writer stores it: always its ring slot, and also the head copy
// when s is a head slab (plane_write_head; buf = nullptr otherwise).
__host__ __device__ inline PlaneRef plane_write(const PlaneStore &st, uint32_t s)
{
    return PlaneRef{st.ring, st.ring_stride, (uint64_t)ring_slot(st, s) * st.slab_stride, st.row_stride};
}

__host__ __device__ inline PlaneRef plane_write_head(const PlaneStore &st, uint32_t s)
{
    const uint32_t h = s - (uint32_t)st.head_lo;
    if (st.head != nullptr && h < st.head_n)
        return PlaneRef{st.head, st.head_stride, (uint64_t)h * st.slab_stride, st.row_stride};
    return PlaneRef{nullptr, 0, 0, 0};
}

// One launch of the blocked sweep: one time step (fused macrovar, mode 0) of
// slabs s0, s0+1, ..., s0+count-1 (taken modulo ns), gathering from `src` and
// storing into `dst` (non-temporal stores when nt).  Thread t works on node
// t % slabp of the launch's slab t / slabp: slabs are padded to whole waves
// (slabp = slab rounded up to a multiple of 64), so every wave lies in one
// slab and the slab's buffers and offsets are the same for all its lanes.
// per_node: the body force is evaluated at every node as in step mode 16
// (fc for the collision, fm for the fused fields u = j + F/2); otherwise it
// is r.fx, r.fy, r.fz for both.
// Rows: a launch computes rows 0 .. nrw - 1 of each slab in its own frame
// (local rows).  Local row rl is row rl + src_r0 (modulo src_nr) of the
// source store, row rl + dst_r0 of the destination store and lattice row
// rl + node_r0 (modulo nr).  Over whole slabs all three frames are the
// lattice's (offsets 0, src_nr = nrw = nr).  A row window of W rows starting
// at lattice row r0 computes step t on rows r0 - 1 .. r0 + W (nrw = W + 2)
// into a ring whose slabs hold those W + 2 rows, and step t+1 on rows
// r0 .. r0 + W - 1 from ring rows 1 .. W.
struct BlockArgs {
    Geom g;
    uint32_t ns, nr;     // slabs along the sweep axis, rows per lattice slab
    uint32_t nrw;        // rows computed per slab (local rows)
    uint32_t src_r0, src_nr, dst_r0, node_r0; // frames of the local rows
    uint32_t slab;       // nodes computed per slab, lx * nrw
    uint32_t slabp;      // slab rounded up to a whole number of waves
    FastDiv dslabp;      // division by slabp
    int32_t s0;
    uint32_t count;
    PlaneStore src, dst;
    RatesD r;
    ForceSpec fc, fm;
    int per_node;
    int nt;
};

constexpr uint32_t WAVE_LANES = 64;

inline uint32_t slab_padded(uint32_t slab) { return (slab + WAVE_LANES - 1) / WAVE_LANES * WAVE_LANES; }

// Slab (0 .. ns-1) of the launch's sl-th slab; -ns <= s0 < ns and
// sl < count <= ns, so one wrap either way suffices.
__host__ __device__ inline uint32_t block_plane(const BlockArgs &a, uint32_t sl)
{
    int32_t s = a.s0 + (int32_t)sl;
    if (s < 0)
        s += (int32_t)a.ns;
    if (s >= (int32_t)a.ns)
        s -= (int32_t)a.ns;
    return (uint32_t)s;
}

// The source slabs of slab s: index c + 1 holds slab s - c, for the sweep-axis
// component c of a velocity (slabs s + 1, s, s - 1).
__host__ __device__ inline void block_planes(const BlockArgs &a, uint32_t s, PlaneRef (&pl)[3])
{
    pl[0] = plane_read(a.src, s + 1 == a.ns ? 0 : s + 1);
    pl[1] = plane_read(a.src, s);
    pl[2] = plane_read(a.src, s == 0 ? a.ns - 1 : s - 1);
}

// A node's own part of its gather addresses: element offsets within a slab
// of column ix of rows r + 1, r, r - 1 (index c + 1 for the row-axis
// component c of a velocity, rows wrapping), with the slab's row stride.  The
// offsets are below nr * row_stride <= the lattice size, so 32 bits hold them.
struct BlockLane {
    uint32_t o[3];
    uint32_t ix, lx;
};

__host__ __device__ inline BlockLane block_lane(uint32_t ix, uint32_t r, uint32_t nr, uint32_t row_stride, uint32_t lx)
{
    BlockLane l;
    const uint32_t o = r * row_stride + ix;
    l.o[0] = r + 1 == nr ? ix : o + row_stride;
    l.o[1] = o;
    l.o[2] = r == 0 ? (nr - 1) * row_stride + ix : o - row_stride;
    l.ix = ix;
    l.lx = lx;
    return l;
}

// Node w (< a.slab) of a launch's slab: its column, local row, row in the
// source store and lattice row.  The schedule keeps rl + src_r0 < 2 src_nr
// and rl + node_r0 < 2 nr (windows are narrower than the lattice slab), so
// one conditional subtraction wraps each.
struct BlockRow {
    uint32_t ix, rl, rs, rg;
};

__host__ __device__ inline BlockRow block_row(const BlockArgs &a, uint32_t w)
{
    BlockRow b;
    b.rl = fdiv(w, a.g.dlx);
    b.ix = w - b.rl * a.g.lx;
    uint32_t rs = b.rl + a.src_r0;
    b.rs = rs >= a.src_nr ? rs - a.src_nr : rs;
    uint32_t rg = b.rl + a.node_r0;
    b.rg = rg >= a.nr ? rg - a.nr : rg;
    return b;
}

// Buffer and element index that F_I of the node is gathered from: P(x - c_I)
// in population I, or at a wall the node's own slot of population OPP[I]
// (halfway bounce-back).
template <int AXIS, int I>
__host__ __device__ inline void block_src(const PlaneRef (&pl)[3], const BlockLane &l, void *&buf, uint64_t &idx)
{
    constexpr int cs = AXIS == 2 ? CZ[I] : CY[I]; // along the sweep axis
    constexpr int cr = AXIS == 2 ? CY[I] : CZ[I]; // along the rows
    const PlaneRef &p = pl[cs + 1];
    if constexpr (CX[I] == 0) {
        buf = p.buf;
        idx = (uint64_t)I * p.stride + p.off + l.o[cr + 1];
    } else {
        const bool walled = CX[I] == 1 ? l.ix == 0 : l.ix + 1 == l.lx;
        buf = walled ? pl[1].buf : p.buf;
        idx = walled ? (uint64_t)OPP[I] * pl[1].stride + pl[1].off + l.o[1]
                     : (uint64_t)I * p.stride + p.off + (l.o[cr + 1] - (uint32_t)CX[I]);
    }
}

// Run monitor of Channel-Flow/saveload.f90:1507-1676 (diag): per block, the
// fluid-node sums of ux, uy, uz, ux^2, uy^2, uz^2 and the count, the largest
// |u| with the smallest node index that reaches it, and the largest and
// smallest rho; partials[q*grid + block].  SRC 0: field arrays; SRC 1:
// macrovar(F) of lattice L with rho - rho_off.
constexpr int DIAG_Q = 11;
struct DiagArgs {
    Geom g;
    const void *L;
    const double *rho, *ux, *uy, *uz;
    const int *solid; // nullable: when set, only fluid nodes count
    ForceSpec f;
    double rho_off;
    double *partials;
};

// dissipation_calc (Channel-Flow/saveload.f90:4644-4882): per wall-normal
// plane the fluid-node sums of S_ij S_ij and S_xy by the two methods and the
// count; partials[(chunk*NDISS + q)*lx + ix].  f is pulled from L (the
// current lattice); the fields come from the arrays (SRC 0) or from the
// same pull (SRC 1).
constexpr int NDISS = 5;
struct DissArgs {
    Geom g;
    const void *L;
    const double *rho, *ux, *uy, *uz;
    const int *solid;
    ForceSpec f;
    double rho_off;
    double s1, s9, coef2, coef3, coef5, ww1, ww2, c15tau; // c15tau = 1.5/tau
    double *partials;
};

template <int P>
struct Launch {
    static hipError_t step(const StepArgs &a, int mode, hipStream_t s);
    // closure inputs (what 0, 10 floats per node) or nu_t/nu (what 1) of nodes [n_begin, n_begin + count)
    static hipError_t closure_eval(const StepArgs &a, uint64_t n_begin, uint64_t count, float *out, int what,
                                   hipStream_t s);
    static hipError_t macro(const MacroArgs &a, int what, hipStream_t s);
    static hipError_t gather(const CopyArgs &a, hipStream_t s);
    static hipError_t scatter(const CopyArgs &a, hipStream_t s);
    static hipError_t initpop(const InitArgs &a, hipStream_t s);
    static hipError_t rho_partials(const SumArgs &a, int src, uint32_t grid, hipStream_t s);
    static hipError_t mass_partials(const SumArgs &a, uint32_t grid, hipStream_t s);
    static hipError_t plane_partials(const PlaneArgs &a, int src, uint32_t chunks, hipStream_t s);
    static hipError_t filtered_plane_partials(const PlaneArgs &a, uint32_t chunks, hipStream_t s);
    static hipError_t diag_partials(const DiagArgs &a, int src, uint32_t grid, hipStream_t s);
    static hipError_t dissipation_partials(const DissArgs &a, int src, uint32_t chunks, hipStream_t s);
    // one launch of the two-step blocked sweep along axis 2 (z) or 1 (y)
    static hipError_t step_planes(const BlockArgs &a, int axis, hipStream_t s);
};

extern template struct Launch<0>;
extern template struct Launch<1>;
extern template struct Launch<2>;
extern template struct Launch<3>;

} // namespace d3q19
