// Kernel argument blocks and per-precision launchers (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19_common.h"

namespace d3q19 {

// Relaxation rates and force as passed by the driver (FP64; the FP32 kernels
// round them once).
struct RatesD {
    double s1, s2, s4, s9, s10, s13, s16;
    double omegepsl, omegepslj, omegxx;
    double fx, fy, fz;
};

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

// Where the z planes of one lattice state live during a two-step blocked
// sweep: plane p in slot p % ring_slots of `ring` (a whole lattice when
// ring_slots = lz), except planes head_lo .. head_lo + head_n - 1, which a
// reader takes from `head` (and a writer also stores there).  Strides count
// elements between consecutive populations of a buffer.
struct PlaneStore {
    void *ring;
    uint64_t ring_stride;
    uint32_t ring_slots;
    void *head; // nullable
    uint64_t head_stride;
    int32_t head_lo;
    uint32_t head_n;
};

// Plane p of store s as a reader sees it: buffer, population stride and the
// element offset of the plane.
struct PlaneRef {
    void *buf;
    uint64_t stride;
    uint64_t off;
};

__host__ __device__ inline PlaneRef plane_read(const PlaneStore &s, uint32_t p, uint64_t plane)
{
    const uint32_t h = p - (uint32_t)s.head_lo; // wraps for p < head_lo
    if (s.head != nullptr && h < s.head_n)
        return PlaneRef{s.head, s.head_stride, (uint64_t)h * plane};
    return PlaneRef{s.ring, s.ring_stride, (uint64_t)(p % s.ring_slots) * plane};
}

// Plane p of store s as a writer stores it: always its ring slot, and also
// the head copy when p is a head plane (plane_write_head; buf = nullptr
// otherwise).
__host__ __device__ inline PlaneRef plane_write(const PlaneStore &s, uint32_t p, uint64_t plane)
{
    return PlaneRef{s.ring, s.ring_stride, (uint64_t)(p % s.ring_slots) * plane};
}

__host__ __device__ inline PlaneRef plane_write_head(const PlaneStore &s, uint32_t p, uint64_t plane)
{
    const uint32_t h = p - (uint32_t)s.head_lo;
    if (s.head != nullptr && h < s.head_n)
        return PlaneRef{s.head, s.head_stride, (uint64_t)h * plane};
    return PlaneRef{nullptr, 0, 0};
}

// One launch of the blocked sweep: one time step (fused macrovar, mode 0) of
// planes z0, z0+1, ..., z0+nz-1 (taken modulo lz), gathering from `src` and
// storing into `dst` (non-temporal stores when nt).
struct BlockArgs {
    Geom g;
    FastDiv dplane; // division by lx*ly
    int32_t z0;
    uint32_t nz;
    PlaneStore src, dst;
    RatesD r;
    int nt;
};

// Plane (0 .. lz-1) of the launch's zl-th plane; z0 >= -lz.
__host__ __device__ inline uint32_t block_plane(const BlockArgs &a, uint32_t zl)
{
    int32_t z = a.z0 + (int32_t)zl;
    if (z < 0)
        z += (int32_t)a.g.lz;
    return (uint32_t)z % a.g.lz;
}

// Gather sources of one node (ix, iy, z) of a blocked launch: the planes
// z + 1, z, z - 1 (index c_z + 1 holds plane z - c_z) and the rows
// iy + 1, iy, iy - 1 (index c_y + 1 holds row iy - c_y, times lx).
struct BlockNbr {
    PlaneRef pl[3];
    uint64_t row[3];
    uint32_t ix, lx;
};

__host__ __device__ inline BlockNbr block_nbr(const BlockArgs &a, uint32_t ix, uint32_t iy, uint32_t z)
{
    BlockNbr b;
    const uint32_t lz = a.g.lz, ly = a.g.ly;
    b.pl[0] = plane_read(a.src, z + 1 == lz ? 0 : z + 1, a.g.plane);
    b.pl[1] = plane_read(a.src, z, a.g.plane);
    b.pl[2] = plane_read(a.src, z == 0 ? lz - 1 : z - 1, a.g.plane);
    b.row[0] = (uint64_t)(iy + 1 == ly ? 0 : iy + 1) * a.g.lx;
    b.row[1] = (uint64_t)iy * a.g.lx;
    b.row[2] = (uint64_t)(iy == 0 ? ly - 1 : iy - 1) * a.g.lx;
    b.ix = ix;
    b.lx = a.g.lx;
    return b;
}

// Buffer and element index that F_I of the node is gathered from: P(x - c_I)
// in population I, or at a wall the node's own slot of population OPP[I]
// (halfway bounce-back).
template <int I>
__host__ __device__ inline void block_src(const BlockNbr &b, void *&buf, uint64_t &idx)
{
    const bool walled = (CX[I] == 1 && b.ix == 0) || (CX[I] == -1 && b.ix + 1 == b.lx);
    if (walled) {
        buf = b.pl[1].buf;
        idx = (uint64_t)OPP[I] * b.pl[1].stride + b.pl[1].off + b.row[1] + b.ix;
    } else {
        const PlaneRef &r = b.pl[CZ[I] + 1];
        buf = r.buf;
        idx = (uint64_t)I * r.stride + r.off + b.row[CY[I] + 1] + (b.ix - CX[I]);
    }
}

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
    // one launch of the two-step blocked sweep
    static hipError_t step_planes(const BlockArgs &a, hipStream_t s);
};

extern template struct Launch<0>;
extern template struct Launch<1>;
extern template struct Launch<2>;

} // namespace d3q19
