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
};

extern template struct Launch<0>;
extern template struct Launch<1>;
extern template struct Launch<2>;

} // namespace d3q19
