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
};

struct MacroArgs {
    Geom g;
    const void *L;
    double *rho, *ux, *uy, *uz;
    double fx, fy, fz, rho_off;
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
    double fx, fy, fz, rho_off;
    double *partials;
};

// Plane sums per wall-normal index: 11 moments and the fluid-node count.
constexpr int NPLANE = 12;

template <int P>
struct Launch {
    static hipError_t step(const StepArgs &a, int mode, hipStream_t s);
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
