// Learned-closure training data and training launchers (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19_common.h"
#include "mlp.h"

namespace d3q19 {

// Box filter of a fine lattice onto the 2x-coarsened grid.
struct FilterArgs {
    Geom gf;              // fine geometry (even lx, ly, lz)
    const void *L;        // fine post-collision lattice
    double fx, fy, fz;    // fine body force (u = sum c f + F/2)
    uint32_t lxc, lyc, lzc;
    float *out;           // 10 fields of Nc: rho, ux, uy, uz, uxux, uyuy, uzuz, uxuy, uyuz, uxuz
};

// Features and targets of the coarse nodes.
struct FeatArgs {
    uint32_t lx, ly, lz;  // coarse grid
    const float *fields;  // FilterArgs::out
    float nu, s9;         // coarse viscosity and shear rate
    float inv_ystar, yscale;
    const float *scale;   // 9 feature scales (device)
    float *feat;          // samples x 10
    float *target;        // samples
    uint64_t offset;      // first sample written
    float *ts, *ss;       // nullable: per-node tau:S and S:S for the per-plane target
};

// Target of every node of a coarse x plane replaced by the plane's
// least-squares eddy viscosity -sum(tau:S) / (2 sum(S:S)) / nu.
hipError_t mlp_plane_targets(uint32_t lx, uint32_t ly, uint32_t lz, const float *ts, const float *ss, float nu,
                             float *target, uint64_t offset, hipStream_t s);

template <int P>
struct MlpLaunch {
    static hipError_t filter(const FilterArgs &a, hipStream_t s);
};

hipError_t mlp_features(const FeatArgs &a, hipStream_t s);
hipError_t mlp_grad(const float *w, const float *feat, const float *target, uint32_t first, uint32_t count,
                    float *partial, float *grad, float *wmut, float *m, float *v, int step, float lr, int adam,
                    hipStream_t s);
hipError_t mlp_fill_plane_targets(float *target, uint64_t first, uint64_t count, uint32_t nx, const float *values,
                                  hipStream_t s);
hipError_t mlp_predict(const float *w, const float *feat, uint32_t first, uint32_t count, float *out, hipStream_t s);

extern template struct MlpLaunch<0>;
extern template struct MlpLaunch<1>;
extern template struct MlpLaunch<2>;
extern template struct MlpLaunch<3>;

} // namespace d3q19
