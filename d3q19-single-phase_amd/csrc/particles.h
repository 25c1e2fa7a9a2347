// Rigid spherical particles in the channel: data blocks and launchers.
//
// The reference keeps the particle subsystem's data structures but not its
// source (Channel-Flow/var_inc.f90:30-38,120-150, para.f90:332-379,449-501;
// routine names beads_links, beads_collision, beads_filling, beads_move at
// var_inc.f90:166-168).  The methods here are the published ones:
//   classification  node inside a sphere if |x - X|^2 < R^2 (minimum image in
//                   y, z); cut link q = smaller root of |x + t c - X|^2 = R^2
//   bounce-back     Bouzidi, Firdaouss & Lallemand, Phys. Fluids 13 (2001)
//                   3452, linear, with the moving-wall term (Lallemand & Luo,
//                   J. Comput. Phys. 184 (2003) 406)
//   force, torque   Galilean-invariant momentum exchange, Wen et al.,
//                   J. Comput. Phys. 266 (2014) 161
//   refill          equilibrium at the local wall velocity and the mean
//                   density of the fluid neighbours (Peng et al., Comput.
//                   Math. Appl. 72 (2016) 349, equilibrium variant)
//   contact         repulsive force (1/eps)((zeta - h)/zeta)^2 for gaps h
//                   below zeta, Glowinski et al., Int. J. Multiphase Flow
//                   25 (1999) 755
//   integration     FP64, Adams-Bashforth 2 for velocities, trapezoid for
//                   positions and orientation
// Node flags: solid[n] = -1 fluid, p >= 0 inside particle p, -2-p fluid node
// uncovered by particle p in the last move and not yet refilled.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19_common.h"

namespace d3q19 {

struct PartGeom {
    Geom g;
    int np;
    int maxlink, maxfill;
    double R, R2;
    double mass, inertia, volume;
    double mingap, mingap_w, stf, stf_w;
    double fx, fy, fz;  // fluid body force (pressure-gradient force on particles)
    int pressure_force; // 1: add volume * body force
};

struct PartState {
    double *X, *V, *W, *TH;        // position, velocity, angular velocity, orientation (3*np)
    double *Aprev, *ALprev;        // previous accelerations (3*np)
    double *FH, *TQ, *FC;          // hydrodynamic force, torque, contact force (3*np)
    int *nsteps_done;              // 0 before the first move (Euler start)
    int *solid;                    // N
    int *link_node;                // np*maxlink
    signed char *link_dir;         // np*maxlink
    double *link_q;                // np*maxlink
    double *link_contrib;          // np*maxlink*6
    int *link_count;               // np
    int *fill_node;                // np*maxfill
    int *fill_count;               // np
    int *overflow;                 // 1 when a per-particle list overflowed
};

template <int P>
struct PartLaunch {
    static hipError_t ibb(const PartGeom &pg, const PartState &ps, void *L, hipStream_t s);
    static hipError_t refill(const PartGeom &pg, const PartState &ps, void *L, hipStream_t s);
};

hipError_t part_classify(const PartGeom &pg, const PartState &ps, int first, hipStream_t s);
hipError_t part_sum(const PartGeom &pg, const PartState &ps, hipStream_t s);
hipError_t part_move(const PartGeom &pg, const PartState &ps, hipStream_t s);

extern template struct PartLaunch<0>;
extern template struct PartLaunch<1>;
extern template struct PartLaunch<2>;

} // namespace d3q19
