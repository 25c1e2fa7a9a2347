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
//   contact         Feng & Michaelides, J. Comput. Phys. 202 (2005) 20,
//                   eq. 28: for the gap h = d - 2R below the range zeta,
//                   stf0 ((zeta - h)/zeta)^2, plus stf1 (-h/zeta) on overlap,
//                   along the line of centres; walls act as mirror-image
//                   particles (stf0_w, stf1_w); the stiffness pairs of
//                   Channel-Flow/para.f90:357-360
//   lubrication     normal correction below the cut-off gap hc, Nguyen &
//                   Ladd, Phys. Rev. E 66 (2002) 046708:
//                   -6 pi mu a^2 (1/h - 1/hc) U_n, a = R/2 between two
//                   spheres, R against a wall (flubp, var_inc.f90:150)
//   integration     FP64, Adams-Bashforth 2 for velocities on the
//                   hydrodynamic force and torque averaged over the last two
//                   steps (Ladd, J. Fluid Mech. 271 (1994) 285), trapezoid for
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
    double mass, inertia, volume, rhopart;
    double mingap, mingap_w, stf0, stf1, stf0_w, stf1_w; // contact
    double visc, lub_cut, lub_min;                        // lubrication (visc 0: off)
    double fx, fy, fz;  // fluid body force (pressure-gradient force on particles)
    int pressure_force; // 1: add volume * body force
    int nave;           // run loop: avedensity after every step that is a multiple of nave (0: never)
};

struct PartState {
    double *X, *V, *W, *TH;        // position, velocity, angular velocity, orientation (3*np)
    double *Aprev, *ALprev;        // previous accelerations (3*np)
    double *FH, *TQ, *FC;          // hydrodynamic force, torque, contact force (3*np)
    double *FL;                    // lubrication force (3*np)
    double *FHprev, *TQprev;       // the previous step's hydrodynamic force and torque (3*np)
    double *VWnew;                 // new velocities, then angular velocities, until committed (6*np)
    int *nsteps_done;              // 0 before the first move (Euler start)
    int *solid;                    // N
    int *link_node;                // np*maxlink
    signed char *link_dir;         // np*maxlink
    double *link_q;                // np*maxlink
    double *link_contrib;          // np*maxlink*6
    int *link_count;               // np
    int *fill_node;                // np*maxfill
    int *fill_count;               // np
    int *overflow;                 // bit 0: a per-particle list overflowed; bit 1: a state became non-finite
};

// Bounding box of a sphere's nodes (node centres at i + 0.5), grown by `pad`
// nodes.  x is clamped to the channel.  y and z start at an index in [0, L]
// and run on from there (the users wrap each index into [0, L)); a box at
// least as wide as the period covers it exactly once, from 0.  Every bound is
// formed in double and clamped before it becomes an int, so a centre that is
// not finite or lies more than R + pad + 1 outside the channel in x gives the
// empty box (x1 < x0, zero nodes) instead of indices out of range.
struct Box {
    int x0, x1, y0, y1, z0, z1;
};

__host__ __device__ inline void box_axis(double c, double R, int pad, double L, int &lo, int &hi)
{
    const double a = floor(c - 0.5 - R) - pad, e = floor(c - 0.5 + R) + 1 + pad;
    if (e - a + 1.0 >= L) {
        lo = 0;
        hi = (int)L - 1;
        return;
    }
    double s = fmod(a, L); // exact, in (-L, L)
    if (s < 0.0)
        s = s + L;
    lo = (int)s;
    hi = lo + (int)(e - a);
}

__host__ __device__ inline Box sphere_box(const PartGeom &pg, const double *X, int pad)
{
    Box b = {0, -1, 0, -1, 0, -1};
    const double lx = (double)pg.g.lx;
    const double reach = pg.R + pad + 1.0;
    // comparisons are false for NaN; y and z need finite values for fmod
    if (!(X[0] > -reach && X[0] < lx + reach) || !(X[1] - X[1] == 0.0) || !(X[2] - X[2] == 0.0))
        return b;
    double x0 = floor(X[0] - 0.5 - pg.R) - pad, x1 = floor(X[0] - 0.5 + pg.R) + 1 + pad;
    if (x0 < 0.0)
        x0 = 0.0;
    if (x1 > lx - 1.0)
        x1 = lx - 1.0;
    if (x1 < x0)
        return b;
    b.x0 = (int)x0;
    b.x1 = (int)x1;
    box_axis(X[1], pg.R, pad, (double)pg.g.ly, b.y0, b.y1);
    box_axis(X[2], pg.R, pad, (double)pg.g.lz, b.z0, b.z1);
    return b;
}

template <int P>
struct PartLaunch {
    static hipError_t ibb(const PartGeom &pg, const PartState &ps, void *L, hipStream_t s);
    static hipError_t refill(const PartGeom &pg, const PartState &ps, void *L, hipStream_t s);
};

hipError_t part_classify(const PartGeom &pg, const PartState &ps, int first, hipStream_t s);
// macrovar's solid branch: fields at nodes inside particles
hipError_t part_solid_fields(const PartGeom &pg, const PartState &ps, double rho_off, double *rho, double *ux,
                             double *uy, double *uz, hipStream_t s);
hipError_t part_sum(const PartGeom &pg, const PartState &ps, hipStream_t s);
hipError_t part_move(const PartGeom &pg, const PartState &ps, hipStream_t s);

extern template struct PartLaunch<0>;
extern template struct PartLaunch<1>;
extern template struct PartLaunch<2>;
extern template struct PartLaunch<3>;

} // namespace d3q19
