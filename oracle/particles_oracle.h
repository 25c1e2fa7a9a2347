/* CPU oracle for the rigid-particle path -- test infrastructure only.
 * See particles_oracle.c. */
#ifndef PARTICLES_ORACLE_H
#define PARTICLES_ORACLE_H

#include "lbm_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_pgeom {
    int lx, ly, lz, np, maxlink, maxfill, pressure_force, reserved;
    double R, R2, mass, inertia, volume, mingap, mingap_w, stf, stf_w, fx, fy, fz;
} orc_pgeom;

typedef struct orc_pstate {
    double *X, *Xold, *V, *W, *TH, *Aprev, *ALprev, *FH, *TQ, *FC; /* 3*np each */
    int *solid;                                                  /* N */
    int *link_node, *link_dir;                                   /* np*maxlink */
    double *link_q;                                              /* np*maxlink */
    int *link_count;                                             /* np */
    int *fill_node;                                              /* np*maxfill */
    int *fill_count;                                             /* np */
    int nsteps_done;
} orc_pstate;

int orc_p_classify(const orc_pgeom *g, orc_pstate *s, int first);
void orc_p_ibb(const orc_pgeom *g, orc_pstate *s, double *L);
void orc_p_step(const orc_params *pp, const orc_pgeom *g, const int *solid, const double *Lin, double *Lout,
                double fx, double fy, double fz);
void orc_p_move(const orc_pgeom *g, orc_pstate *s);
void orc_p_refill(const orc_pgeom *g, orc_pstate *s, double *L);

#ifdef __cplusplus
}
#endif
#endif
