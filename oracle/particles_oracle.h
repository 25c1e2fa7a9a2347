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
    double R, R2, mass, inertia, volume, rhopart;
    double mingap, mingap_w, stf0, stf1, stf0_w, stf1_w; /* contact, Feng & Michaelides (2005) eq. 28 */
    double visc, lub_cut, lub_min;                       /* lubrication correction, Nguyen & Ladd (2002) */
    double fx, fy, fz;
} orc_pgeom;

typedef struct orc_pstate {
    double *X, *Xold, *V, *W, *TH, *Aprev, *ALprev, *FH, *TQ, *FC; /* 3*np each */
    double *FL, *FHprev, *TQprev;                                /* 3*np each */
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
                double fx, double fy, double fz, double rho_off);
void orc_p_move(const orc_pgeom *g, orc_pstate *s);
void orc_p_refill(const orc_pgeom *g, orc_pstate *s, double *L);
double orc_p_rho_mean(const orc_pgeom *g, const int *solid, const double *L);
void orc_p_macrovar(const orc_pgeom *g, const orc_pstate *s, const double *L, double fx, double fy, double fz,
                    double rho_off, double *rho, double *ux, double *uy, double *uz);
double orc_device_sum(const double *v, long long n, const unsigned char *mask);

#ifdef __cplusplus
}
#endif
#endif
