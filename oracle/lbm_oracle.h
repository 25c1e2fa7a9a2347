/* CPU oracle for the D3Q19 MRT channel-flow hot path -- test infrastructure
 * only (tests/, smoke(), bench.py cpu_baseline).  See lbm_oracle.c. */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Scalars of Channel-Flow/var_inc.f90:85-93 as set by para.f90:59-174. */
typedef struct orc_params {
    int lx, ly, lz, nx, mrttype;
    double visc, rstar, ustar, force_in_y, ystar, force_mag, tau;
    double s1, s2, s4, s9, s10, s13, s16;
    double omegepsl, omegepslj, omegxx;
    double coef1, coef2, coef3, coef4, coef5, coef3i, coef4i;
    double val8;
    double val1i, val2i, val3i, val4i, val5i, val6i, val7i, val8i, val9i;
    double ww0, ww1, ww2;
} orc_params;

void orc_para(orc_params *p, int lx, int ly, int lz, int laminar, int mrttype);
void orc_collide_node(const orc_params *p, const double *fin, double rho9, double ux9, double uy9, double uz9,
                      double fx9, double fy9, double fz9, double *post);
void orc_collision_mrt(const orc_params *p, double *f, const double *rho, const double *ux, const double *uy,
                       const double *uz, const double *fx, const double *fy, const double *fz);
void orc_collision_mrt_ab(const orc_params *p, const double *fin, double *fout, const double *rho,
                          const double *ux, const double *uy, const double *uz, const double *fx,
                          const double *fy, const double *fz, int nthreads);
void orc_macrovar(const orc_params *p, const double *f, double *rho, double *ux, double *uy, double *uz,
                  const double *fx, const double *fy, const double *fz);
void orc_rhoupdat(const orc_params *p, const double *f, double *rho);
double orc_avedensity(const orc_params *p, double *rho);
void orc_forcing(const orc_params *p, double *fx, double *fy, double *fz);
void orc_forcingp(const orc_params *p, int istep, double *fx, double *fy, double *fz);
void orc_initpop(const orc_params *p, double *f, double *rho, const double *ux, const double *uy,
                 const double *uz);
void orc_initvel(const orc_params *p, double *ux, double *uy, double *uz);
void orc_statistc2(const orc_params *p, const double *rho, const double *ux, const double *uy,
                   const double *uz, double *out);
void orc_diag(const orc_params *p, const double *rho, const double *ux, const double *uy, const double *uz,
              double *out);
void orc_vortcalc(const orc_params *p, const double *ux, const double *uy, const double *uz, const int *solid,
                  const double *omg, double *ox, double *oy, double *oz);
void orc_dissipation_node(const orc_params *p, const double *f9, double rho9, double ux9, double uy9, double uz9,
                          double *out4);
void orc_dissipation(const orc_params *p, const double *f, const double *rho, const double *ux, const double *uy,
                     const double *uz, const int *solid, double *out);

#ifdef __cplusplus
}
#endif
#endif
