/*
 * CPU oracle for the D3Q19 MRT channel-flow hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  It checks the GPU results and is never the product:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product path (liblbm.so, HIP kernels) never links or
 * calls it.
 *
 * It restates, statement by statement and in the same floating-point
 * evaluation order (left-to-right sums, no FMA contraction: build with
 * -ffp-contract=off), the Fortran routines of UDel-CFD/D3Q19-Single-Phase:
 *
 *   orc_collision_mrt   Channel-Flow/collision.f90:24-273   (swap algorithm,
 *                       in place, faithful to the loop order)
 *   orc_exchange        Channel-Flow/collision.f90:281-372  (one rank: every
 *                       neighbour is self, so it is the periodic y/z wrap)
 *   orc_macrovar        Channel-Flow/collision.f90:378-419  (fluid branch)
 *   orc_rhoupdat        Channel-Flow/collision.f90:469-480
 *   orc_avedensity      Channel-Flow/collision.f90:487-513
 *   orc_forcing         Channel-Flow/collision.f90:515-527
 *   orc_initpop         Channel-Flow/initial.f90:19-46
 *   orc_initvel         Channel-Flow/initial.f90:75-147 (A9 = 0: no perturbation)
 *   orc_para            Channel-Flow/para.f90:59-214
 *   orc_statistc2       Channel-Flow/saveload.f90:1348-1502 (profile moments)
 *   orc_diag            Channel-Flow/saveload.f90:1507-1676 (run monitor)
 *   orc_vortcalc        Channel-Flow/saveload.f90:3929-4023 (vorticity)
 *   orc_dissipation     Channel-Flow/saveload.f90:4644-4882 (dissipation_calc)
 *
 * plus one function that is not a restatement but an equivalent schedule:
 *
 *   orc_collision_mrt_ab  same per-node arithmetic, push streaming from an
 *                       input lattice to a separate output lattice, OpenMP
 *                       parallel over z planes.  Streaming is a permutation of
 *                       freshly computed values (each post-collision value is
 *                       copied exactly once), so its output is bitwise equal
 *                       to orc_collision_mrt; the tests assert that.
 *
 * Arrays use the reference layout: f(0:18,lx,ly,lz) column-major (population
 * fastest), macroscopic fields (lx,ly,lz) with x fastest.  Axes follow the
 * reference: x = wall-normal (halfway bounce-back walls beyond ix=1 and
 * ix=lx), y = streamwise (periodic), z = spanwise (periodic).
 *
 * Pinning (tests/test_oracle_*.py): the collision arithmetic against the
 * d'Humieres (2002) matrix form built independently in numpy; MRTtype 2
 * against textbook BGK with Guo forcing; streaming against numpy np.roll
 * push streaming with bounce-back; mass conservation, y/z translation
 * and x-reflection invariants; the laminar start-up against the analytic
 * Poiseuille series the reference prints (saveload.f90:920-936).
 *
 * Origin and licence: because it restates UDel-CFD/D3Q19-Single-Phase
 * (Channel-Flow/*.f90, GNU General Public License version 3) statement by
 * statement, keeping its variable names (evlm*, eqmc*, tl*, sum*) and its
 * operand order, this file is a derivative of that GPL-3 code and is
 * distributed under the GPL-3 terms.  It is test infrastructure only: the
 * product library does not link, load or call it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "lbm_oracle.h"

#define NPOP 19

/* Channel-Flow/para.f90:201-210 */
static const int cix[NPOP] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int ciy[NPOP] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int ciz[NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int ipopp[NPOP] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
static const int ipswap[9] = {2, 4, 6, 9, 10, 13, 14, 17, 18};
static const int ipstay[10] = {0, 1, 3, 5, 7, 8, 11, 12, 15, 16};

/* 1-based Fortran indexing helpers */
#define FI(p, ip, ix, iy, iz) \
    ((size_t)(ip) + (size_t)NPOP * ((size_t)((ix)-1) + (size_t)(p)->lx * ((size_t)((iy)-1) + (size_t)(p)->ly * (size_t)((iz)-1))))
#define MI(p, ix, iy, iz) \
    ((size_t)((ix)-1) + (size_t)(p)->lx * ((size_t)((iy)-1) + (size_t)(p)->ly * (size_t)((iz)-1)))

/* Channel-Flow/para.f90:59-174.  mrttype < 0 keeps the branch default. */
void orc_para(orc_params *p, int lx, int ly, int lz, int laminar, int mrttype)
{
    memset(p, 0, sizeof(*p));
    p->lx = lx;
    p->ly = ly;
    p->lz = lz;
    p->nx = lx;
    const double rho0 = 1.0;
    int mrt;
    if (!laminar) {
        /* para.f90:60-73 */
        p->visc = 0.0036;
        p->rstar = 180.0;
        p->ustar = 2.0 * p->rstar * p->visc / (double)lx;
        p->force_in_y = 2. * rho0 * p->ustar * p->ustar / (double)lx;
        p->ystar = p->visc / p->ustar;
        p->force_mag = 1.0;
        mrt = 1;
    } else {
        /* para.f90:74-88 */
        p->rstar = 20;
        p->ustar = 0.05;
        p->visc = 2.0 * p->ustar * (double)lx / p->rstar;
        p->force_in_y = 8.0 * p->visc * p->ustar / ((double)lx * (double)lx);
        p->ystar = p->visc / p->ustar;
        p->force_mag = 1.0;
        mrt = 2;
    }
    if (mrttype > 0)
        mrt = mrttype;
    p->mrttype = mrt;

    /* para.f90:106-141 */
    p->tau = 3.0 * p->visc + 0.5;
    p->s9 = 1.0 / p->tau;
    p->s13 = p->s9;
    switch (mrt) {
    case 1:
        p->s1 = 1.5;
        p->s2 = 1.4;
        p->s4 = 1.2;
        p->s10 = 1.4;
        p->s16 = 1.98;
        p->omegepsl = 0.0;
        p->omegepslj = -(475.0 / 63.0);
        p->omegxx = 0.0;
        break;
    case 2:
        p->s1 = p->s9;
        p->s2 = p->s9;
        p->s4 = p->s9;
        p->s10 = p->s9;
        p->s16 = p->s9;
        p->omegepsl = 3.0;
        p->omegepslj = -(11.0 / 2.0);
        p->omegxx = -(1.0 / 2.0);
        break;
    default: /* 3 */
        p->s1 = 1.8;
        p->s2 = p->s1;
        p->s4 = p->s9;
        p->s10 = p->s1;
        p->s16 = p->s1;
        p->omegepsl = 3.0;
        p->omegepslj = -(11.0 / 2.0);
        p->omegxx = -(1.0 / 2.0);
        break;
    }
    /* para.f90:143-174 */
    p->coef1 = -(2.0 / 3.0);
    p->coef2 = -11.0;
    p->coef3 = 8.0;
    p->coef4 = -4.0;
    p->coef5 = 2.0;
    p->coef3i = 1.0 / p->coef3;
    p->coef4i = 1.0 / p->coef4;
    p->val8 = 12.0;
    p->val1i = 1.0 / 19.0;
    p->val2i = 1.0 / 2394.0;
    p->val3i = 1.0 / 252.0;
    p->val4i = 1.0 / 10.0;
    p->val5i = 1.0 / 40.0;
    p->val6i = 1.0 / 36.0;
    p->val7i = 1.0 / 72.0;
    p->val8i = 1.0 / 12.0;
    p->val9i = 1.0 / 24.0;
    p->ww0 = 1.0 / 3.0;
    p->ww1 = 1.0 / 18.0;
    p->ww2 = 1.0 / 36.0;
}

/* Channel-Flow/collision.f90:56-220: forcing, moments, relaxation, inverse.
 * post[ip] receives f9(ip) + 0.5*Fbar(ip), the value the swap loop stores. */
static void collide_node(const orc_params *p, const double *fin, double rho9, double ux9, double uy9,
                         double uz9, double fx9, double fy9, double fz9, double *post)
{
    double f9[NPOP], Fbar[NPOP];
    double ux9s, uy9s, uz9s, G1, G2, G3, t1;
    double eqm1, eqm2, eqm3, eqm4, eqm5, eqm6, eqm7, eqm8, eqm9, eqm10, eqm11, eqm12, eqm13, eqm14, eqm15;
    double sum1, sum2, sum3, sum4, sum5, sum6, sum7, sum8, sum9, sum10, sum11;
    double evlm1, evlm2, evlm3, evlm4, evlm5, evlm6, evlm7, evlm8, evlm9, evlm10, evlm11, evlm12, evlm13,
        evlm14, evlm15;
    double eqmc1, eqmc2, eqmc3, eqmc4, eqmc5, eqmc6, eqmc7, eqmc8, eqmc9, eqmc10, eqmc11, eqmc12, eqmc13,
        eqmc14, eqmc15;
    double tl1, tl2, tl3, tl4, tl5, tl6, tl7, tl8, tl9, tl10, tl11, tl12, tl13, tl14, tl15, tl16, tl17, tl18,
        tl19, tl20, tl21;
    double suma, sumb, sumc, sumd, sume, sumf, sumg, sumh, sumi, sumk, sump, sum67, sum89, sum1011;
    int ip;

    ux9s = ux9 * ux9; /* :60-62 */
    uy9s = uy9 * uy9;
    uz9s = uz9 * uz9;

    G3 = ux9 * fx9 + uy9 * fy9 + uz9 * fz9; /* :68 */
    Fbar[0] = -G3;
    for (ip = 1; ip <= 6; ip++) { /* :72-76 */
        G1 = (double)cix[ip] * fx9 + (double)ciy[ip] * fy9 + (double)ciz[ip] * fz9;
        G2 = (double)cix[ip] * ux9 + (double)ciy[ip] * uy9 + (double)ciz[ip] * uz9;
        Fbar[ip] = p->ww1 * (3. * G1 + 9. * G1 * G2 - 3. * G3);
    }
    for (ip = 7; ip <= 18; ip++) { /* :78-82 */
        G1 = (double)cix[ip] * fx9 + (double)ciy[ip] * fy9 + (double)ciz[ip] * fz9;
        G2 = (double)cix[ip] * ux9 + (double)ciy[ip] * uy9 + (double)ciz[ip] * uz9;
        Fbar[ip] = p->ww2 * (3. * G1 + 9. * G1 * G2 - 3. * G3);
    }
    for (ip = 0; ip < NPOP; ip++) /* :84 */
        f9[ip] = fin[ip] + 0.5 * Fbar[ip];

    t1 = ux9s + uy9s + uz9s; /* :86-101 */
    eqm1 = -(11.0 * rho9) + 19.0 * t1;
    eqm2 = p->omegepsl * rho9 + p->omegepslj * t1;
    eqm3 = p->coef1 * ux9;
    eqm4 = p->coef1 * uy9;
    eqm5 = p->coef1 * uz9;
    eqm6 = 2.0 * ux9s - uy9s - uz9s;
    eqm7 = p->omegxx * eqm6;
    eqm8 = uy9s - uz9s;
    eqm9 = p->omegxx * eqm8;
    eqm10 = ux9 * uy9;
    eqm11 = uy9 * uz9;
    eqm12 = ux9 * uz9;
    eqm13 = 0.0;
    eqm14 = 0.0;
    eqm15 = 0.0;

    /* :103-119 */
    sum1 = f9[1] + f9[2] + f9[3] + f9[4] + f9[5] + f9[6];
    sum2 = f9[7] + f9[8] + f9[9] + f9[10] + f9[11] + f9[12] + f9[13] + f9[14] + f9[15] + f9[16] + f9[17] + f9[18];
    sum3 = f9[7] - f9[8] + f9[9] - f9[10] + f9[11] - f9[12] + f9[13] - f9[14];
    sum4 = f9[7] + f9[8] - f9[9] - f9[10] + f9[15] - f9[16] + f9[17] - f9[18];
    sum5 = f9[11] + f9[12] - f9[13] - f9[14] + f9[15] + f9[16] - f9[17] - f9[18];
    sum6 = f9[1] + f9[2];
    sum7 = f9[3] + f9[4] + f9[5] + f9[6];
    sum8 = f9[7] + f9[8] + f9[9] + f9[10] + f9[11] + f9[12] + f9[13] + f9[14];
    sum9 = f9[15] + f9[16] + f9[17] + f9[18];
    sum10 = f9[3] + f9[4] - f9[5] - f9[6];
    sum11 = f9[7] + f9[8] + f9[9] + f9[10] - f9[11] - f9[12] - f9[13] - f9[14];

    /* :121-138 */
    evlm1 = -(30.0 * f9[0]) + p->coef2 * sum1 + p->coef3 * sum2;
    evlm2 = 12.0 * f9[0] + p->coef4 * sum1 + sum2;
    evlm3 = p->coef4 * (f9[1] - f9[2]) + sum3;
    evlm4 = p->coef4 * (f9[3] - f9[4]) + sum4;
    evlm5 = p->coef4 * (f9[5] - f9[6]) + sum5;
    evlm6 = p->coef5 * sum6 - sum7 + sum8 - p->coef5 * sum9;
    evlm7 = p->coef4 * sum6 + p->coef5 * sum7 + sum8 - p->coef5 * sum9;
    evlm8 = sum10 + sum11;
    evlm9 = -(p->coef5 * sum10) + sum11;
    evlm10 = f9[7] - f9[8] - f9[9] + f9[10];
    evlm11 = f9[15] - f9[16] - f9[17] + f9[18];
    evlm12 = f9[11] - f9[12] - f9[13] + f9[14];
    evlm13 = f9[7] - f9[8] + f9[9] - f9[10] - f9[11] + f9[12] - f9[13] + f9[14];
    evlm14 = -f9[7] - f9[8] + f9[9] + f9[10] + f9[15] - f9[16] + f9[17] - f9[18];
    evlm15 = f9[11] + f9[12] - f9[13] - f9[14] - f9[15] - f9[16] + f9[17] + f9[18];

    /* :140-154 */
    eqmc1 = evlm1 - p->s1 * (evlm1 - eqm1);
    eqmc2 = evlm2 - p->s2 * (evlm2 - eqm2);
    eqmc3 = evlm3 - p->s4 * (evlm3 - eqm3);
    eqmc4 = evlm4 - p->s4 * (evlm4 - eqm4);
    eqmc5 = evlm5 - p->s4 * (evlm5 - eqm5);
    eqmc6 = evlm6 - p->s9 * (evlm6 - eqm6);
    eqmc7 = evlm7 - p->s10 * (evlm7 - eqm7);
    eqmc8 = evlm8 - p->s9 * (evlm8 - eqm8);
    eqmc9 = evlm9 - p->s10 * (evlm9 - eqm9);
    eqmc10 = evlm10 - p->s13 * (evlm10 - eqm10);
    eqmc11 = evlm11 - p->s13 * (evlm11 - eqm11);
    eqmc12 = evlm12 - p->s13 * (evlm12 - eqm12);
    eqmc13 = evlm13 - p->s16 * (evlm13 - eqm13);
    eqmc14 = evlm14 - p->s16 * (evlm14 - eqm14);
    eqmc15 = evlm15 - p->s16 * (evlm15 - eqm15);

    /* :157-177 */
    tl1 = p->val1i * rho9;
    tl2 = p->coef2 * p->val2i * eqmc1;
    tl3 = p->coef3 * p->val2i * eqmc1;
    tl4 = p->coef4 * p->val3i * eqmc2;
    tl5 = p->val3i * eqmc2;
    tl6 = p->val4i * ux9;
    tl7 = p->val5i * eqmc3;
    tl8 = p->val4i * uy9;
    tl9 = p->val5i * eqmc4;
    tl10 = p->val4i * uz9;
    tl11 = p->val5i * eqmc5;
    tl12 = p->val6i * eqmc6;
    tl13 = p->val7i * eqmc7;
    tl14 = p->val8i * eqmc8;
    tl15 = p->val9i * eqmc9;
    tl16 = -(p->coef4i * eqmc10);
    tl17 = -(p->coef4i * eqmc11);
    tl18 = -(p->coef4i * eqmc12);
    tl19 = p->coef3i * eqmc13;
    tl20 = p->coef3i * eqmc14;
    tl21 = p->coef3i * eqmc15;

    /* :180-198 */
    f9[0] = tl1 - 30.0 * p->val2i * eqmc1 + p->val8 * p->val3i * eqmc2;
    suma = tl1 + tl2 + tl4;
    sumb = tl1 + tl3 + tl5;
    sumc = tl6 + p->coef4 * tl7;
    sumd = p->coef5 * tl12 + p->coef4 * tl13;
    sume = tl8 + p->coef4 * tl9;
    sumf = -tl12 + p->coef5 * tl13 + tl14 - p->coef5 * tl15;
    sumg = tl10 + p->coef4 * tl11;
    sumh = -tl12 + p->coef5 * tl13 - tl14 + p->coef5 * tl15;
    sumi = tl12 + tl13 + tl14 + tl15;
    sumk = tl12 + tl13 - tl14 - tl15;
    sump = -(p->coef5 * tl12) - p->coef5 * tl13;
    sum67 = tl6 + tl7;
    sum89 = tl8 + tl9;
    sum1011 = tl10 + tl11;

    /* :200-220 */
    f9[1] = suma + sumc + sumd;
    f9[2] = suma - sumc + sumd;
    f9[3] = suma + sume + sumf;
    f9[4] = suma - sume + sumf;
    f9[5] = suma + sumg + sumh;
    f9[6] = suma - sumg + sumh;
    f9[7] = sumb + sum67 + sum89 + sumi + tl16 + tl19 - tl20;
    f9[8] = sumb - sum67 + sum89 + sumi - tl16 - tl19 - tl20;
    f9[9] = sumb + sum67 - sum89 + sumi - tl16 + tl19 + tl20;
    f9[10] = sumb - sum67 - sum89 + sumi + tl16 - tl19 + tl20;
    f9[11] = sumb + sum67 + sum1011 + sumk + tl18 - tl19 + tl21;
    f9[12] = sumb - sum67 + sum1011 + sumk - tl18 + tl19 + tl21;
    f9[13] = sumb + sum67 - sum1011 + sumk - tl18 - tl19 - tl21;
    f9[14] = sumb - sum67 - sum1011 + sumk + tl18 + tl19 - tl21;
    f9[15] = sumb + sum89 + sum1011 + sump + tl17 + tl20 - tl21;
    f9[16] = sumb - sum89 + sum1011 + sump - tl17 - tl20 - tl21;
    f9[17] = sumb + sum89 - sum1011 + sump - tl17 + tl20 + tl21;
    f9[18] = sumb - sum89 - sum1011 + sump + tl17 - tl20 + tl21;

    for (ip = 0; ip < NPOP; ip++) /* the stored value, :230-262 */
        post[ip] = f9[ip] + 0.5 * Fbar[ip];
}

void orc_collide_node(const orc_params *p, const double *fin, double rho9, double ux9, double uy9, double uz9,
                      double fx9, double fy9, double fz9, double *post)
{
    collide_node(p, fin, rho9, ux9, uy9, uz9, fx9, fy9, fz9, post);
}

/* Send buffers of collision.f90:31-32.  Y buffers are (0:18, lx, 0:lz+1),
 * Z buffers (0:18, lx, ly). */
typedef struct {
    double *ym, *yp, *zm, *zp;
} sendbufs;

#define YB(p, ip, ix, kz) ((size_t)(ip) + (size_t)NPOP * ((size_t)((ix)-1) + (size_t)(p)->lx * (size_t)(kz)))
#define ZB(p, ip, ix, jy) ((size_t)(ip) + (size_t)NPOP * ((size_t)((ix)-1) + (size_t)(p)->lx * (size_t)((jy)-1)))

/* Channel-Flow/collision.f90:281-372 on one rank.  With every neighbour id
 * equal to myid (para.f90:267-276), the receive with tag 0 from mym gets the
 * message this rank sent to myp with tag 0 (tmpypS), and so on. */
static void orc_exchange(const orc_params *p, double *f, sendbufs *b)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
    const int ny5 = 5;
    /* collision.f90:290-291: 5 populations per face */
    size_t ylen = (size_t)ny5 * lx * (size_t)(lz + 2);
    size_t zlen = (size_t)ny5 * lx * (size_t)ly;
    double *ypS = malloc(ylen * sizeof(double)), *ymS = malloc(ylen * sizeof(double));
    double *zpS = malloc(zlen * sizeof(double)), *zmS = malloc(zlen * sizeof(double));
#define Y5(a, k, ix, kz) (a)[(size_t)((k)-1) + 5 * ((size_t)((ix)-1) + (size_t)lx * (size_t)(kz))]
#define Z5(a, k, ix, jy) (a)[(size_t)((k)-1) + 5 * ((size_t)((ix)-1) + (size_t)lx * (size_t)((jy)-1))]
    static const int ypk[5] = {3, 7, 8, 15, 17}, ymk[5] = {4, 9, 10, 16, 18};
    static const int zpk[5] = {5, 11, 12, 15, 16}, zmk[5] = {6, 13, 14, 17, 18};
    int k, ix, iy, iz;

    /* :295-305 */
    for (iz = 0; iz <= lz + 1; iz++)
        for (ix = 1; ix <= lx; ix++)
            for (k = 1; k <= 5; k++) {
                Y5(ypS, k, ix, iz) = b->yp[YB(p, ypk[k - 1], ix, iz)];
                Y5(ymS, k, ix, iz) = b->ym[YB(p, ymk[k - 1], ix, iz)];
            }
    /* :309-314 one rank: tmpymR <- tmpypS, tmpypR <- tmpymS */
    double *ymR = ypS, *ypR = ymS;

    /* :318-328 */
    for (iz = 1; iz <= lz; iz++)
        for (ix = 1; ix <= lx; ix++) {
            f[FI(p, 3, ix, 1, iz)] = Y5(ymR, 1, ix, iz);
            if (ix >= 2)
                f[FI(p, 7, ix, 1, iz)] = Y5(ymR, 2, ix, iz);
            if (ix <= lx - 1)
                f[FI(p, 8, ix, 1, iz)] = Y5(ymR, 3, ix, iz);
            f[FI(p, 15, ix, 1, iz)] = Y5(ymR, 4, ix, iz);
            f[FI(p, 17, ix, 1, iz)] = Y5(ymR, 5, ix, iz);

            f[FI(p, 4, ix, ly, iz)] = Y5(ypR, 1, ix, iz);
            if (ix >= 2)
                f[FI(p, 9, ix, ly, iz)] = Y5(ypR, 2, ix, iz);
            if (ix <= lx - 1)
                f[FI(p, 10, ix, ly, iz)] = Y5(ypR, 3, ix, iz);
            f[FI(p, 16, ix, ly, iz)] = Y5(ypR, 4, ix, iz);
            f[FI(p, 18, ix, ly, iz)] = Y5(ypR, 5, ix, iz);
        }

    /* :331-334 corner forwarding into the Z send buffers */
    for (ix = 1; ix <= lx; ix++) {
        b->zm[ZB(p, 17, ix, 1)] = Y5(ymR, 5, ix, 0);
        b->zm[ZB(p, 18, ix, ly)] = Y5(ypR, 5, ix, 0);
        b->zp[ZB(p, 15, ix, 1)] = Y5(ymR, 4, ix, lz + 1);
        b->zp[ZB(p, 16, ix, ly)] = Y5(ypR, 4, ix, lz + 1);
    }

    /* :337-347 */
    for (iy = 1; iy <= ly; iy++)
        for (ix = 1; ix <= lx; ix++)
            for (k = 1; k <= 5; k++) {
                Z5(zpS, k, ix, iy) = b->zp[ZB(p, zpk[k - 1], ix, iy)];
                Z5(zmS, k, ix, iy) = b->zm[ZB(p, zmk[k - 1], ix, iy)];
            }
    /* :351-356 one rank: tmpzmR <- tmpzpS, tmpzpR <- tmpzmS */
    double *zmR = zpS, *zpR = zmS;

    /* :360-370 */
    for (iy = 1; iy <= ly; iy++)
        for (ix = 1; ix <= lx; ix++) {
            f[FI(p, 5, ix, iy, 1)] = Z5(zmR, 1, ix, iy);
            if (ix >= 2)
                f[FI(p, 11, ix, iy, 1)] = Z5(zmR, 2, ix, iy);
            if (ix <= lx - 1)
                f[FI(p, 12, ix, iy, 1)] = Z5(zmR, 3, ix, iy);
            f[FI(p, 15, ix, iy, 1)] = Z5(zmR, 4, ix, iy);
            f[FI(p, 16, ix, iy, 1)] = Z5(zmR, 5, ix, iy);

            f[FI(p, 6, ix, iy, lz)] = Z5(zpR, 1, ix, iy);
            if (ix >= 2)
                f[FI(p, 13, ix, iy, lz)] = Z5(zpR, 2, ix, iy);
            if (ix <= lx - 1)
                f[FI(p, 14, ix, iy, lz)] = Z5(zpR, 3, ix, iy);
            f[FI(p, 17, ix, iy, lz)] = Z5(zpR, 4, ix, iy);
            f[FI(p, 18, ix, iy, lz)] = Z5(zpR, 5, ix, iy);
        }
#undef Y5
#undef Z5
    free(ypS);
    free(ymS);
    free(zpS);
    free(zmS);
}

/* Channel-Flow/collision.f90:24-273.  The macroscopic arrays are read, not
 * recomputed: the caller decides what they hold (macrovar, rhoupdat plus an
 * external velocity during pre-relaxation, or avedensity's correction). */
void orc_collision_mrt(const orc_params *p, double *f, const double *rho, const double *ux, const double *uy,
                       const double *uz, const double *fx, const double *fy, const double *fz)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
    sendbufs b;
    size_t ylen = (size_t)NPOP * lx * (size_t)(lz + 2), zlen = (size_t)NPOP * lx * (size_t)ly;
    b.ym = calloc(ylen, sizeof(double));
    b.yp = calloc(ylen, sizeof(double));
    b.zm = calloc(zlen, sizeof(double));
    b.zp = calloc(zlen, sizeof(double));
    double post[NPOP];
    int ix, iy, iz, ipi, ip;

    for (iz = 1; iz <= lz; iz++)
        for (iy = 1; iy <= ly; iy++)
            for (ix = 1; ix <= lx; ix++) {
                size_t m = MI(p, ix, iy, iz);
                collide_node(p, &f[FI(p, 0, ix, iy, iz)], rho[m], ux[m], uy[m], uz[m], fx[m], fy[m], fz[m],
                             post);
                /* :224-242 stay set */
                for (ipi = 0; ipi < 10; ipi++) {
                    ip = ipstay[ipi];
                    int im = ix + cix[ip], jm = iy + ciy[ip], km = iz + ciz[ip];
                    if (im < 1 || im > lx)
                        f[FI(p, ipopp[ip], ix, iy, iz)] = post[ip];
                    else if (jm < 1)
                        b.ym[YB(p, ip, im, km)] = post[ip];
                    else if (jm > ly)
                        b.yp[YB(p, ip, im, km)] = post[ip];
                    else if (km < 1)
                        b.zm[ZB(p, ip, im, jm)] = post[ip];
                    else if (km > lz)
                        b.zp[ZB(p, ip, im, jm)] = post[ip];
                    else
                        f[FI(p, ipopp[ip], ix, iy, iz)] = post[ip];
                }
                /* :245-264 swap set */
                for (ipi = 0; ipi < 9; ipi++) {
                    ip = ipswap[ipi];
                    int im = ix + cix[ip], jm = iy + ciy[ip], km = iz + ciz[ip];
                    if (im < 1 || im > lx)
                        f[FI(p, ipopp[ip], ix, iy, iz)] = post[ip];
                    else if (jm < 1)
                        b.ym[YB(p, ip, im, km)] = post[ip];
                    else if (jm > ly)
                        b.yp[YB(p, ip, im, km)] = post[ip];
                    else if (km < 1)
                        b.zm[ZB(p, ip, im, jm)] = post[ip];
                    else if (km > lz)
                        b.zp[ZB(p, ip, im, jm)] = post[ip];
                    else {
                        f[FI(p, ipopp[ip], ix, iy, iz)] = f[FI(p, ip, im, jm, km)];
                        f[FI(p, ip, im, jm, km)] = post[ip];
                    }
                }
            }
    orc_exchange(p, f, &b); /* :271 */
    free(b.ym);
    free(b.yp);
    free(b.zm);
    free(b.zp);
}

/* Same arithmetic; push streaming from fin to a separate fout (x-wall
 * halfway bounce-back, periodic y/z).  Parallel over z planes. */
void orc_collision_mrt_ab(const orc_params *p, const double *fin, double *fout, const double *rho,
                          const double *ux, const double *uy, const double *uz, const double *fx,
                          const double *fy, const double *fz, int nthreads)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int iz = 1; iz <= lz; iz++) {
        double post[NPOP];
        for (int iy = 1; iy <= ly; iy++)
            for (int ix = 1; ix <= lx; ix++) {
                size_t m = MI(p, ix, iy, iz);
                collide_node(p, &fin[FI(p, 0, ix, iy, iz)], rho[m], ux[m], uy[m], uz[m], fx[m], fy[m], fz[m],
                             post);
                for (int ip = 0; ip < NPOP; ip++) {
                    int im = ix + cix[ip];
                    if (im < 1 || im > lx) {
                        fout[FI(p, ipopp[ip], ix, iy, iz)] = post[ip];
                    } else {
                        int jm = iy + ciy[ip], km = iz + ciz[ip];
                        if (jm < 1)
                            jm += ly;
                        if (jm > ly)
                            jm -= ly;
                        if (km < 1)
                            km += lz;
                        if (km > lz)
                            km -= lz;
                        fout[FI(p, ip, im, jm, km)] = post[ip];
                    }
                }
            }
    }
}

/* Channel-Flow/collision.f90:396-418 (fluid nodes; ibnodes = -1 everywhere,
 * para.f90:447). */
void orc_macrovar(const orc_params *p, const double *f, double *rho, double *ux, double *uy, double *uz,
                  const double *fx, const double *fy, const double *fz)
{
    const long n = (long)p->lx * p->ly * p->lz;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long m = 0; m < n; m++) {
        const double *f9 = &f[NPOP * m];
        double sum1 = f9[7] - f9[10];
        double sum2 = f9[9] - f9[8];
        double sum3 = f9[11] - f9[14];
        double sum4 = f9[13] - f9[12];
        double sum5 = f9[15] - f9[18];
        double sum6 = f9[17] - f9[16];
        double ux9 = f9[1] - f9[2] + sum1 + sum2 + sum3 + sum4;
        double uy9 = f9[3] - f9[4] + sum1 - sum2 + sum5 + sum6;
        double uz9 = f9[5] - f9[6] + sum3 - sum4 + sum5 - sum6;
        double rho9 = f9[0] + f9[1] + f9[2] + f9[3] + f9[4] + f9[5] + f9[6] + f9[7] + f9[8] + f9[9] + f9[10] +
                      f9[11] + f9[12] + f9[13] + f9[14] + f9[15] + f9[16] + f9[17] + f9[18];
        ux[m] = ux9 + fx[m] / 2.;
        uy[m] = uy9 + fy[m] / 2.;
        uz[m] = uz9 + fz[m] / 2.;
        rho[m] = rho9;
    }
}

/* Channel-Flow/collision.f90:475-478: rho = f(0); rho = rho + f(ip). */
void orc_rhoupdat(const orc_params *p, const double *f, double *rho)
{
    size_t n = (size_t)p->lx * p->ly * p->lz;
    for (size_t m = 0; m < n; m++) {
        double r = f[NPOP * m];
        for (int ip = 1; ip < NPOP; ip++)
            r = r + f[NPOP * m + ip];
        rho[m] = r;
    }
}

/* Channel-Flow/collision.f90:496-511 (all nodes fluid; sequential sum in
 * array-element order, as Fortran SUM traverses it). */
double orc_avedensity(const orc_params *p, double *rho)
{
    size_t n = (size_t)p->lx * p->ly * p->lz;
    double rhomean0 = 0.0;
    for (size_t m = 0; m < n; m++)
        rhomean0 = rhomean0 + rho[m];
    double rhomean = rhomean0 / (double)n;
    for (size_t m = 0; m < n; m++)
        rho[m] = rho[m] - rhomean;
    return rhomean;
}

/* Channel-Flow/collision.f90:522-524 */
void orc_forcing(const orc_params *p, double *fx, double *fy, double *fz)
{
    size_t n = (size_t)p->lx * p->ly * p->lz;
    double fyv = p->force_in_y * p->force_mag;
    for (size_t m = 0; m < n; m++) {
        fx[m] = 0.0;
        fy[m] = fyv;
        fz[m] = 0.0;
    }
}

/* FORCINGP, Channel-Flow/collision.f90:529-602, one rank (alphay = alphaz = 0,
 * collision.f90:551-559): the mean force plus a divergence-free perturbation
 * in two wall-parallel bands, beta = 3 streamwise and gamma = 2 spanwise
 * waves, amplitude 40*beta/ny*sin(2 pi istep / 2000).  The second band
 * (collision.f90:583-599) is shifted a quarter period in y and overwrites the
 * first where the two meet.  Every expression keeps the reference's operand
 * order (Fortran evaluates a*b/c*d left to right; a leading minus negates
 * the whole product, which equals negating its first factor exactly). */
void orc_forcingp(const orc_params *p, int istep, double *fx, double *fy, double *fz)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz, nx = p->lx, ny = p->ly, nz = p->lz;
    const double pi = 4.0 * atan(1.0); /* var_inc.f90:71-72 */
    const double pi2 = 2.0 * pi;
    const double Tpd = 2000., beta9 = 3.0, gamma9 = 2.0, phase9 = 0.25;
    const int ixs0 = 2;
    const double Amp0 = 40.00 * beta9 / (double)ny * sin(pi2 * (double)istep / Tpd); /* :541 */
    const double force_in_y = p->force_in_y;
    const size_t n = (size_t)lx * ly * lz;
    for (size_t m = 0; m < n; m++) { /* :546-548 */
        fx[m] = 0.0;
        fy[m] = force_in_y;
        fz[m] = 0.0;
    }
    const int lxh = lx / 2; /* var_inc.f90:53 */
    int ixs = ixs0, ihh = lxh / 2;
    for (int k = 1; k <= lz; k++) { /* :563-580 */
        const double z9 = pi2 * ((double)k - 0.5) / (double)nz;
        for (int j = 1; j <= ly; j++) {
            const double y9 = pi2 * ((double)j - 0.5) / (double)ny;
            for (int i = 1; i <= ihh; i++) {
                const double x9 = pi2 * ((double)i - 0.5) / (double)ihh;
                const size_t m = (size_t)(i + ixs - 1) + (size_t)lx * ((size_t)(j - 1) + (size_t)ly * (k - 1));
                fx[m] = force_in_y * 0.5 * Amp0 * (double)ihh * (1. - cos(x9)) * cos(beta9 * y9) * cos(gamma9 * z9);
                fy[m] = force_in_y * (1.0 - Amp0 * (double)ny / beta9 * sin(x9) * sin(beta9 * y9) * cos(gamma9 * z9));
                fz[m] = force_in_y * 0.5 * Amp0 * (double)nz / gamma9 * sin(x9) * cos(beta9 * y9) * sin(gamma9 * z9);
            }
        }
    }
    ihh = lxh / 2;
    ixs = nx - ixs0 - ihh;
    for (int k = 1; k <= lz; k++) { /* :583-599 */
        const double z9 = pi2 * ((double)k - 0.5) / (double)nz;
        for (int j = 1; j <= ly; j++) {
            const double y9 = pi2 * (((double)j - 0.5) / (double)ny + phase9);
            for (int i = 1; i <= ihh; i++) {
                const double x9 = pi2 * ((double)i - 0.5) / (double)ihh;
                const size_t m = (size_t)(i + ixs - 1) + (size_t)lx * ((size_t)(j - 1) + (size_t)ly * (k - 1));
                fx[m] = -(force_in_y * 0.5 * Amp0 * (double)ihh * (1. - cos(x9)) * cos(beta9 * y9) * cos(gamma9 * z9));
                fy[m] = force_in_y * (1.0 + Amp0 * (double)ny / beta9 * sin(x9) * sin(beta9 * y9) * cos(gamma9 * z9));
                fz[m] = -(force_in_y * 0.5 * Amp0 * (double)nz / gamma9 * sin(x9) * cos(beta9 * y9) * sin(gamma9 * z9));
            }
        }
    }
}

/* Channel-Flow/initial.f90:26-44: rho = 0, f = equilibrium. */
void orc_initpop(const orc_params *p, double *f, double *rho, const double *ux, const double *uy,
                 const double *uz)
{
    size_t n = (size_t)p->lx * p->ly * p->lz;
    for (size_t m = 0; m < n; m++) {
        double usqr = ux[m] * ux[m] + uy[m] * uy[m] + uz[m] * uz[m];
        usqr = 1.5 * usqr;
        rho[m] = 0.0;
        f[NPOP * m + 0] = p->ww0 * (rho[m] - usqr);
        for (int ip = 1; ip <= 6; ip++) {
            double G = ((double)cix[ip] * ux[m] + (double)ciy[ip] * uy[m] + (double)ciz[ip] * uz[m]);
            f[NPOP * m + ip] = p->ww1 * (rho[m] + 3.0 * G + 4.5 * G * G - usqr);
        }
        for (int ip = 7; ip < NPOP; ip++) {
            double G = ((double)cix[ip] * ux[m] + (double)ciy[ip] * uy[m] + (double)ciz[ip] * uz[m]);
            f[NPOP * m + ip] = p->ww2 * (rho[m] + 3.0 * G + 4.5 * G * G - usqr);
        }
    }
}

/* Channel-Flow/initial.f90:94-115: log-law mean profile mirrored about the
 * centre plane; the perturbation amplitude A9 is 0 (:84), so the perturbation
 * loop (:119-144) adds exact zeros and is omitted.  nxh = lx/2 as for an even
 * lx = nx7 (var_inc.f90:51-54). */
void orc_initvel(const orc_params *p, double *ux, double *uy, double *uz)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz, nxh = lx / 2;
    size_t n = (size_t)lx * ly * lz;
    for (size_t m = 0; m < n; m++) {
        ux[m] = 0.0;
        uy[m] = 0.0;
        uz[m] = 0.0;
    }
    for (int i = 1; i <= nxh; i++) {
        double yplus = ((double)i - 0.5) / p->ystar;
        double v;
        if (yplus < 10.8) {
            v = yplus * p->ustar;
        } else {
            double u9 = log(yplus) / 0.41 + 5.0;
            v = u9 * p->ustar;
        }
        for (int iz = 1; iz <= lz; iz++)
            for (int iy = 1; iy <= ly; iy++) {
                uy[MI(p, i, iy, iz)] = v;
                uy[MI(p, lx + 1 - i, iy, iz)] = v;
            }
    }
}

/* Channel-Flow/saveload.f90:1382-1470 on one rank, all nodes fluid.
 * out is (15, lx) column-major: per plane i the 15 columns written at :1492. */
void orc_statistc2(const orc_params *p, const double *rho, const double *ux, const double *uy,
                   const double *uz, double *out)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
    for (int i = 1; i <= lx; i++) {
        double vx = 0, vy = 0, vz = 0, sxz = 0, sxy = 0, syz = 0, xx = 0, yy = 0, zz = 0, pr = 0, p2 = 0;
        int nf = 0;
        for (int k = 1; k <= lz; k++)
            for (int j = 1; j <= ly; j++) {
                size_t m = MI(p, i, j, k);
                nf++;
                vx = vx + ux[m];
                vy = vy + uy[m];
                vz = vz + uz[m];
                sxz = sxz + ux[m] * uz[m];
                sxy = sxy + ux[m] * uy[m];
                syz = syz + uy[m] * uz[m];
                xx = xx + ux[m] * ux[m];
                yy = yy + uy[m] * uy[m];
                zz = zz + uz[m] * uz[m];
                pr = pr + rho[m];
                p2 = p2 + rho[m] * rho[m];
            }
        double us = p->ustar, rn = (double)nf;
        double vxa = vx / rn / us, vya = vy / rn / us, vza = vz / rn / us;
        double vxs = xx / rn / (us * us), vys = yy / rn / (us * us), vzs = zz / rn / (us * us);
        double sxzt = sxz / rn / (us * us), sxyt = sxy / rn / (us * us), syzt = syz / rn / (us * us);
        double volf = rn / ((double)ly * (double)lz);
        vxs = vxs - vxa * vxa;
        vys = vys - vya * vya;
        vzs = vzs - vza * vza;
        sxzt = sxzt - vxa * vza;
        sxyt = sxyt - vxa * vya;
        syzt = syzt - vya * vza;
        double pra = pr / rn / (us * us);
        double prs = p2 / rn / ((us * us) * (us * us));
        prs = prs - pra * pra;
        double yplus = ((double)i - 0.5) / p->ystar;
        if (i > lx / 2)
            yplus = ((double)(lx - i) + 0.5) / p->ystar;
        double *o = &out[15 * (size_t)(i - 1)];
        o[0] = (double)i - 0.5;
        o[1] = yplus;
        o[2] = vxa;
        o[3] = vya;
        o[4] = vza;
        o[5] = vxs;
        o[6] = vys;
        o[7] = vzs;
        o[8] = sxzt;
        o[9] = sxyt;
        o[10] = syzt;
        o[11] = pra;
        o[12] = prs;
        o[13] = volf;
        o[14] = 1.0 - volf;
    }
}

/* Channel-Flow/saveload.f90:1535-1660 on one rank, all nodes fluid.
 * out: vmax, imout, jmout, kmout, umean, vmean, wmean, urms, vrms, wrms,
 * volf, rhomax, rhomin (all normalised as written to diag.dat). */
void orc_diag(const orc_params *p, const double *rho, const double *ux, const double *uy, const double *uz,
              double *out)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
    size_t n = (size_t)lx * ly * lz;
    double um = 0, vm = 0, wm = 0, ur = 0, vr = 0, wr = 0;
    for (size_t m = 0; m < n; m++) {
        um = um + ux[m];
        vm = vm + uy[m];
        wm = wm + uz[m];
        ur = ur + ux[m] * ux[m];
        vr = vr + uy[m] * uy[m];
        wr = wr + uz[m] * uz[m];
    }
    double vmax = 0.0;
    int im = 0, jm = 0, km = 0;
    for (int k = 1; k <= lz; k++)
        for (int j = 1; j <= ly; j++)
            for (int i = 1; i <= lx; i++) {
                size_t m = MI(p, i, j, k);
                double vel = sqrt(ux[m] * ux[m] + uy[m] * uy[m] + uz[m] * uz[m]);
                if (vel > vmax) {
                    vmax = vel;
                    im = i;
                    jm = j;
                    km = k;
                }
            }
    double rmax = rho[0], rmin = rho[0];
    for (size_t m = 1; m < n; m++) {
        if (rho[m] > rmax)
            rmax = rho[m];
        if (rho[m] < rmin)
            rmin = rho[m];
    }
    double rn = (double)n;
    um = um / rn;
    vm = vm / rn;
    wm = wm / rn;
    ur = sqrt(ur / rn - um * um);
    vr = sqrt(vr / rn - vm * vm);
    wr = sqrt(wr / rn - wm * wm);
    out[0] = vmax;
    out[1] = im;
    out[2] = jm;
    out[3] = km;
    out[4] = um / p->ustar;
    out[5] = vm / p->ustar;
    out[6] = wm / p->ustar;
    out[7] = ur / p->ustar;
    out[8] = vr / p->ustar;
    out[9] = wr / p->ustar;
    out[10] = 1. - rn / rn;
    out[11] = rmax;
    out[12] = rmin;
}

/* Channel-Flow/saveload.f90:3929-4023 (vortcalc) on one rank: vorticity by
 * central differences in y and z (periodic, the exchng8 images of
 * :4027-4054), in x central inside and the one-sided wall forms of
 * :3970-3975 at ix = 1 and lx.  Nodes with solid[m] >= 0 (nullable) carry
 * twice the angular velocity of particle solid[m] (omg, 3 per particle). */
void orc_vortcalc(const orc_params *p, const double *ux, const double *uy, const double *uz, const int *solid,
                  const double *omg, double *ox, double *oy, double *oz)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
    for (int k = 1; k <= lz; k++)
        for (int j = 1; j <= ly; j++)
            for (int i = 1; i <= lx; i++) {
                const size_t m = MI(p, i, j, k);
                if (solid && solid[m] >= 0) {
                    const double *w = omg + 3 * (size_t)solid[m];
                    ox[m] = 2.0 * w[0];
                    oy[m] = 2.0 * w[1];
                    oz[m] = 2.0 * w[2];
                    continue;
                }
                double pwx, pvx, pwy, puy, puz, pvz;
                if (i == 1) {
                    pwx = (3.0 * uz[m] + uz[MI(p, i + 1, j, k)]) / 3.0;
                    pvx = (3.0 * uy[m] + uy[MI(p, i + 1, j, k)]) / 3.0;
                } else if (i == lx) {
                    pwx = -(3.0 * uz[m] + uz[MI(p, i - 1, j, k)]) / 3.0;
                    pvx = -(3.0 * uy[m] + uy[MI(p, i - 1, j, k)]) / 3.0;
                } else {
                    pwx = (uz[MI(p, i + 1, j, k)] - uz[MI(p, i - 1, j, k)]) / 2.0;
                    pvx = (uy[MI(p, i + 1, j, k)] - uy[MI(p, i - 1, j, k)]) / 2.0;
                }
                const int jp = j == ly ? 1 : j + 1, jm = j == 1 ? ly : j - 1;
                const int kp = k == lz ? 1 : k + 1, km = k == 1 ? lz : k - 1;
                pwy = (uz[MI(p, i, jp, k)] - uz[MI(p, i, jm, k)]) / 2.0;
                puy = (ux[MI(p, i, jp, k)] - ux[MI(p, i, jm, k)]) / 2.0;
                puz = (ux[MI(p, i, j, kp)] - ux[MI(p, i, j, km)]) / 2.0;
                pvz = (uy[MI(p, i, j, kp)] - uy[MI(p, i, j, km)]) / 2.0;
                ox[m] = pwy - pvz;
                oy[m] = puz - pwx;
                oz[m] = pvx - puy;
            }
}

/* The per-node part of Channel-Flow/saveload.f90:4677-4776 (dissipation_calc):
 * the strain-rate magnitude S_ij S_ij and S_xy from the non-equilibrium
 * moments (method 1) and from the non-equilibrium populations (method 2). */
void orc_dissipation_node(const orc_params *p, const double *f9, double rho9, double ux9, double uy9, double uz9,
                          double *out4)
{
    const double ux9s = ux9 * ux9, uy9s = uy9 * uy9, uz9s = uz9 * uz9;
    const double eqm1 = -11.0 * rho9 + 19.0 * (ux9s + uy9s + uz9s);
    const double eqm6 = 2.0 * ux9s - uy9s - uz9s;
    const double eqm8 = uy9s - uz9s;
    const double eqm10 = ux9 * uy9;
    const double eqm11 = uy9 * uz9;
    const double eqm12 = ux9 * uz9;
    const double sum1 = f9[1] + f9[2] + f9[3] + f9[4] + f9[5] + f9[6];
    const double sum2 = f9[7] + f9[8] + f9[9] + f9[10] + f9[11] + f9[12] + f9[13] + f9[14] + f9[15] + f9[16] +
                        f9[17] + f9[18];
    const double sum6 = f9[1] + f9[2];
    const double sum7 = f9[3] + f9[4] + f9[5] + f9[6];
    const double sum8 = f9[7] + f9[8] + f9[9] + f9[10] + f9[11] + f9[12] + f9[13] + f9[14];
    const double sum9 = f9[15] + f9[16] + f9[17] + f9[18];
    const double sum10 = f9[3] + f9[4] - f9[5] - f9[6];
    const double sum11 = f9[7] + f9[8] + f9[9] + f9[10] - f9[11] - f9[12] - f9[13] - f9[14];
    const double evlm1 = -30.0 * f9[0] + p->coef2 * sum1 + p->coef3 * sum2;
    const double evlm6 = p->coef5 * sum6 - sum7 + sum8 - p->coef5 * sum9;
    const double evlm8 = sum10 + sum11;
    const double evlm10 = f9[7] - f9[8] - f9[9] + f9[10];
    const double evlm11 = f9[15] - f9[16] - f9[17] + f9[18];
    const double evlm12 = f9[11] - f9[12] - f9[13] + f9[14];
    const double neqm1 = p->s1 * (evlm1 - eqm1);
    const double neqm9 = p->s9 * (evlm6 - eqm6);
    const double neqm11 = p->s9 * (evlm8 - eqm8);
    const double neqm13 = p->s9 * (evlm10 - eqm10);
    const double neqm14 = p->s9 * (evlm11 - eqm11);
    const double neqm15 = p->s9 * (evlm12 - eqm12);
    double Sxx = -(neqm1 + 19.0 * neqm9) / 38.0;
    double Syy = -(2.0 * neqm1 - 19.0 * (neqm9 - 3.0 * neqm11)) / 76.0;
    double Szz = -(2.0 * neqm1 - 19.0 * (neqm9 + 3.0 * neqm11)) / 76.0;
    double Sxy = -1.5 * neqm13;
    double Syz = -1.5 * neqm14;
    double Szx = -1.5 * neqm15;
    out4[0] = Sxx * Sxx + Syy * Syy + Szz * Szz + 2.0 * (Sxy * Sxy + Syz * Syz + Szx * Szx);
    out4[1] = Sxy;
    Sxx = 0.0;
    Syy = 0.0;
    Szz = 0.0;
    Sxy = 0.0;
    Syz = 0.0;
    Szx = 0.0;
    const double usqr = 1.5 * (ux9s + uy9s + uz9s);
    for (int ip = 1; ip < NPOP; ip++) {
        const double wght = ip <= 6 ? p->ww1 : p->ww2;
        const double edtu = (double)cix[ip] * ux9 + (double)ciy[ip] * uy9 + (double)ciz[ip] * uz9;
        const double feq9 = wght * (rho9 + 3.0 * edtu + 4.5 * (edtu * edtu) - usqr);
        const double fneq9 = f9[ip] - feq9;
        Sxx = Sxx + fneq9 * (double)(cix[ip] * cix[ip]);
        Syy = Syy + fneq9 * (double)(ciy[ip] * ciy[ip]);
        Szz = Szz + fneq9 * (double)(ciz[ip] * ciz[ip]);
        Sxy = Sxy + fneq9 * (double)(cix[ip] * ciy[ip]);
        Syz = Syz + fneq9 * (double)(ciy[ip] * ciz[ip]);
        Szx = Szx + fneq9 * (double)(ciz[ip] * cix[ip]);
    }
    const double Sij2m2_tmp = Sxx * Sxx + Syy * Syy + Szz * Szz + 2.0 * (Sxy * Sxy + Syz * Syz + Szx * Szx);
    out4[2] = Sij2m2_tmp * ((1.5 / p->tau) * (1.5 / p->tau));
    out4[3] = Sxy * (1.5 / p->tau);
}

/* Channel-Flow/saveload.f90:4644-4882 (dissipation_calc) on one rank: per
 * wall-normal plane i the fluid-node means of S_ij S_ij and S_xy by the two
 * methods, from the populations f (reference layout) and the fields.
 * out is (6, lx) column-major: x, y+, <S_ij S_ij>_1, <S_ij S_ij>_2,
 * <S_xy>_1, <S_xy>_2 (the columns written at :4873). */
void orc_dissipation(const orc_params *p, const double *f, const double *rho, const double *ux, const double *uy,
                     const double *uz, const int *solid, double *out)
{
    const int lx = p->lx, ly = p->ly, lz = p->lz;
    for (int i = 1; i <= lx; i++) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        int nf = 0;
        for (int k = 1; k <= lz; k++)
            for (int j = 1; j <= ly; j++) {
                const size_t m = MI(p, i, j, k);
                if (solid && solid[m] >= 0)
                    continue;
                double v[4];
                orc_dissipation_node(p, &f[FI(p, 0, i, j, k)], rho[m], ux[m], uy[m], uz[m], v);
                s[0] = s[0] + v[0];
                s[1] = s[1] + v[2];
                s[2] = s[2] + v[1];
                s[3] = s[3] + v[3];
                nf++;
            }
        double yplus = ((double)i - 0.5) / p->ystar;
        if (i > lx / 2)
            yplus = ((double)(lx - i) + 0.5) / p->ystar;
        double *o = &out[6 * (size_t)(i - 1)];
        o[0] = (double)i - 0.5;
        o[1] = yplus;
        o[2] = s[0] / (double)nf;
        o[3] = s[1] / (double)nf;
        o[4] = s[2] / (double)nf;
        o[5] = s[3] / (double)nf;
    }
}
