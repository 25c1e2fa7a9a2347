/*
 * CPU oracle for the rigid-particle path -- TEST INFRASTRUCTURE ONLY (tests/
 * check the GPU results against it).
 *
 * The reference has no particle source (only its data structures and routine
 * names: Channel-Flow/var_inc.f90:30-38,120-150,166-168), so this file states
 * the published methods directly, sequentially, in the order a single thread
 * would apply them:
 *   classification  |x - X|^2 < R^2 with minimum image in y, z; cut-link
 *                   fraction q = smaller root of |x + t c - X|^2 = R^2
 *   bounce-back     Bouzidi, Firdaouss & Lallemand (2001), linear, moving wall
 *   force/torque    momentum exchange of Wen et al. (2014)
 *   refill          equilibrium at the wall velocity, mean neighbour density
 *   contact         Feng & Michaelides, J. Comput. Phys. 202 (2005) 20, eq. 28,
 *                   with the stiffness pairs stf0/stf1 and stf0_w/stf1_w of
 *                   Channel-Flow/para.f90:356-360 (walls as image particles)
 *   lubrication     normal lubrication correction below a cut-off gap, Nguyen
 *                   & Ladd, Phys. Rev. E 66 (2002) 046708 (flubp,
 *                   Channel-Flow/var_inc.f90:150)
 *   integration     Adams-Bashforth 2 velocities on the hydrodynamic force and
 *                   torque averaged over the last two steps (Ladd, J. Fluid
 *                   Mech. 271 (1994) 285), trapezoidal positions
 * The fluid step is the pull form of collision_MRT with nodes inside
 * particles skipped.  State is the post-collision lattice P (19*N,
 * population-major) in two buffers used alternately, as on the device.
 * Reductions the device makes (avedensity's mean) are restated in the
 * device's fixed summation order (orc_device_sum).
 *
 * Pinned by tests/test_particles_oracle.py: brute-force classification over
 * every node and direction; the q root against the quadratic; the bounce-back
 * against the kinematic value of piecewise-linear populations mirrored at a
 * planar wall (q = 0.2, 0.35, 0.5, 0.65, 0.8, moving and fixed wall, the
 * fallback); a sphere moving with a uniform flow feeling no force (Galilean
 * invariance); the momentum balance between fluid and a fixed and a moving
 * particle; the refilled density and velocity; contact and lubrication forces
 * against hand-evaluated values; Newton's third law; the closed-form AB2
 * trajectory of a particle under a constant force.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"
#include "particles_oracle.h"

#define NPOP 19
#define ORC_PI 3.14159265358979323846

static const int cix[NPOP] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int ciy[NPOP] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int ciz[NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int ipopp[NPOP] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

static int wrapi(int i, int L)
{
    int r = i % L;
    return r < 0 ? r + L : r;
}

static double min_image(double d, double L)
{
    if (d > 0.5 * L)
        d = d - L;
    else if (d < -0.5 * L)
        d = d + L;
    return d;
}

/* Fixed summation order of the per-particle sums (force/torque over links,
 * contact force over the other particles): term k goes to partial k % 64,
 * each partial adds its terms in increasing k, then the 64 partials are
 * added pairwise (l += l + 32, l += l + 16, ..., l += l + 1). */
#define SUMW 64

static void lane_tree(double (*part)[6], int ncomp)
{
    for (int off = SUMW / 2; off > 0; off >>= 1)
        for (int l = 0; l < off; l++)
            for (int j = 0; j < ncomp; j++)
                part[l][j] = part[l][j] + part[l + off][j];
}

static size_t nodeof(const orc_pgeom *g, int ix, int iy, int iz)
{
    return (size_t)ix + (size_t)g->lx * ((size_t)iy + (size_t)g->ly * (size_t)iz);
}

static void offset(const orc_pgeom *g, int ix, int iy, int iz, const double *X, double *dx, double *dy, double *dz)
{
    *dx = ((double)ix + 0.5) - X[0];
    *dy = min_image(((double)iy + 0.5) - X[1], (double)g->ly);
    *dz = min_image(((double)iz + 0.5) - X[2], (double)g->lz);
}

static int inside(const orc_pgeom *g, int ix, int iy, int iz, const double *X)
{
    double dx, dy, dz;
    offset(g, ix, iy, iz, X, &dx, &dy, &dz);
    const double d2 = dx * dx + dy * dy + dz * dz;
    return d2 < g->R2;
}

typedef struct {
    int x0, x1, y0, y1, z0, z1;
} box;

static box sphere_box(const orc_pgeom *g, const double *X, int pad)
{
    box b;
    b.x0 = (int)floor(X[0] - 0.5 - g->R) - pad;
    b.x1 = (int)floor(X[0] - 0.5 + g->R) + 1 + pad;
    b.y0 = (int)floor(X[1] - 0.5 - g->R) - pad;
    b.y1 = (int)floor(X[1] - 0.5 + g->R) + 1 + pad;
    b.z0 = (int)floor(X[2] - 0.5 - g->R) - pad;
    b.z1 = (int)floor(X[2] - 0.5 + g->R) + 1 + pad;
    if (b.x0 < 0)
        b.x0 = 0;
    if (b.x1 > g->lx - 1)
        b.x1 = g->lx - 1;
    if (b.y1 - b.y0 + 1 >= g->ly) {
        b.y0 = 0;
        b.y1 = g->ly - 1;
    }
    if (b.z1 - b.z0 + 1 >= g->lz) {
        b.z0 = 0;
        b.z1 = g->lz - 1;
    }
    return b;
}

int orc_p_classify(const orc_pgeom *g, orc_pstate *s, int first)
{
    int overflow = 0;
    /* uncover: nodes of the previous box still marked p but no longer inside */
    if (!first) {
        for (int p = 0; p < g->np; p++) {
            const double *Xo = s->Xold + 3 * p, *X = s->X + 3 * p;
            box b = sphere_box(g, Xo, 0);
            int cnt = 0;
            for (int iz0 = b.z0; iz0 <= b.z1; iz0++)
                for (int iy0 = b.y0; iy0 <= b.y1; iy0++)
                    for (int ix = b.x0; ix <= b.x1; ix++) {
                        int iy = wrapi(iy0, g->ly), iz = wrapi(iz0, g->lz);
                        size_t n = nodeof(g, ix, iy, iz);
                        if (s->solid[n] == p && !inside(g, ix, iy, iz, X)) {
                            s->solid[n] = -2 - p;
                            if (cnt < g->maxfill)
                                s->fill_node[(size_t)p * g->maxfill + cnt] = (int)n;
                            else
                                overflow = 1;
                            cnt++;
                        }
                    }
            s->fill_count[p] = cnt < g->maxfill ? cnt : g->maxfill;
        }
    }
    /* mark: smallest particle id wins an overlap */
    for (int p = 0; p < g->np; p++) {
        const double *X = s->X + 3 * p;
        box b = sphere_box(g, X, 0);
        for (int iz0 = b.z0; iz0 <= b.z1; iz0++)
            for (int iy0 = b.y0; iy0 <= b.y1; iy0++)
                for (int ix = b.x0; ix <= b.x1; ix++) {
                    int iy = wrapi(iy0, g->ly), iz = wrapi(iz0, g->lz);
                    if (inside(g, ix, iy, iz, X)) {
                        size_t n = nodeof(g, ix, iy, iz);
                        if (s->solid[n] < 0 || s->solid[n] > p)
                            s->solid[n] = p;
                    }
                }
    }
    /* links in canonical order: box node (z, y, x), then direction */
    for (int p = 0; p < g->np; p++) {
        const double *X = s->X + 3 * p;
        box b = sphere_box(g, X, 1);
        int k = 0;
        for (int iz0 = b.z0; iz0 <= b.z1; iz0++)
            for (int iy0 = b.y0; iy0 <= b.y1; iy0++)
                for (int ix = b.x0; ix <= b.x1; ix++) {
                    int iy = wrapi(iy0, g->ly), iz = wrapi(iz0, g->lz);
                    size_t n = nodeof(g, ix, iy, iz);
                    if (s->solid[n] >= 0)
                        continue;
                    double dx, dy, dz;
                    offset(g, ix, iy, iz, X, &dx, &dy, &dz);
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    const double cc = d2 - g->R2;
                    for (int i = 1; i < NPOP; i++) {
                        int jx = ix + cix[i];
                        if (jx < 0 || jx >= g->lx)
                            continue;
                        size_t m = nodeof(g, jx, wrapi(iy + ciy[i], g->ly), wrapi(iz + ciz[i], g->lz));
                        if (s->solid[m] != p)
                            continue;
                        const double cx = (double)cix[i], cy = (double)ciy[i], cz = (double)ciz[i];
                        const double a = cx * cx + cy * cy + cz * cz;
                        const double bb = cx * dx + cy * dy + cz * dz;
                        const double disc = bb * bb - a * cc;
                        double q = (disc >= 0.0) ? (-bb - sqrt(disc)) / a : 0.5;
                        if (q < 0.0)
                            q = 0.0;
                        if (q > 1.0)
                            q = 1.0;
                        if (k < g->maxlink) {
                            size_t sl = (size_t)p * g->maxlink + k;
                            s->link_node[sl] = (int)n;
                            s->link_dir[sl] = i;
                            s->link_q[sl] = q;
                        } else {
                            overflow = 1;
                        }
                        k++;
                    }
                }
        s->link_count[p] = k < g->maxlink ? k : g->maxlink;
    }
    return overflow;
}

void orc_p_ibb(const orc_pgeom *g, orc_pstate *s, double *L)
{
    const size_t N = (size_t)g->lx * g->ly * g->lz;
    for (int p = 0; p < g->np; p++) {
        const double *X = s->X + 3 * p, *V = s->V + 3 * p, *W = s->W + 3 * p;
        double part[SUMW][6];
        memset(part, 0, sizeof(part));
        for (int k = 0; k < s->link_count[p]; k++) {
            size_t sl = (size_t)p * g->maxlink + k;
            size_t n = (size_t)s->link_node[sl];
            int i = s->link_dir[sl], ib = ipopp[i];
            double q = s->link_q[sl];
            int ix = (int)(n % g->lx), iy = (int)((n / g->lx) % g->ly), iz = (int)(n / ((size_t)g->lx * g->ly));
            const double cx = (double)cix[i], cy = (double)ciy[i], cz = (double)ciz[i];
            const double w = i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0;
            double dx, dy, dz;
            offset(g, ix, iy, iz, X, &dx, &dy, &dz);
            const double rx = dx + q * cx, ry = dy + q * cy, rz = dz + q * cz;
            const double uwx = V[0] + (W[1] * rz - W[2] * ry);
            const double uwy = V[1] + (W[2] * rx - W[0] * rz);
            const double uwz = V[2] + (W[0] * ry - W[1] * rx);
            const double cu = cx * uwx + cy * uwy + cz * uwz;
            const double fi = L[(size_t)i * N + n];
            double gval;
            if (q < 0.5) {
                int bx = ix - cix[i];
                int ok = (bx >= 0 && bx < g->lx);
                size_t nb = 0;
                if (ok) {
                    nb = nodeof(g, bx, wrapi(iy - ciy[i], g->ly), wrapi(iz - ciz[i], g->lz));
                    ok = s->solid[nb] < 0;
                }
                if (ok) {
                    const double fb = L[(size_t)i * N + nb];
                    gval = (2.0 * q) * fi + (1.0 - 2.0 * q) * fb - 6.0 * w * cu;
                } else {
                    gval = fi - 6.0 * w * cu;
                }
            } else {
                const double fib = L[(size_t)ib * N + n];
                gval = (1.0 / (2.0 * q)) * fi + ((2.0 * q - 1.0) / (2.0 * q)) * fib - (3.0 / q) * w * cu;
            }
            size_t ns = nodeof(g, ix + cix[i], wrapi(iy + ciy[i], g->ly), wrapi(iz + ciz[i], g->lz));
            L[(size_t)ib * N + ns] = gval;
            const double ff = fi + w, gf = gval + w;
            const double px = ff * (cx - uwx) + gf * (cx + uwx);
            const double py = ff * (cy - uwy) + gf * (cy + uwy);
            const double pz = ff * (cz - uwz) + gf * (cz + uwz);
            const double c[6] = {px, py, pz, ry * pz - rz * py, rz * px - rx * pz, rx * py - ry * px};
            for (int j = 0; j < 6; j++)
                part[k % SUMW][j] = part[k % SUMW][j] + c[j];
        }
        lane_tree(part, 6);
        for (int j = 0; j < 3; j++) {
            s->FH[3 * p + j] = part[0][j];
            s->TQ[3 * p + j] = part[0][3 + j];
        }
    }
}

void orc_p_step(const orc_params *pp, const orc_pgeom *g, const int *solid, const double *Lin, double *Lout,
                double fx, double fy, double fz, double rho_off)
{
    const int lx = g->lx, ly = g->ly, lz = g->lz;
    const size_t N = (size_t)lx * ly * lz;
    for (int iz = 0; iz < lz; iz++)
        for (int iy = 0; iy < ly; iy++)
            for (int ix = 0; ix < lx; ix++) {
                size_t n = nodeof(g, ix, iy, iz);
                if (solid && solid[n] >= 0)
                    continue;
                double f[NPOP], post[NPOP];
                for (int i = 0; i < NPOP; i++) {
                    int sx = ix - cix[i];
                    if (sx < 0 || sx >= lx)
                        f[i] = Lin[(size_t)ipopp[i] * N + n];
                    else
                        f[i] = Lin[(size_t)i * N + nodeof(g, sx, wrapi(iy - ciy[i], ly), wrapi(iz - ciz[i], lz))];
                }
                /* macrovar, Channel-Flow/collision.f90:398-417 */
                double sum1 = f[7] - f[10], sum2 = f[9] - f[8], sum3 = f[11] - f[14], sum4 = f[13] - f[12];
                double sum5 = f[15] - f[18], sum6 = f[17] - f[16];
                double ux9 = f[1] - f[2] + sum1 + sum2 + sum3 + sum4;
                double uy9 = f[3] - f[4] + sum1 - sum2 + sum5 + sum6;
                double uz9 = f[5] - f[6] + sum3 - sum4 + sum5 - sum6;
                double rho9 = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
                              f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
                ux9 = ux9 + fx / 2.;
                uy9 = uy9 + fy / 2.;
                uz9 = uz9 + fz / 2.;
                rho9 = rho9 - rho_off; /* avedensity, Channel-Flow/collision.f90:508 */
                orc_collide_node(pp, f, rho9, ux9, uy9, uz9, fx, fy, fz, post);
                for (int i = 0; i < NPOP; i++)
                    Lout[(size_t)i * N + n] = post[i];
            }
}

/* Contact force of Feng & Michaelides (2005) eq. 28 on a sphere whose centre
 * lies at distance d from the other sphere's (or its mirror image's) centre:
 * with the gap h = d - 2R and the range zeta,
 *   h > zeta:       0
 *   0 <= h <= zeta: stf0 ((zeta - h)/zeta)^2
 *   h < 0:          stf0 ((zeta - h)/zeta)^2 + stf1 (-h/zeta)
 * along the line of centres, away from the other centre.  stf0 and stf1 are
 * the force scales c/eps_P and c/E_P of that equation (Channel-Flow/
 * para.f90:356-360 names them stf0, stf1 and, for walls, stf0_w, stf1_w). */
static double contact_mag(double h, double zeta, double stf0, double stf1)
{
    if (!(h < zeta))
        return 0.0;
    const double r = (zeta - h) / zeta;
    double m = stf0 * (r * r);
    if (h < 0.0)
        m = m + stf1 * ((-h) / zeta);
    return m;
}

/* Normal lubrication correction of Nguyen & Ladd (2002) for gaps h below the
 * cut-off hc: -6 pi mu a^2 (1/h - 1/hc) U_n, with a = R1 R2/(R1 + R2) (R/2
 * for two equal spheres, R for a sphere and a plane wall), U_n the approach
 * velocity along the unit normal, mu = rho0 nu (rho0 = 1) and h no smaller
 * than hmin.  Returns the coefficient of U_n (zero beyond hc or without
 * viscosity). */
static double lub_coef(double h, double a, const orc_pgeom *g)
{
    if (!(g->visc > 0.0) || !(h < g->lub_cut))
        return 0.0;
    const double he = h > g->lub_min ? h : g->lub_min;
    return -6.0 * ORC_PI * g->visc * (a * a) * (1.0 / he - 1.0 / g->lub_cut);
}

void orc_p_move(const orc_pgeom *g, orc_pstate *s)
{
    const double Ly = (double)g->ly, Lz = (double)g->lz, Lx = (double)g->lx;
    /* every particle's forces use the positions and velocities before the
     * move: the new ones are committed after the loop */
    double *Xn = malloc(9 * (size_t)g->np * sizeof(double));
    double *Vn = Xn + 3 * (size_t)g->np, *Wn = Xn + 6 * (size_t)g->np;
    const int first = (s->nsteps_done == 0);
    for (int p = 0; p < g->np; p++) {
        const double *Xp = s->X + 3 * p, *Vp = s->V + 3 * p;
        double part[SUMW][6];
        memset(part, 0, sizeof(part));
        for (int q = 0; q < g->np; q++) {
            if (q == p)
                continue;
            const double *Xq = s->X + 3 * q, *Vq = s->V + 3 * q;
            const double ddx = Xp[0] - Xq[0];
            const double ddy = min_image(Xp[1] - Xq[1], Ly);
            const double ddz = min_image(Xp[2] - Xq[2], Lz);
            const double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            if (!(dist > 0.0))
                continue;
            const double h = dist - 2.0 * g->R;
            const double nx = ddx / dist, ny = ddy / dist, nz = ddz / dist;
            const double mag = contact_mag(h, g->mingap, g->stf0, g->stf1);
            const double un = (Vp[0] - Vq[0]) * nx + (Vp[1] - Vq[1]) * ny + (Vp[2] - Vq[2]) * nz;
            const double lub = lub_coef(h, 0.5 * g->R, g) * un;
            double *pc = part[q % SUMW];
            pc[0] = pc[0] + mag * nx;
            pc[1] = pc[1] + mag * ny;
            pc[2] = pc[2] + mag * nz;
            pc[3] = pc[3] + lub * nx;
            pc[4] = pc[4] + lub * ny;
            pc[5] = pc[5] + lub * nz;
        }
        lane_tree(part, 6);
        double fc[3] = {part[0][0], part[0][1], part[0][2]};
        double fl[3] = {part[0][3], part[0][4], part[0][5]};
        /* walls at x = 0 and x = lx: the mirror image lies 2 (distance) away */
        const double dlo = Xp[0], dhi = Lx - Xp[0];
        fc[0] = fc[0] + contact_mag(2.0 * dlo - 2.0 * g->R, g->mingap_w, g->stf0_w, g->stf1_w);
        fc[0] = fc[0] - contact_mag(2.0 * dhi - 2.0 * g->R, g->mingap_w, g->stf0_w, g->stf1_w);
        fl[0] = fl[0] + lub_coef(dlo - g->R, g->R, g) * Vp[0];
        fl[0] = fl[0] + lub_coef(dhi - g->R, g->R, g) * Vp[0];
        const double fb[3] = {g->fx, g->fy, g->fz};
        for (int j = 0; j < 3; j++) {
            const int k = 3 * p + j;
            s->FC[k] = fc[j];
            s->FL[k] = fl[j];
            const double fh = first ? s->FH[k] : 0.5 * (s->FH[k] + s->FHprev[k]);
            const double tq = first ? s->TQ[k] : 0.5 * (s->TQ[k] + s->TQprev[k]);
            s->FHprev[k] = s->FH[k];
            s->TQprev[k] = s->TQ[k];
            double ftot = fh + fc[j] + fl[j];
            if (g->pressure_force)
                ftot = ftot + g->volume * fb[j];
            const double a = ftot / g->mass;
            const double al = tq / g->inertia;
            const double v0 = s->V[k], w0 = s->W[k];
            const double vn = first ? v0 + a : v0 + (1.5 * a - 0.5 * s->Aprev[k]);
            const double wn = first ? w0 + al : w0 + (1.5 * al - 0.5 * s->ALprev[k]);
            Xn[k] = Xp[j] + 0.5 * (v0 + vn);
            s->TH[k] = s->TH[k] + 0.5 * (w0 + wn);
            s->Aprev[k] = a;
            s->ALprev[k] = al;
            Vn[k] = vn;
            Wn[k] = wn;
        }
        if (Xn[3 * p + 1] >= Ly)
            Xn[3 * p + 1] = Xn[3 * p + 1] - Ly;
        else if (Xn[3 * p + 1] < 0.0)
            Xn[3 * p + 1] = Xn[3 * p + 1] + Ly;
        if (Xn[3 * p + 2] >= Lz)
            Xn[3 * p + 2] = Xn[3 * p + 2] - Lz;
        else if (Xn[3 * p + 2] < 0.0)
            Xn[3 * p + 2] = Xn[3 * p + 2] + Lz;
    }
    memcpy(s->Xold, s->X, 3 * (size_t)g->np * sizeof(double));
    memcpy(s->X, Xn, 3 * (size_t)g->np * sizeof(double));
    memcpy(s->V, Vn, 3 * (size_t)g->np * sizeof(double));
    memcpy(s->W, Wn, 3 * (size_t)g->np * sizeof(double));
    s->nsteps_done++;
    free(Xn);
}

void orc_p_refill(const orc_pgeom *g, orc_pstate *s, double *L)
{
    const size_t N = (size_t)g->lx * g->ly * g->lz;
    for (int p = 0; p < g->np; p++) {
        const double *X = s->X + 3 * p, *V = s->V + 3 * p, *W = s->W + 3 * p;
        for (int k = 0; k < s->fill_count[p]; k++) {
            size_t n = (size_t)s->fill_node[(size_t)p * g->maxfill + k];
            if (s->solid[n] != -2 - p)
                continue;
            int ix = (int)(n % g->lx), iy = (int)((n / g->lx) % g->ly), iz = (int)(n / ((size_t)g->lx * g->ly));
            double rsum = 0.0;
            int cnt = 0;
            for (int j = 1; j < NPOP; j++) {
                int jx = ix + cix[j];
                if (jx < 0 || jx >= g->lx)
                    continue;
                size_t m = nodeof(g, jx, wrapi(iy + ciy[j], g->ly), wrapi(iz + ciz[j], g->lz));
                if (s->solid[m] != -1)
                    continue;
                double r = L[m];
                for (int i = 1; i < NPOP; i++)
                    r = r + L[(size_t)i * N + m];
                rsum = rsum + r;
                cnt++;
            }
            const double rho = cnt > 0 ? rsum / (double)cnt : 0.0;
            double dx, dy, dz;
            offset(g, ix, iy, iz, X, &dx, &dy, &dz);
            const double ux = V[0] + (W[1] * dz - W[2] * dy);
            const double uy = V[1] + (W[2] * dx - W[0] * dz);
            const double uz = V[2] + (W[0] * dy - W[1] * dx);
            double usqr = ux * ux + uy * uy + uz * uz;
            usqr = 1.5 * usqr;
            L[n] = (1.0 / 3.0) * (rho - usqr);
            for (int i = 1; i < NPOP; i++) {
                const double G = (double)cix[i] * ux + (double)ciy[i] * uy + (double)ciz[i] * uz;
                const double w = i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0;
                L[(size_t)i * N + n] = w * (rho + 3.0 * G + 4.5 * G * G - usqr);
            }
        }
    }
    /* uncovered nodes settle after every refill has read its neighbours */
    for (int p = 0; p < g->np; p++)
        for (int k = 0; k < s->fill_count[p]; k++) {
            size_t n = (size_t)s->fill_node[(size_t)p * g->maxfill + k];
            if (s->solid[n] == -2 - p)
                s->solid[n] = -1;
        }
}

/* Sum of v[i] over i < n with mask[i] set (every i without a mask) in the
 * device's order: 256-thread blocks, grid = min(2048, ceil(n/256)); thread t
 * of block b adds i = 256 b + t, 256 (b + grid) + t, ... in increasing order;
 * each 64-lane wave adds its lanes by the butterfly v_l += v_(l xor o) for
 * o = 32, 16, ..., 1 and takes lane 0; the block adds its four wave sums in
 * order; the block sums are added in increasing b from 0.0. */
double orc_device_sum(const double *v, long long n, const unsigned char *mask)
{
    const long long B = 256;
    long long grid = (n + B - 1) / B;
    if (grid > 2048)
        grid = 2048;
    double total = 0.0;
    for (long long b = 0; b < grid; b++) {
        double lane[256];
        for (int t = 0; t < 256; t++) {
            double acc = 0.0;
            for (long long i = b * B + t; i < n; i += grid * B)
                if (!mask || mask[i])
                    acc += v[i];
            lane[t] = acc;
        }
        double ws[4];
        for (int w = 0; w < 4; w++) {
            double x[64], y[64];
            for (int l = 0; l < 64; l++)
                x[l] = lane[64 * w + l];
            for (int o = 32; o > 0; o >>= 1) {
                for (int l = 0; l < 64; l++)
                    y[l] = x[l] + x[l ^ o];
                memcpy(x, y, sizeof(x));
            }
            ws[w] = x[0];
        }
        double sb = ws[0];
        for (int w = 1; w < 4; w++)
            sb += ws[w];
        total += sb;
    }
    return total;
}

/* macrovar's density (Channel-Flow/collision.f90:411-413) of the pulled
 * populations at fluid node (ix, iy, iz) of post-collision lattice L. */
static void pull_node(const orc_pgeom *g, const double *L, int ix, int iy, int iz, double *f)
{
    const size_t N = (size_t)g->lx * g->ly * g->lz;
    const size_t n = nodeof(g, ix, iy, iz);
    for (int i = 0; i < NPOP; i++) {
        int sx = ix - cix[i];
        if (sx < 0 || sx >= g->lx)
            f[i] = L[(size_t)ipopp[i] * N + n];
        else
            f[i] = L[(size_t)i * N + nodeof(g, sx, wrapi(iy - ciy[i], g->ly), wrapi(iz - ciz[i], g->lz))];
    }
}

static double rho_of(const double *f)
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
           f[14] + f[15] + f[16] + f[17] + f[18];
}

/* avedensity's mean (Channel-Flow/collision.f90:487-506) over the fluid nodes
 * (solid < 0) of the populations pulled from L, in the device's order. */
double orc_p_rho_mean(const orc_pgeom *g, const int *solid, const double *L)
{
    const size_t N = (size_t)g->lx * g->ly * g->lz;
    double *r = calloc(N, sizeof(double));
    unsigned char *m = calloc(N, 1);
    long long cnt = 0;
    for (int iz = 0; iz < g->lz; iz++)
        for (int iy = 0; iy < g->ly; iy++)
            for (int ix = 0; ix < g->lx; ix++) {
                const size_t n = nodeof(g, ix, iy, iz);
                m[n] = !solid || solid[n] < 0;
                r[n] = 0.0;
                if (m[n]) {
                    double f[NPOP];
                    pull_node(g, L, ix, iy, iz, f);
                    r[n] = rho_of(f);
                    cnt++;
                }
            }
    const double sum = orc_device_sum(r, (long long)N, m);
    free(r);
    free(m);
    return cnt > 0 ? sum / (double)cnt : 0.0;
}

/* macrovar (Channel-Flow/collision.f90:378-463) of the particle-laden
 * channel: fluid nodes from the populations pulled from L with u + F/2 and
 * rho - rho_off; nodes inside particle p get the rigid-body velocity
 * V + W x r at the node (r from the centre, minimum image in y and z,
 * :424-455) and rho = rhopart (:457), less rho_off as avedensity subtracts
 * its mean from every node (:508). */
void orc_p_macrovar(const orc_pgeom *g, const orc_pstate *s, const double *L, double fx, double fy, double fz,
                    double rho_off, double *rho, double *ux, double *uy, double *uz)
{
    for (int iz = 0; iz < g->lz; iz++)
        for (int iy = 0; iy < g->ly; iy++)
            for (int ix = 0; ix < g->lx; ix++) {
                const size_t n = nodeof(g, ix, iy, iz);
                const int p = s->solid[n];
                if (p < 0) {
                    double f[NPOP];
                    pull_node(g, L, ix, iy, iz, f);
                    double sum1 = f[7] - f[10], sum2 = f[9] - f[8], sum3 = f[11] - f[14], sum4 = f[13] - f[12];
                    double sum5 = f[15] - f[18], sum6 = f[17] - f[16];
                    double ux9 = f[1] - f[2] + sum1 + sum2 + sum3 + sum4;
                    double uy9 = f[3] - f[4] + sum1 - sum2 + sum5 + sum6;
                    double uz9 = f[5] - f[6] + sum3 - sum4 + sum5 - sum6;
                    ux[n] = ux9 + fx / 2.;
                    uy[n] = uy9 + fy / 2.;
                    uz[n] = uz9 + fz / 2.;
                    rho[n] = rho_of(f) - rho_off;
                } else {
                    const double *X = s->X + 3 * p, *V = s->V + 3 * p, *W = s->W + 3 * p;
                    double dx, dy, dz;
                    offset(g, ix, iy, iz, X, &dx, &dy, &dz);
                    ux[n] = V[0] + (W[1] * dz - W[2] * dy);
                    uy[n] = V[1] + (W[2] * dx - W[0] * dz);
                    uz[n] = V[2] + (W[0] * dy - W[1] * dx);
                    rho[n] = g->rhopart - rho_off;
                }
            }
}
