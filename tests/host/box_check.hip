// Host check of sphere_box (particles.h): for centres anywhere -- inside the
// channel, far outside it, huge, infinite, NaN -- every node the particle
// kernels would enumerate from the box lies inside the lattice, and for
// centres inside the channel the box covers every node within R of the
// centre (minimum image in y, z), for pads 0 and 1.  Built with hipcc and run
// on the CPU; launches nothing on a GPU.
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "particles.h"

using namespace d3q19;

static int wrap_i(int i, int L)
{
    int r = i % L;
    return r < 0 ? r + L : r;
}

static bool check_box(const PartGeom &pg, const double *X, int pad)
{
    const Box b = sphere_box(pg, X, pad);
    const long long bx = b.x1 - b.x0 + 1, by = b.y1 - b.y0 + 1, bz = b.z1 - b.z0 + 1;
    if (bx <= 0)
        return true; // the kernels enumerate bx*by*bz nodes: none
    if (by <= 0 || bz <= 0 || by > (long long)pg.g.ly || bz > (long long)pg.g.lz) {
        fprintf(stderr, "box y/z extent out of range for X = (%g, %g, %g)\n", X[0], X[1], X[2]);
        return false;
    }
    if (b.x0 < 0 || b.x1 >= (int)pg.g.lx || b.y0 < 0 || b.y0 > (int)pg.g.ly || b.z0 < 0 || b.z0 > (int)pg.g.lz) {
        fprintf(stderr, "box start out of range for X = (%g, %g, %g): x [%d, %d] y0 %d z0 %d\n", X[0], X[1], X[2],
                b.x0, b.x1, b.y0, b.z0);
        return false;
    }
    // every enumerated node index inside the lattice (the kernels' arithmetic)
    const long long total = bx * by * bz;
    for (long long e = 0; e < total; e++) {
        const int kx = (int)(e % bx), ky = (int)((e / bx) % by), kz = (int)(e / (bx * by));
        const int ix = b.x0 + kx, iy = wrap_i(b.y0 + ky, pg.g.ly), iz = wrap_i(b.z0 + kz, pg.g.lz);
        const uint64_t n = (uint64_t)ix + (uint64_t)pg.g.lx * ((uint64_t)iy + (uint64_t)pg.g.ly * iz);
        if (ix < 0 || ix >= (int)pg.g.lx || n >= pg.g.N) {
            fprintf(stderr, "node out of range for X = (%g, %g, %g)\n", X[0], X[1], X[2]);
            return false;
        }
    }
    // coverage: for a centre inside the channel, every node within R is in the box
    if (X[0] > 0 && X[0] < pg.g.lx && isfinite(X[1]) && isfinite(X[2]) && X[1] >= 0 && X[1] < pg.g.ly &&
        X[2] >= 0 && X[2] < pg.g.lz) {
        const int xlo = (int)floor(X[0]) - (int)pg.R - 2, xhi = (int)floor(X[0]) + (int)pg.R + 2;
        for (int ix = xlo < 0 ? 0 : xlo; ix <= xhi && ix < (int)pg.g.lx; ix++)
            for (int dy = -(int)pg.R - 2; dy <= (int)pg.R + 2; dy++)
                for (int dz = -(int)pg.R - 2; dz <= (int)pg.R + 2; dz++) {
                    const int iy = wrap_i((int)floor(X[1]) + dy, pg.g.ly), iz = wrap_i((int)floor(X[2]) + dz, pg.g.lz);
                    double ddx = ix + 0.5 - X[0], ddy = iy + 0.5 - X[1], ddz = iz + 0.5 - X[2];
                    if (ddy > 0.5 * pg.g.ly) ddy -= pg.g.ly;
                    if (ddy < -0.5 * pg.g.ly) ddy += pg.g.ly;
                    if (ddz > 0.5 * pg.g.lz) ddz -= pg.g.lz;
                    if (ddz < -0.5 * pg.g.lz) ddz += pg.g.lz;
                    if (ddx * ddx + ddy * ddy + ddz * ddz >= pg.R2)
                        continue;
                    const bool inx = ix >= b.x0 && ix <= b.x1;
                    const bool iny = wrap_i(iy - b.y0, pg.g.ly) < by;
                    const bool inz = wrap_i(iz - b.z0, pg.g.lz) < bz;
                    if (!(inx && iny && inz)) {
                        fprintf(stderr, "node (%d, %d, %d) within R of X = (%g, %g, %g) not in the box\n", ix, iy, iz,
                                X[0], X[1], X[2]);
                        return false;
                    }
                }
    }
    return true;
}

int main()
{
    const uint32_t shapes[][3] = {{192, 768, 384}, {32, 64, 32}, {24, 10, 8}, {16, 12, 12}};
    const double specials[] = {0.0,   0.3,    5.5,     -0.7,   -5.0,  -7.0,  -1e6,  1e6,   3e9,
                               -3e9,  1e12,   -1e12,   1e300,  -1e300, INFINITY, -INFINITY, NAN};
    long long boxes = 0;
    for (auto &s : shapes) {
        PartGeom pg = {};
        pg.g.lx = s[0];
        pg.g.ly = s[1];
        pg.g.lz = s[2];
        pg.g.N = (uint64_t)s[0] * s[1] * s[2];
        pg.R = 5.0;
        pg.R2 = 25.0;
        for (int pad = 0; pad <= 1; pad++) {
            // specials in every coordinate, combined with ordinary values
            for (double a : specials)
                for (double c : specials) {
                    const double cases[][3] = {{a, c, c}, {a, 4.2, 3.7}, {10.5, a, c}, {10.5, 4.2, a}, {a, a, a}};
                    for (auto &X : cases) {
                        boxes++;
                        if (!check_box(pg, X, pad))
                            return 1;
                    }
                }
            // a sweep of ordinary centres across and beyond the channel
            for (double x = -12.0; x <= s[0] + 12.0; x += 0.61)
                for (double y = -20.0; y <= s[1] + 20.0; y += 5.3) {
                    const double X[3] = {x, y, 0.5 * s[2] + 0.123};
                    boxes++;
                    if (!check_box(pg, X, pad))
                        return 1;
                }
        }
    }
    printf("ok boxes %lld\n", boxes);
    return 0;
}
