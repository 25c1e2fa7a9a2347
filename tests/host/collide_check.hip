// Host-side check of the node collision arithmetic of d3q19_common.h against
// the CPU oracle's restatement of Channel-Flow/collision.f90:56-220, on random
// near-equilibrium states with all three rate sets of para.f90:110-141:
//
//   1. mrt_collide<double> (the FP64 kernel's arithmetic) equals
//      orc_collide_node bit for bit;
//   2. mrt_collide_delta<double> (the increment form) agrees with it to a few
//      ulps, for fields fused from the same f (zero conserved changes) and
//      for unrelated field values (the array mode);
//   3. mrt_collide_delta<float> stays within FP32 round-off of the FP64
//      result, and its momentum increment is F up to the rounding of the
//      stored populations;
//   4. the closed-form moments of the Guo term that the increment form uses
//      (guo_moments) equal the moments of the 19 Guo terms of the reference
//      (guo_terms, then moments()), with sum Fbar = 0, sum c Fbar = F and
//      vanishing third-order moments.  Dropping a term or a factor of the
//      closed form, or swapping a velocity and a force component, fails it.
//
// Built with hipcc -ffp-contract=off and run on the CPU; launches nothing on a
// GPU.  Prints one "ok" line per check, exits non-zero on a failure.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "d3q19_common.h"
#include "lbm_oracle.h"

using namespace d3q19;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;

static double uniform01()
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static double sym(double a) { return a * (2.0 * uniform01() - 1.0); }

template <typename T>
static Rates<T> rates_of(const orc_params &p, const double F[3])
{
    Rates<T> r;
    r.s1 = (T)p.s1;
    r.s2 = (T)p.s2;
    r.s4 = (T)p.s4;
    r.s9 = (T)p.s9;
    r.s10 = (T)p.s10;
    r.s13 = (T)p.s13;
    r.s16 = (T)p.s16;
    r.omegepsl = (T)p.omegepsl;
    r.omegepslj = (T)p.omegepslj;
    r.omegxx = (T)p.omegxx;
    r.fx = (T)F[0];
    r.fy = (T)F[1];
    r.fz = (T)F[2];
    return r;
}

// A shifted population set near the equilibrium of (rho, u) plus noise.
static void random_state(double f[19], double u[3], double F[3], double &rho)
{
    static const double W[3] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0};
    rho = sym(1e-3);
    for (int a = 0; a < 3; a++)
        u[a] = sym(0.12);
    for (int a = 0; a < 3; a++)
        F[a] = sym(1e-6);
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int i = 0; i < 19; i++) {
        const int c2 = CX[i] * CX[i] + CY[i] * CY[i] + CZ[i] * CZ[i];
        const double w = W[c2];
        const double cu = CX[i] * u[0] + CY[i] * u[1] + CZ[i] * u[2];
        f[i] = w * (rho + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu) + sym(1e-4) * w;
    }
}

int main()
{
    const int trials = 20000;
    long mismatches = 0;
    double worst_delta64 = 0.0, worst_delta64_arrays = 0.0, worst_delta32 = 0.0, worst_mom32 = 0.0;
    for (int mrt = 1; mrt <= 3; mrt++) {
        orc_params p;
        orc_para(&p, 32, 64, 32, 0, mrt);
        for (int t = 0; t < trials; t++) {
            double fin[19], u[3], F[3], rho;
            random_state(fin, u, F, rho);

            // 1. arbitrary fields: reference form vs the oracle, bit for bit
            double ref[19];
            orc_collide_node(&p, fin, rho, u[0], u[1], u[2], F[0], F[1], F[2], ref);
            double f[19];
            memcpy(f, fin, sizeof f);
            const Rates<double> r64 = rates_of<double>(p, F);
            mrt_collide(f, rho, u[0], u[1], u[2], r64);
            if (memcmp(f, ref, sizeof f) != 0)
                mismatches++;

            // 2a. increment form with the same arbitrary fields (array mode)
            double rf, jx, jy, jz;
            macro_sums(fin, rf, jx, jy, jz);
            double g[19];
            memcpy(g, fin, sizeof g);
            mrt_collide_delta(g, rho, u[0], u[1], u[2], rho - rf, u[0] - (jx + F[0] / 2.), u[1] - (jy + F[1] / 2.),
                              u[2] - (jz + F[2] / 2.), r64);
            for (int i = 0; i < 19; i++)
                worst_delta64_arrays = fmax(worst_delta64_arrays, fabs(g[i] - ref[i]));

            // 2b. fields fused from f: reference form vs increment form with
            // zero conserved changes
            const double uf[3] = {jx + F[0] / 2., jy + F[1] / 2., jz + F[2] / 2.};
            double ref_fused[19];
            orc_collide_node(&p, fin, rf, uf[0], uf[1], uf[2], F[0], F[1], F[2], ref_fused);
            memcpy(g, fin, sizeof g);
            mrt_collide_delta(g, rf, uf[0], uf[1], uf[2], 0.0, 0.0, 0.0, 0.0, r64);
            for (int i = 0; i < 19; i++)
                worst_delta64 = fmax(worst_delta64, fabs(g[i] - ref_fused[i]));

            // 3. FP32 increment form from the FP32-rounded populations
            float h[19];
            for (int i = 0; i < 19; i++)
                h[i] = (float)fin[i];
            float rf32, jx32, jy32, jz32;
            macro_sums(h, rf32, jx32, jy32, jz32);
            const Rates<float> r32 = rates_of<float>(p, F);
            const float ux32 = jx32 + r32.fx / 2.f, uy32 = jy32 + r32.fy / 2.f, uz32 = jz32 + r32.fz / 2.f;
            float h0[19];
            memcpy(h0, h, sizeof h);
            mrt_collide_delta(h, rf32, ux32, uy32, uz32, 0.f, 0.f, 0.f, 0.f, r32);
            for (int i = 0; i < 19; i++)
                worst_delta32 = fmax(worst_delta32, fabs((double)h[i] - ref_fused[i]));
            // momentum increment (exact sums in FP64 of the FP32 values) vs F
            double dj[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < 19; i++) {
                const double d = (double)h[i] - (double)h0[i];
                dj[0] += CX[i] * d;
                dj[1] += CY[i] * d;
                dj[2] += CZ[i] * d;
            }
            for (int a = 0; a < 3; a++)
                worst_mom32 = fmax(worst_mom32, fabs(dj[a] - (double)((&r32.fx)[a])));
        }
    }
    // 4. closed-form Guo moments vs the moments of the 19 Guo terms
    double worst_guo = 0.0;
    for (int t = 0; t < trials; t++) {
        double u[3], F[3];
        for (int a = 0; a < 3; a++) {
            u[a] = sym(0.15);
            F[a] = sym(1e-3);
        }
        double Fb[19], m[15], sf[12];
        guo_terms(u[0], u[1], u[2], F[0], F[1], F[2], Fb);
        moments(Fb, m);
        guo_moments(u[0], u[1], u[2], F[0], F[1], F[2], sf);
        double r, jx, jy, jz;
        macro_sums(Fb, r, jx, jy, jz);
        for (int k = 0; k < 12; k++)
            worst_guo = fmax(worst_guo, fabs(m[k] - sf[k]));
        for (int k = 12; k < 15; k++)
            worst_guo = fmax(worst_guo, fabs(m[k]));
        worst_guo = fmax(worst_guo, fmax(fabs(r), fmax(fabs(jx - F[0]), fmax(fabs(jy - F[1]), fabs(jz - F[2])))));
    }
    int fail = 0;
    // the moments are of size |u||F| ~ 1e-4: a dropped term is >= 1e-6
    printf("Guo-term moments, closed form vs moments of the 19 terms: max |diff| %.3e\n", worst_guo);
    if (worst_guo > 1e-17)
        fail = 1;
    else
        printf("ok guo_moments\n");
    printf("reference form vs oracle: %ld of %d states differ\n", mismatches, 3 * trials);
    if (mismatches != 0)
        fail = 1;
    else
        printf("ok bitwise\n");
    printf("increment form FP64: max |diff| %.3e fused fields, %.3e array fields\n", worst_delta64,
           worst_delta64_arrays);
    if (worst_delta64 > 1e-16 || worst_delta64_arrays > 1e-16)
        fail = 1;
    else
        printf("ok increment64\n");
    // populations are below ~0.1: FP32 round-off of one value is ~4e-9
    printf("increment form FP32: max |diff| %.3e, momentum increment error %.3e\n", worst_delta32, worst_mom32);
    if (worst_delta32 > 5e-8 || worst_mom32 > 1e-8)
        fail = 1;
    else
        printf("ok increment32\n");
    return fail;
}
