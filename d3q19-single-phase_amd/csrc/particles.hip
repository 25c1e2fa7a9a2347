// Particle kernels (see particles.h for the methods and their sources).
// Built with -ffp-contract=off: every FP64 operation rounds as written, so
// the CPU oracle's identical expressions give identical bits.
#include <utility>

#include "particles.h"

namespace d3q19 {

namespace {

constexpr int PBLOCK = 256;
constexpr int IBB_BLOCKS = 4; // blocks per particle for the cut links (~1000-2000 each at R = 5)

__host__ __device__ inline int wrap(int i, int L)
{
    int r = i % L;
    return r < 0 ? r + L : r;
}

// Minimum-image offset of a node coordinate from a centre along a periodic axis.
__host__ __device__ inline double min_image(double d, double L)
{
    if (d > 0.5 * L)
        d = d - L;
    else if (d < -0.5 * L)
        d = d + L;
    return d;
}

__host__ __device__ inline void node_offset(const PartGeom &pg, int ix, int iy, int iz, const double *X, double &dx,
                                            double &dy, double &dz)
{
    dx = ((double)ix + 0.5) - X[0];
    dy = min_image(((double)iy + 0.5) - X[1], (double)pg.g.ly);
    dz = min_image(((double)iz + 0.5) - X[2], (double)pg.g.lz);
}

__host__ __device__ inline bool inside(const PartGeom &pg, int ix, int iy, int iz, const double *X)
{
    double dx, dy, dz;
    node_offset(pg, ix, iy, iz, X, dx, dy, dz);
    const double d2 = dx * dx + dy * dy + dz * dz;
    return d2 < pg.R2;
}

__device__ inline uint32_t node_of(const Geom &g, int ix, int iy, int iz)
{
    return (uint32_t)ix + g.lx * ((uint32_t)iy + g.ly * (uint32_t)iz);
}

// Block-wide exclusive scan of one int per thread (PBLOCK threads, thread
// order), and the block total: an inclusive scan within each wave by lane
// shifts, then the wave totals through LDS.  Integer sums, so the result is
// exact and independent of the scheme.
__device__ inline int block_scan(int v, int &total)
{
    constexpr int NW = PBLOCK / 64;
    __shared__ int wsum[NW];
    const int lane = (int)(threadIdx.x & 63u), wid = (int)(threadIdx.x >> 6);
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, (unsigned)o, 64);
        if (lane >= o)
            x += t;
    }
    if (lane == 63)
        wsum[wid] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const int sw = wsum[w];
        off += w < wid ? sw : 0;
        tot += sw;
    }
    __syncthreads(); // wsum is rewritten by the next call
    total = tot;
    return off + x - v;
}

// Phase 1: nodes of particle p's previous box still marked p but no longer
// inside become -2-p and join p's refill list (canonical box order).
__global__ void __launch_bounds__(PBLOCK) k_p_uncover(PartGeom pg, PartState ps, const double *Xold)
{
    const int p = blockIdx.x;
    const Box b = sphere_box(pg, Xold + 3 * p, 0);
    const int bx = b.x1 - b.x0 + 1, by = b.y1 - b.y0 + 1, bz = b.z1 - b.z0 + 1;
    const int total = bx * by * bz;
    const double *X = ps.X + 3 * p;
    int base = 0;
    for (int c0 = 0; c0 < total; c0 += PBLOCK) {
        const int e = c0 + threadIdx.x;
        int flag = 0;
        uint32_t n = 0;
        if (e < total) {
            const int kx = e % bx, ky = (e / bx) % by, kz = e / (bx * by);
            const int ix = b.x0 + kx, iy = wrap(b.y0 + ky, pg.g.ly), iz = wrap(b.z0 + kz, pg.g.lz);
            n = node_of(pg.g, ix, iy, iz);
            if (ps.solid[n] == p && !inside(pg, ix, iy, iz, X))
                flag = 1;
        }
        int tot;
        const int off = block_scan(flag, tot);
        if (flag) {
            ps.solid[n] = -2 - p;
            if (base + off < pg.maxfill)
                ps.fill_node[(size_t)p * pg.maxfill + base + off] = (int)n;
            else
                atomicOr(ps.overflow, 1);
        }
        base += tot;
    }
    if (threadIdx.x == 0)
        ps.fill_count[p] = base < pg.maxfill ? base : pg.maxfill;
}

// Phase 2: claim the nodes inside particle p (smallest id wins an overlap).
__global__ void __launch_bounds__(PBLOCK) k_p_mark(PartGeom pg, PartState ps)
{
    const int p = blockIdx.x;
    const double *X = ps.X + 3 * p;
    const Box b = sphere_box(pg, X, 0);
    const int bx = b.x1 - b.x0 + 1, by = b.y1 - b.y0 + 1, bz = b.z1 - b.z0 + 1;
    const int total = bx * by * bz;
    for (int e = threadIdx.x; e < total; e += PBLOCK) {
        const int kx = e % bx, ky = (e / bx) % by, kz = e / (bx * by);
        const int ix = b.x0 + kx, iy = wrap(b.y0 + ky, pg.g.ly), iz = wrap(b.z0 + kz, pg.g.lz);
        if (inside(pg, ix, iy, iz, X)) {
            // negative (fluid) values compare above every id as unsigned
            atomicMin((unsigned int *)&ps.solid[node_of(pg.g, ix, iy, iz)], (unsigned int)p);
        }
    }
}

// Phase 3: cut links of particle p: fluid node x, direction i with x + c_i
// inside p.  Canonical order: box node (z, y, x), then direction 1..18.
__global__ void __launch_bounds__(PBLOCK) k_p_links(PartGeom pg, PartState ps)
{
    const int p = blockIdx.x;
    const double *X = ps.X + 3 * p;
    const Box b = sphere_box(pg, X, 1);
    const int bx = b.x1 - b.x0 + 1, by = b.y1 - b.y0 + 1, bz = b.z1 - b.z0 + 1;
    const int total = bx * by * bz;
    int base = 0;
    for (int c0 = 0; c0 < total; c0 += PBLOCK) {
        const int e = c0 + threadIdx.x;
        uint32_t mask = 0;
        int ix = 0, iy = 0, iz = 0;
        uint32_t n = 0;
        if (e < total) {
            const int kx = e % bx, ky = (e / bx) % by, kz = e / (bx * by);
            ix = b.x0 + kx;
            iy = wrap(b.y0 + ky, pg.g.ly);
            iz = wrap(b.z0 + kz, pg.g.lz);
            n = node_of(pg.g, ix, iy, iz);
            // a link is at most sqrt(2) long: nodes farther than R + 1.5 from
            // the centre cannot reach the sphere
            double ox, oy, oz;
            node_offset(pg, ix, iy, iz, X, ox, oy, oz);
            const double reach = pg.R + 1.5;
            if (ox * ox + oy * oy + oz * oz < reach * reach) {
                // the node's flag and its 18 neighbours' load together; a
                // neighbour beyond a wall reads the node itself (in bounds)
                // and is masked out
                const int ym = wrap(iy - 1, pg.g.ly), yp = wrap(iy + 1, pg.g.ly);
                const int zm = wrap(iz - 1, pg.g.lz), zp = wrap(iz + 1, pg.g.lz);
                const int own = ps.solid[n];
                int sv[19];
#pragma unroll
                for (int i = 1; i < 19; i++) {
                    const int jx = ix + CX[i];
                    const bool in = jx >= 0 && jx < (int)pg.g.lx;
                    const int jy = CY[i] == 0 ? iy : (CY[i] > 0 ? yp : ym);
                    const int jz = CZ[i] == 0 ? iz : (CZ[i] > 0 ? zp : zm);
                    sv[i] = ps.solid[node_of(pg.g, in ? jx : ix, jy, jz)];
                }
                if (own < 0) {
#pragma unroll
                    for (int i = 1; i < 19; i++) {
                        const int jx = ix + CX[i];
                        if (jx >= 0 && jx < (int)pg.g.lx && sv[i] == p)
                            mask |= 1u << i;
                    }
                }
            }
        }
        int tot;
        const int off = block_scan(__builtin_popcount(mask), tot);
        if (mask) {
            int k = base + off;
            for (uint32_t m = mask; m != 0; m &= m - 1) { // set bits, lowest direction first
                const int i = __builtin_ctz(m);
                if (k < pg.maxlink) {
                    const size_t s = (size_t)p * pg.maxlink + k;
                    ps.link_node[s] = (int)n;
                    ps.link_dir[s] = (signed char)i;
                } else {
                    atomicOr(ps.overflow, 1);
                }
                k++;
            }
        }
        base += tot;
    }
    if (threadIdx.x == 0)
        ps.link_count[p] = base < pg.maxlink ? base : pg.maxlink;
}

// Phase 4: the cut-link fraction q of every link, one lane per link: the
// smaller root of |x + t c_i - X|^2 = R^2, clamped to [0, 1] (1/2 if the
// line misses the sphere).  |c_i|^2 is 1 or 2, so the division by it is the
// exact product with 1 or 1/2.
__global__ void __launch_bounds__(PBLOCK) k_p_linkq(PartGeom pg, PartState ps)
{
    const int p = blockIdx.y;
    const int nl = ps.link_count[p];
    const double *X = ps.X + 3 * p;
    for (int k = blockIdx.x * PBLOCK + threadIdx.x; k < nl; k += gridDim.x * PBLOCK) {
        const size_t s = (size_t)p * pg.maxlink + k;
        const uint32_t n = (uint32_t)ps.link_node[s];
        const int i = ps.link_dir[s];
        const uint32_t qn = fdiv(n, pg.g.dlx);
        const int ix = (int)(n - qn * pg.g.lx);
        const int iz = (int)fdiv(qn, pg.g.dly);
        const int iy = (int)(qn - (uint32_t)iz * pg.g.ly);
        double dx, dy, dz;
        node_offset(pg, ix, iy, iz, X, dx, dy, dz);
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double cc = d2 - pg.R2;
        const double cx = (double)CX[i], cy = (double)CY[i], cz = (double)CZ[i];
        const double a = cx * cx + cy * cy + cz * cz;
        const double inv_a = i <= 6 ? 1.0 : 0.5;
        const double bb = cx * dx + cy * dy + cz * dz;
        const double disc = bb * bb - a * cc;
        double q = (disc >= 0.0) ? (-bb - sqrt(disc)) * inv_a : 0.5;
        if (q < 0.0)
            q = 0.0;
        if (q > 1.0)
            q = 1.0;
        ps.link_q[s] = q;
    }
}

// Storage access with a run-time direction (2-byte storage scales by
// direction class).
template <int P>
__device__ inline double ld_rt(const typename Prec<P>::S *L, uint64_t N, int i, uint32_t n)
{
    const typename Prec<P>::S v = L[(uint64_t)i * N + n];
    if constexpr (half_storage(P))
        return (double)Prec<P>::to_T_rt(v, i);
    else
        return (double)v;
}

template <int P>
__device__ inline void st_rt(typename Prec<P>::S *L, uint64_t N, int i, uint32_t n, double v)
{
    using S = typename Prec<P>::S;
    if constexpr (half_storage(P))
        L[(uint64_t)i * N + n] = Prec<P>::to_S_rt((float)v, i);
    else
        L[(uint64_t)i * N + n] = (S)v;
}

// Interpolated bounce-back on one cut link of particle p (k < link_count[p]),
// written into the solid node's slot that the fluid node pulls next step;
// the link's momentum-exchange contribution.
template <int P>
__device__ inline void ibb_link(const PartGeom &pg, const PartState &ps, typename Prec<P>::S *L, uint64_t N,
                                     int p, int k)
{
    const size_t s = (size_t)p * pg.maxlink + k;
    const uint32_t n = (uint32_t)ps.link_node[s];
    const int i = ps.link_dir[s];
    const int ib = OPP[i];
    const double q = ps.link_q[s];
    const uint32_t qn = fdiv(n, pg.g.dlx);
    const int ix = (int)(n - qn * pg.g.lx);
    const int iz = (int)fdiv(qn, pg.g.dly);
    const int iy = (int)(qn - (uint32_t)iz * pg.g.ly);
    const double cx = (double)CX[i], cy = (double)CY[i], cz = (double)CZ[i];
    const double w = i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0;

    const double *X = ps.X + 3 * p, *V = ps.V + 3 * p, *W = ps.W + 3 * p;
    double dx, dy, dz;
    node_offset(pg, ix, iy, iz, X, dx, dy, dz);
    const double rx = dx + q * cx, ry = dy + q * cy, rz = dz + q * cz; // wall point - centre
    const double uwx = V[0] + (W[1] * rz - W[2] * ry);
    const double uwy = V[1] + (W[2] * rx - W[0] * rz);
    const double uwz = V[2] + (W[0] * ry - W[1] * rx);
    const double cu = cx * uwx + cy * uwy + cz * uwz;

    // Every value either branch may need loads at once, so one memory round
    // trip serves the link: f_i at x, f_opp(i) at x, and flag and f_i of the
    // second fluid node x - c_i behind x (x itself when that lies beyond a
    // wall, so the address stays in bounds; the branch below then ignores it).
    // When x - c_i is a solid node, another link of this launch may be storing
    // into its slots; the value read there is then discarded.
    const int bx_ = ix - CX[i];
    const bool behind_in = (bx_ >= 0 && bx_ < (int)pg.g.lx);
    const uint32_t nb =
        behind_in ? node_of(pg.g, bx_, wrap(iy - CY[i], pg.g.ly), wrap(iz - CZ[i], pg.g.lz)) : n;
    const double fi = ld_rt<P>(L, N, i, n);
    const double fib = ld_rt<P>(L, N, ib, n);
    const int solid_b = ps.solid[nb];
    const double fb = ld_rt<P>(L, N, i, nb);
    double g;
    if (q < 0.5) {
        if (behind_in && solid_b < 0)
            g = (2.0 * q) * fi + (1.0 - 2.0 * q) * fb - 6.0 * w * cu;
        else
            g = fi - 6.0 * w * cu;
    } else {
        g = (1.0 / (2.0 * q)) * fi + ((2.0 * q - 1.0) / (2.0 * q)) * fib - (3.0 / q) * w * cu;
    }
    // the solid node x + c_i: its slot ib is what x pulls for direction ib
    const uint32_t ns = node_of(pg.g, ix + CX[i], wrap(iy + CY[i], pg.g.ly), wrap(iz + CZ[i], pg.g.lz));
    st_rt<P>(L, N, ib, ns, g);

    // momentum exchange with full populations f = shifted + w*rho0
    const double ff = fi + w, gf = g + w;
    const double px = ff * (cx - uwx) + gf * (cx + uwx);
    const double py = ff * (cy - uwy) + gf * (cy + uwy);
    const double pz = ff * (cz - uwz) + gf * (cz + uwz);
    double *c = ps.link_contrib + s * 6;
    c[0] = px;
    c[1] = py;
    c[2] = pz;
    c[3] = ry * pz - rz * py;
    c[4] = rz * px - rx * pz;
    c[5] = rx * py - ry * px;
}

// Every cut link: IBB_BLOCKS blocks per particle stride over its links.
template <int P>
__global__ void __launch_bounds__(PBLOCK) k_p_ibb(PartGeom pg, PartState ps, void *Lv)
{
    using S = typename Prec<P>::S;
    const int p = blockIdx.y;
    const int nl = ps.link_count[p];
    S *L = (S *)Lv;
    const uint64_t N = pg.g.N;
    for (int k = blockIdx.x * PBLOCK + threadIdx.x; k < nl; k += gridDim.x * PBLOCK)
        ibb_link<P>(pg, ps, L, N, p, k);
}

// Sums in a fixed order over one 64-lane block: lane l adds terms l, l+64,
// l+128, ... in increasing order, then the 64 partials are added pairwise
// (l += l + 32, l += l + 16, ..., l += l + 1).  The CPU oracle sums in the
// same order, so results agree bit for bit.
constexpr int SUMW = 64;

template <int C>
__device__ inline void lane_tree(double (&v)[C], double (*s)[SUMW])
{
    const int l = threadIdx.x;
#pragma unroll
    for (int j = 0; j < C; j++)
        s[j][l] = v[j];
    __syncthreads();
    for (int off = SUMW / 2; off > 0; off >>= 1) {
        if (l < off) {
#pragma unroll
            for (int j = 0; j < C; j++)
                s[j][l] = s[j][l] + s[j][l + off];
        }
        __syncthreads();
    }
}

// Per-particle force and torque from the per-link contributions, one block
// of SUMW lanes per particle.
__global__ void __launch_bounds__(SUMW) k_p_sum(PartGeom pg, PartState ps)
{
    const int p = blockIdx.x;
    const int l = threadIdx.x;
    double f[6] = {0, 0, 0, 0, 0, 0};
    const int nl = ps.link_count[p];
    for (int k = l; k < nl; k += SUMW) {
        const double *c = ps.link_contrib + ((size_t)p * pg.maxlink + k) * 6;
#pragma unroll
        for (int j = 0; j < 6; j++)
            f[j] = f[j] + c[j];
    }
    __shared__ double s[6][SUMW];
    lane_tree<6>(f, s);
    if (l == 0) {
        for (int j = 0; j < 3; j++) {
            ps.FH[3 * p + j] = s[j][0];
            ps.TQ[3 * p + j] = s[3 + j][0];
        }
    }
}

// Contact force magnitude of Feng & Michaelides (2005) eq. 28 for the gap h
// and range zeta: stf0 ((zeta - h)/zeta)^2 below zeta, plus stf1 (-h/zeta)
// on overlap (h < 0).
__device__ inline double contact_mag(double h, double zeta, double stf0, double stf1)
{
    if (!(h < zeta))
        return 0.0;
    const double r = (zeta - h) / zeta;
    double m = stf0 * (r * r);
    if (h < 0.0)
        m = m + stf1 * ((-h) / zeta);
    return m;
}

// Coefficient of the approach velocity in the lubrication correction of
// Nguyen & Ladd (2002): -6 pi mu a^2 (1/h - 1/hc), the gap no smaller than
// lub_min, zero at or beyond hc or without viscosity (mu = rho0 nu, rho0 = 1).
__device__ inline double lub_coef(double h, double a, const PartGeom &pg)
{
    if (!(pg.visc > 0.0) || !(h < pg.lub_cut))
        return 0.0;
    const double he = h > pg.lub_min ? h : pg.lub_min;
    return -6.0 * 3.14159265358979323846 * pg.visc * (a * a) * (1.0 / he - 1.0 / pg.lub_cut);
}

// Contact and lubrication forces and FP64 integration, one block of SUMW
// lanes per particle: the lanes split the other particles (q = l, l+64, ...),
// their sums are added by lane_tree, lane 0 adds the walls and integrates.
// Every block reads the positions and velocities before the move; the new
// ones are staged (Xold, VWnew) and swapped in by k_p_commit.
__global__ void __launch_bounds__(SUMW) k_p_move(PartGeom pg, PartState ps, double *Xold)
{
    const int p = blockIdx.x;
    const int l = threadIdx.x;
    const double Ly = (double)pg.g.ly, Lz = (double)pg.g.lz, Lx = (double)pg.g.lx;
    const double *Xp = ps.X + 3 * p, *Vp = ps.V + 3 * p;
    double f6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; // contact, lubrication
    for (int q = l; q < pg.np; q += SUMW) {
        if (q == p)
            continue;
        const double *Xq = ps.X + 3 * q, *Vq = ps.V + 3 * q;
        const double ddx = Xp[0] - Xq[0];
        const double ddy = min_image(Xp[1] - Xq[1], Ly);
        const double ddz = min_image(Xp[2] - Xq[2], Lz);
        const double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        if (!(dist > 0.0))
            continue;
        const double h = dist - 2.0 * pg.R;
        const double nx = ddx / dist, ny = ddy / dist, nz = ddz / dist;
        const double mag = contact_mag(h, pg.mingap, pg.stf0, pg.stf1);
        const double un = (Vp[0] - Vq[0]) * nx + (Vp[1] - Vq[1]) * ny + (Vp[2] - Vq[2]) * nz;
        const double lub = lub_coef(h, 0.5 * pg.R, pg) * un;
        f6[0] = f6[0] + mag * nx;
        f6[1] = f6[1] + mag * ny;
        f6[2] = f6[2] + mag * nz;
        f6[3] = f6[3] + lub * nx;
        f6[4] = f6[4] + lub * ny;
        f6[5] = f6[5] + lub * nz;
    }
    __shared__ double s[6][SUMW];
    lane_tree<6>(f6, s);
    if (l != 0)
        return;
    double fc[3] = {s[0][0], s[1][0], s[2][0]};
    double fl[3] = {s[3][0], s[4][0], s[5][0]};
    // walls at x = 0 and x = lx: the mirror image lies twice the distance away
    const double dlo = Xp[0], dhi = Lx - Xp[0];
    fc[0] = fc[0] + contact_mag(2.0 * dlo - 2.0 * pg.R, pg.mingap_w, pg.stf0_w, pg.stf1_w);
    fc[0] = fc[0] - contact_mag(2.0 * dhi - 2.0 * pg.R, pg.mingap_w, pg.stf0_w, pg.stf1_w);
    fl[0] = fl[0] + lub_coef(dlo - pg.R, pg.R, pg) * Vp[0];
    fl[0] = fl[0] + lub_coef(dhi - pg.R, pg.R, pg) * Vp[0];
    const double fb[3] = {pg.fx, pg.fy, pg.fz};
    const bool first = (*ps.nsteps_done == 0);
    double Xn[3], Vn[3], Wn[3];
    for (int j = 0; j < 3; j++) {
        const int k = 3 * p + j;
        ps.FC[k] = fc[j];
        ps.FL[k] = fl[j];
        const double fh = first ? ps.FH[k] : 0.5 * (ps.FH[k] + ps.FHprev[k]);
        const double tq = first ? ps.TQ[k] : 0.5 * (ps.TQ[k] + ps.TQprev[k]);
        ps.FHprev[k] = ps.FH[k];
        ps.TQprev[k] = ps.TQ[k];
        double ftot = fh + fc[j] + fl[j];
        if (pg.pressure_force)
            ftot = ftot + pg.volume * fb[j];
        const double a = ftot / pg.mass;
        const double al = tq / pg.inertia;
        const double v0 = ps.V[k], w0 = ps.W[k];
        Vn[j] = first ? v0 + a : v0 + (1.5 * a - 0.5 * ps.Aprev[k]);
        Wn[j] = first ? w0 + al : w0 + (1.5 * al - 0.5 * ps.ALprev[k]);
        Xn[j] = Xp[j] + 0.5 * (v0 + Vn[j]);
        ps.TH[k] = ps.TH[k] + 0.5 * (w0 + Wn[j]);
        ps.Aprev[k] = a;
        ps.ALprev[k] = al;
    }
    if (Xn[1] >= Ly)
        Xn[1] = Xn[1] - Ly;
    else if (Xn[1] < 0.0)
        Xn[1] = Xn[1] + Ly;
    if (Xn[2] >= Lz)
        Xn[2] = Xn[2] - Lz;
    else if (Xn[2] < 0.0)
        Xn[2] = Xn[2] + Lz;
    // a diverged run: flag it (sphere_box then gives empty boxes for such a
    // particle, so no kernel indexes from it)
    bool finite = true;
    for (int j = 0; j < 3; j++)
        finite = finite && (Xn[j] - Xn[j] == 0.0) && (Vn[j] - Vn[j] == 0.0) && (Wn[j] - Wn[j] == 0.0);
    if (!finite)
        atomicOr(ps.overflow, 2);
    for (int j = 0; j < 3; j++) {
        Xold[3 * p + j] = Xn[j];
        ps.VWnew[3 * p + j] = Vn[j];
        ps.VWnew[3 * pg.np + 3 * p + j] = Wn[j];
    }
}

// Swap new positions in (the previous positions are kept in Xold) and commit
// the new velocities.
__global__ void k_p_commit(PartGeom pg, PartState ps, double *Xold)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pg.np)
        return;
    for (int j = 0; j < 3; j++) {
        const double xn = Xold[3 * p + j];
        Xold[3 * p + j] = ps.X[3 * p + j];
        ps.X[3 * p + j] = xn;
        ps.V[3 * p + j] = ps.VWnew[3 * p + j];
        ps.W[3 * p + j] = ps.VWnew[3 * pg.np + 3 * p + j];
    }
    if (p == 0)
        *ps.nsteps_done = *ps.nsteps_done + 1;
}

// macrovar's solid branch (Channel-Flow/collision.f90:420-459): nodes inside
// particle p carry the rigid-body velocity V + W x r (r from the centre,
// minimum image in y and z) and rho = rhopart, less the pending avedensity
// offset, which the reference subtracts from every node (:508).  One block
// per particle strides over its box.
__global__ void __launch_bounds__(PBLOCK) k_p_solid_fields(PartGeom pg, PartState ps, double rho_off, double *rho,
                                                            double *ux, double *uy, double *uz)
{
    const int p = blockIdx.x;
    const double *X = ps.X + 3 * p, *V = ps.V + 3 * p, *W = ps.W + 3 * p;
    const Box b = sphere_box(pg, X, 0);
    if (b.x1 < b.x0)
        return;
    const int nx = b.x1 - b.x0 + 1, ny = b.y1 - b.y0 + 1, nz = b.z1 - b.z0 + 1;
    const int total = nx * ny * nz;
    for (int t = threadIdx.x; t < total; t += PBLOCK) {
        const int ix = b.x0 + t % nx;
        const int iy = wrap(b.y0 + (t / nx) % ny, pg.g.ly);
        const int iz = wrap(b.z0 + t / (nx * ny), pg.g.lz);
        const uint32_t n = node_of(pg.g, ix, iy, iz);
        if (ps.solid[n] != p)
            continue;
        double dx, dy, dz;
        node_offset(pg, ix, iy, iz, X, dx, dy, dz);
        ux[n] = V[0] + (W[1] * dz - W[2] * dy);
        uy[n] = V[1] + (W[2] * dx - W[0] * dz);
        uz[n] = V[2] + (W[0] * dy - W[1] * dx);
        rho[n] = pg.rhopart - rho_off;
    }
}

// Refill: nodes uncovered by particle p get the equilibrium of the wall
// velocity there and the mean density of their settled fluid neighbours.
// Refill of the uncovered nodes of particle p, REFILL_NODES nodes per wave
// pass.  Lane 18 t + (j - 1) of a pass serves node t of the pass and its
// neighbour x + c_j (j = 1..18): it loads that neighbour's flag and, for a
// fluid neighbour inside the walls, its 19 populations, and sums them in
// population order (its density).  Every lane of the node then adds the 18
// densities in j order from LDS, as one thread walking the neighbours would,
// so rho, and with it every stored value, does not depend on the lane
// layout.  The node gets the equilibrium at rho and the wall velocity
// V + W x r; lane j stores population j, lane 1 also population 0.  A node
// claimed by another particle since it was listed is left alone.
constexpr int REFILL_NODES = 3; // 3 x 18 = 54 of the 64 lanes

template <int P>
__global__ void __launch_bounds__(64) k_p_refill(PartGeom pg, PartState ps, void *Lv)
{
    using S = typename Prec<P>::S;
    S *L = (S *)Lv;
    const uint64_t N = pg.g.N;
    const int p = blockIdx.y;
    const int nf = ps.fill_count[p];
    const int l = threadIdx.x;
    const int t = l / 18, j = l % 18 + 1;
    __shared__ double rs[REFILL_NODES][18];
    __shared__ int ok[REFILL_NODES][18];
    for (int k0 = 0; k0 < nf; k0 += REFILL_NODES) {
        const int k = k0 + t;
        const bool lane = t < REFILL_NODES && k < nf;
        uint32_t n = 0;
        int ix = 0, iy = 0, iz = 0;
        bool mine = false;
        if (lane) {
            n = (uint32_t)ps.fill_node[(size_t)p * pg.maxfill + k];
            mine = ps.solid[n] == -2 - p;
            const uint32_t qn = fdiv(n, pg.g.dlx);
            ix = (int)(n - qn * pg.g.lx);
            iz = (int)fdiv(qn, pg.g.dly);
            iy = (int)(qn - (uint32_t)iz * pg.g.ly);
        }
        if (lane && mine) {
            double r = 0.0;
            int use = 0;
            const int jx = ix + CX[j];
            if (jx >= 0 && jx < (int)pg.g.lx) {
                const uint32_t m = node_of(pg.g, jx, wrap(iy + CY[j], pg.g.ly), wrap(iz + CZ[j], pg.g.lz));
                if (ps.solid[m] == -1) {
                    r = ld_rt<P>(L, N, 0, m);
                    for (int i = 1; i < 19; i++)
                        r = r + ld_rt<P>(L, N, i, m);
                    use = 1;
                }
            }
            rs[t][j - 1] = r;
            ok[t][j - 1] = use;
        }
        __syncthreads();
        if (lane && mine) {
            double rsum = 0.0;
            int cnt = 0;
            for (int jj = 0; jj < 18; jj++)
                if (ok[t][jj]) {
                    rsum = rsum + rs[t][jj];
                    cnt++;
                }
            const double rho = cnt > 0 ? rsum / (double)cnt : 0.0;
            const double *X = ps.X + 3 * p, *V = ps.V + 3 * p, *W = ps.W + 3 * p;
            double dx, dy, dz;
            node_offset(pg, ix, iy, iz, X, dx, dy, dz);
            const double ux = V[0] + (W[1] * dz - W[2] * dy);
            const double uy = V[1] + (W[2] * dx - W[0] * dz);
            const double uz = V[2] + (W[0] * dy - W[1] * dx);
            double usqr = ux * ux + uy * uy + uz * uz;
            usqr = 1.5 * usqr;
            if (j == 1)
                st_rt<P>(L, N, 0, n, (1.0 / 3.0) * (rho - usqr));
            const double G = (double)CX[j] * ux + (double)CY[j] * uy + (double)CZ[j] * uz;
            const double w = j <= 6 ? 1.0 / 18.0 : 1.0 / 36.0;
            st_rt<P>(L, N, j, n, w * (rho + 3.0 * G + 4.5 * G * G - usqr));
        }
        __syncthreads(); // the LDS slots are reused by the next pass
    }
}

// Uncovered nodes become ordinary fluid once every refill has read its
// neighbours.
__global__ void k_p_settle(PartGeom pg, PartState ps)
{
    const int p = blockIdx.y;
    const int nf = ps.fill_count[p];
    for (int k = blockIdx.x * 64 + threadIdx.x; k < nf; k += gridDim.x * 64) {
        const uint32_t n = (uint32_t)ps.fill_node[(size_t)p * pg.maxfill + k];
        if (ps.solid[n] == -2 - p)
            ps.solid[n] = -1;
    }
}

} // namespace

hipError_t part_classify(const PartGeom &pg, const PartState &ps, int first, hipStream_t s)
{
    // ps.Aprev is not touched; Xold lives after X in the state (see lbm_capi)
    if (!first)
        hipLaunchKernelGGL(k_p_uncover, dim3(pg.np), dim3(PBLOCK), 0, s, pg, ps, ps.X + 3 * (size_t)pg.np);
    hipLaunchKernelGGL(k_p_mark, dim3(pg.np), dim3(PBLOCK), 0, s, pg, ps);
    hipLaunchKernelGGL(k_p_links, dim3(pg.np), dim3(PBLOCK), 0, s, pg, ps);
    hipLaunchKernelGGL(k_p_linkq, dim3(IBB_BLOCKS, pg.np), dim3(PBLOCK), 0, s, pg, ps);
    return hipGetLastError();
}

hipError_t part_solid_fields(const PartGeom &pg, const PartState &ps, double rho_off, double *rho, double *ux,
                             double *uy, double *uz, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_solid_fields, dim3(pg.np), dim3(PBLOCK), 0, s, pg, ps, rho_off, rho, ux, uy, uz);
    return hipGetLastError();
}

hipError_t part_sum(const PartGeom &pg, const PartState &ps, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_sum, dim3(pg.np), dim3(SUMW), 0, s, pg, ps);
    return hipGetLastError();
}

hipError_t part_move(const PartGeom &pg, const PartState &ps, hipStream_t s)
{
    double *Xold = ps.X + 3 * (size_t)pg.np;
    hipLaunchKernelGGL(k_p_move, dim3(pg.np), dim3(SUMW), 0, s, pg, ps, Xold);
    hipLaunchKernelGGL(k_p_commit, dim3((pg.np + 63) / 64), dim3(64), 0, s, pg, ps, Xold);
    return hipGetLastError();
}

template <int P>
hipError_t PartLaunch<P>::ibb(const PartGeom &pg, const PartState &ps, void *L, hipStream_t s)
{
    hipLaunchKernelGGL((k_p_ibb<P>), dim3(IBB_BLOCKS, pg.np), dim3(PBLOCK), 0, s, pg, ps, L);
    return hipGetLastError();
}

template <int P>
hipError_t PartLaunch<P>::refill(const PartGeom &pg, const PartState &ps, void *L, hipStream_t s)
{
    const dim3 grid(1, pg.np); // one 64-lane block per particle passes over its refill list
    hipLaunchKernelGGL((k_p_refill<P>), grid, dim3(64), 0, s, pg, ps, L);
    hipLaunchKernelGGL(k_p_settle, grid, dim3(64), 0, s, pg, ps);
    return hipGetLastError();
}

template struct PartLaunch<0>;
template struct PartLaunch<1>;
template struct PartLaunch<2>;
template struct PartLaunch<3>;

} // namespace d3q19
