// Lattice kernels for gfx950, templated on the storage precision P
// (0 = FP64, 1 = FP32, 2 = FP16 storage / FP32 compute).  Included by one
// translation unit per precision, which instantiates Launch<P>.
//
// Streaming scheme: two lattices (A-B), pull.  The device lattice holds the
// post-collision populations P_k (the values collision.f90:230-262 stores);
// the reference's post-streaming array f is F_k(x)_i = P_k(x - c_i)_i, or
// P_k(x)_opp(i) when x - c_i lies beyond a wall (halfway bounce-back,
// collision.f90:229-230,250-251), with periodic wraps in y and z
// (collisionExchnge on one rank, collision.f90:281-372).  One step reads
// F_k by that gather, forms rho and u as macrovar does (collision.f90:396-418)
// unless the fields come from arrays, collides, and writes P_{k+1} with
// aligned stores.  Every read and write is a permutation of whole values, so
// the result equals the reference's in-place swap algorithm bit for bit.
#pragma once

#include <utility>

#include "d3q19_common.h"
#include "kernels.h"
#include "mlp.h"

namespace d3q19 {

constexpr int BLOCK = 256;
// Threads per block of the step kernels k_step and k_step_closure, measured
// against each other on one box at a time: 512 for FP64 (18 320 against
// 18 150 MLUPS with 256 on the 192x768x384 channel) and FP16 storage (71 600
// against 68 100); 256 for the plain FP32 step (39 220-39 330 against
// 37 990-38 800 with 512 on the 2048x1024x512 channel) and the fused closure
// (96 registers: five 256-thread blocks per CU against two of 512, 248 against
// 255 us per step on the 96x384x192 channel).  128 was slower.
template <int P, int MODE>
constexpr int step_block() { return ((MODE & 8) != 0 || P == 1) ? 256 : 512; }

// Element `node` of a population row, addressed as the row base plus a 32-bit
// byte offset when that cannot wrap: 2-byte storage, whose rows hold fewer than
// 2^31 nodes (lbm_create's limit), i.e. fewer than 2^32 bytes.  A wave-uniform
// row base then stays in scalar registers and the load or store takes the
// lane's offset directly, with no 64-bit address arithmetic per population.
template <int P>
__device__ __forceinline__ typename Prec<P>::S *elem(typename Prec<P>::S *row, uint32_t node)
{
    if constexpr (half_storage(P))
        return (typename Prec<P>::S *)((char *)row + node * (uint32_t)sizeof(typename Prec<P>::S));
    else
        return row + node;
}

template <int P>
__device__ __forceinline__ const typename Prec<P>::S *elem(const typename Prec<P>::S *row, uint32_t node)
{
    return elem<P>(const_cast<typename Prec<P>::S *>(row), node);
}

// F_I(x): streamed population of direction I at node x.  The bounce-back
// choice is made with masks, not branches, so every load issues back to back.
// Loads are plain (cached): the +-x shifted reads of neighbouring waves share
// cache lines.  W32 = 1 (N <= w32_max_n(P)): a pull with c_x != 0 addresses
// both of its rows, I and OPP[I], from the lower of the two with a 32-bit
// byte offset (|OPP[I] - I| <= 3, so below 4N elements: 8N bytes in 2-byte
// storage, 16N in FP32), and the lane selects between two 32-bit offsets
// instead of two 64-bit row bases.  W32 = 2 (2-byte storage, N <= 2^30): only
// the pulls whose rows are adjacent (|OPP[I] - I| = 1: I = 1, 2, 8, 9, 12,
// 13; offsets below 2N elements, 4N bytes) take that form.
constexpr uint64_t W32_MAX_N = 1ull << 29;
constexpr uint64_t w32_max_n(int P) { return half_storage(P) ? W32_MAX_N : (P == 1 ? (1ull << 28) : 0); }
constexpr uint64_t W32_ADJ_MAX_N = 1ull << 30;

template <int P, int I, int W32 = 0>
__device__ __forceinline__ typename Prec<P>::T pull(const typename Prec<P>::S *__restrict__ L, uint64_t N,
                                                    const Nbr &b)
{
    if constexpr (CX[I] != 0) {
        // a walled pull reads the node's own opposite slot (stored with the
        // opposite direction's scale, which equals this direction's); the
        // two row bases are wave-uniform, so the lane selects between them
        static_assert(Prec<P>::template scale_class<I>() == Prec<P>::template scale_class<OPP[I]>());
        const bool w = walled<I>(b);
        const uint32_t node = w ? b.n : src_of<I>(b);
        if constexpr (P == 0) {
            // memory-bound: this form measured 1.4 % faster than the one below
            return Prec<P>::template to_T<I>(L[(w ? (uint64_t)OPP[I] * N : (uint64_t)I * N) + node]);
        } else if constexpr (W32 == 1 || (W32 == 2 && (OPP[I] - I == 1 || I - OPP[I] == 1))) {
            static_assert(half_storage(P) || P == 1, "32-bit pull offsets are for FP32 and 2-byte storage");
            constexpr int M = I < OPP[I] ? I : OPP[I];
            uint64_t row_m = (uint64_t)M * N;
            uint32_t d_i = (uint32_t)(I - M) * (uint32_t)N, d_o = (uint32_t)(OPP[I] - M) * (uint32_t)N;
            // wave-uniform (N is the kernel argument Geom::N), kept in scalar
            // registers as below
            __asm__("" : "+s"(row_m));
            __asm__("" : "+s"(d_i));
            __asm__("" : "+s"(d_o));
            // element offset below 4N, its byte offset below 2^32 (w32_max_n):
            // the load takes the scalar row base and a 32-bit lane offset
            using S = typename Prec<P>::S;
            const uint32_t off = w ? b.n + d_o : src_of<I>(b) + d_i;
            return Prec<P>::template to_T<I>(
                *(const S *)((const char *)(L + row_m) + off * (uint32_t)sizeof(S)));
        } else {
            uint64_t row_i = (uint64_t)I * N, row_o = (uint64_t)OPP[I] * N;
            // opaque to the optimiser, which would otherwise turn the select
            // of two scalar row offsets into a per-lane 64-bit multiply of a
            // selected index (quarter-rate; FP16 storage is issue-bound).
            // The "s" operands require N to be wave-uniform: every caller
            // passes the kernel argument Geom::N.
            __asm__("" : "+s"(row_i));
            __asm__("" : "+s"(row_o));
            return Prec<P>::template to_T<I>(*elem<P>(L + (w ? row_o : row_i), node));
        }
    } else {
        return Prec<P>::template to_T<I>(*elem<P>(L + (uint64_t)I * N, src_of<I>(b)));
    }
}

template <int P, int W32, int... Is>
__device__ __forceinline__ void pull_all_impl(const typename Prec<P>::S *__restrict__ L, uint64_t N, const Nbr &b,
                                              typename Prec<P>::T (&f)[19], std::integer_sequence<int, Is...>)
{
    ((f[Is] = pull<P, Is, W32>(L, N, b)), ...);
}

template <int P, int W32 = 0>
__device__ __forceinline__ void pull_all(const typename Prec<P>::S *__restrict__ L, uint64_t N, const Nbr &b,
                                         typename Prec<P>::T (&f)[19])
{
    pull_all_impl<P, W32>(L, N, b, f, std::make_integer_sequence<int, 19>{});
}

template <int P, int... Is>
__device__ __forceinline__ void store_all_impl(typename Prec<P>::S *__restrict__ L, uint64_t N, uint32_t n,
                                               const typename Prec<P>::T (&f)[19], std::integer_sequence<int, Is...>)
{
    ((Prec<P>::template st<Is>(elem<P>(L + (uint64_t)Is * N, n), f[Is])), ...);
}

template <int P>
__device__ __forceinline__ void store_all(typename Prec<P>::S *__restrict__ L, uint64_t N, uint32_t n,
                                          const typename Prec<P>::T (&f)[19])
{
    store_all_impl<P>(L, N, n, f, std::make_integer_sequence<int, 19>{});
}

// Scatter the streamed value F_I(x) back to the post-collision slot it came
// from (the inverse of pull).
template <int P, int... Is>
__device__ __forceinline__ void scatter_all_impl(typename Prec<P>::S *__restrict__ L, uint64_t N, const Nbr &b,
                                                 const typename Prec<P>::T (&f)[19], std::integer_sequence<int, Is...>)
{
    ((walled<Is>(b) ? Prec<P>::template st<OPP[Is]>(L + (uint64_t)OPP[Is] * N + b.n, f[Is])
                    : Prec<P>::template st<Is>(L + (uint64_t)Is * N + src_of<Is>(b), f[Is])),
     ...);
}

template <int P>
__device__ __forceinline__ void scatter_all(typename Prec<P>::S *__restrict__ L, uint64_t N, const Nbr &b,
                                            const typename Prec<P>::T (&f)[19])
{
    scatter_all_impl<P>(L, N, b, f, std::make_integer_sequence<int, 19>{});
}

__device__ __forceinline__ void node_coords(uint32_t n, const Geom &g, uint32_t &ix, uint32_t &iy, uint32_t &iz)
{
    const uint32_t q = fdiv(n, g.dlx);
    ix = n - q * g.lx;
    iz = fdiv(q, g.dly);
    iy = q - iz * g.ly;
}

// Body force of s at node (ix, iy, iz), the value the reference's force
// arrays hold there.  The FORCINGP products keep the operand order of
// Channel-Flow/collision.f90:571-576 and 591-596; band 2's leading minus is
// carried by the first factor, which changes no bit of the product.
__device__ __forceinline__ void node_force(const ForceSpec &s, const Geom &g, uint32_t ix, uint32_t iy, uint32_t iz,
                                           double &fx, double &fy, double &fz)
{
#pragma clang fp contract(off)
    fx = s.fx;
    fy = s.fy;
    fz = s.fz;
    if (s.tab == nullptr)
        return;
    // every table value loads at once (both bands' y factors), so no load
    // waits for the band flag; the selects pick what the band needs
    const double *xt = s.tab + 3 * ix;
    const double *yt = s.tab + 3 * g.lx + 4 * iy;
    const double *zt = s.tab + 3 * g.lx + 4 * g.ly + 2 * iz;
    const double band = xt[0], x1 = xt[1], x2 = xt[2];
    const double c1 = yt[0], s1 = yt[1], c2 = yt[2], s2 = yt[3];
    const double cz = zt[0], sz = zt[1];
    const bool first = band == 1.0;
    const double sg = first ? 1.0 : -1.0;
    const double cy = first ? c1 : c2, sy = first ? s1 : s2;
    const double gx = (((sg * s.a) * x1) * cy) * cz;
    const double gy = s.fy * (1.0 + (((-sg * s.b) * x2) * sy) * cz);
    const double gz = (((sg * s.c) * x2) * cy) * sz;
    const bool in_band = band != 0.0;
    fx = in_band ? gx : fx;
    fy = in_band ? gy : fy;
    fz = in_band ? gz : fz;
}

template <int NB>
__device__ __forceinline__ uint32_t node_index(const Geom &g)
{
    uint32_t blk = blockIdx.x;
    if (g.remap)
        blk = xcd_remap(blk, gridDim.x);
    return blk * NB + threadIdx.x;
}

// ---------------------------------------------------------------------------
// One time step: gather F_k, macroscopic fields, MRT collision, store P_{k+1}.
// MODE 0: rho, u from F_k (macrovar fused); MODE 1: same with rho - rho_off
// (avedensity's correction); MODE 2: rho, u read from the field arrays.
// MODE + 4: nodes inside a particle are not stored (their slots keep the
// interpolated bounce-back values written before the step, which the fluid
// nodes pull).
// MODE + 8: the learned closure (mlp.h) sets each node's shear rates
// (FP32 compute only).
// MODE + 16: the body force varies over the nodes (FORCINGP), or the fused
// fields were formed with another force than the collision's (the force was
// changed after the last macrovar): a.fc is the collision's force, a.fm the
// fields'.
// MODE + 32: 2-byte storage with N <= W32_MAX_N (and the FP32 closure with
// N <= 2^28): the pulls with c_x != 0 take 32-bit offsets (pull<P, I, 1>);
// the same loads, fewer instructions.  MODE + 64: 2-byte storage with
// N <= 2^30: the pulls between adjacent rows only (pull<P, I, 2>).
template <int P, int MODE>
__device__ __forceinline__ void step_node(const StepArgs &a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    // The closure's network runs on the matrix cores over whole waves: there
    // every lane stays to the end (a lane past the lattice works on the last
    // node, a particle node on its own slots) and only active lanes store.
    constexpr bool WAVE = (MODE & 8) != 0;
    constexpr int NB = step_block<P, MODE>();
    const uint32_t n0 = node_index<NB>(a.g);
    bool active = n0 < a.g.N;
    if constexpr (!WAVE) {
        if (!active)
            return;
    }
    const uint32_t n = active ? n0 : (uint32_t)(a.g.N - 1);
    // a node inside a particle is computed like the others (its gather
    // issues with the flag load instead of waiting for it) but not stored
    if constexpr ((MODE & 4) != 0)
        active = active && a.solid[n] < 0;
    const Nbr b = make_nbr(n, a.g);
    const S *__restrict__ src = (const S *)a.src;
    S *__restrict__ dst = (S *)a.dst;

    // the 19 loads issue first; the closure's weights are staged while they
    // are in flight
    T f[19];
    pull_all<P, (MODE & 32) != 0 ? 1 : ((MODE & 64) != 0 ? 2 : 0)>(src, a.g.N, b, f);
    __shared__ __attribute__((aligned(16))) _Float16 wl[WAVE ? MLP_LDS : 8];
    if constexpr (WAVE) {
        mlp_stage(a.mlp, wl, threadIdx.x, NB);
        __syncthreads();
    }

    Rates<T> r = rates_as<T>(a.r);
    T fmx = r.fx, fmy = r.fy, fmz = r.fz; // force in u = j + F/2 of the fused fields
    if constexpr ((MODE & 16) != 0) {
        uint32_t ix, iy, iz;
        node_coords(n, a.g, ix, iy, iz);
        double x, y, z;
        node_force(a.fc, a.g, ix, iy, iz, x, y, z);
        r.fx = (T)x;
        r.fy = (T)y;
        r.fz = (T)z;
        if constexpr ((MODE & 3) != 2) {
            node_force(a.fm, a.g, ix, iy, iz, x, y, z);
            fmx = (T)x;
            fmy = (T)y;
            fmz = (T)z;
        }
    }

    T rho9, ux9, uy9, uz9;
    T jx = T(0), jy = T(0), jz = T(0);
    if constexpr ((MODE & 3) == 2) {
        rho9 = (T)a.rho[n];
        ux9 = (T)a.ux[n];
        uy9 = (T)a.uy[n];
        uz9 = (T)a.uz[n];
    } else {
        macro_sums(f, rho9, ux9, uy9, uz9);
        jx = ux9;
        jy = uy9;
        jz = uz9;
        ux9 = ux9 + fmx / T(2.); // collision.f90:415-417
        uy9 = uy9 + fmy / T(2.);
        uz9 = uz9 + fmz / T(2.);
        if constexpr ((MODE & 3) == 1)
            rho9 = rho9 - (T)a.rho_off; // collision.f90:508
    }
    if constexpr (P == 0) {
        static_assert((MODE & 8) == 0, "the closure runs in FP32 arithmetic only");
        mrt_collide(f, rho9, ux9, uy9, uz9, r);
    } else {
        // increment form; conserved-moment changes are zero for fields fused
        // from f, -rho_off after avedensity, field - moment for arrays
        T drho = T(0), djx = T(0), djy = T(0), djz = T(0);
        if constexpr ((MODE & 3) == 1)
            drho = -(T)a.rho_off;
        if constexpr ((MODE & 16) != 0 && (MODE & 3) != 2) {
            // fields formed with fm, collision with fc
            djx = ux9 - (jx + r.fx / T(2.));
            djy = uy9 - (jy + r.fy / T(2.));
            djz = uz9 - (jz + r.fz / T(2.));
        }
        if constexpr ((MODE & 3) == 2) {
            T rf, jx, jy, jz;
            macro_sums(f, rf, jx, jy, jz);
            drho = rho9 - rf;
            djx = ux9 - (jx + r.fx / T(2.));
            djy = uy9 - (jy + r.fy / T(2.));
            djz = uz9 - (jz + r.fz / T(2.));
        }
        if constexpr ((MODE & 8) != 0) {
            const uint32_t ix = n - fdiv(n, a.g.dlx) * a.g.lx;
            const float wd = fminf((float)ix + 0.5f, (float)a.g.lx - (float)ix - 0.5f);
            mrt_collide_delta(f, rho9, ux9, uy9, uz9, drho, djx, djy, djz, r,
                              MlpClosureWave{a.mlp, wl, wd * a.inv_ystar / a.yscale});
        } else {
            mrt_collide_delta(f, rho9, ux9, uy9, uz9, drho, djx, djy, djz, r);
        }
    }
    if constexpr (WAVE || (MODE & 4) != 0) {
        if (!active)
            return;
    }
    store_all<P>(dst, a.g.N, n, f);
}

template <int P, int MODE>
__global__ void __launch_bounds__((step_block<P, MODE>())) k_step(StepArgs a)
{
    step_node<P, MODE>(a);
}

// The closure modes, held to 128 registers so that at least four waves per
// SIMD keep the lattice traffic in flight while others run the network.
template <int P, int MODE>
__global__ void __launch_bounds__((step_block<P, MODE>()), 4) k_step_closure(StepArgs a)
{
    step_node<P, MODE>(a);
}

// ---------------------------------------------------------------------------
// Two-step blocked sweep: one time step of a few slabs between slab stores
// (kernels.h: PlaneStore).  The arithmetic per node is k_step<P, 0>'s, so the
// values are the same bit for bit; only where the states live differs.
// NT: non-temporal stores (the whole-lattice output); otherwise cached
// stores, so a ring of slabs written here stays in the last-level cache for
// the next launch, which reads it.
// LNT: non-temporal loads.  The step-t launches read each lattice value once,
// and such loads keep that stream from evicting the ring from the last-level
// cache (FP64 +3.5 % on the 192x768x384 channel, which also let the ring grow
// from 146 to 224 MB; non-temporal loads of the ring itself cost 17 %).
template <int P, int AXIS, int LNT, int... Is>
__device__ __forceinline__ void block_gather(const PlaneRef (&pl)[3], const BlockLane &l,
                                             typename Prec<P>::T (&f)[19], std::integer_sequence<int, Is...>)
{
    using S = typename Prec<P>::S;
    (
        [&] {
            void *buf;
            uint64_t idx;
            block_src<AXIS, Is>(pl, l, buf, idx);
            if constexpr (LNT != 0)
                f[Is] = Prec<P>::template to_T<Is>(__builtin_nontemporal_load((const S *)buf + idx));
            else
                f[Is] = Prec<P>::template to_T<Is>(((const S *)buf)[idx]);
        }(),
        ...);
}

// Store the node's 19 values at element e of slab reference d (population I
// at I * d.stride + e).
template <int P, int NT, int... Is>
__device__ __forceinline__ void block_store(const PlaneRef &d, uint32_t e, const typename Prec<P>::T (&f)[19],
                                            std::integer_sequence<int, Is...>)
{
    using S = typename Prec<P>::S;
    S *base = (S *)d.buf + d.off;
    if constexpr (NT)
        ((Prec<P>::template st<Is>(base + (uint64_t)Is * d.stride + e, f[Is])), ...);
    else
        ((base[(uint64_t)Is * d.stride + e] = Prec<P>::template to_S<Is>(f[Is])), ...);
}

// Threads per block of the sweep, measured on the 192x768x384 channel: 512
// in FP32 (1.4 % faster than 256); in FP64 256 for the step-t launches (NT
// = 0) and 512 for the step-t+1 launches (0.7 % faster than 256 for both;
// 512 for both 1 % slower, 128 3 % slower).
template <int P, int NT>
constexpr int plane_block() { return P == 1 ? 512 : (NT ? 512 : 256); }

// FP: the body force per node (mode 16 of k_step: a.fc for the collision,
// a.fm for the fused fields), with the same arithmetic as k_step<P, 16>.
template <int P, int AXIS, int NT, int FP>
__global__ void __launch_bounds__((plane_block<P, NT>())) k_step_planes(BlockArgs a)
{
    using T = typename Prec<P>::T;
    const uint32_t t = blockIdx.x * plane_block<P, NT>() + threadIdx.x;
    // the wave's slab, uniform across its lanes (slabs are padded to whole
    // waves), so the slab buffers and offsets below stay in scalar registers
    const uint32_t sl = __builtin_amdgcn_readfirstlane(fdiv(t, a.dslabp));
    const uint32_t w = t - sl * a.slabp;
    if (sl >= a.count || w >= a.slab)
        return;
    const BlockRow br = block_row(a, w);
    const uint32_t ix = br.ix;
    const uint32_t sb = block_plane(a, sl);
    PlaneRef pl[3];
    block_planes(a, sb, pl);
    const BlockLane l = block_lane(ix, br.rs, a.src_nr, (uint32_t)a.src.row_stride, a.g.lx);

    T f[19];
    block_gather<P, AXIS, NT == 0>(pl, l, f, std::make_integer_sequence<int, 19>{});

    Rates<T> rt = rates_as<T>(a.r);
    T fmx = rt.fx, fmy = rt.fy, fmz = rt.fz; // force in u = j + F/2 of the fused fields
    if constexpr (FP != 0) {
        const uint32_t iy = AXIS == 2 ? br.rg : sb, iz = AXIS == 2 ? sb : br.rg;
        double x, y, z;
        node_force(a.fc, a.g, ix, iy, iz, x, y, z);
        rt.fx = (T)x;
        rt.fy = (T)y;
        rt.fz = (T)z;
        node_force(a.fm, a.g, ix, iy, iz, x, y, z);
        fmx = (T)x;
        fmy = (T)y;
        fmz = (T)z;
    }
    T rho9, ux9, uy9, uz9;
    macro_sums(f, rho9, ux9, uy9, uz9);
    const T jx = ux9, jy = uy9, jz = uz9;
    ux9 = ux9 + fmx / T(2.); // collision.f90:415-417
    uy9 = uy9 + fmy / T(2.);
    uz9 = uz9 + fmz / T(2.);
    if constexpr (P == 0) {
        mrt_collide(f, rho9, ux9, uy9, uz9, rt);
    } else {
        T djx = T(0), djy = T(0), djz = T(0);
        if constexpr (FP != 0) {
            // fields formed with fm, collision with fc
            djx = ux9 - (jx + rt.fx / T(2.));
            djy = uy9 - (jy + rt.fy / T(2.));
            djz = uz9 - (jz + rt.fz / T(2.));
        }
        mrt_collide_delta(f, rho9, ux9, uy9, uz9, T(0), djx, djy, djz, rt);
    }

    const PlaneRef d = plane_write(a.dst, sb);
    const uint32_t e = (br.rl + a.dst_r0) * (uint32_t)d.row_stride + ix; // head and ring rows share the stride
    block_store<P, NT>(d, e, f, std::make_integer_sequence<int, 19>{});
    const PlaneRef hd = plane_write_head(a.dst, sb);
    if (hd.buf != nullptr)
        block_store<P, 0>(hd, e, f, std::make_integer_sequence<int, 19>{});
}

// Closure diagnostics for nodes [n_begin, n_begin + count) of the current
// lattice, as the next fused collision (mode 0, or 16 with per-node forces)
// sees them.  WHAT 0: the ten scaled network inputs (out[10 i + k] for node
// n_begin + i); WHAT 1: nu_t / nu = softplus(o) from the matrix-core network.
// The preamble is mrt_collide_delta's up to the closure call.
template <int P, int WHAT>
__global__ void __launch_bounds__(BLOCK) k_closure_eval(StepArgs a, uint64_t n_begin, uint64_t count, float *out)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    const uint64_t i0 = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool active = i0 < count;
    const uint32_t n = (uint32_t)(n_begin + (active ? i0 : count - 1));
    __shared__ __attribute__((aligned(16))) _Float16 wl[WHAT == 1 ? MLP_LDS : 8];
    if constexpr (WHAT == 1) {
        mlp_stage(a.mlp, wl, threadIdx.x, BLOCK);
        __syncthreads();
    }
    const Nbr b = make_nbr(n, a.g);
    T f[19];
    pull_all<P>((const S *)a.src, a.g.N, b, f);
    Rates<T> r = rates_as<T>(a.r);
    uint32_t ix, iy, iz;
    node_coords(n, a.g, ix, iy, iz);
    double x, y, z;
    node_force(a.fc, a.g, ix, iy, iz, x, y, z);
    r.fx = (T)x;
    r.fy = (T)y;
    r.fz = (T)z;
    node_force(a.fm, a.g, ix, iy, iz, x, y, z);
    T rho9, ux9, uy9, uz9;
    macro_sums(f, rho9, ux9, uy9, uz9);
    ux9 = ux9 + (T)x / T(2.);
    uy9 = uy9 + (T)y / T(2.);
    uz9 = uz9 + (T)z / T(2.);
    T m[15], sf[12], eq[12];
    delta_moments(f, rho9, ux9, uy9, uz9, r, m, sf, eq);
    T neq[9];
    closure_inputs(m, eq, neq);
    const float wd = fminf((float)ix + 0.5f, (float)a.g.lx - (float)ix - 0.5f);
    float in[MLP_IN];
    mlp_inputs_dev(a.mlp, neq, wd * a.inv_ystar / a.yscale, in);
    if constexpr (WHAT == 0) {
        if (active)
#pragma unroll
            for (int k = 0; k < MLP_IN; k++)
                out[i0 * MLP_IN + k] = in[k];
    } else {
        const float o = mlp_forward_wave(a.mlp, wl, in);
        if (active)
            out[i0] = mlp_nut(a.mlp, o);
    }
}

// ---------------------------------------------------------------------------
// The same step for VW = 2 consecutive nodes of one x row per thread, so
// every population moves as one aligned vector (16 B in FP64, 8 B in FP32).
// A run of VW nodes never straddles a row (lx % VW == 0), so its y/z source
// run is one aligned vector too.  The +-x shift of the ten populations with
// c_x != 0 moves one value across the run boundary: it comes from the
// neighbouring lane's vector (a cross-lane shuffle), or from memory for the
// first/last lane of a wave; a run at a wall takes the halfway bounce-back
// value from its own node's opposite slot instead.  The values and the
// arithmetic per node are those of k_step: FP64 results are identical bit
// for bit; in FP32, built with FMA contraction, the compiler may contract
// the two kernels differently (round-off-level differences).
// Modes 0 and 1 (fields fused from F), FP64 and FP32 storage.
template <int P>
struct Vec;
template <>
struct Vec<0> {
    static constexpr int W = 2;
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec<1> {
    // two nodes (8 B): four would need ~230 VGPRs and spill
    static constexpr int W = 2;
    typedef float type __attribute__((ext_vector_type(2)));
};

// F_I for the run n0 .. n0+VW-1 (b = neighbours of n0).  Every lane of the
// wave calls this (the shuffles need them all); inactive lanes load nothing.
template <int P, int I>
__device__ __forceinline__ void pull_run(const typename Prec<P>::S *__restrict__ L, uint64_t N, const Nbr &b,
                                         bool whi_last, bool active, bool first_lane, bool last_lane,
                                         typename Prec<P>::T (&F)[Vec<P>::W][19])
{
    using S = typename Prec<P>::S;
    using VT = typename Vec<P>::type;
    constexpr int VW = Vec<P>::W;
    const uint64_t s0 = (uint64_t)I * N + src_yz<I>(b);
    VT v = VT(S(0));
    if (active)
        v = *(const VT *)(L + s0);
    if constexpr (CX[I] == 0) {
#pragma unroll
        for (int k = 0; k < VW; k++)
            F[k][I] = Prec<P>::template to_T<I>(v[k]);
    } else if constexpr (CX[I] == 1) {
        // node k pulls x - 1: element k - 1 of the run, element VW-1 of the
        // previous lane's run for k = 0
        S prev = __shfl_up(v[VW - 1], 1);
        if (active && first_lane && !b.wlo)
            prev = L[s0 - 1];
        if (active && b.wlo)
            prev = L[(uint64_t)OPP[I] * N + b.n];
        F[0][I] = Prec<P>::template to_T<I>(prev);
#pragma unroll
        for (int k = 1; k < VW; k++)
            F[k][I] = Prec<P>::template to_T<I>(v[k - 1]);
    } else {
        // node k pulls x + 1: element k + 1, element 0 of the next lane's run
        // for k = VW - 1
        S next = __shfl_down(v[0], 1);
        if (active && last_lane && !whi_last)
            next = L[s0 + VW];
        if (active && whi_last)
            next = L[(uint64_t)OPP[I] * N + b.n + (VW - 1)];
#pragma unroll
        for (int k = 0; k < VW - 1; k++)
            F[k][I] = Prec<P>::template to_T<I>(v[k + 1]);
        F[VW - 1][I] = Prec<P>::template to_T<I>(next);
    }
}

template <int P, int... Is>
__device__ __forceinline__ void pull_run_all(const typename Prec<P>::S *__restrict__ L, uint64_t N, const Nbr &b,
                                             bool whi_last, bool active, bool first_lane, bool last_lane,
                                             typename Prec<P>::T (&F)[Vec<P>::W][19],
                                             std::integer_sequence<int, Is...>)
{
    (pull_run<P, Is>(L, N, b, whi_last, active, first_lane, last_lane, F), ...);
}

template <int P, int... Is>
__device__ __forceinline__ void store_run_all(typename Prec<P>::S *__restrict__ L, uint64_t N, uint32_t n0,
                                              const typename Prec<P>::T (&F)[Vec<P>::W][19],
                                              std::integer_sequence<int, Is...>)
{
    using VT = typename Vec<P>::type;
    constexpr int VW = Vec<P>::W;
    auto one = [&](auto ic) {
        constexpr int I = decltype(ic)::value;
        VT v;
#pragma unroll
        for (int k = 0; k < VW; k++)
            v[k] = F[k][I];
        __builtin_nontemporal_store(v, (VT *)(L + (uint64_t)I * N + n0));
    };
    (one(std::integral_constant<int, Is>{}), ...);
}

template <int P, int MODE>
__global__ void __launch_bounds__(BLOCK) k_step_run(StepArgs a)
{
    static_assert(P == 0 || P == 1, "vector runs for FP64 and FP32 storage");
    static_assert(MODE == 0 || MODE == 1, "fields fused from F only");
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    constexpr int VW = Vec<P>::W;
    const uint32_t nruns = (uint32_t)(a.g.N / VW);
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    const bool active = t < nruns;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n0 = active ? t * VW : 0u;
    const Nbr b = make_nbr(n0, a.g);
    const uint32_t ix0 = n0 - fdiv(n0, a.g.dlx) * a.g.lx;
    const bool whi_last = (ix0 + VW == a.g.lx);
    const bool first_lane = (lane == 0);
    const bool last_lane = (lane == 63u) || (t + 1 >= nruns);

    T F[VW][19];
    pull_run_all<P>((const S *)a.src, a.g.N, b, whi_last, active, first_lane, last_lane, F,
                    std::make_integer_sequence<int, 19>{});
    if (!active)
        return;

    Rates<T> r = rates_as<T>(a.r);

#pragma unroll
    for (int k = 0; k < VW; k++) {
        T rho9, ux9, uy9, uz9;
        macro_sums(F[k], rho9, ux9, uy9, uz9);
        ux9 = ux9 + r.fx / T(2.); // collision.f90:415-417
        uy9 = uy9 + r.fy / T(2.);
        uz9 = uz9 + r.fz / T(2.);
        if constexpr (MODE == 1)
            rho9 = rho9 - (T)a.rho_off; // collision.f90:508
        if constexpr (P == 0) {
            mrt_collide(F[k], rho9, ux9, uy9, uz9, r);
        } else {
            const T drho = MODE == 1 ? -(T)a.rho_off : T(0);
            mrt_collide_delta(F[k], rho9, ux9, uy9, uz9, drho, T(0), T(0), T(0), r);
        }
    }
    store_run_all<P>((S *)a.dst, a.g.N, n0, F, std::make_integer_sequence<int, 19>{});
}

// macrovar / rhoupdat / materialisation of the fields of lattice L.
// WHAT 0: rho, ux, uy, uz = macrovar(F) with rho - rho_off.
// WHAT 1: rho = sum F (rhoupdat); max |rho_new - rho_old| into *errbits.
template <int P, int WHAT>
__global__ void __launch_bounds__(BLOCK) k_macro(MacroArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    const uint32_t n = blockIdx.x * BLOCK + threadIdx.x;
    double err = 0.0;
    if (n < a.g.N) {
        const Nbr b = make_nbr(n, a.g);
        T f[19];
        pull_all<P>((const S *)a.L, a.g.N, b, f);
        T rho9, ux9, uy9, uz9;
        macro_sums(f, rho9, ux9, uy9, uz9);
        if constexpr (WHAT == 0) {
            uint32_t ix, iy, iz;
            node_coords(n, a.g, ix, iy, iz);
            double fx, fy, fz;
            node_force(a.f, a.g, ix, iy, iz, fx, fy, fz);
            a.ux[n] = (double)(ux9 + (T)fx / T(2.));
            a.uy[n] = (double)(uy9 + (T)fy / T(2.));
            a.uz[n] = (double)(uz9 + (T)fz / T(2.));
            a.rho[n] = (double)(rho9 - (T)a.rho_off);
        } else {
            const double rnew = (double)rho9;
            err = fabs(rnew - a.rho[n]);
            a.rho[n] = rnew;
        }
    }
    if constexpr (WHAT == 1) {
        // max over the block, then one atomic per block, all on the bit
        // patterns: non-negative doubles order like their unsigned bit
        // patterns, and a NaN with the sign cleared (fabs) orders above
        // +inf, so a NaN anywhere becomes the result instead of being
        // dropped as fmax would drop it
        unsigned long long u = (unsigned long long)__double_as_longlong(err);
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long v = (unsigned long long)__shfl_xor((long long)u, o);
            u = v > u ? v : u;
        }
        __shared__ unsigned long long wmax[BLOCK / 64];
        if ((threadIdx.x & 63) == 0)
            wmax[threadIdx.x >> 6] = u;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = wmax[0];
            for (int w = 1; w < BLOCK / 64; w++)
                m = wmax[w] > m ? wmax[w] : m;
            atomicMax(a.errbits, m);
        }
    }
}

// Download: F_k of nodes [n0, n0+count) as reference-layout FP64 rows of 19.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_gather(CopyArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    __shared__ double tile[19][BLOCK];
    const uint64_t base = (uint64_t)blockIdx.x * BLOCK;
    const uint64_t n = a.n0 + base + threadIdx.x;
    if (base + threadIdx.x < a.count) {
        const Nbr b = make_nbr((uint32_t)n, a.g);
        T f[19];
        pull_all<P>((const S *)a.L, a.g.N, b, f);
#pragma unroll
        for (int i = 0; i < 19; i++)
            tile[i][threadIdx.x] = (double)f[i];
    }
    __syncthreads();
    const uint64_t nodes = (a.count - base < BLOCK) ? (a.count - base) : BLOCK;
    double *out = a.aos + base * 19;
    for (uint32_t e = threadIdx.x; e < nodes * 19; e += BLOCK)
        out[e] = tile[e % 19][e / 19];
}

// Upload: reference-layout FP64 rows of 19 (F values) for nodes
// [n0, n0+count), scattered to the post-collision slots they stream from.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_scatter(CopyArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    __shared__ double tile[19][BLOCK];
    const uint64_t base = (uint64_t)blockIdx.x * BLOCK;
    const uint64_t nodes = (a.count - base < BLOCK) ? (a.count - base) : BLOCK;
    const double *in = a.aos + base * 19;
    for (uint32_t e = threadIdx.x; e < nodes * 19; e += BLOCK)
        tile[e % 19][e / 19] = in[e];
    __syncthreads();
    if (base + threadIdx.x < a.count) {
        const uint64_t n = a.n0 + base + threadIdx.x;
        const Nbr b = make_nbr((uint32_t)n, a.g);
        T f[19];
#pragma unroll
        for (int i = 0; i < 19; i++)
            f[i] = (T)tile[i][threadIdx.x];
        scatter_all<P>((S *)a.L, a.g.N, b, f);
    }
}

// initpop, Channel-Flow/initial.f90:26-44, with rho = 0: F = equilibrium of
// the velocity arrays, scattered like an upload; rho array set to 0.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_initpop(InitArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    const uint32_t n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= a.g.N)
        return;
    const T ux = (T)a.ux[n], uy = (T)a.uy[n], uz = (T)a.uz[n];
    const T rho = T(0.0);
    T usqr = ux * ux + uy * uy + uz * uz;
    usqr = T(1.5) * usqr;
    T f[19];
    f[0] = T(1.0 / 3.0) * (rho - usqr);
#define D3Q19_EQ(I)                                                                     \
    {                                                                                   \
        constexpr T W = (I) <= 6 ? T(1.0 / 18.0) : T(1.0 / 36.0);                       \
        const T G = cdot<(I)>(ux, uy, uz);                                              \
        f[I] = W * (rho + T(3.0) * G + T(4.5) * G * G - usqr);                          \
    }
    D3Q19_EQ(1) D3Q19_EQ(2) D3Q19_EQ(3) D3Q19_EQ(4) D3Q19_EQ(5) D3Q19_EQ(6) D3Q19_EQ(7) D3Q19_EQ(8)
    D3Q19_EQ(9) D3Q19_EQ(10) D3Q19_EQ(11) D3Q19_EQ(12) D3Q19_EQ(13) D3Q19_EQ(14) D3Q19_EQ(15)
    D3Q19_EQ(16) D3Q19_EQ(17) D3Q19_EQ(18)
#undef D3Q19_EQ
    a.rho[n] = 0.0;
    const Nbr b = make_nbr(n, a.g);
    scatter_all<P>((S *)a.L, a.g.N, b, f);
}

// Deterministic block sum: each thread's partial, reduced in a fixed tree.
__device__ __forceinline__ double block_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    __shared__ double ws[BLOCK / 64];
    if ((threadIdx.x & 63) == 0)
        ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        s = ws[0];
        for (int w = 1; w < BLOCK / 64; w++)
            s += ws[w];
    }
    return s;
}

// Per-block partial sums of rho over nodes.  SRC 0: the rho array;
// SRC 1: rho = macrovar(F) computed from lattice L.  Grid-stride, fixed
// assignment of nodes to threads, so the result is reproducible.
template <int P, int SRC>
__global__ void __launch_bounds__(BLOCK) k_rho_partials(SumArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    double acc = 0.0, cnt = 0.0;
    for (uint64_t n = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; n < a.g.N; n += (uint64_t)gridDim.x * BLOCK) {
        if (a.solid && a.solid[n] >= 0)
            continue;
        cnt += 1.0;
        if constexpr (SRC == 0) {
            acc += a.rho[n];
        } else {
            const Nbr b = make_nbr((uint32_t)n, a.g);
            T f[19];
            pull_all<P>((const S *)a.L, a.g.N, b, f);
            T rho9, ux9, uy9, uz9;
            macro_sums(f, rho9, ux9, uy9, uz9);
            acc += (double)rho9;
        }
    }
    const double s = block_sum(acc);
    __syncthreads();
    const double c = block_sum(cnt);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = s;
        a.partials[gridDim.x + blockIdx.x] = c;
    }
}

// Per-block partial sums of every stored population (total mass).
template <int P, int... Is>
__device__ __forceinline__ double node_mass(const typename Prec<P>::S *__restrict__ L, uint64_t N, uint64_t n,
                                            std::integer_sequence<int, Is...>)
{
    double m = 0.0;
    ((m += (double)Prec<P>::template to_T<Is>(L[(uint64_t)Is * N + n])), ...);
    return m;
}

template <int P>
__global__ void __launch_bounds__(BLOCK) k_mass_partials(SumArgs a)
{
    using S = typename Prec<P>::S;
    double acc = 0.0;
    for (uint64_t n = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; n < a.g.N; n += (uint64_t)gridDim.x * BLOCK)
        if (!a.solid || a.solid[n] < 0)
            acc += node_mass<P>((const S *)a.L, a.g.N, n, std::make_integer_sequence<int, 19>{});
    const double s = block_sum(acc);
    if (threadIdx.x == 0)
        a.partials[blockIdx.x] = s;
}

// Plane sums for the statistics of Channel-Flow/saveload.f90:1382-1409:
// per wall-normal index ix, the sums over the fluid nodes (iy, iz) of
// ux, uy, uz, ux*uz, ux*uy, uy*uz, ux^2, uy^2, uz^2, rho, rho^2, and their count.
// Block (blockIdx.x, blockIdx.y) handles ix = blockIdx.y*blockDim.x + tid and
// rows blockIdx.x, blockIdx.x + gridDim.x, ...; partials[(c*11 + q)*lx + ix].
// SRC 0: field arrays; SRC 1: macrovar(F) of lattice L with rho - rho_off.
template <int P, int SRC>
__global__ void __launch_bounds__(BLOCK) k_plane_partials(PlaneArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    const uint32_t ix = blockIdx.y * blockDim.x + threadIdx.x;
    if (ix >= a.g.lx)
        return;
    double s[NPLANE] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t rows = a.g.ly * a.g.lz;
    for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const uint32_t n = row * a.g.lx + ix;
        if (a.solid && a.solid[n] >= 0)
            continue;
        double rho, ux, uy, uz;
        if constexpr (SRC == 0) {
            rho = a.rho[n];
            ux = a.ux[n];
            uy = a.uy[n];
            uz = a.uz[n];
        } else {
            const Nbr b = make_nbr(n, a.g);
            T f[19];
            pull_all<P>((const S *)a.L, a.g.N, b, f);
            T rho9, ux9, uy9, uz9;
            macro_sums(f, rho9, ux9, uy9, uz9);
            uint32_t cx, cy, cz;
            node_coords(n, a.g, cx, cy, cz);
            double fx, fy, fz;
            node_force(a.f, a.g, cx, cy, cz, fx, fy, fz);
            ux = (double)(ux9 + (T)fx / T(2.));
            uy = (double)(uy9 + (T)fy / T(2.));
            uz = (double)(uz9 + (T)fz / T(2.));
            rho = (double)(rho9 - (T)a.rho_off);
        }
        s[0] += ux;
        s[1] += uy;
        s[2] += uz;
        s[3] += ux * uz;
        s[4] += ux * uy;
        s[5] += uy * uz;
        s[6] += ux * ux;
        s[7] += uy * uy;
        s[8] += uz * uz;
        s[9] += rho;
        s[10] += rho * rho;
        s[11] += 1.0;
    }
    double *out = a.partials + (uint64_t)blockIdx.x * NPLANE * a.g.lx;
#pragma unroll
    for (int q = 0; q < NPLANE; q++)
        out[(uint64_t)q * a.g.lx + ix] = s[q];
}

// The same plane sums for the 2x2x2 box-filtered fields on the grid of half
// the size: coarse node (cx, cy, cz) carries the means of rho and u (macrovar
// of F, with F/2 and rho_off) over fine nodes (2cx+{0,1}, 2cy+{0,1},
// 2cz+{0,1}).  Partials are laid out for lx/2 planes.  Even lx, ly, lz.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_filtered_plane_partials(PlaneArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    const uint32_t lxc = a.g.lx / 2, lyc = a.g.ly / 2;
    const uint32_t ic = blockIdx.y * blockDim.x + threadIdx.x;
    if (ic >= lxc)
        return;
    double s[NPLANE] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t rows = lyc * (a.g.lz / 2);
    for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const uint32_t cy = row % lyc, cz = row / lyc;
        double rho = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
        for (uint32_t dz = 0; dz < 2; dz++)
            for (uint32_t dy = 0; dy < 2; dy++)
                for (uint32_t dx = 0; dx < 2; dx++) {
                    const uint32_t n = (2 * ic + dx) + a.g.lx * ((2 * cy + dy) + a.g.ly * (2 * cz + dz));
                    const Nbr b = make_nbr(n, a.g);
                    T f[19];
                    pull_all<P>((const S *)a.L, a.g.N, b, f);
                    T rho9, ux9, uy9, uz9;
                    macro_sums(f, rho9, ux9, uy9, uz9);
                    double fx, fy, fz;
                    node_force(a.f, a.g, 2 * ic + dx, 2 * cy + dy, 2 * cz + dz, fx, fy, fz);
                    ux += (double)(ux9 + (T)fx / T(2.));
                    uy += (double)(uy9 + (T)fy / T(2.));
                    uz += (double)(uz9 + (T)fz / T(2.));
                    rho += (double)(rho9 - (T)a.rho_off);
                }
        rho *= 0.125;
        ux *= 0.125;
        uy *= 0.125;
        uz *= 0.125;
        s[0] += ux;
        s[1] += uy;
        s[2] += uz;
        s[3] += ux * uz;
        s[4] += ux * uy;
        s[5] += uy * uz;
        s[6] += ux * ux;
        s[7] += uy * uy;
        s[8] += uz * uz;
        s[9] += rho;
        s[10] += rho * rho;
        s[11] += 1.0;
    }
    double *out = a.partials + (uint64_t)blockIdx.x * NPLANE * lxc;
#pragma unroll
    for (int q = 0; q < NPLANE; q++)
        out[(uint64_t)q * lxc + ic] = s[q];
}

// ---------------------------------------------------------------------------
template <int P>
static inline uint32_t nblocks(uint64_t n)
{
    return (uint32_t)((n + BLOCK - 1) / BLOCK);
}

// Does the vector-run step kernel apply to this launch?
// The fields of node n: from the arrays (SRC 0) or macrovar of the pulled
// populations with the fields' force and rho - rho_off (SRC 1), the values
// k_macro stores; f9 receives the pulled populations when wanted.
template <int P, int SRC>
__device__ __forceinline__ void node_fields(const void *L, const double *rho_a, const double *ux_a,
                                            const double *uy_a, const double *uz_a, const ForceSpec &fs,
                                            double rho_off, const Geom &g, uint32_t n, double &rho, double &ux,
                                            double &uy, double &uz, double *f9)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    if constexpr (SRC == 0) {
        rho = rho_a[n];
        ux = ux_a[n];
        uy = uy_a[n];
        uz = uz_a[n];
        if (f9) {
            const Nbr b = make_nbr(n, g);
            T f[19];
            pull_all<P>((const S *)L, g.N, b, f);
#pragma unroll
            for (int i = 0; i < 19; i++)
                f9[i] = (double)f[i];
        }
    } else {
        const Nbr b = make_nbr(n, g);
        T f[19];
        pull_all<P>((const S *)L, g.N, b, f);
        T rho9, ux9, uy9, uz9;
        macro_sums(f, rho9, ux9, uy9, uz9);
        uint32_t cx, cy, cz;
        node_coords(n, g, cx, cy, cz);
        double fx, fy, fz;
        node_force(fs, g, cx, cy, cz, fx, fy, fz);
        ux = (double)(ux9 + (T)fx / T(2.));
        uy = (double)(uy9 + (T)fy / T(2.));
        uz = (double)(uz9 + (T)fz / T(2.));
        rho = (double)(rho9 - (T)rho_off);
        if (f9) {
#pragma unroll
            for (int i = 0; i < 19; i++)
                f9[i] = (double)f[i];
        }
    }
}

// Larger |u| wins; on a tie the smaller node index (the reference's loop
// keeps the first strict maximum in x, y, z order).  idx < 0: none yet.
__device__ __forceinline__ void vmax_merge(double &v, double &idx, double ov, double oidx)
{
    if (ov > v || (ov == v && oidx >= 0.0 && (idx < 0.0 || oidx < idx))) {
        v = ov;
        idx = oidx;
    }
}

template <int P, int SRC>
__global__ void __launch_bounds__(BLOCK) k_diag_partials(DiagArgs a)
{
    double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double vmax = 0.0, vidx = -1.0, rmax = -INFINITY, rmin = INFINITY;
    for (uint64_t n = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; n < a.g.N; n += (uint64_t)gridDim.x * BLOCK) {
        if (a.solid && a.solid[n] >= 0)
            continue;
        double rho, ux, uy, uz;
        node_fields<P, SRC>(a.L, a.rho, a.ux, a.uy, a.uz, a.f, a.rho_off, a.g, (uint32_t)n, rho, ux, uy, uz,
                            nullptr);
        s[0] += ux;
        s[1] += uy;
        s[2] += uz;
        s[3] += ux * ux;
        s[4] += uy * uy;
        s[5] += uz * uz;
        s[6] += 1.0;
        const double vel = sqrt(ux * ux + uy * uy + uz * uz);
        if (vel > vmax) { // nodes come in increasing order: the first maximum stays
            vmax = vel;
            vidx = (double)n;
        }
        rmax = rho > rmax ? rho : rmax;
        rmin = rho < rmin ? rho : rmin;
    }
    double out[DIAG_Q];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        out[q] = block_sum(s[q]);
        __syncthreads();
    }
    // maxima: lanes by xor butterfly, then the waves in order
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(vmax, o), oi = __shfl_xor(vidx, o);
        vmax_merge(vmax, vidx, ov, oi);
        const double orx = __shfl_xor(rmax, o), orn = __shfl_xor(rmin, o);
        rmax = orx > rmax ? orx : rmax;
        rmin = orn < rmin ? orn : rmin;
    }
    __shared__ double wm[4][BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        wm[0][threadIdx.x >> 6] = vmax;
        wm[1][threadIdx.x >> 6] = vidx;
        wm[2][threadIdx.x >> 6] = rmax;
        wm[3][threadIdx.x >> 6] = rmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = wm[0][0], vi = wm[1][0], rx = wm[2][0], rn = wm[3][0];
        for (int w = 1; w < BLOCK / 64; w++) {
            vmax_merge(v, vi, wm[0][w], wm[1][w]);
            rx = wm[2][w] > rx ? wm[2][w] : rx;
            rn = wm[3][w] < rn ? wm[3][w] : rn;
        }
        out[7] = v;
        out[8] = vi;
        out[9] = rx;
        out[10] = rn;
#pragma unroll
        for (int q = 0; q < DIAG_Q; q++)
            a.partials[(uint64_t)q * gridDim.x + blockIdx.x] = out[q];
    }
}

// Per-node strain of dissipation_calc (Channel-Flow/saveload.f90:4677-4776)
// in the reference's operand order: v[0] = S_ij S_ij and v[1] = S_xy from
// the non-equilibrium moments, v[2], v[3] the same from the non-equilibrium
// populations.
__device__ inline void dissipation_node(const DissArgs &a, const double *f9, double rho9, double ux9, double uy9,
                                        double uz9, double *v)
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
    const double evlm1 = -30.0 * f9[0] + a.coef2 * sum1 + a.coef3 * sum2;
    const double evlm6 = a.coef5 * sum6 - sum7 + sum8 - a.coef5 * sum9;
    const double evlm8 = sum10 + sum11;
    const double evlm10 = f9[7] - f9[8] - f9[9] + f9[10];
    const double evlm11 = f9[15] - f9[16] - f9[17] + f9[18];
    const double evlm12 = f9[11] - f9[12] - f9[13] + f9[14];
    const double neqm1 = a.s1 * (evlm1 - eqm1);
    const double neqm9 = a.s9 * (evlm6 - eqm6);
    const double neqm11 = a.s9 * (evlm8 - eqm8);
    const double neqm13 = a.s9 * (evlm10 - eqm10);
    const double neqm14 = a.s9 * (evlm11 - eqm11);
    const double neqm15 = a.s9 * (evlm12 - eqm12);
    double Sxx = -(neqm1 + 19.0 * neqm9) / 38.0;
    double Syy = -(2.0 * neqm1 - 19.0 * (neqm9 - 3.0 * neqm11)) / 76.0;
    double Szz = -(2.0 * neqm1 - 19.0 * (neqm9 + 3.0 * neqm11)) / 76.0;
    double Sxy = -1.5 * neqm13;
    double Syz = -1.5 * neqm14;
    double Szx = -1.5 * neqm15;
    v[0] = Sxx * Sxx + Syy * Syy + Szz * Szz + 2.0 * (Sxy * Sxy + Syz * Syz + Szx * Szx);
    v[1] = Sxy;
    Sxx = 0.0;
    Syy = 0.0;
    Szz = 0.0;
    Sxy = 0.0;
    Syz = 0.0;
    Szx = 0.0;
    const double usqr = 1.5 * (ux9s + uy9s + uz9s);
#pragma unroll
    for (int ip = 1; ip < 19; ip++) {
        const double wght = ip <= 6 ? a.ww1 : a.ww2;
        const double edtu = (double)CX[ip] * ux9 + (double)CY[ip] * uy9 + (double)CZ[ip] * uz9;
        const double feq9 = wght * (rho9 + 3.0 * edtu + 4.5 * (edtu * edtu) - usqr);
        const double fneq9 = f9[ip] - feq9;
        Sxx = Sxx + fneq9 * (double)(CX[ip] * CX[ip]);
        Syy = Syy + fneq9 * (double)(CY[ip] * CY[ip]);
        Szz = Szz + fneq9 * (double)(CZ[ip] * CZ[ip]);
        Sxy = Sxy + fneq9 * (double)(CX[ip] * CY[ip]);
        Syz = Syz + fneq9 * (double)(CY[ip] * CZ[ip]);
        Szx = Szx + fneq9 * (double)(CZ[ip] * CX[ip]);
    }
    const double m2 = Sxx * Sxx + Syy * Syy + Szz * Szz + 2.0 * (Sxy * Sxy + Syz * Syz + Szx * Szx);
    v[2] = m2 * (a.c15tau * a.c15tau);
    v[3] = Sxy * a.c15tau;
}

// Plane sums of dissipation_calc: block (chunk, x tile) strides over rows.
template <int P, int SRC>
__global__ void __launch_bounds__(BLOCK) k_diss_partials(DissArgs a)
{
    const uint32_t ix = blockIdx.y * blockDim.x + threadIdx.x;
    if (ix >= a.g.lx)
        return;
    double s[NDISS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t rows = a.g.ly * a.g.lz;
    for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const uint32_t n = row * a.g.lx + ix;
        if (a.solid && a.solid[n] >= 0)
            continue;
        double f9[19], rho, ux, uy, uz, v[4];
        node_fields<P, SRC>(a.L, a.rho, a.ux, a.uy, a.uz, a.f, a.rho_off, a.g, n, rho, ux, uy, uz, f9);
        dissipation_node(a, f9, rho, ux, uy, uz, v);
        s[0] += v[0];
        s[1] += v[2];
        s[2] += v[1];
        s[3] += v[3];
        s[4] += 1.0;
    }
    double *out = a.partials + (uint64_t)blockIdx.x * NDISS * a.g.lx;
#pragma unroll
    for (int q = 0; q < NDISS; q++)
        out[(uint64_t)q * a.g.lx + ix] = s[q];
}

template <int P>
static inline bool use_runs(const StepArgs &a, int mode)
{
    if constexpr (half_storage(P)) {
        return false;
    } else {
        return a.runs && (mode == 0 || mode == 1) && a.g.lx % Vec<P>::W == 0 && a.g.N / Vec<P>::W < 0xffffffffull;
    }
}

template <int P>
hipError_t Launch<P>::step(const StepArgs &a, int mode, hipStream_t s)
{
    if constexpr (!half_storage(P)) {
        if (use_runs<P>(a, mode)) {
            const dim3 grid(nblocks<P>(a.g.N / Vec<P>::W)), blk(BLOCK);
            if (mode == 0)
                hipLaunchKernelGGL((k_step_run<P, 0>), grid, blk, 0, s, a);
            else
                hipLaunchKernelGGL((k_step_run<P, 1>), grid, blk, 0, s, a);
            return hipGetLastError();
        }
    }
    // one launch geometry per kernel's block size (step_block)
    auto geo = [&](int nb) { return std::make_pair(dim3((uint32_t)((a.g.N + nb - 1) / nb)), dim3(nb)); };
    if constexpr (half_storage(P) || P == 1) {
        // 32-bit pull offsets where the lattice allows them: the main loop's
        // modes in 2-byte storage (FP16 1.8-2.5 % faster, FP16C 2.7-2.9 %, on
        // the 192x768x384 channel; FP32's memory-bound step measured no
        // change), and the closure modes, which are issue-bound, in FP32 and
        // 2-byte storage
        if (a.w32 == 1 && a.g.N <= w32_max_n(P)) {
            switch (mode) {
            case 0:
            case 1:
            case 4:
            case 16:
                if constexpr (half_storage(P)) {
                    if (mode == 0)
                        hipLaunchKernelGGL((k_step<P, 32>), geo(step_block<P, 32>()).first,
                                           geo(step_block<P, 32>()).second, 0, s, a);
                    else if (mode == 1)
                        hipLaunchKernelGGL((k_step<P, 33>), geo(step_block<P, 33>()).first,
                                           geo(step_block<P, 33>()).second, 0, s, a);
                    else if (mode == 4)
                        hipLaunchKernelGGL((k_step<P, 36>), geo(step_block<P, 36>()).first,
                                           geo(step_block<P, 36>()).second, 0, s, a);
                    else
                        hipLaunchKernelGGL((k_step<P, 48>), geo(step_block<P, 48>()).first,
                                           geo(step_block<P, 48>()).second, 0, s, a);
                    return hipGetLastError();
                }
                break;
            case 8:
                hipLaunchKernelGGL((k_step_closure<P, 40>), geo(step_block<P, 40>()).first,
                                   geo(step_block<P, 40>()).second, 0, s, a);
                return hipGetLastError();
            case 12:
                hipLaunchKernelGGL((k_step_closure<P, 44>), geo(step_block<P, 44>()).first,
                                   geo(step_block<P, 44>()).second, 0, s, a);
                return hipGetLastError();
            case 24:
                hipLaunchKernelGGL((k_step_closure<P, 56>), geo(step_block<P, 56>()).first,
                                   geo(step_block<P, 56>()).second, 0, s, a);
                return hipGetLastError();
            default:
                break;
            }
        }
    }
    if constexpr (half_storage(P)) {
        // above 2^29 nodes (the 2048x1024x512 channel), the pulls between
        // adjacent rows only (step mode + 64)
        if (a.w32 && (a.g.N > w32_max_n(P) || a.w32 == 2) && a.g.N <= W32_ADJ_MAX_N) {
            switch (mode) {
            case 0:
                hipLaunchKernelGGL((k_step<P, 64>), geo(step_block<P, 64>()).first, geo(step_block<P, 64>()).second, 0,
                                   s, a);
                return hipGetLastError();
            case 1:
                hipLaunchKernelGGL((k_step<P, 65>), geo(step_block<P, 65>()).first, geo(step_block<P, 65>()).second, 0,
                                   s, a);
                return hipGetLastError();
            case 4:
                hipLaunchKernelGGL((k_step<P, 68>), geo(step_block<P, 68>()).first, geo(step_block<P, 68>()).second, 0,
                                   s, a);
                return hipGetLastError();
            case 16:
                hipLaunchKernelGGL((k_step<P, 80>), geo(step_block<P, 80>()).first, geo(step_block<P, 80>()).second, 0,
                                   s, a);
                return hipGetLastError();
            default:
                break;
            }
        }
    }
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_step<P, 0>), geo(step_block<P, 0>()).first, geo(step_block<P, 0>()).second, 0, s, a);
        break;
    case 1:
        hipLaunchKernelGGL((k_step<P, 1>), geo(step_block<P, 1>()).first, geo(step_block<P, 1>()).second, 0, s, a);
        break;
    case 2:
        hipLaunchKernelGGL((k_step<P, 2>), geo(step_block<P, 2>()).first, geo(step_block<P, 2>()).second, 0, s, a);
        break;
    case 4:
        hipLaunchKernelGGL((k_step<P, 4>), geo(step_block<P, 4>()).first, geo(step_block<P, 4>()).second, 0, s, a);
        break;
    case 5:
        hipLaunchKernelGGL((k_step<P, 5>), geo(step_block<P, 5>()).first, geo(step_block<P, 5>()).second, 0, s, a);
        break;
    case 6:
        hipLaunchKernelGGL((k_step<P, 6>), geo(step_block<P, 6>()).first, geo(step_block<P, 6>()).second, 0, s, a);
        break;
    case 8:
        if constexpr (P == 0)
            return hipErrorInvalidValue;
        else
            hipLaunchKernelGGL((k_step_closure<P, 8>), geo(step_block<P, 8>()).first, geo(step_block<P, 8>()).second, 0, s, a);
        break;
    case 12:
        if constexpr (P == 0)
            return hipErrorInvalidValue;
        else
            hipLaunchKernelGGL((k_step_closure<P, 12>), geo(step_block<P, 12>()).first, geo(step_block<P, 12>()).second, 0, s, a);
        break;
    case 16:
        hipLaunchKernelGGL((k_step<P, 16>), geo(step_block<P, 16>()).first, geo(step_block<P, 16>()).second, 0, s, a);
        break;
    case 17:
        hipLaunchKernelGGL((k_step<P, 17>), geo(step_block<P, 17>()).first, geo(step_block<P, 17>()).second, 0, s, a);
        break;
    case 18:
        hipLaunchKernelGGL((k_step<P, 18>), geo(step_block<P, 18>()).first, geo(step_block<P, 18>()).second, 0, s, a);
        break;
    case 24:
        if constexpr (P == 0)
            return hipErrorInvalidValue;
        else
            hipLaunchKernelGGL((k_step_closure<P, 24>), geo(step_block<P, 24>()).first, geo(step_block<P, 24>()).second, 0, s, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::step_planes(const BlockArgs &a, int axis, hipStream_t s)
{
    const uint32_t PB = a.nt ? plane_block<P, 1>() : plane_block<P, 0>();
    const dim3 grid((uint32_t)(((uint64_t)a.slabp * a.count + PB - 1) / PB)), blk(PB);
    const int v = (axis == 2 ? 4 : axis == 1 ? 0 : -8) + (a.nt ? 2 : 0) + (a.per_node ? 1 : 0);
    switch (v) {
    case 0: hipLaunchKernelGGL((k_step_planes<P, 1, 0, 0>), grid, blk, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_step_planes<P, 1, 0, 1>), grid, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_step_planes<P, 1, 1, 0>), grid, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_step_planes<P, 1, 1, 1>), grid, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_step_planes<P, 2, 0, 0>), grid, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_step_planes<P, 2, 0, 1>), grid, blk, 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_step_planes<P, 2, 1, 0>), grid, blk, 0, s, a); break;
    case 7: hipLaunchKernelGGL((k_step_planes<P, 2, 1, 1>), grid, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::closure_eval(const StepArgs &a, uint64_t n_begin, uint64_t count, float *out, int what,
                                   hipStream_t s)
{
    if constexpr (P == 0) {
        return hipErrorInvalidValue; // the closure runs in FP32 arithmetic only
    } else {
        const dim3 grid(nblocks<P>(count)), blk(BLOCK);
        if (what == 0)
            hipLaunchKernelGGL((k_closure_eval<P, 0>), grid, blk, 0, s, a, n_begin, count, out);
        else
            hipLaunchKernelGGL((k_closure_eval<P, 1>), grid, blk, 0, s, a, n_begin, count, out);
        return hipGetLastError();
    }
}

template <int P>
hipError_t Launch<P>::macro(const MacroArgs &a, int what, hipStream_t s)
{
    const dim3 grid(nblocks<P>(a.g.N)), blk(BLOCK);
    if (what == 0)
        hipLaunchKernelGGL((k_macro<P, 0>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((k_macro<P, 1>), grid, blk, 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::gather(const CopyArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_gather<P>), dim3(nblocks<P>(a.count)), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::scatter(const CopyArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_scatter<P>), dim3(nblocks<P>(a.count)), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::initpop(const InitArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_initpop<P>), dim3(nblocks<P>(a.g.N)), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::rho_partials(const SumArgs &a, int src, uint32_t grid, hipStream_t s)
{
    if (src == 0)
        hipLaunchKernelGGL((k_rho_partials<P, 0>), dim3(grid), dim3(BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL((k_rho_partials<P, 1>), dim3(grid), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::mass_partials(const SumArgs &a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL((k_mass_partials<P>), dim3(grid), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::plane_partials(const PlaneArgs &a, int src, uint32_t chunks, hipStream_t s)
{
    const uint32_t bx = a.g.lx >= BLOCK ? BLOCK : ((a.g.lx + 63) / 64) * 64;
    const dim3 grid(chunks, (a.g.lx + bx - 1) / bx), blk(bx);
    if (src == 0)
        hipLaunchKernelGGL((k_plane_partials<P, 0>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((k_plane_partials<P, 1>), grid, blk, 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::filtered_plane_partials(const PlaneArgs &a, uint32_t chunks, hipStream_t s)
{
    const uint32_t lxc = a.g.lx / 2;
    const uint32_t bx = lxc >= BLOCK ? BLOCK : ((lxc + 63) / 64) * 64;
    const dim3 grid(chunks, (lxc + bx - 1) / bx), blk(bx);
    hipLaunchKernelGGL((k_filtered_plane_partials<P>), grid, blk, 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::diag_partials(const DiagArgs &a, int src, uint32_t grid, hipStream_t s)
{
    if (src == 0)
        hipLaunchKernelGGL((k_diag_partials<P, 0>), dim3(grid), dim3(BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL((k_diag_partials<P, 1>), dim3(grid), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int P>
hipError_t Launch<P>::dissipation_partials(const DissArgs &a, int src, uint32_t chunks, hipStream_t s)
{
    const uint32_t bx = a.g.lx >= BLOCK ? BLOCK : ((a.g.lx + 63) / 64) * 64;
    const dim3 grid(chunks, (a.g.lx + bx - 1) / bx), blk(bx);
    if (src == 0)
        hipLaunchKernelGGL((k_diss_partials<P, 0>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((k_diss_partials<P, 1>), grid, blk, 0, s, a);
    return hipGetLastError();
}

} // namespace d3q19
