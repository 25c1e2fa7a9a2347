// D3Q19 MRT node arithmetic for gfx950, shared by the FP64 and
// FP32/FP16/FP16C kernel translation units.
//
// mrt_collide() follows Channel-Flow/collision.f90:60-220 operation by
// operation: every sum is evaluated left to right as written there and every
// product keeps the reference's operand grouping, so with FMA contraction off
// (the FP64 translation unit is built with -ffp-contract=off) the FP64 result
// is the reference's bit for bit.  Two value-preserving simplifications are
// made: products of a population or force with the lattice integers
// cix/ciy/ciz in {-1,0,1} keep only the nonzero terms (x*1 = x, x*(-1) = -x and
// adding an exact zero leave every nonzero value unchanged), and products of
// two structural constants of para.f90:143-174 (e.g. coef2*val2i) are folded
// at compile time, where IEEE round-to-nearest gives the same double the
// reference forms at run time.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace d3q19 {

// Velocity set in d'Humieres et al. (2002) order, Channel-Flow/para.f90:201-206.
constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

// Unsigned division by a run-time constant d via a multiply-high, valid for
// numerators below 2^31 (every node index here is below 2^31).
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    uint32_t s = 0;
    while (s < 32 && (1ull << s) < d)
        s++;
    uint64_t m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
    return FastDiv{d, (uint32_t)m, s};
}

__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f)
{
    const uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
    return (t + n) >> f.s;
}

// Grid geometry of one rank's block: x wall-normal (fastest in memory),
// y streamwise, z spanwise.  Populations are stored structure-of-arrays:
// element (i, n) at i*N + n with n = ix + lx*(iy + ly*iz).
struct Geom {
    uint32_t lx, ly, lz, plane; // plane = lx*ly
    uint64_t N;
    FastDiv dlx, dly;
    uint32_t remap; // 1: map consecutive node blocks to the same XCD (measured
                    // slower for the step kernel: off by default)
};

// Block index -> logical block index so that blocks dispatched to the same
// XCD (round-robin, blockIdx % 8) cover one contiguous range of nodes.
__host__ __device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Per-node neighbour offsets (mod 2^32) and wall flags.
struct Nbr {
    uint32_t n;
    uint32_t ym, yp, zm, zp; // offsets to the iy-1, iy+1, iz-1, iz+1 neighbours (wrapped)
    bool wlo, whi;           // ix == 0, ix == lx-1
};

__host__ __device__ __forceinline__ Nbr make_nbr(uint32_t n, const Geom &g)
{
    Nbr b;
    b.n = n;
    const uint32_t q = fdiv(n, g.dlx);
    const uint32_t ix = n - q * g.lx;
    const uint32_t iz = fdiv(q, g.dly);
    const uint32_t iy = q - iz * g.ly;
    b.ym = (iy == 0) ? (g.ly - 1) * g.lx : (uint32_t)(-(int32_t)g.lx);
    b.yp = (iy == g.ly - 1) ? (uint32_t)(-(int32_t)((g.ly - 1) * g.lx)) : g.lx;
    b.zm = (iz == 0) ? (g.lz - 1) * g.plane : (uint32_t)(-(int32_t)g.plane);
    b.zp = (iz == g.lz - 1) ? (uint32_t)(-(int32_t)((g.lz - 1) * g.plane)) : g.plane;
    b.wlo = (ix == 0);
    b.whi = (ix == g.lx - 1);
    return b;
}

// Index of x - c_I (the pull source of direction I), ignoring walls.
template <int I>
__host__ __device__ __forceinline__ uint32_t src_of(const Nbr &b)
{
    uint32_t s = b.n;
    if constexpr (CX[I] == 1)
        s -= 1u;
    else if constexpr (CX[I] == -1)
        s += 1u;
    if constexpr (CY[I] == 1)
        s += b.ym;
    else if constexpr (CY[I] == -1)
        s += b.yp;
    if constexpr (CZ[I] == 1)
        s += b.zm;
    else if constexpr (CZ[I] == -1)
        s += b.zp;
    return s;
}

// Index of the node shifted by -c_I in y and z only (the start of the
// aligned source run that the vector kernels load for direction I).
template <int I>
__host__ __device__ __forceinline__ uint32_t src_yz(const Nbr &b)
{
    uint32_t s = b.n;
    if constexpr (CY[I] == 1)
        s += b.ym;
    else if constexpr (CY[I] == -1)
        s += b.yp;
    if constexpr (CZ[I] == 1)
        s += b.zm;
    else if constexpr (CZ[I] == -1)
        s += b.zp;
    return s;
}

// Is x - c_I beyond a wall?
template <int I>
__host__ __device__ __forceinline__ bool walled(const Nbr &b)
{
    if constexpr (CX[I] == 1)
        return b.wlo;
    else if constexpr (CX[I] == -1)
        return b.whi;
    else
        return false;
}

// Relaxation rates and body force: the run-time constants of
// Channel-Flow/para.f90:106-141 and collision.f90:522-524.
template <typename T>
struct Rates {
    T s1, s2, s4, s9, s10, s13, s16;
    T omegepsl, omegepslj, omegxx;
    T fx, fy, fz;
};

// Storage precision traits.  Stores are non-temporal (the next read of a
// written value is one full lattice sweep later).  FP16 keeps f_i * 2^k with 2^k close to 1/w_i
// (4, 16, 32), so all directions use the same part of the half range;
// scaling by a power of two is exact.
template <int P>
struct Prec;

template <>
struct Prec<0> {
    using S = double;
    using T = double;
    template <int I>
    static constexpr int scale_class() { return 0; }
    template <int I>
    static __device__ __forceinline__ void st(S *p, T v) { __builtin_nontemporal_store(v, p); }
    template <int I>
    static __device__ __forceinline__ T to_T(S v) { return v; }
    template <int I>
    static __device__ __forceinline__ S to_S(T v) { return v; }
};

template <>
struct Prec<1> {
    using S = float;
    using T = float;
    template <int I>
    static constexpr int scale_class() { return 0; }
    template <int I>
    static __device__ __forceinline__ void st(S *p, T v) { __builtin_nontemporal_store(v, p); }
    template <int I>
    static __device__ __forceinline__ T to_T(S v) { return v; }
    template <int I>
    static __device__ __forceinline__ S to_S(T v) { return v; }
};

template <int I>
constexpr float fp16_scale() { return I == 0 ? 4.0f : (I <= 6 ? 16.0f : 32.0f); }

// fp16_scale<I>() for a direction known only at run time.
__host__ __device__ __forceinline__ float fp16_scale_rt(int i) { return i == 0 ? 4.0f : (i <= 6 ? 16.0f : 32.0f); }

template <>
struct Prec<2> {
    using S = _Float16;
    using T = float;
    template <int I>
    static __device__ __forceinline__ T to_T(S v) { return (float)v * (1.0f / fp16_scale<I>()); }
    template <int I>
    static __device__ __forceinline__ S to_S(T v) { return (_Float16)(v * fp16_scale<I>()); }
    template <int I>
    static constexpr int scale_class() { return I == 0 ? 0 : (I <= 6 ? 1 : 2); }
    template <int I>
    static __device__ __forceinline__ void st(S *p, T v) { __builtin_nontemporal_store(to_S<I>(v), p); }
    // run-time direction (particle links, refill)
    static __device__ __forceinline__ T to_T_rt(S v, int i)
    {
        return (float)v * (i == 0 ? 0.25f : (i <= 6 ? 1.0f / 16.0f : 1.0f / 32.0f));
    }
    static __device__ __forceinline__ S to_S_rt(T v, int i) { return (S)(v * fp16_scale_rt(i)); }
};

// FP16C storage (P = 3): the same scaled values as FP16 storage (f_i 2^k) in
// a 16-bit floating-point format with 1 sign, 4 exponent and 11 significand
// bits ("FP16C" of Lehmann et al., Phys. Rev. E 106, 015308, 2022).  Normal
// numbers span [2^-13, 4 - 2^-10] with one significand bit more than IEEE
// FP16; below 2^-13 they continue as denormals down to 2^-24, as FP16's do.
// The conversions go through FP32's own denormal range: a value times
// 2^-113 carries the format's exponent in its FP32 exponent field, so the
// 16-bit word is the FP32 bits shifted right by 12, rounded to nearest even,
// and decoding shifts them back and multiplies by 2^113 (both exact).
// Magnitudes above the largest value, and NaN, saturate to it (sign kept).
constexpr float FP16C_DOWN = 9.629649721936179e-35f; // 2^-113
constexpr float FP16C_UP = 1.0384593717069655e34f;   // 2^113
constexpr uint32_t FP16C_MAXBITS = 0x07FFF000u;      // FP32 bits of the largest value times 2^-113

// 16-bit word of y, a value already multiplied by 2^-113.  The magnitude is
// saturated as a float (one v_min_f32 with the |y| source modifier; FP32
// denormals are kept, and minNum turns NaN into the largest value); the
// rounded 15 bits are formed in the upper half of a 32-bit word beside the
// sign, so the store takes that half (global_store_short_d16_hi).
__host__ __device__ __forceinline__ uint16_t fp16c_pack(float y)
{
    const uint32_t bits = __builtin_bit_cast(uint32_t, y);
    const uint32_t a = __builtin_bit_cast(uint32_t, fminf(fabsf(y), __builtin_bit_cast(float, FP16C_MAXBITS)));
    const uint32_t t = a + 0x7ffu + ((a >> 12) & 1u);
    return (uint16_t)(((t << 4) | (bits & 0x80000000u)) >> 16);
}

// The value of a 16-bit word times 2^-113 (multiply by FP16C_UP for it).  The
// word is sign-extended (a global_load_sshort), so after the shift by 12 the
// sign is in bit 31 and its copies in bits 27-30 are masked off.
__host__ __device__ __forceinline__ float fp16c_unpack(uint16_t h)
{
    const uint32_t s = (uint32_t)(int32_t)(int16_t)h;
    return __builtin_bit_cast(float, (s << 12) & 0x87ffffffu);
}

template <>
struct Prec<3> {
    using S = uint16_t;
    using T = float;
    template <int I>
    static __device__ __forceinline__ T to_T(S v) { return fp16c_unpack(v) * (FP16C_UP / fp16_scale<I>()); }
    template <int I>
    static __device__ __forceinline__ S to_S(T v) { return fp16c_pack(v * (fp16_scale<I>() * FP16C_DOWN)); }
    template <int I>
    static constexpr int scale_class() { return I == 0 ? 0 : (I <= 6 ? 1 : 2); }
    template <int I>
    static __device__ __forceinline__ void st(S *p, T v) { __builtin_nontemporal_store(to_S<I>(v), p); }
    static __device__ __forceinline__ T to_T_rt(S v, int i) { return fp16c_unpack(v) * (FP16C_UP / fp16_scale_rt(i)); }
    static __device__ __forceinline__ S to_S_rt(T v, int i) { return fp16c_pack(v * (fp16_scale_rt(i) * FP16C_DOWN)); }
};

// 2-byte storage (FP16, FP16C).
constexpr bool half_storage(int P) { return P == 2 || P == 3; }

// Lattice velocity component times a value: c*a for c in {-1,0,1}.
template <int C, typename T>
__host__ __device__ __forceinline__ T cmul(T a) { return C > 0 ? a : -a; }

// cix*a + ciy*b + ciz*c (collision.f90:73-74, 79-80) keeping only the
// nonzero terms, summed in x, y, z order.
template <int I, typename T>
__host__ __device__ __forceinline__ T cdot(T a, T b, T c)
{
    constexpr int cx = CX[I], cy = CY[I], cz = CZ[I];
    if constexpr (cx != 0 && cy != 0)
        return cmul<cx>(a) + cmul<cy>(b);
    else if constexpr (cx != 0 && cz != 0)
        return cmul<cx>(a) + cmul<cz>(c);
    else if constexpr (cy != 0 && cz != 0)
        return cmul<cy>(b) + cmul<cz>(c);
    else if constexpr (cx != 0)
        return cmul<cx>(a);
    else if constexpr (cy != 0)
        return cmul<cy>(b);
    else if constexpr (cz != 0)
        return cmul<cz>(c);
    else
        return T(0);
}

// Guo forcing term Fbar_i, collision.f90:70-82.
template <int I, typename T>
__host__ __device__ __forceinline__ T fbar(T ux9, T uy9, T uz9, T fx9, T fy9, T fz9, T G3)
{
    if constexpr (I == 0) {
        return -G3;
    } else {
        constexpr T W = I <= 6 ? T(1.0 / 18.0) : T(1.0 / 36.0); // ww1, ww2: para.f90:173-174
        T G1 = cdot<I>(fx9, fy9, fz9);
        T G2 = cdot<I>(ux9, uy9, uz9);
        return W * (T(3.) * G1 + T(9.) * G1 * G2 - T(3.) * G3);
    }
}

// macrovar, collision.f90:398-413: momentum sums and density in the
// reference's order (before adding F/2).
template <typename T>
__host__ __device__ __forceinline__ void macro_sums(const T (&f9)[19], T &rho9, T &ux9, T &uy9, T &uz9)
{
    T sum1 = f9[7] - f9[10];
    T sum2 = f9[9] - f9[8];
    T sum3 = f9[11] - f9[14];
    T sum4 = f9[13] - f9[12];
    T sum5 = f9[15] - f9[18];
    T sum6 = f9[17] - f9[16];
    ux9 = f9[1] - f9[2] + sum1 + sum2 + sum3 + sum4;
    uy9 = f9[3] - f9[4] + sum1 - sum2 + sum5 + sum6;
    uz9 = f9[5] - f9[6] + sum3 - sum4 + sum5 - sum6;
    rho9 = f9[0] + f9[1] + f9[2] + f9[3] + f9[4] + f9[5] + f9[6] + f9[7] + f9[8] + f9[9] + f9[10] + f9[11] +
           f9[12] + f9[13] + f9[14] + f9[15] + f9[16] + f9[17] + f9[18];
}

// No closure: the shear rates stay s9 and s13 of para.f90:106-108.
struct NoClosure {
    template <typename T>
    __host__ __device__ __forceinline__ void operator()(const T (&)[9], T &, T &) const {}
};
// Guo forcing terms Fbar_0..18 of one node, collision.f90:60-82.
template <typename T>
__host__ __device__ __forceinline__ void guo_terms(T ux9, T uy9, T uz9, T fx9, T fy9, T fz9, T (&Fb)[19])
{
    const T G3 = ux9 * fx9 + uy9 * fy9 + uz9 * fz9; // :68
    Fb[0] = fbar<0>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[1] = fbar<1>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[2] = fbar<2>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[3] = fbar<3>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[4] = fbar<4>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[5] = fbar<5>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[6] = fbar<6>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[7] = fbar<7>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[8] = fbar<8>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[9] = fbar<9>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[10] = fbar<10>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[11] = fbar<11>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[12] = fbar<12>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[13] = fbar<13>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[14] = fbar<14>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[15] = fbar<15>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[16] = fbar<16>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[17] = fbar<17>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
    Fb[18] = fbar<18>(ux9, uy9, uz9, fx9, fy9, fz9, G3);
}

// The 15 non-conserved moments evlm1..15 of f9 (collision.f90:103-138), in
// d'Humieres et al. (2002) order: e, eps, qx, qy, qz, 3pxx, 3pixx, pww, piww,
// pxy, pyz, pxz, mx, my, mz.
template <typename T>
__host__ __device__ __forceinline__ void moments(const T (&f9)[19], T (&m)[15])
{
    // :103-119
    const T sum1 = f9[1] + f9[2] + f9[3] + f9[4] + f9[5] + f9[6];
    const T sum2 = f9[7] + f9[8] + f9[9] + f9[10] + f9[11] + f9[12] + f9[13] + f9[14] + f9[15] + f9[16] +
                   f9[17] + f9[18];
    const T sum3 = f9[7] - f9[8] + f9[9] - f9[10] + f9[11] - f9[12] + f9[13] - f9[14];
    const T sum4 = f9[7] + f9[8] - f9[9] - f9[10] + f9[15] - f9[16] + f9[17] - f9[18];
    const T sum5 = f9[11] + f9[12] - f9[13] - f9[14] + f9[15] + f9[16] - f9[17] - f9[18];
    const T sum6 = f9[1] + f9[2];
    const T sum7 = f9[3] + f9[4] + f9[5] + f9[6];
    const T sum8 = f9[7] + f9[8] + f9[9] + f9[10] + f9[11] + f9[12] + f9[13] + f9[14];
    const T sum9 = f9[15] + f9[16] + f9[17] + f9[18];
    const T sum10 = f9[3] + f9[4] - f9[5] - f9[6];
    const T sum11 = f9[7] + f9[8] + f9[9] + f9[10] - f9[11] - f9[12] - f9[13] - f9[14];

    // :121-138 (coef2 = -11, coef3 = 8, coef4 = -4, coef5 = 2)
    m[0] = -(T(30.0) * f9[0]) + T(-11.0) * sum1 + T(8.0) * sum2;
    m[1] = T(12.0) * f9[0] + T(-4.0) * sum1 + sum2;
    m[2] = T(-4.0) * (f9[1] - f9[2]) + sum3;
    m[3] = T(-4.0) * (f9[3] - f9[4]) + sum4;
    m[4] = T(-4.0) * (f9[5] - f9[6]) + sum5;
    m[5] = T(2.0) * sum6 - sum7 + sum8 - T(2.0) * sum9;
    m[6] = T(-4.0) * sum6 + T(2.0) * sum7 + sum8 - T(2.0) * sum9;
    m[7] = sum10 + sum11;
    m[8] = -(T(2.0) * sum10) + sum11;
    m[9] = f9[7] - f9[8] - f9[9] + f9[10];
    m[10] = f9[15] - f9[16] - f9[17] + f9[18];
    m[11] = f9[11] - f9[12] - f9[13] + f9[14];
    m[12] = f9[7] - f9[8] + f9[9] - f9[10] - f9[11] + f9[12] - f9[13] + f9[14];
    m[13] = -f9[7] - f9[8] + f9[9] + f9[10] + f9[15] - f9[16] + f9[17] - f9[18];
    m[14] = f9[11] + f9[12] - f9[13] - f9[14] - f9[15] - f9[16] + f9[17] + f9[18];
}

// Equilibria eqm1..12 of the non-conserved moments, collision.f90:86-101
// (incompressible: no rho divisor; eqm13..15 = 0).
template <typename T>
__host__ __device__ __forceinline__ void equilibria(T rho9, T ux9, T uy9, T uz9, const Rates<T> &r, T (&e)[12])
{
    const T ux9s = ux9 * ux9; // :60-62
    const T uy9s = uy9 * uy9;
    const T uz9s = uz9 * uz9;
    const T t1 = ux9s + uy9s + uz9s;
    constexpr T coef1 = T(-(2.0 / 3.0));
    e[0] = -(T(11.0) * rho9) + T(19.0) * t1;
    e[1] = r.omegepsl * rho9 + r.omegepslj * t1;
    e[2] = coef1 * ux9;
    e[3] = coef1 * uy9;
    e[4] = coef1 * uz9;
    e[5] = T(2.0) * ux9s - uy9s - uz9s;
    e[6] = r.omegxx * e[5];
    e[7] = uy9s - uz9s;
    e[8] = r.omegxx * e[7];
    e[9] = ux9 * uy9;
    e[10] = uy9 * uz9;
    e[11] = ux9 * uz9;
}

// Inverse transform, collision.f90:157-220: populations from the conserved
// moments c = (rho, jx, jy, jz) and the 15 others x (same order as
// moments()), through the orthogonal basis (val*i = 1/val*, coef3i = 1/8,
// coef4i = -1/4).
template <typename T>
__host__ __device__ __forceinline__ void inverse(T rho, T jx, T jy, T jz, const T (&x)[15], T (&out)[19])
{
    const T tl1 = T(1.0 / 19.0) * rho;
    const T tl2 = T(-11.0 * (1.0 / 2394.0)) * x[0];
    const T tl3 = T(8.0 * (1.0 / 2394.0)) * x[0];
    const T tl4 = T(-4.0 * (1.0 / 252.0)) * x[1];
    const T tl5 = T(1.0 / 252.0) * x[1];
    const T tl6 = T(1.0 / 10.0) * jx;
    const T tl7 = T(1.0 / 40.0) * x[2];
    const T tl8 = T(1.0 / 10.0) * jy;
    const T tl9 = T(1.0 / 40.0) * x[3];
    const T tl10 = T(1.0 / 10.0) * jz;
    const T tl11 = T(1.0 / 40.0) * x[4];
    const T tl12 = T(1.0 / 36.0) * x[5];
    const T tl13 = T(1.0 / 72.0) * x[6];
    const T tl14 = T(1.0 / 12.0) * x[7];
    const T tl15 = T(1.0 / 24.0) * x[8];
    const T tl16 = -(T(-0.25) * x[9]);
    const T tl17 = -(T(-0.25) * x[10]);
    const T tl18 = -(T(-0.25) * x[11]);
    const T tl19 = T(0.125) * x[12];
    const T tl20 = T(0.125) * x[13];
    const T tl21 = T(0.125) * x[14];

    // :180-198
    out[0] = tl1 - T(30.0 * (1.0 / 2394.0)) * x[0] + T(12.0 * (1.0 / 252.0)) * x[1];
    const T suma = tl1 + tl2 + tl4;
    const T sumb = tl1 + tl3 + tl5;
    const T sumc = tl6 + T(-4.0) * tl7;
    const T sumd = T(2.0) * tl12 + T(-4.0) * tl13;
    const T sume = tl8 + T(-4.0) * tl9;
    const T sumf = -tl12 + T(2.0) * tl13 + tl14 - T(2.0) * tl15;
    const T sumg = tl10 + T(-4.0) * tl11;
    const T sumh = -tl12 + T(2.0) * tl13 - tl14 + T(2.0) * tl15;
    const T sumi = tl12 + tl13 + tl14 + tl15;
    const T sumk = tl12 + tl13 - tl14 - tl15;
    const T sump = -(T(2.0) * tl12) - T(2.0) * tl13;
    const T sum67 = tl6 + tl7;
    const T sum89 = tl8 + tl9;
    const T sum1011 = tl10 + tl11;

    // :200-220
    out[1] = suma + sumc + sumd;
    out[2] = suma - sumc + sumd;
    out[3] = suma + sume + sumf;
    out[4] = suma - sume + sumf;
    out[5] = suma + sumg + sumh;
    out[6] = suma - sumg + sumh;
    out[7] = sumb + sum67 + sum89 + sumi + tl16 + tl19 - tl20;
    out[8] = sumb - sum67 + sum89 + sumi - tl16 - tl19 - tl20;
    out[9] = sumb + sum67 - sum89 + sumi - tl16 + tl19 + tl20;
    out[10] = sumb - sum67 - sum89 + sumi + tl16 - tl19 + tl20;
    out[11] = sumb + sum67 + sum1011 + sumk + tl18 - tl19 + tl21;
    out[12] = sumb - sum67 + sum1011 + sumk - tl18 + tl19 + tl21;
    out[13] = sumb + sum67 - sum1011 + sumk - tl18 - tl19 - tl21;
    out[14] = sumb - sum67 - sum1011 + sumk + tl18 + tl19 - tl21;
    out[15] = sumb + sum89 + sum1011 + sump + tl17 + tl20 - tl21;
    out[16] = sumb - sum89 + sum1011 + sump - tl17 - tl20 - tl21;
    out[17] = sumb + sum89 - sum1011 + sump - tl17 + tl20 + tl21;
    out[18] = sumb - sum89 - sum1011 + sump + tl17 - tl20 + tl21;
}

// Relaxation rate of each of the 15 non-conserved moments (moments() order).
template <typename T>
__host__ __device__ __forceinline__ T rate_of(int k, const Rates<T> &r, T s9, T s13)
{
    switch (k) {
    case 0: return r.s1;
    case 1: return r.s2;
    case 2: case 3: case 4: return r.s4;
    case 5: case 7: return s9;
    case 6: case 8: return r.s10;
    case 9: case 10: case 11: return s13;
    default: return r.s16;
    }
}

// Non-equilibrium parts of the nine moments a closure sees (relaxed with s1,
// s2, s9, s10, s13): e, eps, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz.
template <typename T>
__host__ __device__ __forceinline__ void closure_inputs(const T (&m)[15], const T (&e)[12], T (&neq)[9])
{
    neq[0] = m[0] - e[0];
    neq[1] = m[1] - e[1];
    neq[2] = m[5] - e[5];
    neq[3] = m[6] - e[6];
    neq[4] = m[7] - e[7];
    neq[5] = m[8] - e[8];
    neq[6] = m[9] - e[9];
    neq[7] = m[10] - e[10];
    neq[8] = m[11] - e[11];
}

// One MRT collision with Guo forcing, collision.f90:60-220, in the
// reference's form.  On entry f holds the node's populations; on exit the
// stored post-collision values f9(ip) + 0.5*Fbar(ip) (the values the swap
// loop writes, :230-262).  A closure sees the non-equilibrium parts of the
// nine moments relaxed with s1, s2, s9, s10, s13 and may replace this node's
// s9 and s13.
template <typename T, typename Closure = NoClosure>
__host__ __device__ __forceinline__ void mrt_collide(T (&f)[19], T rho9, T ux9, T uy9, T uz9, const Rates<T> &r,
                                            const Closure &closure = Closure())
{
    T Fb[19];
    guo_terms(ux9, uy9, uz9, r.fx, r.fy, r.fz, Fb);

    T f9[19];
#pragma unroll
    for (int i = 0; i < 19; i++) // :84
        f9[i] = f[i] + T(0.5) * Fb[i];

    T eq[12];
    equilibria(rho9, ux9, uy9, uz9, r, eq);
    T m[15];
    moments(f9, m);

    T s9 = r.s9, s13 = r.s13;
    if constexpr (!std::is_same<Closure, NoClosure>::value) {
        T neq[9];
        closure_inputs(m, eq, neq);
        closure(neq, s9, s13);
    }

    // :140-154 relaxation (eqm13..15 = 0)
    T post[15];
#pragma unroll
    for (int k = 0; k < 12; k++)
        post[k] = m[k] - rate_of(k, r, s9, s13) * (m[k] - eq[k]);
#pragma unroll
    for (int k = 12; k < 15; k++)
        post[k] = m[k] - r.s16 * (m[k] - T(0.0));

    // :157-166: the conserved moments come from the field values
    inverse(rho9, ux9, uy9, uz9, post, f9);

#pragma unroll
    for (int i = 0; i < 19; i++) // the stored value, :230-262
        f[i] = f9[i] + T(0.5) * Fb[i];
}

// Moments of the Guo term Fbar_i = w_i (3 c.F + 9 (c.F)(c.u) - 3 u.F)
// (guo_terms), sum_i M_ki Fbar_i in moments() order for k < 12 (the third-
// order moments mx, my, mz of Fbar vanish; sum_i Fbar_i = 0 and
// sum_i c_i Fbar_i = F).  Lattice sums of w_i times even powers of c_i give
//   e 38 u.F, eps -11 u.F, q -2F/3, 3pxx 2 G, 3pixx -G, pww 2 H, piww -H,
//   pxy ux Fy + uy Fx, pyz uy Fz + uz Fy, pxz ux Fz + uz Fx
// with G = 2 ux Fx - uy Fy - uz Fz and H = uy Fy - uz Fz.
template <typename T>
__host__ __device__ __forceinline__ void guo_moments(T ux9, T uy9, T uz9, T fx9, T fy9, T fz9, T (&s)[12])
{
    const T gx = ux9 * fx9, gy = uy9 * fy9, gz = uz9 * fz9;
    const T uf = gx + gy + gz;
    const T G = T(2.0) * gx - gy - gz;
    const T H = gy - gz;
    constexpr T coef1 = T(-(2.0 / 3.0));
    s[0] = T(38.0) * uf;
    s[1] = T(-11.0) * uf;
    s[2] = coef1 * fx9;
    s[3] = coef1 * fy9;
    s[4] = coef1 * fz9;
    s[5] = T(2.0) * G;
    s[6] = -G;
    s[7] = T(2.0) * H;
    s[8] = -H;
    s[9] = ux9 * fy9 + uy9 * fx9;
    s[10] = uy9 * fz9 + uz9 * fy9;
    s[11] = ux9 * fz9 + uz9 * fx9;
}

// The moments a collision relaxes, for the increment form below: m = the 15
// non-conserved moments of f9 = f + Fbar/2, formed as moments(f) + S/2 with S
// the moments of Fbar (guo_moments), and the equilibria.
template <typename T>
__host__ __device__ __forceinline__ void delta_moments(const T (&f)[19], T rho9, T ux9, T uy9, T uz9, const Rates<T> &r,
                                                       T (&m)[15], T (&sf)[12], T (&eq)[12])
{
    moments(f, m);
    guo_moments(ux9, uy9, uz9, r.fx, r.fy, r.fz, sf);
#pragma unroll
    for (int k = 0; k < 12; k++)
        m[k] = m[k] + T(0.5) * sf[k];
    equilibria(rho9, ux9, uy9, uz9, r, eq);
}

// The same collision written as an increment, in moment space, for FP32
// arithmetic:
//
//     P_i = f_i + [M^-1 dm]_i,   dm = (drho, dj + F, S - s (m - m_eq))
//
// with m the moments of f9 = f + Fbar/2, S the moments of Fbar, s the rates
// and (drho, djx, djy, djz) = field value - conserved moment of f9.  Since
// Fbar = M^-1 (0, F, S), this equals the reference form
// f9 + M^-1 (m* - m) + Fbar/2 in exact arithmetic, and the 19 Guo terms are
// never formed.  In FP32 the reference form rebuilds every population from
// moments of size |u| and adds F/2 ~ 1e-7 to sums of size 0.1, which rounds
// a few per cent of the per-step momentum input away, the same way on every
// node of a flow that is uniform along the walls: the mean profile then
// drifts by ~1e-4 of U per thousand steps.  Here the small increments are
// formed first and f_i is touched by one addition, so momentum enters
// exactly up to one rounding of each stored value.  When the fields are
// macrovar of the same f the conserved changes vanish and the caller passes
// zeros.
template <typename T, typename Closure = NoClosure>
__host__ __device__ __forceinline__ void mrt_collide_delta(T (&f)[19], T rho9, T ux9, T uy9, T uz9, T drho, T djx, T djy,
                                                  T djz, const Rates<T> &r, const Closure &closure = Closure())
{
    T m[15], sf[12], eq[12];
    delta_moments(f, rho9, ux9, uy9, uz9, r, m, sf, eq);

    T s9 = r.s9, s13 = r.s13;
    if constexpr (!std::is_same<Closure, NoClosure>::value) {
        T neq[9];
        closure_inputs(m, eq, neq);
        closure(neq, s9, s13);
    }

    T dm[15];
#pragma unroll
    for (int k = 0; k < 12; k++)
        dm[k] = sf[k] - rate_of(k, r, s9, s13) * (m[k] - eq[k]);
#pragma unroll
    for (int k = 12; k < 15; k++)
        dm[k] = -(r.s16 * m[k]);

    T d[19];
    inverse(drho, djx + r.fx, djy + r.fy, djz + r.fz, dm, d);
#pragma unroll
    for (int i = 0; i < 19; i++)
        f[i] = f[i] + d[i];
}

} // namespace d3q19
