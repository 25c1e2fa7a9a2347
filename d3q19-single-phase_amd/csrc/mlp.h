// Learned relaxation-rate closure: a 10-32-32-1 tanh network evaluated per
// node inside the collision (on the matrix cores, FP16 operands with FP32
// sums) and trained by the kernels of mlp.hip (FP32).
//
// Inputs (10): the non-equilibrium parts m_k - m_k^eq of the nine moments the
// MRT relaxes with s1, s2, s9, s10, s13 (e, eps, 3p_xx, 3pi_xx, p_ww, pi_ww,
// p_xy, p_yz, p_xz: evlm1,2,6,7,8,9,10,11,12 minus eqm of
// Channel-Flow/collision.f90:86-138), each divided by a stored scale, and the
// wall distance y+ / 100.  Output o gives an eddy viscosity nu_t = nu *
// (softplus(o) + m) with the floor m > -1 (0: nu_t >= 0; negative: the
// closure may also lower the viscosity, to nu (1 + m) at least); the node's
// shear rates become s9 = s13 = 1 / (3 (nu + nu_t) + 1/2)  (tau = 3 nu + 0.5,
// para.f90:106-108), so tau stays above 1/2.
//
// Parameters, FP32, in this order (NPARAM = 1441): W1[32][10], b1[32],
// W2[32][32], b2[32], W3[32], b3; then (not trained) the 10 input scales, nu
// and the floor m.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace d3q19 {

constexpr int MLP_IN = 10, MLP_H = 32;
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_IN;
constexpr int MLP_W2 = MLP_B1 + MLP_H;
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H;
constexpr int MLP_B3 = MLP_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_B3 + 1; // 1441
constexpr int MLP_SCALE = MLP_NPARAM;  // 10 input scales
constexpr int MLP_NU = MLP_SCALE + MLP_IN;
constexpr int MLP_NUTMIN = MLP_NU + 1;
constexpr int MLP_TOTAL = MLP_NUTMIN + 1;

// tanh(x) = 1 - 2 / (exp(2x) + 1) with the hardware exp2 and reciprocal; the
// argument is clamped so exp cannot overflow.
__device__ __forceinline__ float fast_tanh(float x)
{
    x = fminf(fmaxf(x, -15.0f), 15.0f);
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f); // exp(2x)
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ __forceinline__ float softplus(float o)
{
    return o > 20.0f ? o : log1pf(__expf(o));
}

// nu_t / nu from the network output.
__device__ __forceinline__ float mlp_nut(const float *__restrict__ w, float o)
{
    return softplus(o) + w[MLP_NUTMIN];
}

// Network output o for one node's inputs x (already scaled), FP32 on the
// vector ALUs: the training and prediction kernels' forward pass.
__device__ __forceinline__ float mlp_forward(const float *__restrict__ w, const float (&x)[MLP_IN])
{
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
        float a = w[MLP_B1 + j];
#pragma unroll
        for (int k = 0; k < MLP_IN; k++)
            a = fmaf(w[MLP_W1 + j * MLP_IN + k], x[k], a);
        h1[j] = fast_tanh(a);
    }
    float o = w[MLP_B3];
#pragma unroll 4
    for (int j = 0; j < MLP_H; j++) {
        float a = w[MLP_B2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; k++)
            a = fmaf(w[MLP_W2 + j * MLP_H + k], h1[k], a);
        o = fmaf(w[MLP_W3 + j], fast_tanh(a), o);
    }
    return o;
}

// Shear relaxation rate from the network output.
__device__ __forceinline__ float mlp_rate(const float *__restrict__ w, float o)
{
    const float nu = w[MLP_NU];
    return 1.0f / (3.0f * nu * (1.0f + mlp_nut(w, o)) + 0.5f);
}

// ---------------------------------------------------------------------------
// The same network on the matrix cores, one wave of 64 nodes at a time.
//
// v_mfma_f32_4x4x4_16b_f16 computes 16 independent 4x4x4 products per wave:
// block b = lane / 4 takes A[i][k] from lane 4b + i, B[k][j] from lane 4b + j
// (four halves each) and returns D[i][j] (+ C) to lane 4b + j (four floats).
// With the weights as A (W[4 mt + i][4 kc + k], the same in every block) and
// each node's own inputs as B, D holds that node's hidden units 4 mt .. 4 mt
// + 3 in its own lane: both layers run with no data exchange between lanes.
// Inputs and hidden values are rounded to FP16, products accumulate in FP32;
// the bias factors, tanh, the output layer and softplus stay in FP32.
//
// tanh(a + b) = 1 - 2 / (2^(c a) 2^(c b) + 1) with c = 2 log2(e) and a the
// bias-free sum: the weights are stored pre-multiplied by c and each bias as
// the factor 2^(c b), so the products start from zero and each hidden unit
// costs one exp2, one fma and one reciprocal (IEEE overflow gives exactly
// +-1 for large |a|, no clamp is needed).  The output layer folds the same form:
//   o = b3 + sum_j W3_j tanh(h_j) = (b3 + sum_j W3_j) + sum_j (-2 W3_j) / (2^(c h_j) + 1).
//
// The device copy of the parameters (lbm_closure_set) carries, after the
// MLP_TOTAL parameters, the constants of that form (mlp_derive) and the FP16
// LDS image of c W1 and c W2, row-major by (row, chunk):
//   layer 1: ((4 mt + i) * 3 + kc) * 4 + k   (10 inputs padded to 12)
//   layer 2: 384 + ((4 mt + i) * 8 + kc) * 4 + k
typedef _Float16 mlp_half4 __attribute__((ext_vector_type(4)));
typedef float mlp_float4 __attribute__((ext_vector_type(4)));

constexpr int MLP_MT = MLP_H / 4;                  // row tiles of 4 hidden units
constexpr int MLP_K1 = (MLP_IN + 3) / 4;           // input chunks of 4
constexpr int MLP_K2 = MLP_H / 4;                  // hidden chunks of 4
constexpr int MLP_LDS1 = MLP_H * MLP_K1 * 4;       // 384 halves
constexpr int MLP_LDS = MLP_LDS1 + MLP_H * MLP_K2 * 4; // + 1024

constexpr float MLP_C = 2.8853900817779268f; // 2 log2(e)
// derived block, at MLP_TOTAL
constexpr int MLPX_E1 = 0;            // 2^(c b1) (32)
constexpr int MLPX_E2 = MLPX_E1 + 32; // 2^(c b2) (32)
constexpr int MLPX_W3 = MLPX_E2 + 32; // -2 W3 (32)
constexpr int MLPX_C0 = MLPX_W3 + 32; // b3 + sum W3
constexpr int MLPX_INV = MLPX_C0 + 1; // 1 / scale_k, k < 9
constexpr int MLPX_FLOATS = 112;      // padded: the LDS image starts 16-byte aligned
constexpr int MLP_DEV_X = MLP_TOTAL;                   // derived block offset (floats)
constexpr int MLP_DEV_LDS = MLP_TOTAL + MLPX_FLOATS;   // LDS image offset (floats)
constexpr int MLP_DEV_TOTAL = MLP_DEV_LDS + MLP_LDS / 2;

// Largest hidden bias |b| the fused form takes.  The bias enters as the
// factor 2^(c b) on 2^(c a) (a: the bias-free sum), which is exact up to
// rounding while 2^(c a) is a normal float; where 2^(c a) overflows or
// underflows (|c a| > 126), |c (a + b)| > 126 - 32 c > 34 and tanh(a + b)
// is +-1 in FP32, as the product's inf or near-zero gives.
constexpr float MLP_BIAS_MAX = 32.0f;

// The derived block and the LDS image from the parameters (host, at
// lbm_closure_set; the biases are checked against MLP_BIAS_MAX first).
inline void mlp_derive(const float *p, float *dev)
{
    float *x = dev + MLP_DEV_X;
    for (int k = 0; k < MLPX_FLOATS; k++)
        x[k] = 0.0f;
    float sw3 = 0.0f;
    for (int j = 0; j < MLP_H; j++) {
        x[MLPX_E1 + j] = (float)exp2((double)MLP_C * p[MLP_B1 + j]);
        x[MLPX_E2 + j] = (float)exp2((double)MLP_C * p[MLP_B2 + j]);
        x[MLPX_W3 + j] = -2.0f * p[MLP_W3 + j];
        sw3 = sw3 + p[MLP_W3 + j];
    }
    x[MLPX_C0] = p[MLP_B3] + sw3;
    for (int k = 0; k < 9; k++)
        x[MLPX_INV + k] = 1.0f / p[MLP_SCALE + k];
    _Float16 *wl = (_Float16 *)(dev + MLP_DEV_LDS);
    for (int e = 0; e < MLP_LDS; e++) {
        float v;
        if (e < MLP_LDS1) {
            const int k = e & 3, kc = (e >> 2) % MLP_K1, row = (e >> 2) / MLP_K1, in = 4 * kc + k;
            v = in < MLP_IN ? MLP_C * p[MLP_W1 + row * MLP_IN + in] : 0.0f;
        } else {
            const int e2 = e - MLP_LDS1;
            const int k = e2 & 3, kc = (e2 >> 2) % MLP_K2, row = (e2 >> 2) / MLP_K2;
            v = MLP_C * p[MLP_W2 + row * MLP_H + 4 * kc + k];
        }
        wl[e] = (_Float16)v;
    }
}

// Copy the LDS image into the block's LDS (all threads of the block).
__device__ __forceinline__ void mlp_stage(const float *__restrict__ w, _Float16 *wl, uint32_t tid, uint32_t nthreads)
{
    const uint32_t *src = (const uint32_t *)(w + MLP_DEV_LDS);
    uint32_t *dst = (uint32_t *)wl;
    for (uint32_t e = tid; e < (uint32_t)(MLP_LDS / 2); e += nthreads)
        dst[e] = src[e];
}

// 1 / (2^v e + 1): tanh(a + b) = 1 - 2 * that at v = c a, e = 2^(c b).
__device__ __forceinline__ float mlp_rcp1p_exp2(float v, float e)
{
    return __builtin_amdgcn_rcpf(fmaf(__builtin_amdgcn_exp2f(v), e, 1.0f));
}

// Network output o for this lane's node; every lane of the wave must call it
// together (the matrix instructions act on the whole wave).  w is the device
// copy of the parameters with its derived block.
__device__ __forceinline__ float mlp_forward_wave(const float *__restrict__ w, const _Float16 *wl,
                                                  const float (&x)[MLP_IN])
{
    const float *__restrict__ d = w + MLP_DEV_X;
    const uint32_t i = threadIdx.x & 3u; // this lane's A row within a block
    mlp_half4 xb[MLP_K1];
#pragma unroll
    for (int kc = 0; kc < MLP_K1; kc++)
#pragma unroll
        for (int k = 0; k < 4; k++)
            xb[kc][k] = (_Float16)(4 * kc + k < MLP_IN ? x[4 * kc + k] : 0.0f);
    mlp_half4 hb[MLP_K2];
#pragma unroll
    for (int mt = 0; mt < MLP_MT; mt++) {
        mlp_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < MLP_K1; kc++) {
            const mlp_half4 A = *(const mlp_half4 *)(wl + ((4 * mt + i) * MLP_K1 + kc) * 4);
            acc = __builtin_amdgcn_mfma_f32_4x4x4f16(A, xb[kc], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            hb[mt][c] = (_Float16)fmaf(-2.0f, mlp_rcp1p_exp2(acc[c], d[MLPX_E1 + 4 * mt + c]), 1.0f);
    }
    float o = d[MLPX_C0];
#pragma unroll
    for (int mt = 0; mt < MLP_MT; mt++) {
        mlp_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < MLP_K2; kc++) {
            const mlp_half4 A = *(const mlp_half4 *)(wl + MLP_LDS1 + ((4 * mt + i) * MLP_K2 + kc) * 4);
            acc = __builtin_amdgcn_mfma_f32_4x4x4f16(A, hb[kc], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            o = fmaf(d[MLPX_W3 + 4 * mt + c], mlp_rcp1p_exp2(acc[c], d[MLPX_E2 + 4 * mt + c]), o);
    }
    return o;
}

// Scaled network inputs of a node: the nine non-equilibrium moments over
// their scales, then the wall-distance feature.  w: the raw parameters.
template <typename T>
__device__ __forceinline__ void mlp_inputs(const float *__restrict__ w, const T (&neq)[9], float yfeat,
                                           float (&x)[MLP_IN])
{
#pragma unroll
    for (int k = 0; k < 9; k++)
        x[k] = (float)neq[k] * (1.0f / w[MLP_SCALE + k]);
    x[9] = yfeat;
}

// The same from the device copy's inverse scales (equal values: the host
// rounds 1.0f / scale as the division above does).
template <typename T>
__device__ __forceinline__ void mlp_inputs_dev(const float *__restrict__ w, const T (&neq)[9], float yfeat,
                                               float (&x)[MLP_IN])
{
#pragma unroll
    for (int k = 0; k < 9; k++)
        x[k] = (float)neq[k] * w[MLP_DEV_X + MLPX_INV + k];
    x[9] = yfeat;
}

// Closure functor for mrt_collide_delta in the step kernels: the
// matrix-core network above sets the node's shear rates.
struct MlpClosureWave {
    const float *w; // device copy (parameters + derived block)
    const _Float16 *wl; // staged weights (LDS)
    float yfeat;        // y+ / scale
    template <typename T>
    __device__ __forceinline__ void operator()(const T (&neq)[9], T &s9, T &s13) const
    {
        float x[MLP_IN];
        mlp_inputs_dev(w, neq, yfeat, x);
        const float s = mlp_rate(w, mlp_forward_wave(w, wl, x));
        s9 = (T)s;
        s13 = (T)s;
    }
};

} // namespace d3q19
