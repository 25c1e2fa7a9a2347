// Learned relaxation-rate closure: a 10-32-32-1 tanh network evaluated per
// node inside the collision (on the matrix cores, FP16 operands with FP32
// sums) and trained by the kernels of mlp.hip (FP32).
//
// Inputs (10): the non-equilibrium parts m_k - m_k^eq of the nine moments the
// MRT relaxes with s1, s2, s9, s10, s13 (e, eps, 3p_xx, 3pi_xx, p_ww, pi_ww,
// p_xy, p_yz, p_xz: evlm1,2,6,7,8,9,10,11,12 minus eqm of
// Channel-Flow/collision.f90:86-138), each divided by a stored scale, and the
// wall distance y+ / 100.  Output o gives an eddy viscosity nu_t = nu *
// softplus(o) >= 0; the node's shear rates become
// s9 = s13 = 1 / (3 (nu + nu_t) + 1/2)  (tau = 3 nu + 0.5, para.f90:106-108).
//
// Parameters, FP32, in this order (NPARAM = 1441): W1[32][10], b1[32],
// W2[32][32], b2[32], W3[32], b3; then (not trained) the 10 input scales and
// nu.
#pragma once

#include <hip/hip_runtime.h>
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
constexpr int MLP_TOTAL = MLP_NU + 1;

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
    return 1.0f / (3.0f * nu * (1.0f + softplus(o)) + 0.5f);
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
// the biases, tanh, the output layer and softplus stay in FP32.
//
// Weights are staged per block in LDS as FP16, row-major by (row, chunk):
//   layer 1: ((4 mt + i) * 3 + kc) * 4 + k   (10 inputs padded to 12)
//   layer 2: 384 + ((4 mt + i) * 8 + kc) * 4 + k
typedef _Float16 mlp_half4 __attribute__((ext_vector_type(4)));
typedef float mlp_float4 __attribute__((ext_vector_type(4)));

constexpr int MLP_MT = MLP_H / 4;                  // row tiles of 4 hidden units
constexpr int MLP_K1 = (MLP_IN + 3) / 4;           // input chunks of 4
constexpr int MLP_K2 = MLP_H / 4;                  // hidden chunks of 4
constexpr int MLP_LDS1 = MLP_H * MLP_K1 * 4;       // 384 halves
constexpr int MLP_LDS = MLP_LDS1 + MLP_H * MLP_K2 * 4; // + 1024

// Copy W1 and W2 into the LDS layout above (all threads of the block).
__device__ __forceinline__ void mlp_stage(const float *__restrict__ w, _Float16 *wl, uint32_t tid, uint32_t nthreads)
{
    for (uint32_t e = tid; e < (uint32_t)MLP_LDS; e += nthreads) {
        float v;
        if (e < (uint32_t)MLP_LDS1) {
            const uint32_t k = e & 3u, kc = (e >> 2) % MLP_K1, row = (e >> 2) / MLP_K1, in = 4 * kc + k;
            v = in < (uint32_t)MLP_IN ? w[MLP_W1 + row * MLP_IN + in] : 0.0f;
        } else {
            const uint32_t e2 = e - MLP_LDS1;
            const uint32_t k = e2 & 3u, kc = (e2 >> 2) % MLP_K2, row = (e2 >> 2) / MLP_K2;
            v = w[MLP_W2 + row * MLP_H + 4 * kc + k];
        }
        wl[e] = (_Float16)v;
    }
}

// Network output o for this lane's node; every lane of the wave must call it
// together (the matrix instructions act on the whole wave).
__device__ __forceinline__ float mlp_forward_wave(const float *__restrict__ w, const _Float16 *wl,
                                                  const float (&x)[MLP_IN])
{
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
        mlp_float4 acc = {w[MLP_B1 + 4 * mt], w[MLP_B1 + 4 * mt + 1], w[MLP_B1 + 4 * mt + 2], w[MLP_B1 + 4 * mt + 3]};
#pragma unroll
        for (int kc = 0; kc < MLP_K1; kc++) {
            const mlp_half4 A = *(const mlp_half4 *)(wl + ((4 * mt + i) * MLP_K1 + kc) * 4);
            acc = __builtin_amdgcn_mfma_f32_4x4x4f16(A, xb[kc], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            hb[mt][c] = (_Float16)fast_tanh(acc[c]);
    }
    float o = w[MLP_B3];
#pragma unroll
    for (int mt = 0; mt < MLP_MT; mt++) {
        mlp_float4 acc = {w[MLP_B2 + 4 * mt], w[MLP_B2 + 4 * mt + 1], w[MLP_B2 + 4 * mt + 2], w[MLP_B2 + 4 * mt + 3]};
#pragma unroll
        for (int kc = 0; kc < MLP_K2; kc++) {
            const mlp_half4 A = *(const mlp_half4 *)(wl + MLP_LDS1 + ((4 * mt + i) * MLP_K2 + kc) * 4);
            acc = __builtin_amdgcn_mfma_f32_4x4x4f16(A, hb[kc], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            o = fmaf(w[MLP_W3 + 4 * mt + c], fast_tanh(acc[c]), o);
    }
    return o;
}

// Scaled network inputs of a node: the nine non-equilibrium moments over
// their scales, then the wall-distance feature.
template <typename T>
__device__ __forceinline__ void mlp_inputs(const float *__restrict__ w, const T (&neq)[9], float yfeat,
                                           float (&x)[MLP_IN])
{
#pragma unroll
    for (int k = 0; k < 9; k++)
        x[k] = (float)neq[k] * (1.0f / w[MLP_SCALE + k]);
    x[9] = yfeat;
}

// Closure functor for mrt_collide_delta in the step kernels: the
// matrix-core network above sets the node's shear rates.
struct MlpClosureWave {
    const float *w;
    const _Float16 *wl; // staged weights (LDS)
    float yfeat;        // y+ / scale
    template <typename T>
    __device__ __forceinline__ void operator()(const T (&neq)[9], T &s9, T &s13) const
    {
        float x[MLP_IN];
        mlp_inputs(w, neq, yfeat, x);
        const float s = mlp_rate(w, mlp_forward_wave(w, wl, x));
        s9 = (T)s;
        s13 = (T)s;
    }
};

} // namespace d3q19
