// Learned relaxation-rate closure: a 10-32-32-1 tanh network evaluated per
// node inside the collision.
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

// Network output o for one node's inputs x (already scaled).
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

// Closure functor for mrt_collide: scales the nine non-equilibrium moments,
// appends the wall-distance feature, and sets the node's shear rates.
struct MlpClosure {
    const float *w;
    float yfeat; // y+ / scale
    template <typename T>
    __device__ __forceinline__ void operator()(const T (&neq)[9], T &s9, T &s13) const
    {
        float x[MLP_IN];
#pragma unroll
        for (int k = 0; k < 9; k++)
            x[k] = (float)neq[k] * (1.0f / w[MLP_SCALE + k]);
        x[9] = yfeat;
        const float s = mlp_rate(w, mlp_forward(w, x));
        s9 = (T)s;
        s13 = (T)s;
    }
};

} // namespace d3q19
