// Learned closure: training-data kernels and training kernels (see mlp.h).
//
// Training data come from the solver's own fine runs.  A fine lattice is box
// filtered 2x2x2 onto the coarse grid (k_filter: means of rho, u and of the
// products u_a u_b).  For each coarse node (k_features) the resolved strain
// S is taken by central differences of the filtered velocity (no-slip walls
// half a cell beyond the first and last nodes), the subgrid stress is
// tau_ab = <u_a u_b> - <u_a><u_b>, and the target is the eddy viscosity that
// best carries it, nu_t = -tau:S / (2 S:S), as nu_t/nu clipped to [-10, 10]
// (backscatter is kept: the network's non-negative output then fits the
// conditional mean instead of the mean of the positive part).
// The features are the non-equilibrium moments the coarse MRT would carry
// for that strain to first order (Chapman-Enskog: m_neq = -(2/(3 s)) times
// the strain combination of the moment; zero for e, eps, 3pi_xx, pi_ww), in
// the same scaling the fused collision uses, plus y+.
//
// Training: minibatch gradient of the mean squared error of softplus(o)
// against the target, one wave per 64 samples, per-wave sums in a fixed
// order, then a fixed-order sum over waves and an Adam step.  The forward
// pass is the same device code the fused collision runs.
#include <utility>

#include "kernels_impl.h"
#include "mlp_train.h"

namespace d3q19 {

namespace {

constexpr int WAVE = 64;

template <int P>
__global__ void __launch_bounds__(256) k_filter(FilterArgs a)
{
    using S = typename Prec<P>::S;
    using T = typename Prec<P>::T;
    const uint32_t nc = blockIdx.x * 256 + threadIdx.x;
    const uint32_t Nc = a.lxc * a.lyc * a.lzc;
    if (nc >= Nc)
        return;
    const uint32_t cx = nc % a.lxc, cy = (nc / a.lxc) % a.lyc, cz = nc / (a.lxc * a.lyc);
    double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
                const uint32_t n = (2 * cx + dx) + a.gf.lx * ((2 * cy + dy) + a.gf.ly * (2 * cz + dz));
                const Nbr b = make_nbr(n, a.gf);
                T f[19];
                pull_all<P>((const S *)a.L, a.gf.N, b, f);
                T rho9, ux9, uy9, uz9;
                macro_sums(f, rho9, ux9, uy9, uz9);
                const double ux = (double)ux9 + 0.5 * a.fx, uy = (double)uy9 + 0.5 * a.fy,
                             uz = (double)uz9 + 0.5 * a.fz;
                s[0] += (double)rho9;
                s[1] += ux;
                s[2] += uy;
                s[3] += uz;
                s[4] += ux * ux;
                s[5] += uy * uy;
                s[6] += uz * uz;
                s[7] += ux * uy;
                s[8] += uy * uz;
                s[9] += ux * uz;
            }
#pragma unroll
    for (int q = 0; q < 10; q++)
        a.out[(uint64_t)q * Nc + nc] = (float)(s[q] * 0.125);
}

__global__ void __launch_bounds__(256) k_features(FeatArgs a)
{
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    const uint32_t N = a.lx * a.ly * a.lz;
    if (n >= N)
        return;
    const uint32_t ix = n % a.lx, iy = (n / a.lx) % a.ly, iz = n / (a.lx * a.ly);
    const float *U[3] = {a.fields + 1 * (uint64_t)N, a.fields + 2 * (uint64_t)N, a.fields + 3 * (uint64_t)N};
    // d/dx with no-slip walls half a cell outside: ghost value -u at x = -1, lx
    auto at = [&](int c, int x, uint32_t y, uint32_t z) -> float {
        if (x < 0 || x >= (int)a.lx)
            return -U[c][(x < 0 ? 0 : a.lx - 1) + a.lx * (y + a.ly * z)];
        return U[c][x + a.lx * (y + a.ly * z)];
    };
    float g[3][3]; // g[c][d] = d u_c / d x_d
    const uint32_t yp = (iy + 1) % a.ly, ym = (iy + a.ly - 1) % a.ly;
    const uint32_t zp = (iz + 1) % a.lz, zm = (iz + a.lz - 1) % a.lz;
    for (int c = 0; c < 3; c++) {
        g[c][0] = 0.5f * (at(c, (int)ix + 1, iy, iz) - at(c, (int)ix - 1, iy, iz));
        g[c][1] = 0.5f * (at(c, (int)ix, yp, iz) - at(c, (int)ix, ym, iz));
        g[c][2] = 0.5f * (at(c, (int)ix, iy, zp) - at(c, (int)ix, iy, zm));
    }
    float S[3][3];
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++)
            S[c][d] = 0.5f * (g[c][d] + g[d][c]);
    const float u[3] = {U[0][n], U[1][n], U[2][n]};
    // filtered products: uxux, uyuy, uzuz, uxuy, uyuz, uxuz
    const float *F = a.fields;
    float tau[3][3];
    tau[0][0] = F[4 * (uint64_t)N + n] - u[0] * u[0];
    tau[1][1] = F[5 * (uint64_t)N + n] - u[1] * u[1];
    tau[2][2] = F[6 * (uint64_t)N + n] - u[2] * u[2];
    tau[0][1] = tau[1][0] = F[7 * (uint64_t)N + n] - u[0] * u[1];
    tau[1][2] = tau[2][1] = F[8 * (uint64_t)N + n] - u[1] * u[2];
    tau[0][2] = tau[2][0] = F[9 * (uint64_t)N + n] - u[0] * u[2];
    float ts = 0.f, ss = 0.f;
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            ts += tau[c][d] * S[c][d];
            ss += S[c][d] * S[c][d];
        }
    float y = ss > 1e-30f ? -ts / (2.0f * ss) / a.nu : 0.0f;
    y = fminf(fmaxf(y, -10.0f), 10.0f);
    if (a.ts) {
        a.ts[n] = ts;
        a.ss[n] = ss;
    }
    // first-order non-equilibrium moments of the coarse MRT for strain S
    const float k = -2.0f / (3.0f * a.s9);
    const float neq[9] = {0.0f,
                          0.0f,
                          k * (2.0f * S[0][0] - S[1][1] - S[2][2]),
                          0.0f,
                          k * (S[1][1] - S[2][2]),
                          0.0f,
                          k * S[0][1],
                          k * S[1][2],
                          k * S[0][2]};
    float *o = a.feat + (uint64_t)(a.offset + n) * MLP_IN;
    for (int q = 0; q < 9; q++)
        o[q] = neq[q] / a.scale[q];
    const float wd = fminf((float)ix + 0.5f, (float)a.lx - (float)ix - 0.5f);
    o[9] = wd * a.inv_ystar / a.yscale;
    a.target[a.offset + n] = y;
}

// Gradient of mean((softplus(o) - y)^2) over samples [first, first+count):
// one wave per 64 samples; partial[w * (NPARAM+1) + i] (i = NPARAM: loss).
__global__ void __launch_bounds__(WAVE) k_mlp_grad(const float *__restrict__ w, const float *__restrict__ feat,
                                                   const float *__restrict__ target, uint32_t first, uint32_t count,
                                                   float *__restrict__ partial)
{
    __shared__ float xs[WAVE][MLP_IN];
    __shared__ float h1s[WAVE][MLP_H];
    __shared__ float h2s[WAVE][MLP_H];
    __shared__ float ga1[WAVE][MLP_H];
    __shared__ float ga2[WAVE][MLP_H];
    __shared__ float go[WAVE];
    __shared__ float ls[WAVE];
    const int t = threadIdx.x;
    const uint32_t sidx = blockIdx.x * WAVE + t;
    const bool valid = sidx < count;
    const float inv_count = 1.0f / (float)count;
    float x[MLP_IN];
    for (int k = 0; k < MLP_IN; k++)
        x[k] = valid ? feat[(uint64_t)(first + sidx) * MLP_IN + k] : 0.0f;
    float h1[MLP_H], h2[MLP_H];
    for (int j = 0; j < MLP_H; j++) {
        float a = w[MLP_B1 + j];
        for (int k = 0; k < MLP_IN; k++)
            a = fmaf(w[MLP_W1 + j * MLP_IN + k], x[k], a);
        h1[j] = fast_tanh(a);
    }
    float o = w[MLP_B3];
    for (int j = 0; j < MLP_H; j++) {
        float a = w[MLP_B2 + j];
        for (int k = 0; k < MLP_H; k++)
            a = fmaf(w[MLP_W2 + j * MLP_H + k], h1[k], a);
        h2[j] = fast_tanh(a);
        o = fmaf(w[MLP_W3 + j], h2[j], o);
    }
    const float pred = mlp_nut(w, o);
    const float y = valid ? target[first + sidx] : 0.0f;
    const float err = pred - y;
    const float sig = 1.0f / (1.0f + __expf(-o));
    const float g_o = valid ? 2.0f * err * sig * inv_count : 0.0f;
    float g2[MLP_H];
    for (int j = 0; j < MLP_H; j++)
        g2[j] = g_o * w[MLP_W3 + j] * (1.0f - h2[j] * h2[j]);
    for (int k = 0; k < MLP_H; k++) {
        float gh = 0.0f;
        for (int j = 0; j < MLP_H; j++)
            gh = fmaf(g2[j], w[MLP_W2 + j * MLP_H + k], gh);
        ga1[t][k] = gh * (1.0f - h1[k] * h1[k]);
    }
    for (int k = 0; k < MLP_IN; k++)
        xs[t][k] = x[k];
    for (int j = 0; j < MLP_H; j++) {
        h1s[t][j] = h1[j];
        h2s[t][j] = h2[j];
        ga2[t][j] = g2[j];
    }
    go[t] = g_o;
    ls[t] = valid ? err * err * inv_count : 0.0f;
    __syncthreads();
    float *out = partial + (uint64_t)blockIdx.x * (MLP_NPARAM + 1);
    for (int i = t; i < MLP_NPARAM + 1; i += WAVE) {
        float acc = 0.0f;
        if (i < MLP_B1) { // W1[j][k]
            const int j = i / MLP_IN, k = i % MLP_IN;
            for (int s = 0; s < WAVE; s++)
                acc = fmaf(ga1[s][j], xs[s][k], acc);
        } else if (i < MLP_W2) { // b1
            for (int s = 0; s < WAVE; s++)
                acc += ga1[s][i - MLP_B1];
        } else if (i < MLP_B2) { // W2[j][k]
            const int j = (i - MLP_W2) / MLP_H, k = (i - MLP_W2) % MLP_H;
            for (int s = 0; s < WAVE; s++)
                acc = fmaf(ga2[s][j], h1s[s][k], acc);
        } else if (i < MLP_W3) { // b2
            for (int s = 0; s < WAVE; s++)
                acc += ga2[s][i - MLP_B2];
        } else if (i < MLP_B3) { // W3
            for (int s = 0; s < WAVE; s++)
                acc = fmaf(go[s], h2s[s][i - MLP_W3], acc);
        } else if (i == MLP_B3) {
            for (int s = 0; s < WAVE; s++)
                acc += go[s];
        } else { // loss
            for (int s = 0; s < WAVE; s++)
                acc += ls[s];
        }
        out[i] = acc;
    }
}

// Fixed-order sum of the per-wave partials; optional Adam step.
__global__ void k_mlp_sum_adam(const float *__restrict__ partial, uint32_t nwaves, float *__restrict__ grad,
                               float *w, float *m, float *v, int step, float lr, int adam)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM + 1)
        return;
    float g = 0.0f;
    for (uint32_t b = 0; b < nwaves; b++)
        g += partial[(uint64_t)b * (MLP_NPARAM + 1) + i];
    grad[i] = g;
    if (adam && i < MLP_NPARAM) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / (1.0f - powf(b1, (float)step));
        const float vh = vi / (1.0f - powf(b2, (float)step));
        w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
    }
}

// target[first + i] = values[i mod nx] for i < count.
__global__ void k_fill_plane_targets(float *__restrict__ target, uint64_t first, uint64_t count, uint32_t nx,
                                     const float *__restrict__ values)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        target[first + i] = values[i % nx];
}

// Network outputs for samples [first, first+count) (evaluation).
__global__ void k_mlp_predict(const float *__restrict__ w, const float *__restrict__ feat, uint32_t first,
                              uint32_t count, float *__restrict__ out)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count)
        return;
    float x[MLP_IN];
    for (int k = 0; k < MLP_IN; k++)
        x[k] = feat[(uint64_t)(first + s) * MLP_IN + k];
    out[s] = mlp_nut(w, mlp_forward(w, x));
}

} // namespace

template <int P>
hipError_t MlpLaunch<P>::filter(const FilterArgs &a, hipStream_t s)
{
    const uint32_t Nc = a.lxc * a.lyc * a.lzc;
    hipLaunchKernelGGL((k_filter<P>), dim3((Nc + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t mlp_features(const FeatArgs &a, hipStream_t s)
{
    const uint32_t N = a.lx * a.ly * a.lz;
    hipLaunchKernelGGL(k_features, dim3((N + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// One block per coarse x plane: FP64 sums of tau:S and S:S over the plane in
// a fixed order (strided per thread, then a fixed LDS tree), the plane's
// least-squares eddy viscosity, written to every node of the plane.
__global__ void __launch_bounds__(256) k_plane_targets(uint32_t lx, uint32_t ly, uint32_t lz,
                                                       const float *__restrict__ ts, const float *__restrict__ ss,
                                                       float nu, float *__restrict__ target, uint64_t offset)
{
    const uint32_t ix = blockIdx.x;
    const uint32_t rows = ly * lz;
    double a = 0.0, b = 0.0;
    for (uint32_t r = threadIdx.x; r < rows; r += 256) {
        const uint64_t n = ix + (uint64_t)lx * r;
        a += (double)ts[n];
        b += (double)ss[n];
    }
    __shared__ double sa[256], sb[256];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sa[threadIdx.x] += sa[threadIdx.x + h];
            sb[threadIdx.x] += sb[threadIdx.x + h];
        }
        __syncthreads();
    }
    float y = sb[0] > 1e-30 ? (float)(-sa[0] / (2.0 * sb[0]) / (double)nu) : 0.0f;
    y = fminf(fmaxf(y, -10.0f), 10.0f);
    for (uint32_t r = threadIdx.x; r < rows; r += 256)
        target[offset + ix + (uint64_t)lx * r] = y;
}

hipError_t mlp_plane_targets(uint32_t lx, uint32_t ly, uint32_t lz, const float *ts, const float *ss, float nu,
                             float *target, uint64_t offset, hipStream_t s)
{
    hipLaunchKernelGGL(k_plane_targets, dim3(lx), dim3(256), 0, s, lx, ly, lz, ts, ss, nu, target, offset);
    return hipGetLastError();
}

hipError_t mlp_grad(const float *w, const float *feat, const float *target, uint32_t first, uint32_t count,
                    float *partial, float *grad, float *wmut, float *m, float *v, int step, float lr, int adam,
                    hipStream_t s)
{
    const uint32_t nw = (count + WAVE - 1) / WAVE;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nw), dim3(WAVE), 0, s, w, feat, target, first, count, partial);
    hipLaunchKernelGGL(k_mlp_sum_adam, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, s, partial, nw, grad, wmut,
                       m, v, step, lr, adam);
    return hipGetLastError();
}

hipError_t mlp_fill_plane_targets(float *target, uint64_t first, uint64_t count, uint32_t nx, const float *values,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_plane_targets, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, s, target, first,
                       count, nx, values);
    return hipGetLastError();
}

hipError_t mlp_predict(const float *w, const float *feat, uint32_t first, uint32_t count, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_predict, dim3((count + 255) / 256), dim3(256), 0, s, w, feat, first, count, out);
    return hipGetLastError();
}

template struct MlpLaunch<0>;
template struct MlpLaunch<1>;
template struct MlpLaunch<2>;
template struct MlpLaunch<3>;

} // namespace d3q19
