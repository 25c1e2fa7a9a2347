// FP64 step-kernel variants on the 192x768x384 channel (56.6 M nodes),
// timed with HIP events: bounce-back select by branch or by mask, cached or
// non-temporal loads and stores, XCD block remap on or off, block size,
// waves-per-EU limits, and compute-free copies with the same access pattern.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <utility>
#include <vector>
#include "kernels_impl.h"

using namespace d3q19;
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

template <int LD> __device__ __forceinline__ double ldv(const double* p) {
    if constexpr (LD == 0) return __builtin_nontemporal_load(p); else return *p;
}
template <int ST> __device__ __forceinline__ void stv(double* p, double v) {
    if constexpr (ST == 0) __builtin_nontemporal_store(v, p); else *p = v;
}

// branch-free pull: choose node and row with masks
template <int I, int LD>
__device__ __forceinline__ double pull_bf(const double* __restrict__ L, uint64_t N, const Nbr& b) {
    if constexpr (CX[I] != 0) {
        const uint32_t w = walled<I>(b) ? 0xffffffffu : 0u;
        const uint32_t node = (b.n & w) | (src_of<I>(b) & ~w);
        const uint64_t row = w ? (uint64_t)OPP[I] * N : (uint64_t)I * N;
        return ldv<LD>(L + row + node);
    } else {
        return ldv<LD>(L + (uint64_t)I * N + src_of<I>(b));
    }
}

template <int LD, int... Is>
__device__ __forceinline__ void pull_bf_all(const double* __restrict__ L, uint64_t N, const Nbr& b, double (&f)[19], std::integer_sequence<int, Is...>) {
    ((f[Is] = pull_bf<Is, LD>(L, N, b)), ...);
}

template <int VAR, int LD, int ST, int REMAP, int BS>
__global__ void __launch_bounds__(BS) k_var(StepArgs a) {
    uint32_t blk = blockIdx.x;
    if (REMAP) blk = xcd_remap(blk, gridDim.x);
    const uint32_t n = blk * BS + threadIdx.x;
    if (n >= a.g.N) return;
    const Nbr b = make_nbr(n, a.g);
    const double* __restrict__ src = (const double*)a.src;
    double* __restrict__ dst = (double*)a.dst;
    double f[19];
    if constexpr (VAR == 0) pull_all<0>(src, a.g.N, b, f);
    else pull_bf_all<LD>(src, a.g.N, b, f, std::make_integer_sequence<int, 19>{});
    Rates<double> r;
    r.s1 = a.r.s1; r.s2 = a.r.s2; r.s4 = a.r.s4; r.s9 = a.r.s9; r.s10 = a.r.s10; r.s13 = a.r.s13; r.s16 = a.r.s16;
    r.omegepsl = a.r.omegepsl; r.omegepslj = a.r.omegepslj; r.omegxx = a.r.omegxx; r.fx = a.r.fx; r.fy = a.r.fy; r.fz = a.r.fz;
    double rho9, ux9, uy9, uz9;
    macro_sums(f, rho9, ux9, uy9, uz9);
    ux9 = ux9 + r.fx / 2.; uy9 = uy9 + r.fy / 2.; uz9 = uz9 + r.fz / 2.;
    mrt_collide(f, rho9, ux9, uy9, uz9, r);
#pragma unroll
    for (int i = 0; i < 19; i++) stv<ST>(dst + (uint64_t)i * a.g.N + n, f[i]);
}

template <int VAR, int LD, int ST, int REMAP, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6, 8))) k_var_w6(StepArgs a) {
    uint32_t blk = blockIdx.x;
    if (REMAP) blk = xcd_remap(blk, gridDim.x);
    const uint32_t n = blk * BS + threadIdx.x;
    if (n >= a.g.N) return;
    const Nbr b = make_nbr(n, a.g);
    const double* __restrict__ src = (const double*)a.src;
    double* __restrict__ dst = (double*)a.dst;
    double f[19];
    pull_bf_all<LD>(src, a.g.N, b, f, std::make_integer_sequence<int, 19>{});
    Rates<double> r;
    r.s1 = a.r.s1; r.s2 = a.r.s2; r.s4 = a.r.s4; r.s9 = a.r.s9; r.s10 = a.r.s10; r.s13 = a.r.s13; r.s16 = a.r.s16;
    r.omegepsl = a.r.omegepsl; r.omegepslj = a.r.omegepslj; r.omegxx = a.r.omegxx; r.fx = a.r.fx; r.fy = a.r.fy; r.fz = a.r.fz;
    double rho9, ux9, uy9, uz9;
    macro_sums(f, rho9, ux9, uy9, uz9);
    ux9 = ux9 + r.fx / 2.; uy9 = uy9 + r.fy / 2.; uz9 = uz9 + r.fz / 2.;
    mrt_collide(f, rho9, ux9, uy9, uz9, r);
#pragma unroll
    for (int i = 0; i < 19; i++) stv<ST>(dst + (uint64_t)i * a.g.N + n, f[i]);
}

// copy with the same pull pattern (no arithmetic)
template <int LD, int ST>
__global__ void __launch_bounds__(256) k_pullcopy(StepArgs a) {
    const uint32_t n = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (n >= a.g.N) return;
    const Nbr b = make_nbr(n, a.g);
    double f[19];
    pull_bf_all<LD>((const double*)a.src, a.g.N, b, f, std::make_integer_sequence<int, 19>{});
#pragma unroll
    for (int i = 0; i < 19; i++) stv<ST>((double*)a.dst + (uint64_t)i * a.g.N + n, f[i]);
}

// aligned copy of 19 arrays
template <int LD, int ST>
__global__ void __launch_bounds__(256) k_copy19(StepArgs a) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n >= a.g.N) return;
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++) f[i] = ldv<LD>((const double*)a.src + (uint64_t)i * a.g.N + n);
#pragma unroll
    for (int i = 0; i < 19; i++) stv<ST>((double*)a.dst + (uint64_t)i * a.g.N + n, f[i]);
}

// plain 1-stream copy (vectorised dwordx4)
__global__ void __launch_bounds__(256) k_copy1(const double2* __restrict__ s, double2* __restrict__ d, uint64_t n2) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * 256) d[i] = s[i];
}

typedef void (*KF)(StepArgs);

int main(int argc, char** argv) {
    const uint32_t lx = 192, ly = 768, lz = 384;
    const int reps = argc > 1 ? atoi(argv[1]) : 20;
    StepArgs a; memset(&a, 0, sizeof(a));
    a.g.lx = lx; a.g.ly = ly; a.g.lz = lz; a.g.plane = lx * ly; a.g.N = (uint64_t)lx * ly * lz;
    a.g.dlx = make_fastdiv(lx); a.g.dly = make_fastdiv(ly); a.g.remap = 1;
    a.r.s1 = 1.5; a.r.s2 = 1.4; a.r.s4 = 1.2; a.r.s10 = 1.4; a.r.s16 = 1.98; a.r.s9 = 1.0 / (3 * 0.0036 + 0.5); a.r.s13 = a.r.s9;
    a.r.omegepslj = -475.0 / 63.0; a.r.fy = 4.7e-7;
    const size_t bytes = a.g.N * 19 * sizeof(double);
    void *A, *B;
    CK(hipMalloc(&A, bytes)); CK(hipMalloc(&B, bytes));
    // small smooth initial state: f = 0.01 * w-ish pattern
    std::vector<double> h(a.g.N);
    for (int i = 0; i < 19; i++) {
        for (uint64_t n = 0; n < a.g.N; n++) h[n] = (i == 0 ? 1.0/3 : (i <= 6 ? 1.0/18 : 1.0/36)) * 1e-3 * ((n * 2654435761u) % 1000) / 1000.0;
        CK(hipMemcpy((double*)A + i * a.g.N, h.data(), a.g.N * 8, hipMemcpyHostToDevice));
    }
    CK(hipMemset(B, 0, bytes));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    const double algo = (double)a.g.N * 304.0;
    struct V { const char* name; KF k; int bs; } vs[] = {
        {"base(select-branch,nt,nt,remap)", k_var<0, 0, 0, 1, 256>, 256},
        {"bf,nt-ld,nt-st,remap", k_var<1, 0, 0, 1, 256>, 256},
        {"bf,plain-ld,nt-st,remap", k_var<1, 1, 0, 1, 256>, 256},
        {"bf,plain-ld,nt-st,noremap", k_var<1, 1, 0, 0, 256>, 256},
        {"bf,plain-ld,plain-st,remap", k_var<1, 1, 1, 1, 256>, 256},
        {"bf,nt-ld,plain-st,remap", k_var<1, 0, 1, 1, 256>, 256},
        {"bf,plain-ld,nt-st,remap,bs512", k_var<1, 1, 0, 1, 512>, 512},
        {"bf,plain-ld,nt-st,remap,bs128", k_var<1, 1, 0, 1, 128>, 128},
        {"bf,plain-ld,nt-st,remap,bs1024", k_var<1, 1, 0, 1, 1024>, 1024},
        {"bf,plain-ld,nt-st,remap,w6", k_var_w6<1, 1, 0, 1, 256>, 256},
        {"bf,nt-ld,nt-st,remap,w6", k_var_w6<1, 0, 0, 1, 256>, 256},
        {"pullcopy plain-ld nt-st", k_pullcopy<1, 0>, 256},
        {"pullcopy nt-ld nt-st", k_pullcopy<0, 0>, 256},
        {"pullcopy plain-ld plain-st", k_pullcopy<1, 1>, 256},
        {"copy19 plain-ld nt-st", k_copy19<1, 0>, 256},
        {"copy19 plain plain", k_copy19<1, 1>, 256},
    };
    for (auto& v : vs) {
        const uint32_t grid = (uint32_t)((a.g.N + v.bs - 1) / v.bs);
        for (int w = 0; w < 3; w++) { a.src = A; a.dst = B; hipLaunchKernelGGL(v.k, dim3(grid), dim3(v.bs), 0, 0, a); }
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) {
            a.src = (r & 1) ? B : A; a.dst = (r & 1) ? A : B;
            hipLaunchKernelGGL(v.k, dim3(grid), dim3(v.bs), 0, 0, a);
        }
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        const double t = ms / reps;
        printf("%-36s %8.3f ms  %7.1f GB/s  %5.1f%%  %8.0f MLUPS\n", v.name, t, algo / t / 1e6, algo / t / 1e6 / 80.0, a.g.N / t / 1e3);
        fflush(stdout);
    }
    // single-stream copy of 2*8.6 GB
    {
        const uint64_t n2 = bytes / 16;
        for (int w = 0; w < 2; w++) hipLaunchKernelGGL(k_copy1, dim3(8192), dim3(256), 0, 0, (const double2*)A, (double2*)B, n2);
        CK(hipDeviceSynchronize()); CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy1, dim3(8192), dim3(256), 0, 0, (const double2*)((r&1)?B:A), (double2*)((r&1)?A:B), n2);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        printf("%-36s %8.3f ms  %7.1f GB/s  %5.1f%%\n", "copy1 dwordx4 grid-stride", ms / reps, 2.0 * bytes / (ms / reps) / 1e6, 2.0 * bytes / (ms / reps) / 1e6 / 80.0);
    }
    CK(hipFree(A)); CK(hipFree(B));
    return 0;
}
