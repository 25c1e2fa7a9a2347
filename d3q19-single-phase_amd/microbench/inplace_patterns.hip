// HBM bandwidth of in-place lattice updates against the two-lattice copy
// (FP64, 56.6 M nodes = the 192x768x384 channel).  The in-place patterns are
// the memory traffic of the AA streaming scheme: every thread reads and
// writes the same 19 locations, so one lattice suffices.
//   even   read slot i of node n, write slot opp(i) of node n
//   odd    read slot opp(i) of node n - c_i, write slot i of node n + c_i
//          (node offsets c_x + lx c_y + lx ly c_z, wrapped modulo N, so every
//          index stays in [0, 19 N))
// Each location is read and written by one thread only, in either step.
// Timed with HIP events; prints GB/s and % of 8 TB/s (traffic = 2 x 19 x 8 B
// per node).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define CK(x)                                                                                 \
    do {                                                                                      \
        hipError_t e = (x);                                                                   \
        if (e != hipSuccess) {                                                                \
            printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__);                   \
            exit(1);                                                                          \
        }                                                                                     \
    } while (0)

__constant__ int CXc[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
__constant__ int CYc[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
__constant__ int CZc[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPPh[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

template <int NT>
__device__ __forceinline__ void st(double *p, double v)
{
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// two lattices: the reference point
template <int NT>
__global__ void __launch_bounds__(256) soa_copy(const double *__restrict__ s, double *__restrict__ d, uint64_t N)
{
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N)
        return;
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++)
        f[i] = s[i * N + n];
#pragma unroll
    for (int i = 0; i < 19; i++)
        st<NT>(d + i * N + n, f[i] * 1.0000001);
}

template <int NT>
__global__ void __launch_bounds__(256) aa_even(double *L, uint64_t N)
{
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N)
        return;
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++)
        f[i] = L[i * N + n];
#pragma unroll
    for (int i = 0; i < 19; i++)
        st<NT>(L + OPPh[i] * N + n, f[i] * 1.0000001);
}

template <int NT>
__global__ void __launch_bounds__(256) aa_odd(double *L, uint64_t N, int64_t lx, int64_t plane)
{
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N)
        return;
    double f[19];
    uint64_t dst[19];
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const int64_t off = CXc[i] + lx * CYc[i] + plane * CZc[i];
        const uint64_t src = (uint64_t)(((int64_t)n - off + (int64_t)N) % (int64_t)N);
        dst[i] = (uint64_t)(((int64_t)n + off + (int64_t)N) % (int64_t)N);
        f[i] = L[OPPh[i] * N + src];
    }
#pragma unroll
    for (int i = 0; i < 19; i++)
        st<NT>(L + i * N + dst[i], f[i] * 1.0000001);
}

// pull from x - c_i into a second lattice (the current step kernel's
// pattern without the collision), same wrap
__global__ void __launch_bounds__(256) pull_copy(const double *__restrict__ s, double *__restrict__ d, uint64_t N,
                                                 int64_t lx, int64_t plane)
{
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N)
        return;
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const int64_t off = CXc[i] + lx * CYc[i] + plane * CZc[i];
        const uint64_t src = (uint64_t)(((int64_t)n - off + (int64_t)N) % (int64_t)N);
        f[i] = s[i * N + src];
    }
#pragma unroll
    for (int i = 0; i < 19; i++)
        __builtin_nontemporal_store(f[i] * 1.0000001, d + i * N + n);
}

int main()
{
    const int64_t lx = 192, ly = 768, lz = 384;
    const uint64_t N = (uint64_t)(lx * ly * lz);
    const size_t bytes = N * 19 * 8;
    double *A, *B;
    CK(hipMalloc(&A, bytes));
    CK(hipMalloc(&B, bytes));
    CK(hipMemset(A, 0, bytes));
    CK(hipMemset(B, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int reps = 20;
    const uint32_t grid = (uint32_t)((N + 255) / 256);
    const double traffic = 2.0 * bytes;
    auto report = [&](const char *name, float ms) {
        const double t = ms / reps;
        printf("%-34s %8.3f ms %8.1f GB/s %5.1f%%\n", name, t, traffic / t / 1e6, traffic / t / 1e6 / 80.0);
        fflush(stdout);
    };
    for (int variant = 0; variant < 7; variant++) {
        auto launch = [&](int r) {
            double *s = (r & 1) ? B : A, *d = (r & 1) ? A : B;
            switch (variant) {
            case 0: hipLaunchKernelGGL(soa_copy<1>, dim3(grid), dim3(256), 0, 0, s, d, N); break;
            case 1: hipLaunchKernelGGL(pull_copy, dim3(grid), dim3(256), 0, 0, s, d, N, lx, lx * ly); break;
            case 2: hipLaunchKernelGGL(aa_even<1>, dim3(grid), dim3(256), 0, 0, A, N); break;
            case 3: hipLaunchKernelGGL(aa_even<0>, dim3(grid), dim3(256), 0, 0, A, N); break;
            case 4: hipLaunchKernelGGL(aa_odd<1>, dim3(grid), dim3(256), 0, 0, A, N, lx, lx * ly); break;
            case 5: hipLaunchKernelGGL(aa_odd<0>, dim3(grid), dim3(256), 0, 0, A, N, lx, lx * ly); break;
            default:
                if (r & 1)
                    hipLaunchKernelGGL(aa_odd<0>, dim3(grid), dim3(256), 0, 0, A, N, lx, lx * ly);
                else
                    hipLaunchKernelGGL(aa_even<0>, dim3(grid), dim3(256), 0, 0, A, N);
                break;
            }
        };
        static const char *names[] = {"two lattices, copy, nt stores", "two lattices, pull, nt stores",
                                      "in place, AA even, nt stores", "in place, AA even, cached stores",
                                      "in place, AA odd, nt stores", "in place, AA odd, cached stores",
                                      "in place, AA even/odd alternating"};
        for (int w = 0; w < 2; w++)
            launch(w);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            launch(r);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        report(names[variant], ms);
    }
    CK(hipFree(A));
    CK(hipFree(B));
    return 0;
}
