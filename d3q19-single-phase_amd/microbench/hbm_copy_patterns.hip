// HBM bandwidth of copy patterns relevant to the 19-population lattice
// (FP64, 56.6 M nodes = the 192x768x384 channel): 19-stream SoA copies,
// AoSoA tiles, read-only and write-only streams, flat copies with 8 and 16
// bytes per lane.  Timed with HIP events; prints GB/s and % of 8 TB/s.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
typedef double dvec2 __attribute__((ext_vector_type(2)));
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

template <int NT> __device__ __forceinline__ void st(double* p, double v) { if (NT) __builtin_nontemporal_store(v, p); else *p = v; }

template <int NT>
__global__ void __launch_bounds__(256) soa_copy(const double* __restrict__ s, double* __restrict__ d, uint64_t N) {
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++) f[i] = s[i * N + n];
#pragma unroll
    for (int i = 0; i < 19; i++) st<NT>(d + i * N + n, f[i] * 1.0000001);
}
// AoSoA: tile of T nodes, 19 rows of T
template <int T, int NT>
__global__ void __launch_bounds__(256) aosoa_copy(const double* __restrict__ s, double* __restrict__ d, uint64_t N) {
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const uint64_t base = (n / T) * (19 * T) + (n % T);
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++) f[i] = s[base + i * T];
#pragma unroll
    for (int i = 0; i < 19; i++) st<NT>(d + base + i * T, f[i] * 1.0000001);
}
__global__ void __launch_bounds__(256) soa_read(const double* __restrict__ s, double* __restrict__ d, uint64_t N) {
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    double a = 0;
#pragma unroll
    for (int i = 0; i < 19; i++) a += s[i * N + n];
    if (a == 12345.678) d[n] = a;
}
template <int NT>
__global__ void __launch_bounds__(256) soa_write(const double* __restrict__ s, double* __restrict__ d, uint64_t N) {
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
#pragma unroll
    for (int i = 0; i < 19; i++) st<NT>(d + i * N + n, (double)n);
}
template <int NT>
__global__ void __launch_bounds__(256) flat_copy2(const dvec2* __restrict__ s, dvec2* __restrict__ d, uint64_t n2) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n2) return;
    dvec2 v = s[i];
    if (NT) __builtin_nontemporal_store(v, &d[i]); else d[i] = v;
}
__global__ void __launch_bounds__(256) flat_copy1(const double* __restrict__ s, double* __restrict__ d, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    __builtin_nontemporal_store(s[i], &d[i]);
}
// block handles 4096 contiguous doubles (16 per thread in 4 dwordx4 waves)
__global__ void __launch_bounds__(256) flat_copy_chunk(const dvec2* __restrict__ s, dvec2* __restrict__ d, uint64_t n2) {
    const uint64_t b = (uint64_t)blockIdx.x * 2048;
    dvec2 v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) { uint64_t i = b + k * 256 + threadIdx.x; v[k] = i < n2 ? s[i] : dvec2{0, 0}; }
#pragma unroll
    for (int k = 0; k < 8; k++) { uint64_t i = b + k * 256 + threadIdx.x; if (i < n2) __builtin_nontemporal_store(v[k], &d[i]); }
}

// 19-stream SoA copy with two nodes per thread (16-byte aligned loads and
// stores).  With SHIFT, the ten populations with c_x = +-1 are pulled from
// x -+ 1 as the step kernel pulls them: the value that crosses the lane's
// pair comes from the neighbouring lane (ds_bpermute), and the first / last
// lane of a wave loads it from memory.  Every index is checked against N.
template <int SHIFT>
__global__ void __launch_bounds__(256) soa_pair(const double* __restrict__ s, double* __restrict__ d, uint64_t N) {
    const uint64_t n2 = N / 2;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = i < n2;
    const uint32_t lane = threadIdx.x & 63;
    dvec2 f[19];
#pragma unroll
    for (int p = 0; p < 19; p++) f[p] = active ? ((const dvec2*)(s + p * N))[i] : dvec2{0, 0};
    if (SHIFT) {
        const bool first = (lane == 0);
        const bool last = (lane == 63) || (i + 1 >= n2);
#pragma unroll
        for (int p = 0; p < 19; p++) {
            const int cx = (p == 1 || p == 7 || p == 9 || p == 11 || p == 13) ? 1
                         : (p == 2 || p == 8 || p == 10 || p == 12 || p == 14) ? -1 : 0;
            if (cx == 1) {
                double prev = __shfl_up(f[p].y, 1);
                if (first) prev = (active && i > 0) ? s[p * N + 2 * i - 1] : 0.0;
                f[p] = dvec2{prev, f[p].x};
            } else if (cx == -1) {
                double next = __shfl_down(f[p].x, 1);
                if (last) next = (active && 2 * i + 2 < N) ? s[p * N + 2 * i + 2] : 0.0;
                f[p] = dvec2{f[p].y, next};
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int p = 0; p < 19; p++) __builtin_nontemporal_store(f[p] * 1.0000001, ((dvec2*)(d + p * N)) + i);
}

// AoSoA with tiles of T nodes (T = lx: one wall-normal row per tile, its 19
// populations contiguous) and two nodes per thread (16-byte accesses).  With
// SHIFT the ten c_x != 0 populations take their +-1 neighbour across the pair
// boundary from the adjacent lane (or from memory on the first/last lane of
// a wave), as a row-tiled pull would.  N % T == 0 and T even.
template <int T, int SHIFT>
__global__ void __launch_bounds__(256) aosoa_pair(const double* __restrict__ s, double* __restrict__ d, uint64_t N) {
    const uint64_t n2 = N / 2;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = i < n2;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t n = active ? 2 * i : 0;
    const uint64_t base = (n / T) * (19ull * T) + (n % T);
    dvec2 f[19];
#pragma unroll
    for (int p = 0; p < 19; p++) f[p] = active ? *(const dvec2*)(s + base + (uint64_t)p * T) : dvec2{0, 0};
    if (SHIFT) {
        const bool first = (lane == 0) || (n % T == 0);
        const bool last = (lane == 63) || (i + 1 >= n2) || (n % T == T - 2);
#pragma unroll
        for (int p = 0; p < 19; p++) {
            const int cx = (p == 1 || p == 7 || p == 9 || p == 11 || p == 13) ? 1
                         : (p == 2 || p == 8 || p == 10 || p == 12 || p == 14) ? -1 : 0;
            if (cx == 1) {
                double prev = __shfl_up(f[p].y, 1);
                // first node of a row: the wall value (own opposite slot)
                if (first) prev = active ? s[base + (uint64_t)(p + 1 == 19 ? 0 : p + 1) * T] : 0.0;
                f[p] = dvec2{prev, f[p].x};
            } else if (cx == -1) {
                double next = __shfl_down(f[p].x, 1);
                if (last) next = active ? s[base + (uint64_t)(p + 1 == 19 ? 0 : p + 1) * T + 1] : 0.0;
                f[p] = dvec2{f[p].y, next};
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int p = 0; p < 19; p++) __builtin_nontemporal_store(f[p] * 1.0000001, (dvec2*)(d + base + (uint64_t)p * T));
}

// 19-stream SoA copy with population stride S >= N (padding between the
// population arrays), to see whether the arrays' relative alignment matters.
__global__ void __launch_bounds__(256) soa_copy_stride(const double* __restrict__ s, double* __restrict__ d,
                                                       uint64_t N, uint64_t S) {
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    double f[19];
#pragma unroll
    for (int i = 0; i < 19; i++) f[i] = s[i * S + n];
#pragma unroll
    for (int i = 0; i < 19; i++) __builtin_nontemporal_store(f[i] * 1.0000001, d + i * S + n);
}

typedef void (*KF)(const double*, double*, uint64_t);

int main() {
    const uint64_t N = 192ull * 768 * 384;
    const size_t bytes = N * 19 * 8;
    double *A, *B;
    CK(hipMalloc(&A, bytes)); CK(hipMalloc(&B, bytes));
    CK(hipMemset(A, 0, bytes)); CK(hipMemset(B, 0, bytes));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    const int reps = 20;
    struct V { const char* name; KF k; double traffic; } vs[] = {
        {"soa_copy19 nt-st", soa_copy<1>, 2.0 * bytes},
        {"soa_copy19 plain-st", soa_copy<0>, 2.0 * bytes},
        {"aosoa T=64 nt-st", aosoa_copy<64, 1>, 2.0 * bytes},
        {"aosoa T=128 nt-st", aosoa_copy<128, 1>, 2.0 * bytes},
        {"aosoa T=192 nt-st", aosoa_copy<192, 1>, 2.0 * bytes},
        {"aosoa T=256 nt-st", aosoa_copy<256, 1>, 2.0 * bytes},
        {"aosoa T=512 nt-st", aosoa_copy<512, 1>, 2.0 * bytes},
        {"aosoa T=64 plain-st", aosoa_copy<64, 0>, 2.0 * bytes},
        {"soa_read19", soa_read, 1.0 * bytes},
        {"soa_write19 nt", soa_write<1>, 1.0 * bytes},
        {"soa_write19 plain", soa_write<0>, 1.0 * bytes},
        {"soa_pair19 16B nt", soa_pair<0>, 2.0 * bytes},
        {"soa_pair19 16B x-shift nt", soa_pair<1>, 2.0 * bytes},
        {"aosoa T=192 pair 16B nt", aosoa_pair<192, 0>, 2.0 * bytes},
        {"aosoa T=192 pair 16B x-shift", aosoa_pair<192, 1>, 2.0 * bytes},
        {"aosoa T=64 pair 16B nt", aosoa_pair<64, 0>, 2.0 * bytes},
    };
    for (auto& v : vs) {
        // the pair kernels cover two nodes per thread
        const bool pair = v.k == soa_pair<0> || v.k == soa_pair<1> || v.k == aosoa_pair<192, 0> ||
                          v.k == aosoa_pair<192, 1> || v.k == aosoa_pair<64, 0>;
        const uint64_t threads = pair ? N / 2 : N;
        const uint32_t grid = (uint32_t)((threads + 255) / 256);
        for (int w = 0; w < 2; w++) hipLaunchKernelGGL(v.k, dim3(grid), dim3(256), 0, 0, A, B, N);
        CK(hipDeviceSynchronize()); CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(v.k, dim3(grid), dim3(256), 0, 0, (r & 1) ? B : A, (r & 1) ? A : B, N);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        printf("%-28s %8.3f ms %8.1f GB/s %5.1f%%\n", v.name, ms / reps, v.traffic / (ms / reps) / 1e6, v.traffic / (ms / reps) / 1e6 / 80.0);
        fflush(stdout);
    }
    const uint64_t n2 = bytes / 16;
    struct F { const char* name; int kind; } fs[] = {{"flat copy dwordx4 nt", 0}, {"flat copy dwordx4 plain", 1}, {"flat copy dwordx2 nt", 2}, {"flat copy chunk2048 nt", 3}};
    for (auto& f : fs) {
        auto launch = [&](int r) {
            const double* s = (r & 1) ? B : A; double* d = (r & 1) ? A : B;
            if (f.kind == 0) hipLaunchKernelGGL(flat_copy2<1>, dim3((n2 + 255) / 256), dim3(256), 0, 0, (const dvec2*)s, (dvec2*)d, n2);
            else if (f.kind == 1) hipLaunchKernelGGL(flat_copy2<0>, dim3((n2 + 255) / 256), dim3(256), 0, 0, (const dvec2*)s, (dvec2*)d, n2);
            else if (f.kind == 2) hipLaunchKernelGGL(flat_copy1, dim3((2 * n2 + 255) / 256), dim3(256), 0, 0, s, d, 2 * n2);
            else hipLaunchKernelGGL(flat_copy_chunk, dim3((n2 + 2047) / 2048), dim3(256), 0, 0, (const dvec2*)s, (dvec2*)d, n2);
        };
        for (int w = 0; w < 2; w++) launch(w);
        CK(hipDeviceSynchronize()); CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) launch(r);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        printf("%-28s %8.3f ms %8.1f GB/s %5.1f%%\n", f.name, ms / reps, 2.0 * bytes / (ms / reps) / 1e6, 2.0 * bytes / (ms / reps) / 1e6 / 80.0);
        fflush(stdout);
    }
    // population stride padding: strides from N up to the next multiple of 2^24
    CK(hipFree(A)); CK(hipFree(B));
    const uint64_t smax = (N + (1ull << 24) - 1) & ~((1ull << 24) - 1);
    CK(hipMalloc(&A, smax * 19 * 8)); CK(hipMalloc(&B, smax * 19 * 8));
    CK(hipMemset(A, 0, smax * 19 * 8)); CK(hipMemset(B, 0, smax * 19 * 8));
    const uint64_t strides[] = {N, N + 32, N + 256, N + 2048, N + 16384, N + 131072,
                                (N + (1ull << 20) - 1) & ~((1ull << 20) - 1), smax};
    const uint32_t grid = (uint32_t)((N + 255) / 256);
    for (uint64_t S : strides) {
        for (int w = 0; w < 2; w++) hipLaunchKernelGGL(soa_copy_stride, dim3(grid), dim3(256), 0, 0, A, B, N, S);
        CK(hipDeviceSynchronize()); CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(soa_copy_stride, dim3(grid), dim3(256), 0, 0, (r & 1) ? B : A, (r & 1) ? A : B, N, S);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        printf("soa_copy19 stride N+%-9llu %8.3f ms %8.1f GB/s %5.1f%%\n", (unsigned long long)(S - N), ms / reps,
               2.0 * bytes / (ms / reps) / 1e6, 2.0 * bytes / (ms / reps) / 1e6 / 80.0);
        fflush(stdout);
    }
    return 0;
}
