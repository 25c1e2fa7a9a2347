// Can two time steps share one pass over HBM?  The step-t results of a few
// z planes go to a small ring buffer that the 256 MB last-level cache
// (Infinity Cache) can hold; step t+1 of the planes behind them reads that
// ring and writes the output lattice.  Per two steps HBM then carries one
// lattice read and one write instead of two of each.
//
// This measures the memory pattern only: 19-population pull copies (y/z
// shifts wrapped, x shifts clamped) without the collision arithmetic, FP64
// and FP32, on the 192 x 768 x 384 channel (x fastest).  Sequence per two
// steps, for chunks of K planes: A_i (step t of planes [iK, iK+K), lattice ->
// ring) then B_i (step t+1 of planes [iK-1, iK+K-1), ring -> lattice); the
// ring has K+2 plane slots.  Every index is reduced into its buffer.
// Reference: two full-lattice pull copies.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define CK(x)                                                                       \
    do {                                                                            \
        hipError_t e = (x);                                                         \
        if (e != hipSuccess) {                                                      \
            printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__);         \
            exit(1);                                                                \
        }                                                                           \
    } while (0)

constexpr int CXc[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CYc[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZc[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

struct Dims {
    int lx, ly, lz;
    uint64_t plane; // lx * ly
};

// out[i][slot_out(z) * plane + xy] = in[i][slot_in(z - cz) * plane + shifted xy]
// for planes z in [z0, z0 + nz) (z taken modulo lz).  in_slots / out_slots:
// the number of plane slots of each buffer (lz for a lattice, K+2 for the ring);
// a plane z lives in slot z mod slots.
template <typename S, int NT>
__global__ void __launch_bounds__(256) pull_planes(const S *__restrict__ in, uint64_t in_stride, int in_slots,
                                                   S *__restrict__ out, uint64_t out_stride, int out_slots, Dims d,
                                                   int z0, int nz)
{
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= d.plane * (uint64_t)nz)
        return;
    const int zl = (int)(t / d.plane);
    const uint64_t xy = t - (uint64_t)zl * d.plane;
    const int ix = (int)(xy % d.lx), iy = (int)(xy / d.lx);
    const int z = ((z0 + zl) % d.lz + d.lz) % d.lz;
    // source rows for c = -1, 0, +1 (index c + 1), wrapped once per node
    const int ys[3] = {iy + 1 == d.ly ? 0 : iy + 1, iy, iy == 0 ? d.ly - 1 : iy - 1};
    const int zs[3] = {(z + 1 == d.lz ? 0 : z + 1) % in_slots, z % in_slots, (z == 0 ? d.lz - 1 : z - 1) % in_slots};
    const int xs[3] = {ix + 1 == d.lx ? ix : ix + 1, ix, ix == 0 ? 0 : ix - 1};
    S f[19];
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const uint64_t src = (uint64_t)zs[CZc[i] + 1] * d.plane + (uint64_t)ys[CYc[i] + 1] * d.lx + xs[CXc[i] + 1];
        f[i] = in[(uint64_t)i * in_stride + src];
    }
    const uint64_t dst = (uint64_t)(z % out_slots) * d.plane + xy;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        if (NT)
            __builtin_nontemporal_store((S)(f[i] * (S)1.0000001), out + (uint64_t)i * out_stride + dst);
        else
            out[(uint64_t)i * out_stride + dst] = (S)(f[i] * (S)1.0000001);
    }
}

template <typename S>
void run(const char *name, Dims d, const int *Ks, int nK)
{
    const uint64_t N = d.plane * d.lz;
    S *L0, *L1;
    CK(hipMalloc(&L0, N * 19 * sizeof(S)));
    CK(hipMalloc(&L1, N * 19 * sizeof(S)));
    CK(hipMemset(L0, 0, N * 19 * sizeof(S)));
    CK(hipMemset(L1, 0, N * 19 * sizeof(S)));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int reps = 10;
    auto grid = [&](int nz) { return dim3((unsigned)((d.plane * (uint64_t)nz + 255) / 256)); };
    // reference: two full-lattice steps
    for (int w = 0; w < 2; w++)
        hipLaunchKernelGGL((pull_planes<S, 1>), grid(d.lz), dim3(256), 0, 0, L0, N, d.lz, L1, N, d.lz, d, 0, d.lz);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) {
        hipLaunchKernelGGL((pull_planes<S, 1>), grid(d.lz), dim3(256), 0, 0, L0, N, d.lz, L1, N, d.lz, d, 0, d.lz);
        hipLaunchKernelGGL((pull_planes<S, 1>), grid(d.lz), dim3(256), 0, 0, L1, N, d.lz, L0, N, d.lz, d, 0, d.lz);
    }
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double ref = ms / reps;
    const double bytes2 = 2.0 * 2.0 * 19.0 * sizeof(S) * (double)N; // two steps, read + write
    printf("%s two full steps           %8.3f ms  (%.1f%% of 8 TB/s algorithmic)\n", name, ref,
           bytes2 / (ref * 1e-3) / 8e12 * 100.0);
    fflush(stdout);
    for (int k = 0; k < 2 * nK; k++) {
        const int K = Ks[k % nK];
        const bool cached_ring = k >= nK; // second round: ring written with cached stores
        const int slots = K + 2;
        S *R;
        CK(hipMalloc(&R, d.plane * slots * 19 * sizeof(S)));
        CK(hipMemset(R, 0, d.plane * slots * 19 * sizeof(S)));
        const uint64_t rs = d.plane * slots;
        auto pass = [&]() {
            for (int z0 = 0; z0 < d.lz; z0 += K) {
                const int nz = (z0 + K <= d.lz) ? K : d.lz - z0;
                if (cached_ring)
                    hipLaunchKernelGGL((pull_planes<S, 0>), grid(nz), dim3(256), 0, 0, L0, N, d.lz, R, rs, slots, d, z0,
                                       nz);
                else
                    hipLaunchKernelGGL((pull_planes<S, 1>), grid(nz), dim3(256), 0, 0, L0, N, d.lz, R, rs, slots, d, z0,
                                       nz);
                hipLaunchKernelGGL((pull_planes<S, 1>), grid(nz), dim3(256), 0, 0, R, rs, slots, L1, N, d.lz, d, z0 - 1,
                                   nz);
            }
        };
        pass();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            pass();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double t = ms / reps;
        printf("%s blocked K=%-3d %s ring %6.1f MB %8.3f ms  speed-up %.2f  (%d launches)\n", name, K,
               cached_ring ? "cached" : "nt    ",
               d.plane * slots * 19.0 * sizeof(S) / 1e6, t, ref / t, 2 * ((d.lz + K - 1) / K));
        fflush(stdout);
        CK(hipFree(R));
    }
    CK(hipFree(L0));
    CK(hipFree(L1));
}

// Three steps per pass: two rings (step t and step t+1 planes), launches
// A_i (lattice -> ring 1), B_i (ring 1 -> ring 2, one plane behind), C_i
// (ring 2 -> lattice, two planes behind); ring stores cached.
template <typename S>
void run3(const char *name, Dims d, const int *Ks, int nK)
{
    const uint64_t N = d.plane * d.lz;
    S *L0, *L1;
    CK(hipMalloc(&L0, N * 19 * sizeof(S)));
    CK(hipMalloc(&L1, N * 19 * sizeof(S)));
    CK(hipMemset(L0, 0, N * 19 * sizeof(S)));
    CK(hipMemset(L1, 0, N * 19 * sizeof(S)));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int reps = 6;
    auto grid = [&](int nz) { return dim3((unsigned)((d.plane * (uint64_t)nz + 255) / 256)); };
    for (int w = 0; w < 3; w++)
        hipLaunchKernelGGL((pull_planes<S, 1>), grid(d.lz), dim3(256), 0, 0, L0, N, d.lz, L1, N, d.lz, d, 0, d.lz);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++)
        for (int q = 0; q < 3; q++)
            hipLaunchKernelGGL((pull_planes<S, 1>), grid(d.lz), dim3(256), 0, 0, (q & 1) ? L1 : L0, N, d.lz,
                               (q & 1) ? L0 : L1, N, d.lz, d, 0, d.lz);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double ref = ms / reps;
    printf("%s three full steps         %8.3f ms\n", name, ref);
    fflush(stdout);
    for (int k = 0; k < nK; k++) {
        const int K = Ks[k], slots = K + 2;
        S *R1, *R2;
        CK(hipMalloc(&R1, d.plane * slots * 19 * sizeof(S)));
        CK(hipMalloc(&R2, d.plane * slots * 19 * sizeof(S)));
        CK(hipMemset(R1, 0, d.plane * slots * 19 * sizeof(S)));
        CK(hipMemset(R2, 0, d.plane * slots * 19 * sizeof(S)));
        const uint64_t rs = d.plane * slots;
        auto pass = [&]() {
            for (int z0 = 0; z0 < d.lz; z0 += K) {
                const int nz = (z0 + K <= d.lz) ? K : d.lz - z0;
                hipLaunchKernelGGL((pull_planes<S, 0>), grid(nz), dim3(256), 0, 0, L0, N, d.lz, R1, rs, slots, d,
                                   z0, nz);
                hipLaunchKernelGGL((pull_planes<S, 0>), grid(nz), dim3(256), 0, 0, R1, rs, slots, R2, rs, slots, d,
                                   z0 - 1, nz);
                hipLaunchKernelGGL((pull_planes<S, 1>), grid(nz), dim3(256), 0, 0, R2, rs, slots, L1, N, d.lz, d,
                                   z0 - 2, nz);
            }
        };
        pass();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            pass();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double t = ms / reps;
        printf("%s 3-step K=%-3d rings %6.1f MB %8.3f ms  speed-up %.2f\n", name, K,
               2.0 * d.plane * slots * 19.0 * sizeof(S) / 1e6, t, ref / t);
        fflush(stdout);
        CK(hipFree(R1));
        CK(hipFree(R2));
    }
    CK(hipFree(L0));
    CK(hipFree(L1));
}

int main()
{
    Dims d{192, 768, 384, 192ull * 768};
    const int K64[] = {1, 2, 4, 8};
    const int K32[] = {1, 2, 4, 8, 16};
    const char *mode = getenv("TB_MODE");
    if (mode && mode[0] == '3') {
        const int K3a[] = {1, 2, 3};
        const int K3b[] = {2, 4, 6, 8};
        run3<double>("fp64", d, K3a, 3);
        run3<float>("fp32", d, K3b, 4);
        return 0;
    }
    run<double>("fp64", d, K64, 4);
    run<float>("fp32", d, K32, 5);
    return 0;
}
