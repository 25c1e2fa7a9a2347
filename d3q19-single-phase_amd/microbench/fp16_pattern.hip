// 19-stream pull copy pattern with FP16 elements: 1 node per lane (2 B loads)
// vs 2 nodes per lane (4 B loads).  x shift ignored (pattern only), y/z
// neighbours wrapped, so every index is in [0, 19 N).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)
__constant__ int CYc[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
__constant__ int CZc[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr uint32_t LX = 192, LY = 768, LZ = 384;
constexpr uint64_t N = (uint64_t)LX * LY * LZ;

__device__ __forceinline__ uint32_t src(uint32_t n, int i)
{
    uint32_t ix = n % LX, q = n / LX, iy = q % LY, iz = q / LY;
    iy = (iy + LY - CYc[i]) % LY;
    iz = (iz + LZ - CZc[i]) % LZ;
    return ix + LX * (iy + LY * iz);
}

template <typename E, int VW>
__global__ void __launch_bounds__(256) k_copy(const E *__restrict__ in, E *__restrict__ out)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t nruns = (uint32_t)(N / VW);
    if (t >= nruns) return;
    const uint32_t n = t * VW;
    E v[19];
#pragma unroll
    for (int i = 0; i < 19; i++)
        v[i] = in[((uint64_t)i * N + src(n, i)) / VW];
#pragma unroll
    for (int i = 0; i < 19; i++)
        __builtin_nontemporal_store(v[i], &out[((uint64_t)i * N + n) / VW]);
}

template <typename E, int VW>
static void run(const char *name, void *a, void *b)
{
    const uint32_t nruns = (uint32_t)(N / VW);
    dim3 grid((nruns + 255) / 256), blk(256);
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    for (int w = 0; w < 3; w++) hipLaunchKernelGGL((k_copy<E, VW>), grid, blk, 0, 0, (const E *)a, (E *)b);
    CK(hipGetLastError());
    CK(hipEventRecord(e0));
    const int reps = 20;
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL((k_copy<E, VW>), grid, blk, 0, 0, (const E *)a, (E *)b);
    CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
    float ms; CK(hipEventElapsedTime(&ms, e0, e1)); ms /= reps;
    const double gbs = 2.0 * 19 * 2 * (double)N / (ms * 1e-3) / 1e9;
    printf("%-28s %.3f ms  %.1f GB/s  %.1f%%  %.0f MLUPS\n", name, ms, gbs, gbs / 80.0, N / (ms * 1e-3) / 1e6);
}

int main()
{
    void *a, *b;
    const size_t bytes = 19 * N * 2;
    CK(hipMalloc(&a, bytes)); CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 0, bytes)); CK(hipMemset(b, 0, bytes));
    for (int k = 0; k < 2; k++) {
        run<uint16_t, 1>("fp16 1 node/lane (2 B)", a, b);
        run<uint32_t, 2>("fp16 2 nodes/lane (4 B)", a, b);
        run<uint64_t, 4>("fp16 4 nodes/lane (8 B)", a, b);
    }
    CK(hipFree(a)); CK(hipFree(b));
    return 0;
}
