// LDS read throughput of this GPU, every CU busy: the ceiling of the sieve's dense stage.
// Prints one JSON line.  Two read forms: the sieve's (two 32-bit words per lane per row,
// rows 64 words apart, folded with v_alignbit, one ds_read2st64_b32 per row pair and
// word) and the widest (ds_read_b128, 16 consecutive bytes per lane).  Each block of 16
// waves fills 64 KiB of LDS and reads it for `iters` steps; every address stays inside it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

typedef uint32_t u32;
constexpr u32 kWords = 16384;   // 64 KiB per block

template <int MODE>
__global__ __launch_bounds__(1024) void lds_read_kernel(u32* out, int iters, u32 one) {
    extern __shared__ uint4 lds_raw[];   // 16-byte aligned
    u32* lds = (u32*)lds_raw;
    for (u32 i = threadIdx.x; i < kWords; i += blockDim.x) lds[i] = i * 2654435761u;
    __syncthreads();
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    u32 acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    u32 base = (w * 977u) & 4095u;
    for (int it = 0; it < iters; ++it) {
        base = (base + 1237u) & 4095u;   // every read below stays below 4096 + 1024 words
        if (MODE == 0) {
            const u32* src = lds + base + lane;
            const u32* srch = src + one;
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] |= __builtin_amdgcn_alignbit(srch[64 * r], src[64 * r], (base + r) & 31u);
        } else {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const uint4 v = *(const uint4*)(lds + ((base & ~3u) + 256 * r + 4 * lane));
                acc[4 * r] ^= v.x;
                acc[4 * r + 1] ^= v.y;
                acc[4 * r + 2] ^= v.z;
                acc[4 * r + 3] ^= v.w;
            }
        }
    }
    u32 x = 0;
    for (int r = 0; r < 8; ++r) x ^= acc[r];
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

template <int MODE>
double bytes_per_s(int grid, int iters, u32* d_out, hipEvent_t e0, hipEvent_t e1) {
    const int block = 1024;
    const double per_thread = MODE == 0 ? 64.0 : 32.0;   // bytes read per thread per step
    hipLaunchKernelGGL(lds_read_kernel<MODE>, dim3(grid), dim3(block), kWords * 4, 0, d_out, 8, 1u);
    if (hipEventRecord(e0, 0) != hipSuccess) return 0;
    hipLaunchKernelGGL(lds_read_kernel<MODE>, dim3(grid), dim3(block), kWords * 4, 0, d_out, iters, 1u);
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 0;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0) return 0;
    return (double)grid * block * iters * per_thread / (ms * 1e-3);
}

}  // namespace

int main() {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 1;
    const int cus = p.multiProcessorCount, grid = 2 * cus, iters = 200000;
    u32* d_out = nullptr;
    hipEvent_t e0, e1;
    if (hipMalloc(&d_out, (size_t)grid * 1024 * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        return 1;
    const double pair = bytes_per_s<0>(grid, iters, d_out, e0, e1);
    const double wide = bytes_per_s<1>(grid, iters, d_out, e0, e1);
    (void)hipFree(d_out);
    if (pair <= 0 || wide <= 0) return 1;
    std::printf("{\"arch\": \"%s\", \"cus\": %d, \"clock_khz\": %d, \"b32_pair_bytes_per_s\": %.6e, "
                "\"b128_bytes_per_s\": %.6e}\n",
                p.gcnArchName, cus, p.clockRate, pair, wide);
    return 0;
}
