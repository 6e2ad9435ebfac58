// Achievable HBM bandwidth of the device: the denominator beside the vendor
// peak in roofline fractions.  Three streaming kernels over buffers far larger
// than the caches (copy, read-only reduction, write-only fill), 16-byte
// accesses, each block a contiguous span so consecutive lanes touch
// consecutive 16-byte words.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hh"

namespace fl {
namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kStreamBlock = 256, kStreamUnroll = 4;

// copy: every thread moves kStreamUnroll 16-byte words per pass, all loads first
__global__ void __launch_bounds__(kStreamBlock) k_stream_copy(const f4v *__restrict__ src, f4v *__restrict__ dst,
                                                              size_t n) {
    const size_t stride = (size_t)gridDim.x * kStreamBlock * kStreamUnroll;
    for (size_t base = (size_t)blockIdx.x * kStreamBlock * kStreamUnroll + threadIdx.x; base < n; base += stride) {
        f4v v[kStreamUnroll];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; u++) {
            const size_t i = base + (size_t)u * kStreamBlock;
            if (i < n) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int u = 0; u < kStreamUnroll; u++) {
            const size_t i = base + (size_t)u * kStreamBlock;
            if (i < n) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

// read: every word is summed; one partial per block keeps the loads alive
__global__ void __launch_bounds__(kStreamBlock) k_stream_read(const f4v *__restrict__ src, size_t n,
                                                              float *__restrict__ part) {
    __shared__ float sh[kStreamBlock / 64];
    const size_t stride = (size_t)gridDim.x * kStreamBlock * kStreamUnroll;
    float acc = 0.0f;
    for (size_t base = (size_t)blockIdx.x * kStreamBlock * kStreamUnroll + threadIdx.x; base < n; base += stride) {
        f4v v[kStreamUnroll];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; u++) {
            const size_t i = base + (size_t)u * kStreamBlock;
            v[u] = i < n ? __builtin_nontemporal_load(src + i) : f4v{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int u = 0; u < kStreamUnroll; u++) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// write: a constant fill
__global__ void __launch_bounds__(kStreamBlock) k_stream_fill(f4v *__restrict__ dst, size_t n, float v0) {
    const size_t stride = (size_t)gridDim.x * kStreamBlock * kStreamUnroll;
    const f4v v{v0, v0, v0, v0};
    for (size_t base = (size_t)blockIdx.x * kStreamBlock * kStreamUnroll + threadIdx.x; base < n; base += stride) {
#pragma unroll
        for (int u = 0; u < kStreamUnroll; u++) {
            const size_t i = base + (size_t)u * kStreamBlock;
            if (i < n) __builtin_nontemporal_store(v, dst + i);
        }
    }
}

}  // namespace

// GB/s of copy (read + write counted), read and write over `bytes` per buffer,
// the best of `reps` timed launches each after one warm-up launch.
void stream_bandwidth(size_t bytes, int reps, double *gbs) {
    if (bytes < (1u << 20) || bytes % 16 != 0) throw Error(-3, "stream buffer: at least 1 MiB, a multiple of 16 bytes");
    if (reps < 1 || reps > 1000) throw Error(-3, "stream repetitions: 1..1000");
    const size_t n = bytes / 16;
    int dev = 0, ncu = 0;
    FL_HIP(hipGetDevice(&dev));
    FL_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::max(1, std::min<int>(ncu * 8, (int)((n + kStreamBlock * kStreamUnroll - 1) /
                                                              (kStreamBlock * kStreamUnroll))));
    f4v *a = nullptr, *b = nullptr;
    float *part = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = nullptr;
    // released on every exit
    struct Release {
        f4v *&a, *&b;
        float *&part;
        hipEvent_t &e0, &e1;
        hipStream_t &s;
        ~Release() {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (part) (void)hipFree(part);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        }
    } rel{a, b, part, e0, e1, s};
    FL_HIP(hipMalloc(&a, bytes));
    FL_HIP(hipMalloc(&b, bytes));
    FL_HIP(hipMalloc(&part, grid * sizeof(float)));
    FL_HIP(hipStreamCreate(&s));
    FL_HIP(hipEventCreate(&e0));
    FL_HIP(hipEventCreate(&e1));
    for (int mode = 0; mode < 3; mode++) {
        float best = 0.0f;
        for (int r = 0; r <= reps; r++) {
            FL_HIP(hipEventRecord(e0, s));
            if (mode == 0) hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kStreamBlock), 0, s, a, b, n);
            else if (mode == 1) hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kStreamBlock), 0, s, a, n, part);
            else hipLaunchKernelGGL(k_stream_fill, dim3(grid), dim3(kStreamBlock), 0, s, a, n, 1.0f);
            FL_HIP(hipGetLastError());
            FL_HIP(hipEventRecord(e1, s));
            FL_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            FL_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0 && (best == 0.0f || ms < best)) best = ms;  // launch 0 warms up
        }
        const double moved = (mode == 0 ? 2.0 : 1.0) * (double)bytes;
        gbs[mode] = moved / (best * 1e-3) / 1e9;
    }
}

}  // namespace fl
