// Shared helpers for the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>

namespace fl {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FL_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess)                                                                          \
            throw ::fl::Error(-2, std::string(#call) + " failed: " + hipGetErrorString(_e) + " at " + \
                                      __FILE__ + ":" + std::to_string(__LINE__));                      \
    } while (0)

template <class T>
T *dev_alloc(size_t n) {
    T *p = nullptr;
    if (n == 0) n = 1;
    FL_HIP(hipMalloc((void **)&p, n * sizeof(T)));
    return p;
}
template <class T>
T *dev_upload(const T *h, size_t n, hipStream_t s) {
    T *p = dev_alloc<T>(n);
    if (n) FL_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return p;
}
inline void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

constexpr int kBlock = 256;

inline int grid_for(long n, int block = kBlock) {
    long g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : g);
}

// Wave-then-block reduction of one double; result valid in thread 0.
__device__ inline double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
    __syncthreads();
    return r;
}

}  // namespace fl
