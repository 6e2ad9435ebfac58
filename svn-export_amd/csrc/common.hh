// Shared helpers for the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace fl {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Status codes of the C ABI beyond the generic ones (-1 input, -2 HIP, ...).
// A positive status is a warning: the call completed and its outputs are
// valid, but a solver stopped before its tolerance (the reference prints
// PETSc's convergence reason and continues, solve.cc:454-458).
enum Status : int {
    kStatusNotConverged = 1,
    kErrSingularElement = -10,  // |det| of a local RT0 matrix below NUM_ZERO (mh_fe_values.cc:54-59)
    kErrBadField = -11,         // conductivity / cross-section not finite and positive
    kErrDiverged = -12,         // CG residual became NaN or infinite
    kErrCallback = -13,         // a caller's data callback reported a failure
    kErrSorption = -14,         // sorption: negative total mass or an unbracketed isotherm root
};

// NUM_ZERO of the reference (math_fce.h:51): the singularity threshold of
// the small-matrix inverses (math_fce.cc:147,168,211).
constexpr double kNumZero = DBL_MIN / DBL_EPSILON;

// Index of the first value that is not finite and positive, or -1.
inline long first_not_positive(const double *v, long n) {
    for (long i = 0; i < n; i++)
        if (!(v[i] > 0.0 && v[i] <= DBL_MAX)) return i;
    return -1;
}

// The same for a device array (one check kernel + a 4-byte copy back;
// synchronizes the stream).  `flag` is one device int of scratch.
long first_not_positive_device(const double *v, long n, int *flag, hipStream_t s);

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

// Grid-wide reduction of NV doubles per thread without a second launch: every
// block writes its NV block totals to `partial`, the last block to arrive
// (atomic ticket) folds all partials with all of its threads in a fixed order
// and a fixed tree, so the result is deterministic.  Returns true in the last
// block only, where v[] then holds the grid totals in every thread.  The
// ticket counter is reset for the next launch.  Requires blockDim.x == 256.
struct SumOp {
    __device__ static double id() { return 0.0; }
    __device__ static double op(double a, double b) { return a + b; }
};
struct MaxOp {
    __device__ static double id() { return 0.0; }  // used on non-negative values
    __device__ static double op(double a, double b) { return fmax(a, b); }
};

template <int NV, class Op = SumOp>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], double *partial, unsigned *count) {
    __shared__ double sh[NV][4];
    __shared__ bool is_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < NV; k++) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x = Op::op(x, __shfl_down(x, o, 64));
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++)
            partial[(size_t)NV * blockIdx.x + k] = Op::op(Op::op(sh[k][0], sh[k][1]), Op::op(sh[k][2], sh[k][3]));
        __threadfence();
        is_last = (atomicAdd(count, 1u) == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    const volatile double *pv = partial;
    double acc[NV];
    for (int k = 0; k < NV; k++) acc[k] = Op::id();
    for (unsigned g = threadIdx.x; g < gridDim.x; g += blockDim.x)
        for (int k = 0; k < NV; k++) acc[k] = Op::op(acc[k], pv[(size_t)NV * g + k]);
    __syncthreads();
    for (int k = 0; k < NV; k++) {
        double x = acc[k];
        for (int o = 32; o > 0; o >>= 1) x = Op::op(x, __shfl_down(x, o, 64));
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    for (int k = 0; k < NV; k++) v[k] = Op::op(Op::op(sh[k][0], sh[k][1]), Op::op(sh[k][2], sh[k][3]));
    if (threadIdx.x == 0) *count = 0;
    return true;
}

// Achievable HBM bandwidth (stream.hip): gbs[0] copy, gbs[1] read, gbs[2] write, GB/s.
void stream_bandwidth(size_t bytes, int reps, double *gbs);

}  // namespace fl
