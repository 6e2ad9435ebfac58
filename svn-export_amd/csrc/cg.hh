// Jacobi-preconditioned conjugate gradients on a SELL-64 matrix, the GPU
// replacement of the KSPSolve CG loop (src/la/solve.cc:316-463) on the
// condensed edge system.
//
// Layout: rows are grouped in slices of 64 (one wavefront); slice s holds
// w_s = max row length entries per row, stored column-major
// (val[slice_ptr[s] + k*64 + lane]), so every load of a wavefront is one
// contiguous 512 B (FP64) segment.  Padding entries have value 0 and point
// at their own row.
//
// Three kernels per iteration (two for systems that stay in the last-level
// cache, see below), no atomics and no device-scope fences: every
// block writes its partial dot products to a per-block array, and every block
// of a later kernel folds the whole (small) partial array itself in a fixed
// order.  All blocks therefore see bitwise identical scalars (alpha, beta, the
// stop test), and the result does not depend on block scheduling.
//   K_D(k): rz_k, rr_k, rz_{k-1} folded; stop test; beta; p = z + beta p_old
//   K_A(k): q = A p; partial p.q
//   K_B(k): alpha = rz_k / p.q; x += alpha p; r -= alpha q; z = D^-1 r;
//           partial r.z, r.r into the (k+1) parity buffers
// Systems whose matrix and vectors exceed the last-level cache (kCgSplitBytes
// of 12 B per stored entry and 56 B per row): the direction is formed by its
// own streaming pass, so the operator gathers one vector per entry (forming
// z + beta p_old at every gathered column instead gathers two), and the
// matrix is read with non-temporal loads, so the vectors stay in the caches
// between the kernels.  Smaller systems stay in the cache from one iteration
// to the next; a launch costs more there than the second gather, so they run
// K_D and K_A as one kernel (k_cg_dir_spmv) with ordinary loads.  Both forms
// give bitwise the same iterates.
#pragma once
#include <hip/hip_runtime.h>

namespace fl {

constexpr int kCgBlock = 512;
// grid of the iteration kernels: kCgGridLarge blocks in the split form,
// kCgGridSmall in the fused one (one wave of resident blocks on 256 CUs, and
// half the partial sums to fold), kCgGridTiny below kCgTinyRows rows
constexpr int kCgGridSmall = 512, kCgGridLarge = 1024;
constexpr int kCgGridTiny = 128, kCgTinyRows = 1 << 17;
// working set above which the split form runs (MI355X: 256 MB Infinity Cache)
constexpr double kCgSplitBytes = 224.0 * (1 << 20);

// done: 0 running, 1 stop test met, 2 paused for a reliable update (mixed precision)
// norm: 0 the stop test is on ||r||_2 / ||b||_2; 1 on the natural norm of the
// Jacobi preconditioner, sqrt(r^T D^-1 r) / sqrt(b^T D^-1 b) (PETSc's
// KSP_NORM_NATURAL); rr, bnorm2, rr_ru and tol2 are squares in that norm.
struct CGState {
    double tol2, rr, bnorm2, rr_ru;
    int done, it, max_it, norm;
};

template <class V>
struct CGDev {
    int n, nslices;
    const int *slice_ptr;  // nslices+1, entry offsets (multiples of 64)
    const int *col;
    const V *val;
    const V *dinv;
    const V *b;
    V *x, *r, *z, *q;
    double *part_pq, *part_rz0, *part_rz1, *part_rr0, *part_rr1;
    CGState *st;
};

// Sum NV per-block partial arrays (each of length G) in a fixed order; the
// totals are returned to every thread of the block.
template <int NV>
__device__ __forceinline__ void fold_partials(const double *const (&parts)[NV], int G, double (&out)[NV]) {
    __shared__ double sh[NV][kCgBlock / 64];
    double acc[NV];
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x)
        for (int k = 0; k < NV; k++) acc[k] += parts[k][g];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < NV; k++) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) sh[k][wid] = v;
    }
    __syncthreads();
    for (int k = 0; k < NV; k++) {
        double t = 0.0;
        for (int w = 0; w < kCgBlock / 64; w++) t += sh[k][w];
        out[k] = t;
    }
    __syncthreads();
}

// Block total of NV values, written by thread 0 to dst[k][blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *const (&dst)[NV]) {
    __shared__ double sh[NV][kCgBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < NV; k++) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) {
            double t = 0.0;
            for (int w = 0; w < kCgBlock / 64; w++) t += sh[k][w];
            dst[k][blockIdx.x] = t;
        }
}

// Slices walked by one wavefront.  Workgroups are dispatched round-robin over
// the 8 XCDs (workgroup b runs on XCD b % 8), each XCD with its own L2.  When
// the grid is a multiple of 8, XCD x walks the x-th eighth of the slices: the
// rows a slice gathers from (its mesh neighbours, numbered a few layers away
// at most) then sit in the same L2 instead of being fetched by all eight.
struct SliceWalk {
    int first, end, step;
};

__device__ __forceinline__ SliceWalk slice_walk(int nslices) {
    const int wid = threadIdx.x >> 6, wpb = blockDim.x >> 6;
    if ((gridDim.x & 7) == 0) {
        const int per = (nslices + 7) >> 3;
        const int s0 = min((int)(blockIdx.x & 7) * per, nslices);
        return SliceWalk{s0 + (int)(blockIdx.x >> 3) * wpb + wid, min(s0 + per, nslices),
                         (int)(gridDim.x >> 3) * wpb};
    }
    return SliceWalk{(int)blockIdx.x * wpb + wid, nslices, (int)gridDim.x * wpb};
}

// Matrix entries: non-temporal loads (NT) where the matrix and vectors exceed
// the last-level cache, so the streamed matrix does not evict the vectors;
// ordinary loads where everything stays cached between iterations.
template <bool NT, class T>
__device__ __forceinline__ T ld_mat(const T *p) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <class V>
__global__ void __launch_bounds__(kCgBlock) k_cg_init(CGDev<V> a, V *p0) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        V b = a.b[i];
        a.x[i] = V(0);
        a.r[i] = b;
        V z = a.dinv[i] * b;
        a.z[i] = z;
        p0[i] = V(0);
        v[0] += (double)b * (double)z;
        v[1] += (double)b * (double)b;
    }
    double *const dst[2] = {a.part_rz0, a.part_rr0};
    block_partials<2>(v, dst);
}

// G: number of blocks of the main kernels (length of the partial arrays)
// bnorm2 > 0: norm^2 of b given (warm start, the initial residual is not b)
template <class V>
__global__ void __launch_bounds__(kCgBlock) k_cg_setup(CGDev<V> a, int G, double rtol2, double atol2, int max_it,
                                                     double bnorm2, int norm) {
    const double *const parts[2] = {a.part_rr0, a.part_rz0};
    double t[2];
    fold_partials<2>(parts, G, t);
    if (threadIdx.x == 0) {
        const double r0 = norm ? t[1] : t[0];
        double bb = bnorm2 > 0.0 ? bnorm2 : r0;
        a.st->bnorm2 = bb;
        a.st->rr = r0;
        a.st->tol2 = fmax(rtol2 * bb, atol2);
        a.st->max_it = max_it;
        a.st->it = 0;
        a.st->done = 0;
        a.st->norm = norm;
        a.st->rr_ru = r0;
    }
}

// Folds rz_k, rr_k and rz_{k-1}; false (and the state set by block 0) when the
// iteration stops or pauses at step k, else true with beta_k.
// is_ru: the residual of step k is a freshly recomputed true residual;
// delta2 > 0: pause (done = 2) once the recursive residual has dropped by
// sqrt(delta2) since the last true residual (reliable-update criterion).
template <class V>
__device__ __forceinline__ bool cg_beta(const CGDev<V> &a, int k, int G, int is_ru, double delta2, V &beta) {
    CGState *st = a.st;
    const double *cur_rz = (k & 1) ? a.part_rz1 : a.part_rz0;
    const double *cur_rr = (k & 1) ? a.part_rr1 : a.part_rr0;
    const double *old_rz = (k & 1) ? a.part_rz0 : a.part_rz1;
    const double *const parts[3] = {cur_rz, cur_rr, old_rz};
    double s[3];
    fold_partials<3>(parts, G, s);
    double rz = s[0], rr = st->norm ? s[0] : s[1];
    if (rr <= st->tol2 || k >= st->max_it || !(rr == rr)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->done = 1;
            st->it = k;
            st->rr = rr;
        }
        return false;
    }
    if (is_ru) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->rr_ru = rr;
    } else if (delta2 > 0.0 && rr < delta2 * st->rr_ru) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->done = 2;
            st->it = k;
            st->rr = rr;
        }
        return false;
    }
    beta = (V)(k == 0 ? 0.0 : rz / s[2]);
    return true;
}

template <class V>
__global__ void __launch_bounds__(kCgBlock) k_cg_dir(CGDev<V> a, int k, int G, const V *__restrict__ p_in,
                                                   V *__restrict__ p_out, int is_ru, double delta2) {
    if (a.st->done) return;
    V beta;
    if (!cg_beta(a, k, G, is_ru, delta2, beta)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x)
        p_out[i] = a.z[i] + beta * p_in[i];
}

// q = A p and the partial p.q; skipped once K_D has stopped the iteration
template <class V, bool NT>
__global__ void __launch_bounds__(kCgBlock) k_cg_spmv(CGDev<V> a, const V *__restrict__ p) {
    if (a.st->done) return;
    double v[1] = {0.0};
    const int lane = threadIdx.x & 63;
    const SliceWalk sw = slice_walk(a.nslices);
    for (int sl = sw.first; sl < sw.end; sl += sw.step) {
        const int row = sl * 64 + lane;
        const int e0 = a.slice_ptr[sl] + lane;
        const int w = (a.slice_ptr[sl + 1] - a.slice_ptr[sl]) >> 6;
        V acc = V(0);
        int k = 0;
        // 4 entries per step: all index / value loads issue before the gathers
        for (; k + 4 <= w; k += 4) {
            const int i0 = e0 + k * 64;
            const int c0 = ld_mat<NT>(a.col + i0), c1 = ld_mat<NT>(a.col + i0 + 64);
            const int c2 = ld_mat<NT>(a.col + i0 + 128), c3 = ld_mat<NT>(a.col + i0 + 192);
            const V v0 = ld_mat<NT>(a.val + i0), v1 = ld_mat<NT>(a.val + i0 + 64);
            const V v2 = ld_mat<NT>(a.val + i0 + 128), v3 = ld_mat<NT>(a.val + i0 + 192);
            const V g0 = p[c0], g1 = p[c1], g2 = p[c2], g3 = p[c3];
            acc += v0 * g0;
            acc += v1 * g1;
            acc += v2 * g2;
            acc += v3 * g3;
        }
        for (; k < w; k++) {
            const int i0 = e0 + k * 64;
            acc += ld_mat<NT>(a.val + i0) * p[ld_mat<NT>(a.col + i0)];
        }
        if (row < a.n) {
            a.q[row] = acc;
            v[0] += (double)p[row] * (double)acc;
        }
    }
    double *const dst[1] = {a.part_pq};
    block_partials<1>(v, dst);
}

// K_D and K_A in one kernel: the direction z + beta p_old is formed at every
// gathered column (two gathers per entry) and written for the own rows.
template <class V, bool NT>
__global__ void __launch_bounds__(kCgBlock) k_cg_dir_spmv(CGDev<V> a, int k, int G, const V *__restrict__ p_in,
                                                        V *__restrict__ p_out, int is_ru, double delta2) {
    if (a.st->done) return;
    V beta;
    if (!cg_beta(a, k, G, is_ru, delta2, beta)) return;
    double v[1] = {0.0};
    const int lane = threadIdx.x & 63;
    const SliceWalk sw = slice_walk(a.nslices);
    for (int sl = sw.first; sl < sw.end; sl += sw.step) {
        const int row = sl * 64 + lane;
        const int e0 = a.slice_ptr[sl] + lane;
        const int w = (a.slice_ptr[sl + 1] - a.slice_ptr[sl]) >> 6;
        V acc = V(0);
        int j = 0;
        for (; j + 4 <= w; j += 4) {
            const int i0 = e0 + j * 64;
            const int c0 = ld_mat<NT>(a.col + i0), c1 = ld_mat<NT>(a.col + i0 + 64);
            const int c2 = ld_mat<NT>(a.col + i0 + 128), c3 = ld_mat<NT>(a.col + i0 + 192);
            const V v0 = ld_mat<NT>(a.val + i0), v1 = ld_mat<NT>(a.val + i0 + 64);
            const V v2 = ld_mat<NT>(a.val + i0 + 128), v3 = ld_mat<NT>(a.val + i0 + 192);
            const V g0 = a.z[c0] + beta * p_in[c0], g1 = a.z[c1] + beta * p_in[c1];
            const V g2 = a.z[c2] + beta * p_in[c2], g3 = a.z[c3] + beta * p_in[c3];
            acc += v0 * g0;
            acc += v1 * g1;
            acc += v2 * g2;
            acc += v3 * g3;
        }
        for (; j < w; j++) {
            const int i0 = e0 + j * 64;
            const int c = ld_mat<NT>(a.col + i0);
            acc += ld_mat<NT>(a.val + i0) * (a.z[c] + beta * p_in[c]);
        }
        if (row < a.n) {
            const V p = a.z[row] + beta * p_in[row];
            p_out[row] = p;
            a.q[row] = acc;
            v[0] += (double)p * (double)acc;
        }
    }
    double *const dst[1] = {a.part_pq};
    block_partials<1>(v, dst);
}

template <class V>
__global__ void __launch_bounds__(kCgBlock) k_cg_update(CGDev<V> a, int k, int G, const V *__restrict__ p) {
    CGState *st = a.st;
    if (st->done) return;
    const double *cur_rz = (k & 1) ? a.part_rz1 : a.part_rz0;
    const double *const parts[2] = {cur_rz, a.part_pq};
    double s[2];
    fold_partials<2>(parts, G, s);
    const V alpha = (V)(s[0] / s[1]);
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        a.x[i] += alpha * p[i];
        V r = a.r[i] - alpha * a.q[i];
        a.r[i] = r;
        V z = a.dinv[i] * r;
        a.z[i] = z;
        v[0] += (double)r * (double)z;
        v[1] += (double)r * (double)r;
    }
    double *const dst[2] = {(k & 1) ? a.part_rz0 : a.part_rz1, (k & 1) ? a.part_rr0 : a.part_rr1};
    block_partials<2>(v, dst);
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 CG with FP64 reliable updates.  The FP32 iteration
// runs the same two kernels with V = float (half the bytes); every few
// iterations, and before accepting convergence, the correction accumulated in
// x32 is added to the FP64 iterate, the true residual b - A x is recomputed in
// FP64 from the FP64 matrix and replaces the recursive FP32 residual (the
// search direction is kept, so the Krylov progress is not lost).  The stop test
// of the step after an update therefore uses the true FP64 residual.
// ---------------------------------------------------------------------------
static __global__ void k_to_float(int n, const double *__restrict__ src, float *__restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = (float)src[i];
}

static __global__ void k_ru_accum(int n, double *__restrict__ x64, float *__restrict__ x32) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x64[i] += (double)x32[i];
        x32[i] = 0.0f;
    }
}

// r64 = b - A64 x64 ; r32 = r64 ; z32 = dinv32 r32 ; partials r.z, r.r (FP64)
static __global__ void __launch_bounds__(kCgBlock) k_ru_resid(CGDev<float> a, const double *__restrict__ b64,
                                                     const double *__restrict__ val64,
                                                     const double *__restrict__ x64, double *part_rz,
                                                     double *part_rr) {
    double v[2] = {0.0, 0.0};
    const int lane = threadIdx.x & 63;
    const SliceWalk sw = slice_walk(a.nslices);
    for (int sl = sw.first; sl < sw.end; sl += sw.step) {
        const int row = sl * 64 + lane;
        const int e0 = a.slice_ptr[sl] + lane;
        const int w = (a.slice_ptr[sl + 1] - a.slice_ptr[sl]) >> 6;
        double acc = 0.0;
        for (int k = 0; k < w; k++) {
            const int i0 = e0 + k * 64;
            acc += val64[i0] * x64[a.col[i0]];
        }
        if (row < a.n) {
            double r = b64[row] - acc;
            float r32 = (float)r;
            float z32 = a.dinv[row] * r32;
            a.r[row] = r32;
            a.z[row] = z32;
            v[0] += (double)r32 * (double)z32;
            v[1] += r * r;
        }
    }
    double *const dst[2] = {part_rz, part_rr};
    block_partials<2>(v, dst);
}

static __global__ void k_cg_clear_done(CGState *st) { st->done = 0; }

// FP64 warm start: r = b - A x, z = D^-1 r, p0 = 0, partials r.z, r.r
static __global__ void __launch_bounds__(kCgBlock) k_cg_init_warm(CGDev<double> a, double *p0) {
    double v[2] = {0.0, 0.0};
    const int lane = threadIdx.x & 63;
    const SliceWalk sw = slice_walk(a.nslices);
    for (int sl = sw.first; sl < sw.end; sl += sw.step) {
        const int row = sl * 64 + lane;
        const int e0 = a.slice_ptr[sl] + lane;
        const int w = (a.slice_ptr[sl + 1] - a.slice_ptr[sl]) >> 6;
        double acc = 0.0;
        for (int k = 0; k < w; k++) {
            const int i0 = e0 + k * 64;
            acc += a.val[i0] * a.x[a.col[i0]];
        }
        if (row < a.n) {
            double r = a.b[row] - acc;
            double z = a.dinv[row] * r;
            a.r[row] = r;
            a.z[row] = z;
            a.q[row] = 0.0;
            p0[row] = 0.0;
            v[0] += r * z;
            v[1] += r * r;
        }
    }
    double *const dst[2] = {a.part_rz0, a.part_rr0};
    block_partials<2>(v, dst);
}

}  // namespace fl
