// Batched Monte-Carlo Darcy solver (matrix-free condensed operator,
// realization-interleaved CG).  The closed form per element is the one of
// darcy.hip (assembly_steady_mh_matrix darcy_flow_mh.cc:525-694, make_schur1/2
// :1009-1131, SchurComplement::resolve schur.cc:714-773), split into its
// K-independent factors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <condition_variable>
#include <cstdlib>
#include <deque>
#include <exception>
#include <functional>
#include <mutex>
#include <numeric>
#include <thread>
#include <unordered_map>


#include "batched.hh"
#include "coarse.hh"
#include "common.hh"
#include "condense.hh"

namespace fl {

namespace {

constexpr int kWaves = 4;            // waves per block of the row kernels
constexpr int kAggWaves = 8;         // waves per block of the aggregate-blocked kernels (one aggregate per block)
constexpr int kFoldThreads = 1024;   // 16 waves per fold block
constexpr int kMaxGrid = 16384;      // partial arrays are sized for this many blocks
constexpr int kMaxLaunchGrid = 2048; // blocks of the row-strided kernels
constexpr int kMaxConn = 16;         // element index packing e * 16 + local index

// Per-slot preconditioner arrays of the FP32 path (Jacobi diagonal D^-1, the
// coarse column phi) are stored as bfloat16: they define the preconditioner,
// not the operator, so rounding them changes M slightly but consistently
// (every kernel reads the same stored values) and halves their traffic.
__host__ __device__ inline unsigned short f_to_bf16(float f) {  // round to nearest even, finite inputs
    unsigned b;
    __builtin_memcpy(&b, &f, 4);
    return (unsigned short)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short u) { return __uint_as_float((unsigned)u << 16); }
template <class V>
struct Pre {  // storage of D^-1 in the V path
    using T = V;
    __device__ static V ld(const T *p, size_t o) { return p[o]; }
};
template <>
struct Pre<float> {
    using T = unsigned short;
    __device__ static float ld(const T *p, size_t o) { return bf16_to_f(p[o]); }
};

// Element table layouts (offsets in values):
//  operator table, element without vb neighbours (nvb = 0):
//      P0[n0*n0] (= c_i c_j (inv0_ij - a0_i a0_j / s0), condensed in FP64), dlt[n0]
//      so that S_e(K) x = K P0 x + dlt.x
//  operator table, element with vb neighbours (fractures, channels):
//      C0[n0*n0] (= c_i c_j inv0_ij), w0[n0] (= c_i a0_i), dlt[n0], s0, V, v[nvb]
//  resolve table   inv0[n0*n0], a0[n0], Ag0[n0], rf[n0], dlt[n0], s0, sAg0, hc, V, v[nvb]
// meta[e] = n0 | nvb << 4 | cmask << 12
// Condensing P0 in FP64 keeps the FP32 operator as accurate as an FP32
// assembled matrix: forming K C0 x - K w0 (w0^T x) / s0 at run time in FP32
// cancels catastrophically on smooth x (the element block annihilates
// constants).  Elements with vb neighbours are evaluated in FP64 always.
__host__ __device__ inline int op_size(int n0, int nvb) {
    return nvb ? n0 * n0 + 2 * n0 + 2 + nvb : n0 * n0 + n0;
}
__host__ __device__ inline int rs_size(int n0, int nvb) { return n0 * n0 + 4 * n0 + 4 + nvb; }

template <class V>
struct BArgs {
    int F, Rp, G, xcd;
    const int *row_ptr, *row_c, *meta, *op_off, *conn_off, *conn;
    const int *rec;  // packed row records (16 ints per row)
    const int *frec; // FP32 tetrahedron-pair row records (20 words per row) or null
    const V *opT, *K;
    const typename Pre<V>::T *dinv;
    const double *opT64, *K64;  // FP64 table and K (elements with vb neighbours)
    V *x, *r, *z, *p0, *p1, *q;
    double *part0, *part1, *part2;
    BatchScal sc;
};

// Rows of this wave: start, stride, end.
//  xcd 0: rows strided over the whole grid
//  xcd 1: block b runs on XCD b % 8 (round-robin dispatch); each XCD owns one
//         contiguous eighth of the rows, strided over its blocks
//  xcd 2: as 1, and each block owns one contiguous run of rows inside its
//         XCD's eighth, its 4 waves stepping through it together, so the
//         neighbour gathers of consecutive rows are reused from L2
__device__ __forceinline__ void row_range(int F, int G, int xcd, int &start, int &step, int &end) {
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: rows index scalar loads
    if (xcd == 0) {
        start = blockIdx.x * kWaves + wid;
        step = G * kWaves;
        end = F;
    } else if (xcd == 2) {
        const int x = blockIdx.x & 7, lb = blockIdx.x >> 3, nb = G >> 3;
        const int per = (F + 7) >> 3;
        const int rb = (per + nb - 1) / nb;
        const int lo = x * per + lb * rb;
        start = lo + wid;
        step = kWaves;
        end = min(min(lo + rb, (x + 1) * per), F);
    } else {
        // block b runs on XCD b % 8 (round-robin dispatch); give each XCD one
        // contiguous eighth of the rows so its gathers stay in its own L2
        const int x = blockIdx.x & 7, lb = blockIdx.x >> 3, nb = G >> 3;
        const int per = (F + 7) >> 3;
        const int x0 = x * per;
        start = x0 + lb * kWaves + wid;
        step = nb * kWaves;
        end = min(F, x0 + per);
    }
}

// per-block partial sums of NV values per realization lane, written to
// dst[k][blockIdx.x * Rp + blockIdx.y * 64 + lane] in a fixed order
template <int NV>
__device__ __forceinline__ void block_part(const double (&v)[NV], double *const (&dst)[NV], int Rp) {
    __shared__ double sh[NV][kWaves][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < NV; k++) sh[k][wid][lane] = v[k];
    __syncthreads();
    if (wid == 0) {
        const size_t o = (size_t)blockIdx.x * Rp + blockIdx.y * 64 + lane;
        for (int k = 0; k < NV; k++) {
            double t = sh[k][0][lane];
            for (int w = 1; w < kWaves; w++) t += sh[k][w][lane];
            dst[k][o] = t;
        }
    }
}

// Sources of the vector the operator is applied to, value at edge row j
// (idx = j * Rp + slot):
//   XPlain: x itself;
//   XFly:   the new search direction z + beta p, formed on the fly;
//   XTwoLevel: the same with z = D^-1 r + Z y formed on the fly from the
//           residual, the Jacobi diagonal and the coarse correction y of the
//           two-level preconditioner (row j's aggregate from a wave-uniform
//           table; Z = [1, phi] per aggregate).
template <class V>
struct XPlain {
    const V *__restrict__ x;
    __device__ __forceinline__ V operator()(int, size_t idx) const { return x[idx]; }
};
template <class V>
struct XFly {
    const V *__restrict__ z;
    const V *__restrict__ p;
    V beta;
    __device__ __forceinline__ V operator()(int, size_t idx) const { return z[idx] + beta * p[idx]; }
};
template <class V>
struct XTwoLevel {
    const V *__restrict__ r;
    const typename Pre<V>::T *__restrict__ dinv;
    const V *__restrict__ p;
    const unsigned short *__restrict__ phi;  // [F][Rp] bfloat16, or null (constant column only)
    const double *__restrict__ y;   // [nc][Rp]
    const int *__restrict__ agg;    // row -> aggregate
    V beta;
    size_t Rp;
    int slot;
    __device__ __forceinline__ V z(int j, size_t idx) const {
        const int a = agg[j];
        if (phi) {
            const double c =
                y[(size_t)(2 * a) * Rp + slot] + (double)bf16_to_f(phi[idx]) * y[(size_t)(2 * a + 1) * Rp + slot];
            return Pre<V>::ld(dinv, idx) * r[idx] + (V)c;
        }
        return Pre<V>::ld(dinv, idx) * r[idx] + (V)y[(size_t)a * Rp + slot];
    }
    __device__ __forceinline__ V operator()(int j, size_t idx) const { return z(j, idx) + beta * p[idx]; }
};

// Row i of S_e(K) x for an element without vb neighbours.
template <class V, int N0, class X>
__device__ __forceinline__ V elem_row_plain(const V *__restrict__ T, V K, const int *__restrict__ cn, int i,
                                            const X &xs, size_t Rp, int r) {
    V t = V(0), xi = V(0);
#pragma unroll
    for (int j = 0; j < N0; j++) {
        const V xj = xs(cn[j], (size_t)cn[j] * Rp + r);
        t += T[i * N0 + j] * xj;
        if (j == i) xi = xj;
    }
    return K * t + T[N0 * N0 + i] * xi;
}

// Row i of S_e(K) x for an element with vb neighbours, in FP64.
template <int N0, class X>
__device__ __forceinline__ double elem_row_vb(const double *__restrict__ T, int nvb, double K,
                                              const int *__restrict__ cn, int i, const X &xs, size_t Rp, int r) {
    double xv[N0];
#pragma unroll
    for (int j = 0; j < N0; j++) xv[j] = (double)xs(cn[j], (size_t)cn[j] * Rp + r);
    const double *C0 = T, *w0 = T + N0 * N0, *dlt = w0 + N0, *sv = dlt + N0, *vv = sv + 2;
    double ux = 0.0;
#pragma unroll
    for (int j = 0; j < N0; j++) ux += w0[j] * xv[j];
    ux *= K;
    double xi = 0.0;
    for (int k = 0; k < nvb; k++) {
        const double xk = (double)xs(cn[N0 + k], (size_t)cn[N0 + k] * Rp + r);
        ux += vv[k] * xk;
        if (N0 + k == i) xi = xk;
    }
    const double g = ux / (K * sv[0] + sv[1]);
    if (i < N0) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < N0; j++) {
            t += C0[i * N0 + j] * xv[j];
            if (j == i) xi = xv[j];
        }
        return K * (t - w0[i] * g) + dlt[i] * xi;
    }
    const double vk = vv[i - N0];
    return vk * (xi - g);
}

// Packed row records: 16 ints per row, two element contributions of 8 ints
// {e, i, n0, op_off, conn[4]}; e = -1 marks an absent second contribution,
// e = -2 in the first marks a row that needs the general path (a contribution
// from an element with vb neighbours, or more than two contributions).  The
// record is read with wave-uniform scalar loads, so one row costs one scalar
// load followed by independent gathers.
constexpr int kRecInts = 16;

template <class V, int N0, class X>
__device__ __forceinline__ V rec_contrib(const int *__restrict__ R, const V *__restrict__ opT,
                                         const V *__restrict__ Kv, const X &xs, size_t Rp, int r) {
    const int e = R[0], i = R[1];
    const V *T = opT + R[3];
    const V K = Kv[(size_t)e * Rp + r];
    V t = V(0), xi = V(0);
#pragma unroll
    for (int j = 0; j < N0; j++) {
        const V xj = xs(R[4 + j], (size_t)R[4 + j] * Rp + r);
        t += T[i * N0 + j] * xj;
        if (j == i) xi = xj;
    }
    return K * t + T[N0 * N0 + i] * xi;
}

template <class V, class X>
__device__ __forceinline__ V rec_any(const int *__restrict__ R, const V *__restrict__ opT, const V *__restrict__ Kv,
                                     const X &xs, size_t Rp, int r) {
    const int n0 = R[2];
    if (n0 == 4) return rec_contrib<V, 4>(R, opT, Kv, xs, Rp, r);
    if (n0 == 3) return rec_contrib<V, 3>(R, opT, Kv, xs, Rp, r);
    return rec_contrib<V, 2>(R, opT, Kv, xs, Rp, r);
}

template <class V, class X>
__device__ __forceinline__ V op_row(const int *__restrict__ rec, const int *__restrict__ row_ptr,
                                    const int *__restrict__ row_c, const int *__restrict__ meta,
                                    const int *__restrict__ op_off, const int *__restrict__ conn_off,
                                    const int *__restrict__ conn, const V *__restrict__ opT,
                                    const V *__restrict__ Kv, const double *__restrict__ opT64,
                                    const double *__restrict__ K64, int row, const X &xs, size_t Rp, int r) {
    const int *R = rec + (size_t)row * kRecInts;
    if (R[0] != -2) {
        V y = rec_any<V>(R, opT, Kv, xs, Rp, r);
        if (R[8] >= 0) y += rec_any<V>(R + 8, opT, Kv, xs, Rp, r);
        return y;
    }
    V y = V(0);
    for (int c = row_ptr[row]; c < row_ptr[row + 1]; c++) {
        const int code = row_c[c];
        const int e = code >> 4, i = code & 15;
        const int mt = meta[e];
        const int n0 = mt & 15, nvb = (mt >> 4) & 255;
        const int *cn = conn + conn_off[e];
        if (nvb == 0) {
            const V *T = opT + op_off[e];
            const V K = Kv[(size_t)e * Rp + r];
            if (n0 == 4) y += elem_row_plain<V, 4>(T, K, cn, i, xs, Rp, r);
            else if (n0 == 3) y += elem_row_plain<V, 3>(T, K, cn, i, xs, Rp, r);
            else y += elem_row_plain<V, 2>(T, K, cn, i, xs, Rp, r);
        } else {
            const double *T = opT64 + op_off[e];
            const double K = K64[(size_t)e * Rp + r];
            double t;
            if (n0 == 4) t = elem_row_vb<4>(T, nvb, K, cn, i, xs, Rp, r);
            else if (n0 == 3) t = elem_row_vb<3>(T, nvb, K, cn, i, xs, Rp, r);
            else t = elem_row_vb<2>(T, nvb, K, cn, i, xs, Rp, r);
            y += (V)t;
        }
    }
    return y;
}

// The common row of a tetrahedral mesh: two tetrahedra without vb
// neighbours (record contributions with n0 = 4).  Same arithmetic and order
// as op_row; both contributions' K and x gathers are issued before any of
// them is used.
template <class V, class X>
__device__ __forceinline__ V op_row_pair(const int *__restrict__ R, const V *__restrict__ opT,
                                         const V *__restrict__ Kv, const X &xs, size_t Rp, int r) {
    const int i1 = R[1], i2 = R[9];
    const V *T1 = opT + R[3], *T2 = opT + R[11];
    const V K1 = Kv[(size_t)R[0] * Rp + r], K2 = Kv[(size_t)R[8] * Rp + r];
    V a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        a[j] = xs(R[4 + j], (size_t)R[4 + j] * Rp + r);
        b[j] = xs(R[12 + j], (size_t)R[12 + j] * Rp + r);
    }
    V t1 = V(0), t2 = V(0), x1 = V(0), x2 = V(0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        t1 += T1[i1 * 4 + j] * a[j];
        t2 += T2[i2 * 4 + j] * b[j];
        if (j == i1) x1 = a[j];
        if (j == i2) x2 = b[j];
    }
    V y = K1 * t1 + T1[16 + i1] * x1;
    y += K2 * t2 + T2[16 + i2] * x2;
    return y;
}

// Row records of the FP32 operator for tetrahedron-pair rows, with the row of
// each element block inlined so the operator needs no table load:
//   [0] e1 (-1: not a pair row, use the general path), [1] e2,
//   [2..5] rows of e1's edges, [6..9] rows of e2's edges,
//   [10..13] P0_e1 row i1, [14..17] P0_e2 row i2 (FP32 bits), [18] dlt_e1[i1], [19] dlt_e2[i2]
// (the diagonal term multiplies the row's own value).
constexpr int kFastRec = 20;
// rows of one or two tetrahedra without vb neighbours take the inlined record
__host__ __device__ inline bool op_fast_row(const int *R) {
    return R[0] >= 0 && R[2] == 4 && (R[8] < 0 || R[10] == 4);
}
__global__ void k_b_fastrec(int F, const int *__restrict__ rec, const float *__restrict__ opT,
                            int *__restrict__ out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= F) return;
    const int *R = rec + (size_t)row * kRecInts;
    int *O = out + (size_t)row * kFastRec;
    if (!op_fast_row(R)) {
        O[0] = -1;
        return;
    }
    for (int k = 0; k < 2; k++) {
        // a row with one tetrahedron repeats it with a zero block row and a
        // zero diagonal term: K * (0 x) + 0 x_row adds +0 to the sum
        const bool one = k == 1 && R[8] < 0;
        const int *Q = R + (one ? 0 : 8 * k);
        const float *T = opT + Q[3];
        const int i = Q[1];
        O[k] = Q[0];
        for (int j = 0; j < 4; j++) {
            O[2 + 4 * k + j] = Q[4 + j];
            O[10 + 4 * k + j] = one ? __float_as_int(0.0f) : __float_as_int(T[i * 4 + j]);
        }
        O[18 + k] = one ? __float_as_int(0.0f) : __float_as_int(T[16 + i]);
    }
}

// ---------------------------------------------------------------------------
// Table of K-independent element factors (element_local at K = 1)
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) k_b_table(const DarcyDev dv, int *meta, const int *op_off, const int *rs_off,
                                                 double *opT, double *rsT) {
    int idx = dv.dim_start[D - 1] + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= dv.dim_start[D]) return;
    int e = dv.perm[idx];
    Local<D> L;
    element_local<D>(dv, e, 1.0, L);
    const int n0 = D + 1;
    int vb0 = dv.el_vb_ptr[e], nvb = dv.el_vb_ptr[e + 1] - vb0;
    double V = 0.0;
    for (int k = 0; k < nvb; k++) V += vb_coef(dv, e, dv.el_vb[vb0 + k]);
    int cmask = 0;
    for (int i = 0; i < n0; i++)
        if (L.c[i] != 0.0) cmask |= 1 << i;
    meta[e] = n0 | (nvb << 4) | (cmask << 12);
    double *O = opT + op_off[e];
    double *Q = rsT + rs_off[e];
    double sAg = 0.0;
    for (int i = 0; i < n0; i++) {
        for (int j = 0; j < n0; j++) {
            const double cij = L.c[i] * L.c[j];
            if (nvb == 0) O[i * n0 + j] = cij * (L.inv[i * n0 + j] - L.a[i] * L.a[j] / L.s);
            else O[i * n0 + j] = cij * L.inv[i * n0 + j];
            Q[i * n0 + j] = L.inv[i * n0 + j];
        }
        if (nvb == 0) {
            O[n0 * n0 + i] = L.dlt[i];
        } else {
            O[n0 * n0 + i] = L.c[i] * L.a[i];
            O[n0 * n0 + n0 + i] = L.dlt[i];
        }
        Q[n0 * n0 + i] = L.a[i];
        Q[n0 * n0 + n0 + i] = L.Ag[i];
        Q[n0 * n0 + 2 * n0 + i] = L.rf[i];
        Q[n0 * n0 + 3 * n0 + i] = L.dlt[i];
        sAg += L.Ag[i];
    }
    if (nvb) {
        O[n0 * n0 + 2 * n0] = L.s;
        O[n0 * n0 + 2 * n0 + 1] = V;
    }
    double *qs = Q + n0 * n0 + 4 * n0;
    qs[0] = L.s;
    qs[1] = sAg;
    qs[2] = L.meas * dv.cs[e] * dv.f[e];
    qs[3] = V;
    for (int k = 0; k < nvb; k++) {
        double vk = vb_coef(dv, e, dv.el_vb[vb0 + k]);
        O[n0 * n0 + 2 * n0 + 2 + k] = vk;
        qs[4 + k] = vk;
    }
}

__global__ void k_b_to_float(size_t n, const double *__restrict__ src, float *__restrict__ dst) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// Load realizations into slots: K64 / K32 [E][Sp] column slot <- K[real][E]
// for the listed (slot, real) pairs, through an LDS transpose (64 elements x
// 64 pairs per block).
__global__ void k_b_slot_K(const double *__restrict__ K, int E, int Sp, const int *__restrict__ pairs, int npairs,
                           double *__restrict__ K64, float *__restrict__ K32) {
    __shared__ double t[64][65];
    const int e0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = wid; j < 64; j += 4) {
        const int pi = p0 + j, e = e0 + lane;
        t[j][lane] = (pi < npairs && e < E) ? K[(size_t)pairs[2 * pi + 1] * E + e] : 0.0;
    }
    __syncthreads();
    const int pi = p0 + lane;
    if (pi >= npairs) return;
    const int slot = pairs[2 * pi];
    for (int j = wid; j < 64; j += 4) {
        const int e = e0 + j;
        if (e < E) {
            const double v = t[lane][j];
            K64[(size_t)e * Sp + slot] = v;
            K32[(size_t)e * Sp + slot] = (float)v;
        }
    }
}

// Per-realization FP32 rows [real][n] into slot columns dst[n][slot] for the
// listed (slot, real) pairs (LDS transpose, 64 entries x 64 pairs per block)
__global__ void k_b_slot_col(const float *__restrict__ src, int n, int Sp, const int *__restrict__ pairs, int npairs,
                             unsigned short *__restrict__ dst) {
    __shared__ float t[64][65];
    const int e0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = wid; j < 64; j += 4) {
        const int pi = p0 + j, e = e0 + lane;
        t[j][lane] = (pi < npairs && e < n) ? src[(size_t)pairs[2 * pi + 1] * n + e] : 0.0f;
    }
    __syncthreads();
    const int pi = p0 + lane;
    if (pi >= npairs) return;
    const int slot = pairs[2 * pi];
    for (int j = wid; j < 64; j += 4) {
        const int e = e0 + j;
        if (e < n) dst[(size_t)e * Sp + slot] = f_to_bf16(t[lane][j]);
    }
}

// Difficulty proxy per realization: variance of ln K over the elements (the
// CG iteration count grows with the log-conductivity spread).
__global__ void k_b_logk_var(const double *__restrict__ K, int E, double *__restrict__ out) {
    __shared__ double sh[2][4];
    const double *k = K + (size_t)blockIdx.x * E;
    double s1 = 0.0, s2 = 0.0;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        const double l = log(k[e]);
        s1 += l;
        s2 += l * l;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wid] = s1;
        sh[1][wid] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < 4; w++) {
            a += sh[0][w];
            b += sh[1][w];
        }
        a /= E;
        out[blockIdx.x] = b / E - a * a;
    }
}

__global__ void k_b_status_fill(int *status, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) status[i] = v;
}

// right-hand side and Jacobi diagonal of the selected slots (FP64)
__global__ void __launch_bounds__(256) k_b_rhs(int F, size_t Rp, const int *__restrict__ row_ptr,
                                               const int *__restrict__ row_c, const int *__restrict__ meta,
                                               const int *__restrict__ rs_off, const double *__restrict__ rsT,
                                               const double *__restrict__ Kv, double *__restrict__ b,
                                               double *__restrict__ dinv, unsigned short *__restrict__ dinv32,
                                               const int *__restrict__ lsel) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    if (lsel && !lsel[r]) return;  // no barrier in this kernel
    for (int row = blockIdx.x * kWaves + (threadIdx.x >> 6); row < F; row += gridDim.x * kWaves) {
        double sb = 0.0, sd = 0.0;
        for (int c = row_ptr[row]; c < row_ptr[row + 1]; c++) {
            const int code = row_c[c];
            const int e = code >> 4, i = code & 15;
            const int mt = meta[e];
            const int n0 = mt & 15, cm = mt >> 12;
            const double *Q = rsT + rs_off[e];
            const double *qs = Q + n0 * n0 + 4 * n0;
            const double K = Kv[(size_t)e * Rp + r];
            const double D = K * qs[0] + qs[3];
            const double hD = (qs[2] - K * qs[1]) / D;
            if (i < n0) {
                const bool ci = (cm >> i) & 1;
                const double a = K * Q[n0 * n0 + i];
                double vb = Q[n0 * n0 + 2 * n0 + i];
                double vd = Q[n0 * n0 + 3 * n0 + i];
                if (ci) {
                    vb = K * Q[n0 * n0 + n0 + i] + vb + a * hD;
                    vd = K * Q[i * n0 + i] - a * a / D + vd;
                }
                sb += vb;
                sd += vd;
            } else {
                const double vk = qs[4 + i - n0];
                sb += vk * hD;
                sd += vk - vk * vk / D;
            }
        }
        const size_t o = (size_t)row * Rp + r;
        b[o] = sb;
        const double di = sd != 0.0 ? 1.0 / sd : 1.0;
        dinv[o] = di;
        dinv32[o] = f_to_bf16((float)di);
    }
}

// cold start of the selected slots: x = 0, r = b, z = D^-1 b, p = 0 (both
// buffers); partials r.z, r.r
template <class V>
__global__ void __launch_bounds__(256) k_b_init(BArgs<V> a, const double *__restrict__ b64, double *x64,
                                                const int *__restrict__ lsel) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    int s0, st, en;
    row_range(a.F, a.G, a.xcd, s0, st, en);
    double v[2] = {0.0, 0.0};
    if (lsel && !lsel[r]) en = s0;  // unselected slot: partials 0, still joins the block reduction
    for (int row = s0; row < en; row += st) {
        const size_t o = (size_t)row * a.Rp + r;
        const double b = b64[o];
        const V bv = (V)b;
        const V z = Pre<V>::ld(a.dinv, o) * bv;
        a.x[o] = V(0);
        if (x64) x64[o] = 0.0;
        a.r[o] = bv;
        a.z[o] = z;
        a.p0[o] = V(0);
        a.p1[o] = V(0);
        v[0] += (double)bv * (double)z;
        v[1] += b * b;
    }
    double *const dst[2] = {a.part1, a.part2};
    block_part<2>(v, dst, a.Rp);
}

// q = S p with p = z + beta p_old formed on the fly; partial p.q
template <class V>
__global__ void __launch_bounds__(256) k_b_op(BArgs<V> a) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    const bool act = a.sc.status[r] == kRun;
    double v[1] = {0.0};
    if (__any(act)) {
        const V beta = act ? (V)a.sc.beta[r] : V(0);
        const int par = a.sc.par[r];
        const V *pin = par ? a.p1 : a.p0;
        V *pout = par ? a.p0 : a.p1;
        int s0, st, en;
        row_range(a.F, a.G, a.xcd, s0, st, en);
        for (int row = s0; row < en; row += st) {
            if (act) {
                const XFly<V> xs{a.z, pin, beta};
                const V q = op_row<V>(a.rec, a.row_ptr, a.row_c, a.meta, a.op_off, a.conn_off, a.conn, a.opT, a.K,
                                      a.opT64, a.K64, row, xs, a.Rp, r);
                const size_t o = (size_t)row * a.Rp + r;
                const V p = a.z[o] + beta * pin[o];
                pout[o] = p;
                a.q[o] = q;
                v[0] += (double)p * (double)q;
            }
        }
    }
    double *const dst[1] = {a.part0};
    block_part<1>(v, dst, a.Rp);
}

// x += alpha p; r -= alpha q; z = D^-1 r; partials r.z, r.r
template <class V>
__global__ void __launch_bounds__(256) k_b_update(BArgs<V> a) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    const bool act = a.sc.status[r] == kRun;
    double v[2] = {0.0, 0.0};
    if (__any(act)) {
        const V alpha = act ? (V)a.sc.alpha[r] : V(0);
        const V *p = a.sc.par[r] ? a.p0 : a.p1;  // the buffer the operator just wrote
        int s0, st, en;
        row_range(a.F, a.G, a.xcd, s0, st, en);
        for (int row = s0; row < en; row += st) {
            if (act) {
                const size_t o = (size_t)row * a.Rp + r;
                a.x[o] += alpha * p[o];
                const V rv = a.r[o] - alpha * a.q[o];
                a.r[o] = rv;
                const V z = Pre<V>::ld(a.dinv, o) * rv;
                a.z[o] = z;
                v[0] += (double)rv * (double)z;
                v[1] += (double)rv * (double)rv;
            }
        }
    }
    double *const dst[2] = {a.part1, a.part2};
    block_part<2>(v, dst, a.Rp);
}

// Reliable-update selection: a realization whose FP32 recursion claims
// convergence, or whose recursive residual norm dropped 10x since its last
// true residual.  Updating only then keeps the recurrence undisturbed while
// the FP32 residual is still accurate.
// Default reliable-update trigger (squared residual drop since the last one):
// 1 replaces the residual of every running realization at every chunk end.
constexpr double kRuDrop2 = 1.0;
__device__ __forceinline__ bool ru_pick(const BatchScal &sc, int r, int st) {
    return st == kCandidate || (st == kRun && sc.rr[r] < sc.ru_drop2 * sc.rr_ru[r]);
}

// Lanes per edge row of the two-level iteration kernels.  W = 64: one row per
// wave, lane = slot of the block's 64-slot chunk.  Once the queue is empty
// and at most W realizations are left, all in slots [0, W), the kernels run
// with W = 16 or 4: 64 / W rows per wave, lane = row offset * W + slot, so the
// tail of a batch does not sweep every row with idle lanes.
template <int W>
struct LaneMap {
    static constexpr int P = 64 / W;  // rows per wave
    __device__ static int slot(int lane) { return W == 64 ? (int)blockIdx.y * 64 + lane : (lane & (W - 1)); }
    __device__ static int sub(int lane) { return W == 64 ? 0 : lane / W; }
    // sum over the lanes of one slot (every such lane gets the same sum)
    __device__ static double fold(double x) {
        for (int m = W; m < 64; m <<= 1) x += __shfl_xor(x, m, 64);
        return x;
    }
};

// mixed precision: x64 += x32, x32 = 0 for the selected realizations
// (lsel: the listed slots instead, when they leave the FP32 phase)
template <int W>
__global__ void __launch_bounds__(256) k_b_accum(int F, int G, int xcd, size_t Rp, BatchScal sc,
                                                 const int *__restrict__ lsel, double *__restrict__ x64,
                                                 float *__restrict__ x32) {
    using L = LaneMap<W>;
    const int lane = threadIdx.x & 63;
    const int r = L::slot(lane), sub = L::sub(lane);
    const bool act = lsel ? lsel[r] != 0 : ru_pick(sc, r, sc.status[r]);
    if (!__any(act)) return;
    int s0, st, en;
    row_range((F + L::P - 1) / L::P, G, xcd, s0, st, en);
    for (int g = s0; g < en; g += st) {
        const int row = g * L::P + sub;
        if (act && row < F) {
            const size_t o = (size_t)row * Rp + r;
            x64[o] += (double)x32[o];
            x32[o] = 0.0f;
        }
    }
}

// True FP64 residual r64 = b - S64 x64 for the realizations whose status bit
// is set in sel (sel = -1: the reliable-update selection; sel = -2: the slots
// flagged in lsel).  MODE 0: store r and
// z = D^-1 r in V; MODE 1: also zero both search-direction buffers (FP64 warm
// start); MODE 2: partial r.r only.
template <class V, int MODE>
__global__ void __launch_bounds__(256) k_b_resid(BArgs<V> a, int sel, const int *__restrict__ lsel,
                                                 const int *__restrict__ op_off64,
                                                 const double *__restrict__ opT64, const double *__restrict__ K64,
                                                 const double *__restrict__ b64, const double *__restrict__ x64) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    const int st0 = a.sc.status[r];
    const bool act = sel == -2 ? lsel[r] != 0 : (sel == -1 ? ru_pick(a.sc, r, st0) : ((sel >> st0) & 1) != 0);
    double v[2] = {0.0, 0.0};
    if (__any(act)) {
        int s0, st, en;
        row_range(a.F, a.G, a.xcd, s0, st, en);
        for (int row = s0; row < en; row += st) {
            if (act) {
                const XPlain<double> xs{x64};
                const double ax = op_row<double>(a.rec, a.row_ptr, a.row_c, a.meta, op_off64, a.conn_off, a.conn,
                                                 opT64, K64, opT64, K64, row, xs, a.Rp, r);
                const size_t o = (size_t)row * a.Rp + r;
                const double rr = b64[o] - ax;
                v[1] += rr * rr;
                if (MODE == 2) v[0] += (double)Pre<V>::ld(a.dinv, o) * rr * rr;
                if (MODE != 2) {
                    const V rv = (V)rr;
                    const V z = Pre<V>::ld(a.dinv, o) * rv;
                    a.r[o] = rv;
                    a.z[o] = z;
                    v[0] += (double)rv * (double)z;
                    if (MODE == 1) {
                        a.p0[o] = V(0);
                        a.p1[o] = V(0);
                    }
                }
            }
        }
    }
    double *const dst[2] = {a.part1, a.part2};
    block_part<2>(v, dst, a.Rp);
}

enum FoldMode { kFoldInit = 0, kFoldAlpha = 1, kFoldBeta = 2, kFoldRU = 3, kFoldWarm = 4, kFoldFinal = 5 };

// Per-realization fold of the block partials in a fixed order, then the
// scalar step of CG for that realization.  One lane per realization.
template <int MODE>
__global__ void __launch_bounds__(kFoldThreads) k_b_fold(BatchScal sc, const double *__restrict__ part0,
                                                         const double *__restrict__ part1,
                                                         const double *__restrict__ part2, int G, int Rp, int R,
                                                         double rtol2, double atol2, int max_it, int mixed,
                                                         const int *__restrict__ lsel) {
    __shared__ double sh[3][kFoldThreads / 64][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    double s[3] = {0.0, 0.0, 0.0};
    // kFoldChunk rows of the wave's share loaded together (rows past the end
    // re-read row wid and are dropped), then summed in row order
    constexpr int kW = kFoldThreads / 64, kFoldChunk = 8;
    for (int g0 = wid; g0 < G; g0 += kFoldChunk * kW) {
        double v1[kFoldChunk], v2[kFoldChunk];
#pragma unroll
        for (int u = 0; u < kFoldChunk; u++) {
            const int g = g0 + u * kW;
            const size_t o = (size_t)(g < G ? g : wid) * Rp + r;
            if (MODE == kFoldAlpha) {
                v1[u] = part0[o];
                v2[u] = 0.0;
            } else {
                v1[u] = part1[o];
                v2[u] = part2[o];
            }
        }
#pragma unroll
        for (int u = 0; u < kFoldChunk; u++) {
            if (g0 + u * kW >= G) break;
            if (MODE == kFoldAlpha) s[0] += v1[u];
            else {
                s[1] += v1[u];
                s[2] += v2[u];
            }
        }
    }
    for (int k = 0; k < 3; k++) sh[k][wid][lane] = s[k];
    __syncthreads();
    if (wid != 0) return;
    for (int k = 0; k < 3; k++) {
        double t = sh[k][0][lane];
        for (int w = 1; w < kFoldThreads / 64; w++) t += sh[k][w][lane];
        s[k] = t;
    }
    int st = sc.status[r];
    const bool picked = !lsel || lsel[r] != 0;  // Init / Final touch only the selected slots
    // s[1] = r.z = r^T D^-1 r, s[2] = r.r; nr: the square of the stop-test norm
    const double nr = sc.nat ? s[1] : s[2];
    if (MODE == kFoldInit) {
        if (picked) {
            sc.bnorm2[r] = nr;
            sc.tol2[r] = fmax(rtol2 * nr, atol2);
            sc.rz[r] = s[1];
            sc.rz_prev[r] = 1.0;
            sc.rr[r] = nr;
            sc.beta[r] = 0.0;
            sc.alpha[r] = 0.0;
            sc.it[r] = 0;
            sc.n_ru[r] = 0;
            sc.it_ru[r] = 0;
            sc.par[r] = 0;
            sc.rr_hist[r] = INFINITY;
            sc.rr_ru[r] = nr;
            st = nr <= sc.tol2[r] ? kConverged : kRun;
            sc.status[r] = st;
        }
    } else if (MODE == kFoldAlpha) {
        if (st == kRun) {
            const double pq = s[0];
            if (pq > 0.0 && pq == pq) sc.alpha[r] = sc.rz[r] / pq;
            else {
                sc.alpha[r] = 0.0;
                sc.status[r] = st = kBreakdown;
            }
        }
    } else if (MODE == kFoldBeta) {
        if (st == kRun) {
            const int it = sc.it[r] + 1;
            sc.it[r] = it;
            sc.par[r] ^= 1;
            const double rz = s[1], rr = nr;
            sc.rz_prev[r] = sc.rz[r];
            sc.rz[r] = rz;
            sc.beta[r] = rz / sc.rz_prev[r];
            sc.rr[r] = rr;
            if (!(rr == rr)) st = kBreakdown;
            else if (rr <= sc.tol2[r]) st = mixed ? kCandidate : kConverged;
            else if (it >= max_it) st = kMaxIt;
            else if (mixed && it - sc.it_ru[r] > 1000 + sc.it_ru[r]) st = kFp64;  // FP32 recursion stalled
            sc.status[r] = st;
        }
    } else if (MODE == kFoldRU) {
        if (ru_pick(sc, r, st)) {
            const double rz = s[1], rr = nr;
            sc.n_ru[r] += 1;
            sc.rr_ru[r] = rr;
            sc.it_ru[r] = sc.it[r];
            sc.rz[r] = rz;
            sc.beta[r] = rz / sc.rz_prev[r];
            sc.rr[r] = rr;
            if (!(rr == rr)) st = kBreakdown;
            else if (st == kCandidate) {
                if (rr <= sc.tol2[r]) st = kConverged;
                else {
                    // the FP32 recursion claimed convergence but the true
                    // residual disagrees: a second refusal without a 10x
                    // drop of the true residual hands the realization to FP64
                    st = (rr > 0.01 * sc.rr_hist[r]) ? kFp64 : kRun;
                    sc.rr_hist[r] = rr;
                }
            } else if (sc.it[r] >= max_it) st = kMaxIt;
            sc.status[r] = st;
        }
    } else if (MODE == kFoldWarm) {
        if (st == kFp64) {
            sc.rz[r] = s[1];
            sc.rz_prev[r] = 1.0;
            sc.rr[r] = nr;
            sc.beta[r] = 0.0;
            sc.par[r] = 0;
            st = nr <= sc.tol2[r] ? kConverged : kRun;
            sc.status[r] = st;
        }
    } else {  // final true residual
        if (picked) sc.rr[r] = nr;
    }
    const unsigned long long run = __ballot(st == kRun), f64 = __ballot(st == kFp64);
    if (lane == 0) {
        sc.active_count[2 * blockIdx.x] = __popcll(run);
        sc.active_count[2 * blockIdx.x + 1] = __popcll(f64);
    }
}

// First stage of a fold over many block partials: block (s, chunk) sums rows
// [s G / kPre, (s + 1) G / kPre) of each given array in a fixed order, so the
// per-slot fold reads kPre rows instead of G (a G-row chain on one CU per
// chunk is latency-bound).
constexpr int kPre = 128;
__global__ void __launch_bounds__(256) k_b_prefold(const double *__restrict__ p0, const double *__restrict__ p1,
                                                   const double *__restrict__ p2, int G, int Rp,
                                                   double *__restrict__ o0, double *__restrict__ o1,
                                                   double *__restrict__ o2) {
    __shared__ double sh[3][kWaves][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = blockIdx.y * 64 + lane;
    const int g0 = blockIdx.x * G / kPre, g1 = (blockIdx.x + 1) * G / kPre;
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
    for (int g = g0 + wid; g < g1; g += kWaves) {
        const size_t o = (size_t)g * Rp + slot;
        if (p0) s[0] += p0[o];
        if (p1) s[1] += p1[o];
        if (p2) s[2] += p2[o];
    }
    for (int k = 0; k < 3; k++) sh[k][wid][lane] = s[k];
    __syncthreads();
    if (wid != 0) return;
    double *const out[3] = {o0, o1, o2};
    const double *const in[3] = {p0, p1, p2};
    for (int k = 0; k < 3; k++) {
        if (!in[k]) continue;
        double t = sh[k][0][lane];
        for (int w = 1; w < kWaves; w++) t += sh[k][w][lane];
        out[k][(size_t)blockIdx.x * Rp + slot] = t;
    }
}

// Reconstruction of element pressure and side fluxes (schur.cc:714-773) for
// the listed (slot, realization) pairs: edge traces lam and K read from the
// slot, output written to the realization's column, sol[idx * Rout + real].
template <int D>
__global__ void __launch_bounds__(256) k_b_resolve(int e_begin, int e_end, const int *__restrict__ perm,
                                                   const int *__restrict__ side_off, int S, size_t Rp,
                                                   const int *__restrict__ meta, const int *__restrict__ rs_off,
                                                   const double *__restrict__ rsT, const int *__restrict__ conn_off,
                                                   const int *__restrict__ conn, const double *__restrict__ Kv,
                                                   const double *__restrict__ lam, double *__restrict__ sol,
                                                   const int *__restrict__ pairs, int npairs, size_t Rout) {
    const int lane = threadIdx.x & 63;
    const int pi = blockIdx.y * 64 + lane;
    const int idx = e_begin + blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (idx >= e_end || pi >= npairs) return;
    const int r = pairs[2 * pi], real = pairs[2 * pi + 1];
    constexpr int n0 = D + 1;
    const int e = perm[idx];
    const int mt = meta[e];
    const int nvb = (mt >> 4) & 255, cm = mt >> 12;
    const double *Q = rsT + rs_off[e];
    const double *qs = Q + n0 * n0 + 4 * n0;
    const int *cn = conn + conn_off[e];
    const double K = Kv[(size_t)e * Rp + r];
    double cl[n0];
    double own = 0.0;
#pragma unroll
    for (int i = 0; i < n0; i++) {
        const double l = ((cm >> i) & 1) ? lam[(size_t)cn[i] * Rp + r] : 0.0;
        cl[i] = l;
        own += (K * Q[n0 * n0 + i]) * l;
    }
    double wl = 0.0;
    for (int k = 0; k < nvb; k++) wl += qs[4 + k] * lam[(size_t)cn[n0 + k] * Rp + r];
    const double Dd = K * qs[0] + qs[3];
    const double h = qs[2] - K * qs[1];
    const double pe = (h + (own + wl)) / Dd;
    const int so = side_off[e];
#pragma unroll
    for (int i = 0; i < n0; i++) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < n0; j++) t += Q[i * n0 + j] * cl[j];
        sol[(size_t)(so + i) * Rout + real] = K * Q[n0 * n0 + n0 + i] + pe * (K * Q[n0 * n0 + i]) - K * t;
    }
    sol[(size_t)(S + e) * Rout + real] = pe;
}

__global__ void k_b_scatter_lambda(int F, size_t Rp, const int *__restrict__ perm, const double *__restrict__ x,
                                   double *__restrict__ lam_out, const int *__restrict__ pairs, int npairs,
                                   size_t Rout) {
    const int lane = threadIdx.x & 63;
    const int pi = blockIdx.y * 64 + lane;
    if (pi >= npairs) return;
    const int r = pairs[2 * pi], real = pairs[2 * pi + 1];
    for (int row = blockIdx.x * kWaves + (threadIdx.x >> 6); row < F; row += gridDim.x * kWaves)
        lam_out[(size_t)perm[row] * Rout + real] = x[(size_t)row * Rp + r];
}

// edge traces of one slot, rows -> edge order
__global__ void k_b_extract_x(int F, size_t Sp, int slot, const int *__restrict__ perm, const double *__restrict__ x,
                              double *__restrict__ out) {
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < F; row += gridDim.x * blockDim.x)
        out[perm[row]] = x[(size_t)row * Sp + slot];
}

// a single-solve [q | p | lambda] vector into a realization's output column
__global__ void k_b_put_col(size_t n, size_t Rout, int real, const double *__restrict__ src,
                            double *__restrict__ sol) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        sol[i * Rout + real] = src[i];
}

__global__ void k_b_lambda_out(int R, size_t Rout, int F, size_t off, const double *__restrict__ sol,
                               double *__restrict__ out) {
    const int r = blockIdx.y;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x)
        out[(size_t)r * F + f] = sol[(off + f) * Rout + r];
}

__global__ void k_b_extract(size_t n, size_t Rp, int r, const double *__restrict__ sol, double *__restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = sol[i * Rp + r];
}

// ---------------------------------------------------------------------------
// Two-level preconditioner: M^-1 r = D^-1 r + Z (Z^T S Z)^-1 Z^T r with Z one
// (constant) or two (constant, phi) columns per aggregate.  Aggregates are
// contiguous row ranges (cells of the row ordering's Z-curve), so the
// restriction Z^T r is a block reduction of the update kernel, and z is never
// stored: the direction kernel forms z = D^-1 r + Z y on the fly from r and
// the coarse correction y.  M is one fixed SPD operator per realization (the
// FP32 coarse inverse is symmetric and every kernel reads the same stored
// values), so the conjugate-gradient step uses the Fletcher-Reeves beta
// z_new . r_new / z_old . r_old.
// ---------------------------------------------------------------------------
constexpr int kNbMax = 32;  // neighbour aggregates per aggregate (27 on a regular Z-curve level)

struct CArgs {
    int ncol, nagg, nc;
    int nt;                                      // 64-row tiles of the padded coarse order
    const int *agg, *aptr, *nbptr, *nb, *lbmap;  // lbmap[a * nagg + b]: local index of b among a's neighbours
    const unsigned short *phi;                   // [F][Rp] bfloat16 (ncol 2) or null
    const float *ainv;                           // [Rp][nt (nt + 1) / 2][64][64]: packed lower tiles of the inverse
    double *rc, *y;                              // [nt * 64][Rp]: Z^T r and the coarse solution
};

// entry (i, j) of the condensed element block S_e(K) (FP64 operator table)
__device__ __forceinline__ double elem_entry(const double *__restrict__ T, int n0, int nvb, double K, int i, int j) {
    if (nvb == 0) return K * T[i * n0 + j] + (i == j ? T[n0 * n0 + i] : 0.0);
    const double *C0 = T, *w0 = T + n0 * n0, *dlt = w0 + n0, *sv = dlt + n0, *vv = sv + 2;
    const double ui = i < n0 ? K * w0[i] : vv[i - n0];
    const double uj = j < n0 ? K * w0[j] : vv[j - n0];
    double v = 0.0;
    if (i < n0 && j < n0) v = K * C0[i * n0 + j] + (i == j ? dlt[i] : 0.0);
    else if (i == j) v = vv[i - n0];
    return v - ui * uj / (K * sv[0] + sv[1]);
}

// Coarse operator Z^T S Z of the selected slots, as per-aggregate blocks
// Asp[((a * kNbMax + lb) * ncol^2 + s * ncol + t) * Rp + slot] (lb: local
// index of the neighbour aggregate b).  Each block is the sum, over the
// precomputed terms (row of a, element contribution (e, i) of the row, local
// column j of e whose edge lies in b), of z_s(row) S_e(K)_ij z_t(edge j), in
// row / contribution / column order.  The term lists are cut into chunks of
// at most kAsmChunk terms, one wave per (chunk, 64 slots) accumulating its
// chunk in registers (a block of the diagonal holds thousands of terms, and
// one wave walking them all set the kernel's length); k_c_assemble_sum adds
// the chunks of each block in chunk order.
constexpr int kAsmChunk = 256;

template <int NCOL>
__global__ void __launch_bounds__(64) k_c_assemble(BArgs<double> a, CArgs c, const int *__restrict__ lsel,
                                                   const int *__restrict__ cbeg, const int *__restrict__ cblk,
                                                   const int *__restrict__ tptr, const int2 *__restrict__ terms,
                                                   double *__restrict__ Apart) {
    const int lane = threadIdx.x;
    const int slot = blockIdx.y * 64 + lane;
    const int ch = blockIdx.x;
    const bool act = lsel[slot] != 0;
    if (!__any(act)) return;
    const int k0 = cbeg[ch], k1 = min(k0 + kAsmChunk, tptr[cblk[ch] + 1]);
    double acc[NCOL * NCOL];
#pragma unroll
    for (int q = 0; q < NCOL * NCOL; q++) acc[q] = 0.0;
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const int2 u = terms[k];
        const int row = u.x, e = u.y >> 8, i = (u.y >> 4) & 15, j = u.y & 15;
        const int mt = a.meta[e];
        const int n0 = mt & 15, nvb = (mt >> 4) & 255;
        const double *T = a.opT64 + a.op_off[e];
        const double K = a.K64[(size_t)e * a.Rp + slot];
        const double sij = elem_entry(T, n0, nvb, K, i, j);
        if constexpr (NCOL == 1) {
            acc[0] += sij;
        } else {
            const int g = a.conn[a.conn_off[e] + j];
            const double zf1 = (double)bf16_to_f(c.phi[(size_t)row * a.Rp + slot]);
            const double zg1 = (double)bf16_to_f(c.phi[(size_t)g * a.Rp + slot]);
            acc[0] += sij;
            acc[1] += sij * zg1;
            acc[2] += zf1 * sij;
            acc[3] += zf1 * sij * zg1;
        }
    }
    if (!act) return;
#pragma unroll
    for (int q = 0; q < NCOL * NCOL; q++) Apart[((size_t)ch * (NCOL * NCOL) + q) * a.Rp + slot] = acc[q];
}

// Asp of block tg = the sum of its chunks cfirst[tg] .. cfirst[tg + 1] - 1 in order
template <int NCOL>
__global__ void __launch_bounds__(64) k_c_assemble_sum(CArgs c, size_t Rp, const int *__restrict__ lsel,
                                                       const int *__restrict__ cfirst, const int *__restrict__ tagg,
                                                       const double *__restrict__ Apart, double *__restrict__ Asp) {
    const int slot = blockIdx.y * 64 + threadIdx.x;
    if (lsel[slot] == 0) return;
    const int tg = blockIdx.x, ag = tagg[tg], lb = tg - c.nbptr[ag];
    double acc[NCOL * NCOL];
#pragma unroll
    for (int q = 0; q < NCOL * NCOL; q++) acc[q] = 0.0;
    for (int ch = cfirst[tg]; ch < cfirst[tg + 1]; ch++)
#pragma unroll
        for (int q = 0; q < NCOL * NCOL; q++) acc[q] += Apart[((size_t)ch * (NCOL * NCOL) + q) * Rp + slot];
#pragma unroll
    for (int q = 0; q < NCOL * NCOL; q++) Asp[((size_t)(ag * kNbMax + lb) * (NCOL * NCOL) + q) * Rp + slot] = acc[q];
}

// Dense coarse matrices of the listed slots, packed [np][ncp][ncp] (ncp = 64
// nt) for the batched Cholesky; an empty coarse column (phi constant on an
// aggregate) becomes a decoupled unit diagonal, which Z maps to zero, and so
// do the padding rows past nc.
template <int NCOL>
__global__ void k_c_pack(CArgs c, const double *__restrict__ Asp, size_t Rp, const int *__restrict__ pairs,
                         double *__restrict__ Apack) {
    const int i = blockIdx.x, k = blockIdx.y, ncp = c.nt * 64;
    const int slot = pairs[2 * k];
    const int ai = i / NCOL, s = i % NCOL;
    double *row = Apack + ((size_t)k * ncp + i) * ncp;
    for (int j = threadIdx.x; j < ncp; j += blockDim.x) {
        double v = 0.0;
        if (i < c.nc && j < c.nc) {
            const int bj = j / NCOL, t = j % NCOL;
            const int lb = c.lbmap[(size_t)ai * c.nagg + bj];
            if (lb >= 0) v = Asp[((size_t)(ai * kNbMax + lb) * (NCOL * NCOL) + s * NCOL + t) * Rp + slot];
        }
        if (i == j && !(v > 0.0)) v = 1.0;
        row[j] = v;
    }
}

// phi of the selected slots: per aggregate zero mean and unit mean square
// (keeps the two coarse columns independent and scaled alike)
__global__ void k_c_phi_norm(CArgs c, size_t Rp, const int *__restrict__ lsel, unsigned short *__restrict__ phi) {
    const int lane = threadIdx.x;
    const int slot = blockIdx.y * 64 + lane;
    const int ag = blockIdx.x;
    if (!lsel[slot]) return;
    const int r0 = c.aptr[ag], r1 = c.aptr[ag + 1];
    double s1 = 0.0, s2 = 0.0;
    for (int row = r0; row < r1; row++) {
        const double v = bf16_to_f(phi[(size_t)row * Rp + slot]);
        s1 += v;
        s2 += v * v;
    }
    const double n = (double)(r1 - r0);
    const double mean = s1 / n, var = s2 / n - mean * mean;
    const double sc = var > 1e-24 * (s2 / n + 1e-300) ? 1.0 / sqrt(var) : 0.0;
    for (int row = r0; row < r1; row++) {
        const size_t o = (size_t)row * Rp + slot;
        phi[o] = f_to_bf16((float)(((double)bf16_to_f(phi[o]) - mean) * sc));
    }
}

// Shared tail of the aggregate-blocked kernels: block reduction of two
// partials and the ncol restrictions, written per aggregate.  With W < 64
// the lanes of one slot are summed first; the other lanes write zeros into the
// partials of slots that are idle in that mode.
template <int NCOL, int W = 64>
__device__ __forceinline__ void agg_part(const double (&vin)[2], const double (&rcin)[2], double *p1, double *p2,
                                         CArgs &c, size_t Rp) {
    __shared__ double sh[2 + NCOL][kAggWaves][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double v[2] = {vin[0], vin[1]}, rc[2] = {rcin[0], rcin[1]};
    if (W < 64) {
        const bool keep = LaneMap<W>::sub(lane) == 0;
        // every lane takes part in the shuffles; only then the selection
        for (int k = 0; k < 2; k++) {
            const double f = LaneMap<W>::fold(v[k]);
            v[k] = keep ? f : 0.0;
        }
        for (int s = 0; s < NCOL; s++) {
            const double f = LaneMap<W>::fold(rc[s]);
            rc[s] = keep ? f : 0.0;
        }
    }
    const int slot = blockIdx.y * 64 + lane;
    for (int k = 0; k < 2; k++) sh[k][wid][lane] = v[k];
    for (int s = 0; s < NCOL; s++) sh[2 + s][wid][lane] = rc[s];
    __syncthreads();
    if (wid == 0) {
        double t[2 + NCOL];
        for (int k = 0; k < 2 + NCOL; k++) {
            double x = sh[k][0][lane];
            for (int w = 1; w < kAggWaves; w++) x += sh[k][w][lane];
            t[k] = x;
        }
        const size_t o = (size_t)blockIdx.x * Rp + slot;
        p1[o] = t[0];
        p2[o] = t[1];
        for (int s = 0; s < NCOL; s++) c.rc[(size_t)(blockIdx.x * NCOL + s) * Rp + slot] = t[2 + s];
    }
}

// r -= alpha q; partials D^-1 r.r, r.r and the restriction Z^T r, one block
// per (aggregate, 64 slots).  x += alpha p is done by the
// direction kernel k_c_zp, which reads p anyway.
template <class V, int NCOL, int W, int U>
__global__ void __launch_bounds__(kAggWaves * 64) k_c_update(BArgs<V> a, CArgs c) {
    using L = LaneMap<W>;
    const int lane = threadIdx.x & 63;
    const int slot = L::slot(lane), sub = L::sub(lane);
    const bool act = a.sc.status[slot] == kRun;
    double v[2] = {0.0, 0.0}, rc[2] = {0.0, 0.0};
    if (U == 2 && W == 64 && __any(act)) {
        // two rows per step, all loads first (row and row + kAggWaves)
        const V alpha = act ? (V)a.sc.alpha[slot] : V(0);
        const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int r1 = c.aptr[blockIdx.x + 1];
        for (int row = c.aptr[blockIdx.x] + wid; row < r1; row += 2 * kAggWaves) {
            if (!act) continue;
            const bool two = row + kAggWaves < r1;
            const size_t o0 = (size_t)row * a.Rp + slot, o1 = two ? o0 + (size_t)kAggWaves * a.Rp : o0;
            const V q0 = a.q[o0], q1 = a.q[o1];
            const V r0 = a.r[o0], r1v = a.r[o1];
            const V d0 = Pre<V>::ld(a.dinv, o0), d1 = Pre<V>::ld(a.dinv, o1);
            double f0 = 0.0, f1 = 0.0;
            if constexpr (NCOL == 2) {
                f0 = (double)bf16_to_f(c.phi[o0]);
                f1 = (double)bf16_to_f(c.phi[o1]);
            }
            const V rv0 = r0 - alpha * q0, rv1 = r1v - alpha * q1;
            a.r[o0] = rv0;
            const V dr0 = d0 * rv0, dr1 = d1 * rv1;
            v[0] += (double)dr0 * (double)rv0;
            v[1] += (double)rv0 * (double)rv0;
            rc[0] += (double)rv0;
            if constexpr (NCOL == 2) rc[1] += f0 * (double)rv0;
            if (two) {
                a.r[o1] = rv1;
                v[0] += (double)dr1 * (double)rv1;
                v[1] += (double)rv1 * (double)rv1;
                rc[0] += (double)rv1;
                if constexpr (NCOL == 2) rc[1] += f1 * (double)rv1;
            }
        }
    } else if (__any(act)) {
        const V alpha = act ? (V)a.sc.alpha[slot] : V(0);
        const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int r1 = c.aptr[blockIdx.x + 1];
        for (int g = c.aptr[blockIdx.x] + wid * L::P; g < r1; g += kAggWaves * L::P) {
            const int row = g + sub;
            if (!act || row >= r1) continue;
            const size_t o = (size_t)row * a.Rp + slot;
            const V q = a.q[o];
            const V rv = a.r[o] - alpha * q;
            a.r[o] = rv;
            const V dr = Pre<V>::ld(a.dinv, o) * rv;
            v[0] += (double)dr * (double)rv;
            v[1] += (double)rv * (double)rv;
            rc[0] += (double)rv;
            if constexpr (NCOL == 2) rc[1] += (double)bf16_to_f(c.phi[o]) * (double)rv;
        }
    }
    agg_part<NCOL, W>(v, rc, a.part1, a.part2, c, a.Rp);
}

// cold start of the selected slots: x = 0, r = b, p = 0; partials and Z^T b
template <class V, int NCOL>
__global__ void __launch_bounds__(kAggWaves * 64) k_c_init(BArgs<V> a, CArgs c, const double *__restrict__ b64, double *x64,
                                                const int *__restrict__ lsel) {
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.y * 64 + lane;
    const bool act = lsel[slot] != 0;
    double v[2] = {0.0, 0.0}, rc[2] = {0.0, 0.0};
    if (__any(act)) {
        const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        for (int row = c.aptr[blockIdx.x] + wid; row < c.aptr[blockIdx.x + 1]; row += kAggWaves) {
            if (!act) continue;
            const size_t o = (size_t)row * a.Rp + slot;
            const double b = b64[o];
            const V bv = (V)b;
            a.x[o] = V(0);
            if (x64) x64[o] = 0.0;
            a.r[o] = bv;
            a.p0[o] = V(0);
            a.p1[o] = V(0);
            v[0] += (double)(Pre<V>::ld(a.dinv, o) * bv) * (double)bv;
            v[1] += b * b;
            rc[0] += (double)bv;
            if (NCOL == 2) rc[1] += (double)bf16_to_f(c.phi[o]) * (double)bv;
        }
    }
    agg_part<NCOL>(v, rc, a.part1, a.part2, c, a.Rp);
}

// True FP64 residual of the reliable-update selection (mixed precision):
// r = b - S64 x64 stored in V, partials and Z^T r
template <class V, int NCOL, int W>
__global__ void __launch_bounds__(kAggWaves * 64) k_c_resid(BArgs<V> a, CArgs c, const double *__restrict__ b64,
                                                 const double *__restrict__ x64) {
    using L = LaneMap<W>;
    const int lane = threadIdx.x & 63;
    const int slot = L::slot(lane), sub = L::sub(lane);
    const bool act = ru_pick(a.sc, slot, a.sc.status[slot]);
    double v[2] = {0.0, 0.0}, rc[2] = {0.0, 0.0};
    if (__any(act)) {
        const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const XPlain<double> xs{x64};
        const int r1 = c.aptr[blockIdx.x + 1];
        for (int g = c.aptr[blockIdx.x] + wid * L::P; g < r1; g += kAggWaves * L::P) {
            const int row = g + sub;
            if (!act || row >= r1) continue;
            const double ax = op_row<double>(a.rec, a.row_ptr, a.row_c, a.meta, a.op_off, a.conn_off, a.conn,
                                             a.opT64, a.K64, a.opT64, a.K64, row, xs, a.Rp, slot);
            const size_t o = (size_t)row * a.Rp + slot;
            const double rr = b64[o] - ax;
            const V rv = (V)rr;
            a.r[o] = rv;
            v[0] += (double)(Pre<V>::ld(a.dinv, o) * rv) * (double)rv;
            v[1] += rr * rr;
            rc[0] += (double)rv;
            if (NCOL == 2) rc[1] += (double)bf16_to_f(c.phi[o]) * (double)rv;
        }
    }
    agg_part<NCOL, W>(v, rc, a.part1, a.part2, c, a.Rp);
}

// q = S p for the two-level path (the direction is formed by k_c_zp); partial p.q
// PAIR 0: general row path; 1: tetrahedron-pair rows with all gathers issued
// first.
template <class V, int W, int PAIR>
__global__ void __launch_bounds__(256) k_c_op(BArgs<V> a, CArgs c) {
    using L = LaneMap<W>;
    const int lane = threadIdx.x & 63;
    const int r = L::slot(lane), sub = L::sub(lane);
    const bool act = a.sc.status[r] == kRun;
    double v[1] = {0.0};
    if (__any(act)) {
        const XPlain<V> xs{a.p0};
        int s0, st, en;
        row_range((a.F + L::P - 1) / L::P, a.G, a.xcd, s0, st, en);
        for (int g = s0; g < en; g += st) {
            const int row = g * L::P + sub;
            if (act && row < a.F) {
                const int *R = a.rec + (size_t)row * kRecInts;
                V q;
                if (PAIR != 0 && R[0] >= 0 && R[2] == 4 && R[8] >= 0 && R[10] == 4)
                    q = op_row_pair<V>(R, a.opT, a.K, xs, a.Rp, r);
                else
                    q = op_row<V>(a.rec, a.row_ptr, a.row_c, a.meta, a.op_off, a.conn_off, a.conn, a.opT, a.K,
                                  a.opT64, a.K64, row, xs, a.Rp, r);
                const size_t o = (size_t)row * a.Rp + r;
                a.q[o] = q;
                v[0] += (double)a.p0[o] * (double)q;
            }
        }
    }
    if (W < 64) {
        const double f = L::fold(v[0]);
        v[0] = sub == 0 ? f : 0.0;
    }
    double *const dst[1] = {a.part0};
    block_part<1>(v, dst, a.Rp);
}

// Tiles of the FP32 two-level operator (k_c_opl).  Rows are cut into tiles of
// T consecutive rows (the row order is a Z-curve, so a tile is compact in
// space).  Per tile, the segments staged in LDS: its own rows (local index
// row - r0), then its halo rows (rows outside the tile that its rows' elements
// touch, halo[hptr[t] ..]), then the K of its rows' elements (els[eptr[t] ..]).
// lrec[row][10]: local indices of the row's two elements and of their 2 x 4
// edge rows (0xFFFF in [0]: a general row, done by the general blocks).
struct OpTiles {
    const int *gen;                  // general rows (elements with vb neighbours)
    const int *grec;                 // their flattened records [ngen][kGenRec]
    const int *hptr, *halo, *eptr, *els;
    const unsigned short *lrec;
    int ngen, gen_blocks, ntiles, T, nseg_max;
};
constexpr int kLrec = 10;
// flattened general-row records: [nC, kGenC x {e, i, n0, nvb, op_off, rows[kGenR]}]
constexpr int kGenC = 4, kGenR = 6, kGenStride = 5 + kGenR, kGenRec = 48;
static_assert(1 + kGenC * kGenStride <= kGenRec, "general record size");
constexpr int kSlotGroup = 32;  // slots per LDS pass (128 B per staged row)

// LDS bytes of k_c_opl: staged segments, tile records (indices, values) and
// the partial-sum table
constexpr int kStageMax = 7;  // 16-byte staging loads per thread (512) and slot group (nseg_max <= 448)
// the staging buffer holds all kStageMax loads of every thread, so the loads
// and stores need no bounds test (a test per load would serialize them)
constexpr int kRecWords = 16;  // one row's record in LDS: 5 words of packed indices, 10 values, 1 pad
__host__ __device__ inline size_t opl_lds_bytes(int nseg_max, int T) {
    return (size_t)kStageMax * 512 * 16 + (size_t)T * kRecWords * 4 + 8 * kSlotGroup * 8 + (size_t)nseg_max * 8;
}

// FP32 operator q = S p of the two-level path, in two kernels that run
// concurrently (k_c_opg on a second stream).  Partial p.q per (block, slot)
// over all slots of the live chunks: rows [0, gen_blocks) of part0 from
// k_c_opg, rows [gen_blocks, gen_blocks + ntiles) from k_c_opl.

// contribution (S_e(K) x)_i of an element without vb neighbours from gathered
// x (elem_row_plain's arithmetic and order)
__device__ __forceinline__ float gen_plain(const float *__restrict__ T, int n0, int i, float K,
                                           const float (&x)[kGenR]) {
    float t = 0.0f, xi = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j < n0) {
            t += T[i * n0 + j] * x[j];
            if (j == i) xi = x[j];
        }
    return K * t + T[n0 * n0 + i] * xi;
}

// the same for an element with vb neighbours, in FP64 (elem_row_vb's arithmetic)
__device__ __forceinline__ double gen_vb(const double *__restrict__ T, int n0, int nvb, int i, double K,
                                         const float (&x)[kGenR]) {
    const double *C0 = T, *w0 = T + n0 * n0, *dlt = w0 + n0, *sv = dlt + n0, *vv = sv + 2;
    double ux = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j < n0) ux += w0[j] * (double)x[j];
    ux *= K;
    // vb neighbours are x[n0 .. n0 + nvb): a static index over x keeps it in
    // registers, the table entry is vv[j - n0] (same order as over k)
    double xi = 0.0;
#pragma unroll
    for (int j = 0; j < kGenR; j++)
        if (j >= n0 && j < n0 + nvb) {
            const double xk = (double)x[j];
            ux += vv[j - n0] * xk;
            if (j == i) xi = xk;
        }
    const double g = ux / (K * sv[0] + sv[1]);
    if (i < n0) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j < n0) {
                t += C0[i * n0 + j] * (double)x[j];
                if (j == i) xi = (double)x[j];
            }
        return K * (t - w0[i] * g) + dlt[i] * xi;
    }
    return vv[i - n0] * (xi - g);
}

// The record and operator tables are separate restrict arguments: the compiler
// then knows the q stores do not alias them and reads them, wave-uniform, with
// scalar loads.
// One general row for CH slots (64 apart) from its flattened record: every
// gather of the first NC contributions first (records past the count hold
// valid rows), then the contributions in order.  The FP32 conductivity of a
// plain element is the FP64 one rounded, as the slot loader stores it.
template <int NC, int CH>
__device__ __forceinline__ void gen_row(const int (&Gr)[kGenRec], int nc, const float *__restrict__ p,
                                        const double *__restrict__ K64, const float *__restrict__ opT,
                                        const double *__restrict__ opT64, size_t Rp, int lane, float (&y)[CH]) {
    float x[CH][NC][kGenR];
    double k64[CH][NC];
#pragma unroll
    for (int h = 0; h < CH; h++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
#pragma unroll
            for (int j = 0; j < kGenR; j++) x[h][c][j] = p[(size_t)Gr[1 + c * kGenStride + 5 + j] * Rp + h * 64 + lane];
            k64[h][c] = K64[(size_t)Gr[1 + c * kGenStride] * Rp + h * 64 + lane];
        }
#pragma unroll
    for (int h = 0; h < CH; h++) {
        float t = 0.0f;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (c >= nc) break;
            const int *Q = Gr + 1 + c * kGenStride;
            const int i = Q[1], n0 = Q[2], nvb = Q[3], off = Q[4];
            if (nvb == 0) t += gen_plain(opT + off, n0, i, (float)k64[h][c], x[h][c]);
            else t += (float)gen_vb(opT64 + off, n0, nvb, i, k64[h][c], x[h][c]);
        }
        y[h] = t;
    }
}

// k_c_opg: the general rows, one row per wave and one slot per lane of each
// of the CH live chunks (all chunks' gathers issued together).  The record and
// operator tables are separate restrict arguments: the compiler then knows the
// q stores do not alias them and reads them, wave-uniform, with scalar loads.
template <int CH>
__global__ void __launch_bounds__(256) k_c_opg(BArgs<float> a, OpTiles ot, const int *__restrict__ gen,
                                               const int *__restrict__ grec, const float *__restrict__ opT,
                                               const double *__restrict__ opT64) {
    __shared__ double sh[kWaves][64];
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t Rp = a.Rp;
    const XPlain<float> xs{a.p0};
    const float *__restrict__ p = a.p0;
    bool act[CH];
    double v[CH];
#pragma unroll
    for (int h = 0; h < CH; h++) {
        act[h] = a.sc.status[h * 64 + lane] == kRun;
        v[h] = 0.0;
    }
    // XCD-aware: the blocks of XCD x (block b runs on XCD b % 8) take the x-th
    // eighth of the general rows, so one L2 holds the neighbour rows a
    // stretch of the (Z-ordered) general rows shares
    const int xcd = blockIdx.x & 7, nb8 = ot.gen_blocks >> 3;
    const int k0 = (int)((long)ot.ngen * xcd / 8), k1 = (int)((long)ot.ngen * (xcd + 1) / 8);
    for (int k = k0 + (int)(blockIdx.x >> 3) * kWaves + wid; k < k1; k += nb8 * kWaves) {
        const int row = gen[k];
        int Gr[kGenRec];
#pragma unroll
        for (int w = 0; w < kGenRec; w++) Gr[w] = grec[(size_t)k * kGenRec + w];
        const int nc = Gr[0];
        float q[CH];
        if (nc < 0) {
#pragma unroll
            for (int h = 0; h < CH; h++)
                q[h] = op_row<float>(a.rec, a.row_ptr, a.row_c, a.meta, a.op_off, a.conn_off, a.conn, a.opT, a.K,
                                     a.opT64, a.K64, row, xs, Rp, h * 64 + lane);
        } else if (nc <= 2) {
            gen_row<2, CH>(Gr, nc, p, a.K64, opT, opT64, Rp, lane, q);
        } else {
            gen_row<kGenC, CH>(Gr, nc, p, a.K64, opT, opT64, Rp, lane, q);
        }
#pragma unroll
        for (int h = 0; h < CH; h++) {
            if (!act[h]) continue;
            const size_t o = (size_t)row * Rp + h * 64 + lane;
            a.q[o] = q[h];
            v[h] += (double)p[o] * (double)q[h];
        }
    }
#pragma unroll
    for (int h = 0; h < CH; h++) {
        __syncthreads();
        sh[wid][lane] = v[h];
        __syncthreads();
        if (wid == 0) {
            double t = sh[0][lane];
            for (int w = 1; w < kWaves; w++) t += sh[w][lane];
            a.part0[(size_t)blockIdx.x * Rp + h * 64 + lane] = t;
        }
    }
}

// k_c_opl: one tile per block (512 threads).  Tile t of XCD x (block b runs on
// XCD b % 8) is tile x * (ntiles / 8) + b / 8, so tiles in flight on one XCD
// are neighbours and share halo lines in its L2.  The tile's records and
// segment list are staged in LDS once; then per group of 32 slots its own rows,
// halo rows and element conductivities are loaded with coalesced 16-byte loads
// (every value of the tile's neighbourhood read once per slot group; the next
// group's loads are in flight while the current group is computed), stored to
// LDS, and 8 lanes x 4 slots compute each row from LDS.  Per slot the
// arithmetic and its order are fixed by the pair-row record (k_b_fastrec):
// t1 and t2 over the two elements' four rows in order, then
// q = (K1 t1 + d1 x) + (K2 t2 + d2 x).
constexpr int kOplThreads = 512;
// One slot group's segments of a tile into registers: 8 lanes x 16 bytes per
// segment row (32 slots), segments past the end clamped to the last one.  The
// segment row addresses come from LDS each time (kept in registers across
// the pipeline they would spill).
// (global address space: a generic load would also count as an LDS access
// and make every LDS wait of the compute wait for the staged loads)
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) f4v *gf4v_ptr;
__device__ __forceinline__ void opl_stage(const float *const *segp, int nseg, int tid, int sb,
                                          f4v (&vv)[kStageMax]) {
    const float *ptr[kStageMax];
#pragma unroll
    for (int u = 0; u < kStageMax; u++) ptr[u] = segp[min((tid >> 3) + u * (kOplThreads / 8), nseg - 1)];
#pragma unroll
    for (int u = 0; u < kStageMax; u++) vv[u] = *(gf4v_ptr)(ptr[u] + sb + 4 * (tid & 7));
}

// nsg: 32-slot groups computed (2 nch, or 1 for a narrowed tail of at most 32
// live slots at the front); partials of the other slots of the nch chunks are
// zeroed.
__global__ void __launch_bounds__(kOplThreads, 4) k_c_opl(BArgs<float> a, OpTiles ot, int nch, int nsg) {
    extern __shared__ double lds_d[];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const size_t Rp = a.Rp;
    const int tb = blockIdx.x;
    const int tile = (tb & 7) * (ot.ntiles >> 3) + (tb >> 3);
    double *const part = a.part0 + (size_t)(ot.gen_blocks + blockIdx.x) * Rp;
    const int r0 = tile * ot.T, nr = max(0, min(ot.T, a.F - r0));
    const int h0 = ot.hptr[tile], nh = ot.hptr[tile + 1] - h0;
    const int e0 = ot.eptr[tile], ne = ot.eptr[tile + 1] - e0;
    const int nseg = nr + nh + ne, eo = nr + nh;
    if (nr == 0) {  // padding tile: zero partials
        for (int k = tid; k < nch * 64; k += blockDim.x) part[k] = 0.0;
        return;
    }
    float4 *buf = reinterpret_cast<float4 *>(lds_d);  // [kStageMax * kOplThreads] float4
    // per row 64 bytes: the 10 local indices as 16-bit pairs in words 0-4, the
    // 10 values in words 5-14, read with four 16-byte LDS loads
    unsigned *lrw = reinterpret_cast<unsigned *>(buf + kStageMax * kOplThreads);
    double *red = reinterpret_cast<double *>(lrw + (size_t)ot.T * kRecWords);  // [8 waves][32 slots]
    const float **segp = reinterpret_cast<const float **>(red + 8 * kSlotGroup);  // row of each segment
    {
        // tile set-up; indices past the end are clamped to valid ones and their
        // values dropped, so no load waits for another
        constexpr int kRecLd = (128 * kLrec + kOplThreads - 1) / kOplThreads;
        constexpr int kSegLd = (512 + kOplThreads - 1) / kOplThreads;
        unsigned short rl[kRecLd];
        int rf[kRecLd];
        const float *sg[kSegLd];
#pragma unroll
        for (int u = 0; u < kRecLd; u++) {
            const int w = min(tid + u * kOplThreads, nr * kLrec - 1);
            const int rr = w / kLrec, k = w - rr * kLrec;
            rl[u] = ot.lrec[(size_t)r0 * kLrec + w];
            rf[u] = a.frec[(size_t)(r0 + rr) * kFastRec + 10 + k];
        }
#pragma unroll
        for (int u = 0; u < kSegLd; u++) {
            const int g = min(tid + u * kOplThreads, nseg - 1);
            const int hi = min(max(h0 + g - nr, h0), max(h0, h0 + nh - 1));
            const int ei = min(max(e0 + g - eo, e0), max(e0, e0 + ne - 1));
            const int vh = ot.halo[hi], ve = ot.els[ei];
            sg[u] = g < eo ? a.p0 + (size_t)(g < nr ? r0 + g : vh) * Rp : a.K + (size_t)ve * Rp;
        }
#pragma unroll
        for (int u = 0; u < kRecLd; u++) {
            const int w = tid + u * kOplThreads;
            if (w < nr * kLrec) {
                const int rr = w / kLrec, k = w - rr * kLrec;
                reinterpret_cast<unsigned short *>(lrw + rr * kRecWords)[k] = rl[u];
                lrw[rr * kRecWords + 5 + k] = (unsigned)rf[u];
            }
        }
#pragma unroll
        for (int u = 0; u < kSegLd; u++) {
            const int g = tid + u * kOplThreads;
            if (g < nseg) segp[g] = sg[u];
        }
    }
    __syncthreads();
    const int quad = tid & 7, rsub = tid >> 3;  // 64 rows per pass, 4 slots per lane
    for (int k = nsg * kSlotGroup + tid; k < nch * 64; k += blockDim.x) part[k] = 0.0;
    // software pipeline: the next slot group's segments are loaded into
    // registers while the current group is computed from LDS.  The run flags
    // of all groups (nch <= 2 here) are read first, as bits, so no wait inside
    // the loop waits for the staged loads too.
    unsigned long long actm = 0;
    for (int sg = 0; sg < nsg; sg++)
#pragma unroll
        for (int c = 0; c < 4; c++)
            actm |= (unsigned long long)(a.sc.status[sg * kSlotGroup + 4 * quad + c] == kRun) << (4 * sg + c);
    f4v vv[kStageMax];
    opl_stage(segp, nseg, tid, 0, vv);
    for (int sg = 0; sg < nsg; sg++) {
        const int sb = sg * kSlotGroup;
        bool act[4];
#pragma unroll
        for (int c = 0; c < 4; c++) act[c] = (actm >> (4 * sg + c)) & 1;
        __syncthreads();  // the previous group's rows are computed
#pragma unroll
        for (int u = 0; u < kStageMax; u++) reinterpret_cast<f4v *>(buf)[tid + u * kOplThreads] = vv[u];
        __syncthreads();
        if (sg + 1 < nsg) opl_stage(segp, nseg, tid, sb + kSlotGroup, vv);
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r = rsub; r < nr; r += kOplThreads / 8) {
            const uint4 *rq = reinterpret_cast<const uint4 *>(lrw + r * kRecWords);
            const uint4 q0 = rq[0], q1 = rq[1], q2 = rq[2], q3 = rq[3];
            if ((q0.x & 0xFFFFu) == 0xFFFFu) continue;
            const unsigned pk[5] = {q0.x, q0.y, q0.z, q0.w, q1.x};
            int R[kLrec];
#pragma unroll
            for (int k = 0; k < kLrec; k++) R[k] = (int)((pk[k >> 1] >> (16 * (k & 1))) & 0xFFFFu);
            const float Fv[kLrec] = {__uint_as_float(q1.y), __uint_as_float(q1.z), __uint_as_float(q1.w),
                                     __uint_as_float(q2.x), __uint_as_float(q2.y), __uint_as_float(q2.z),
                                     __uint_as_float(q2.w), __uint_as_float(q3.x), __uint_as_float(q3.y),
                                     __uint_as_float(q3.z)};
            const float4 xr = buf[r * 8 + quad];
            const float4 k1 = buf[(eo + R[0]) * 8 + quad], k2 = buf[(eo + R[1]) * 8 + quad];
            float4 xa[4], xb[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                xa[j] = buf[R[2 + j] * 8 + quad];
                xb[j] = buf[R[6 + j] * 8 + quad];
            }
            const float d1 = Fv[8], d2 = Fv[9];
            float q[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float x0 = (&xr.x)[c];
                float t1 = 0.0f, t2 = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    t1 += Fv[j] * (&xa[j].x)[c];
                    t2 += Fv[4 + j] * (&xb[j].x)[c];
                }
                float qq = (&k1.x)[c] * t1 + d1 * x0;
                qq += (&k2.x)[c] * t2 + d2 * x0;
                q[c] = qq;
                if (act[c]) v[c] += (double)x0 * (double)qq;
            }
            float *dst = a.q + (size_t)(r0 + r) * Rp + sb + 4 * quad;
            if (act[0] && act[1] && act[2] && act[3]) {
                *reinterpret_cast<float4 *>(dst) = make_float4(q[0], q[1], q[2], q[3]);
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++)
                    if (act[c]) dst[c] = q[c];
            }
        }
        // sum over the 8 rows of the wave (lanes 8 apart), then over the 8 waves
#pragma unroll
        for (int c = 0; c < 4; c++) {
            double x = v[c];
            x += __shfl_xor(x, 8, 64);
            x += __shfl_xor(x, 16, 64);
            x += __shfl_xor(x, 32, 64);
            v[c] = x;
        }
        if (lane < 8) {
#pragma unroll
            for (int c = 0; c < 4; c++) red[wid * kSlotGroup + 4 * quad + c] = v[c];
        }
        __syncthreads();
        if (tid < kSlotGroup) {
            double t = red[tid];
            for (int w = 1; w < kOplThreads / 64; w++) t += red[w * kSlotGroup + tid];
            part[sb + tid] = t;
        }
    }
}

// FP32 operator of the narrowed tail (W lanes per row, 64 / W rows per wave)
// from the inlined pair records: each lane group reads its row's record
// (identical addresses within the group), the next group row's record is
// loaded while the current gathers fly.
template <int W>
__global__ void __launch_bounds__(256) k_c_opn(BArgs<float> a, CArgs c) {
    using L = LaneMap<W>;
    const int lane = threadIdx.x & 63;
    const int r = L::slot(lane), sub = L::sub(lane);
    const bool act = a.sc.status[r] == kRun;
    double v[1] = {0.0};
    if (__any(act)) {
        const float *__restrict__ p = a.p0;
        const float *__restrict__ Kv = a.K;
        const size_t Rp = a.Rp;
        int s0, st, en;
        row_range((a.F + L::P - 1) / L::P, a.G, a.xcd, s0, st, en);
        int C[kFastRec], N[kFastRec];
        {
            const int row = s0 * L::P + sub;
            const bool ok = s0 < en && row < a.F;
#pragma unroll
            for (int k = 0; k < kFastRec; k++) C[k] = ok ? a.frec[(size_t)row * kFastRec + k] : -1;
        }
        for (int g = s0; g < en; g += st) {
            const int row = g * L::P + sub;
            {
                const int nrow = (g + st) * L::P + sub;
                const bool ok = g + st < en && nrow < a.F;
#pragma unroll
                for (int k = 0; k < kFastRec; k++) N[k] = ok ? a.frec[(size_t)nrow * kFastRec + k] : -1;
            }
            if (act && row < a.F) {
                const size_t o = (size_t)row * Rp + r;
                const float xr = p[o];
                float q;
                if (C[0] >= 0) {
                    const float K1 = Kv[(size_t)C[0] * Rp + r], K2 = Kv[(size_t)C[1] * Rp + r];
                    float xa[4], xb[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        xa[j] = p[(size_t)C[2 + j] * Rp + r];
                        xb[j] = p[(size_t)C[6 + j] * Rp + r];
                    }
                    float t1 = 0.0f, t2 = 0.0f;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        t1 += __int_as_float(C[10 + j]) * xa[j];
                        t2 += __int_as_float(C[14 + j]) * xb[j];
                    }
                    q = K1 * t1 + __int_as_float(C[18]) * xr;
                    q += K2 * t2 + __int_as_float(C[19]) * xr;
                } else {
                    const XPlain<float> xs{p};
                    q = op_row<float>(a.rec, a.row_ptr, a.row_c, a.meta, a.op_off, a.conn_off, a.conn, a.opT, a.K,
                                      a.opT64, a.K64, row, xs, Rp, r);
                }
                a.q[o] = q;
                v[0] += (double)xr * (double)q;
            }
#pragma unroll
            for (int k = 0; k < kFastRec; k++) C[k] = N[k];
        }
    }
    {
        const double f = L::fold(v[0]);
        v[0] = sub == 0 ? f : 0.0;
    }
    double *const dst[1] = {a.part0};
    block_part<1>(v, dst, a.Rp);
}

// Per slot, by the zp flags of the last fold: bit 1 x += alpha p (the
// iteration's own step), bit 0 the new direction in place, p = z + beta p
// with z = D^-1 r + Z y (two-level preconditioner)
template <class V, int NCOL, int W>
__global__ void __launch_bounds__(256) k_c_zp(BArgs<V> a, CArgs c) {
    using L = LaneMap<W>;
    const int lane = threadIdx.x & 63;
    const int slot = L::slot(lane), sub = L::sub(lane);
    const int fl = a.sc.zp[slot];
    if (!__any(fl != 0)) return;
    const V beta = (fl & 1) ? (V)a.sc.beta[slot] : V(0);
    const V alpha = (fl & 2) ? (V)a.sc.alpha[slot] : V(0);
    const XTwoLevel<V> xs{a.r, a.dinv, a.p0, NCOL == 2 ? c.phi : nullptr, c.y, c.agg, beta, (size_t)a.Rp, slot};
    int s0, st, en;
    row_range((a.F + L::P - 1) / L::P, a.G, a.xcd, s0, st, en);
    for (int g = s0; g < en; g += st) {
        const int row = g * L::P + sub;
        if (fl == 0 || row >= a.F) continue;
        const size_t o = (size_t)row * a.Rp + slot;
        const V pold = a.p0[o];
        if (fl & 2) a.x[o] += alpha * pold;
        if (fl & 1) a.p0[o] = xs.z(row, o) + beta * pold;
    }
}

// Coarse solve y = (Z^T S Z)^-1 Z^T r of the slots the fold will step, from
// the packed lower tiles of the FP32 inverse (FP64 accumulation), in two
// kernels.  k_c_solve_t: block (kSolveTiles tiles (I, J), slot) streams its 64 x 64 tiles
// once and forms the tile's row product X_IJ r_J for y_I and, off the
// diagonal, its transposed product X_IJ^T r_I for y_J, into the per-slot
// table py[slot][block row][block column][64].  k_c_solve_y: block (I, slot)
// sums y_I over the block columns in order (deterministic) and writes its
// share of the per-slot sums
//   0: D^-1 r . r   1: r . r   2: y . Z^T r
// to cpart[(k * nt + I) * Rp + slot].  MODE: kFoldInit (selected slots),
// kFoldBeta (running slots), kFoldRU (reliable-update selection).
template <int MODE>
__device__ __forceinline__ bool coarse_act(const BatchScal &sc, int slot, const int *__restrict__ lsel) {
    const int st0 = sc.status[slot];
    if (MODE == kFoldInit) return lsel[slot] != 0;
    if (MODE == kFoldBeta) return st0 == kRun;
    return ru_pick(sc, slot, st0);
}

// consecutive tiles of one slot per k_c_solve_t block, loaded together (2 and
// 4 ran slower, as did 4 slots' tiles per block: the blocks are short and
// many, and more work per block lengthens each one's serial part)
constexpr int kSolveTiles = 1;

template <int MODE>
__global__ void __launch_bounds__(256) k_c_solve_t(BatchScal sc, CArgs c, size_t Rp, const int *__restrict__ lsel,
                                                   double *__restrict__ py) {
    const int slot = blockIdx.y;
    if (!coarse_act<MODE>(sc, slot, lsel)) return;  // block-uniform
    const int nt = c.nt, ntl = coarse::ntiles_lower(nt), t0 = blockIdx.x * kSolveTiles;
    __shared__ double xc[kSolveTiles][64], xr[kSolveTiles][64], red[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int I[kSolveTiles], J[kSolveTiles];
#pragma unroll
    for (int u = 0; u < kSolveTiles; u++) coarse::tile_of(min(t0 + u, ntl - 1), I[u], J[u]);
    // tile rows (tid >> 4) + 16 q, columns 4 (tid & 15) .. + 3
    const float4 *T = reinterpret_cast<const float4 *>(c.ainv + ((size_t)slot * ntl + t0) * 4096);
    float4 v[kSolveTiles][4];
#pragma unroll
    for (int u = 0; u < kSolveTiles; u++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            v[u][q] = t0 + u < ntl ? T[u * 1024 + tid + 256 * q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kSolveTiles; u++) {
        if (tid < 64) xc[u][tid] = c.rc[(size_t)(J[u] * 64 + tid) * Rp + slot];
        else if (tid < 128) xr[u][tid - 64] = c.rc[(size_t)(I[u] * 64 + tid - 64) * Rp + slot];
    }
    __syncthreads();
    const int c0 = 4 * (tid & 15), rb = tid >> 4;
#pragma unroll
    for (int u = 0; u < kSolveTiles; u++) {
        if (t0 + u >= ntl) break;  // block-uniform
        double cp[4] = {0.0, 0.0, 0.0, 0.0};
        double *row_out = py + (((size_t)slot * nt + I[u]) * nt + J[u]) * 64;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = rb + 16 * q;
            double s = (double)v[u][q].x * xc[u][c0];
            s += (double)v[u][q].y * xc[u][c0 + 1];
            s += (double)v[u][q].z * xc[u][c0 + 2];
            s += (double)v[u][q].w * xc[u][c0 + 3];
            s += __shfl_xor(s, 1, 64);
            s += __shfl_xor(s, 2, 64);
            s += __shfl_xor(s, 4, 64);
            s += __shfl_xor(s, 8, 64);
            if ((tid & 15) == 0) row_out[r] = s;
            const double xi = xr[u][r];
            cp[0] += (double)v[u][q].x * xi;
            cp[1] += (double)v[u][q].y * xi;
            cp[2] += (double)v[u][q].z * xi;
            cp[3] += (double)v[u][q].w * xi;
        }
        if (I[u] == J[u]) continue;  // block-uniform
#pragma unroll
        for (int q = 0; q < 4; q++) {
            cp[q] += __shfl_xor(cp[q], 16, 64);
            cp[q] += __shfl_xor(cp[q], 32, 64);
        }
        if (lane < 16)
#pragma unroll
            for (int q = 0; q < 4; q++) red[wid][c0 + q] = cp[q];
        __syncthreads();
        if (tid < 64)
            py[(((size_t)slot * nt + J[u]) * nt + I[u]) * 64 + tid] =
                ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
        __syncthreads();  // red is reused by the next tile
    }
}

template <int MODE>
__global__ void __launch_bounds__(64) k_c_solve_y(BatchScal sc, CArgs c, const double *__restrict__ part1,
                                                  const double *__restrict__ part2, size_t Rp,
                                                  const int *__restrict__ lsel, const double *__restrict__ py,
                                                  double *__restrict__ cpart) {
    const int I = blockIdx.x, slot = blockIdx.y, r = threadIdx.x;
    if (!coarse_act<MODE>(sc, slot, lsel)) return;
    const int nt = c.nt;
    const double *p = py + ((size_t)slot * nt + I) * nt * 64 + r;
    double y = 0.0;
    constexpr int kChunk = 16;  // loads in flight together; summed in J order
    for (int J0 = 0; J0 < nt; J0 += kChunk) {
        double v[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; u++) v[u] = p[(size_t)min(J0 + u, nt - 1) * 64];
#pragma unroll
        for (int u = 0; u < kChunk; u++)
            if (J0 + u < nt) y += v[u];
    }
    double s[3] = {0.0, 0.0, 0.0};
    const int row = I * 64 + r;
    if (row < c.nc) {
        c.y[(size_t)row * Rp + slot] = y;
        s[2] = y * c.rc[(size_t)row * Rp + slot];
    }
    // this block's share of the aggregate partials
    const int g0 = I * c.nagg / nt, g1 = (I + 1) * c.nagg / nt;
    for (int g = g0 + r; g < g1; g += 64) {
        s[0] += part1[(size_t)g * Rp + slot];
        s[1] += part2[(size_t)g * Rp + slot];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double x = s[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (r == 0) cpart[((size_t)k * nt + I) * Rp + slot] = x;
    }
}

// Scalar step of CG per slot from the coarse-solve sums (one lane per slot).
// zp flags for the direction kernel: bit 0 re-form the direction, bit 1 add
// alpha p to x (the update of x is deferred to the direction kernel, which
// reads p anyway).
template <int MODE>
__global__ void k_c_fin(BatchScal sc, const double *__restrict__ cpart, int nparts, size_t Rp, double rtol2,
                        double atol2, int max_it, int mixed, const int *__restrict__ lsel) {
    const int slot = blockIdx.x * 64 + threadIdx.x;
    const int st0 = sc.status[slot];
    if (!coarse_act<MODE>(sc, slot, lsel)) {
        sc.zp[slot] = 0;
        return;
    }
    double t[3] = {0.0, 0.0, 0.0};
    constexpr int kChunk = 16;  // loads in flight together; each sum in part order
    for (int p0 = 0; p0 < nparts; p0 += kChunk) {
        double v[3][kChunk];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int u = 0; u < kChunk; u++)
                v[k][u] = cpart[((size_t)k * nparts + min(p0 + u, nparts - 1)) * Rp + slot];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int u = 0; u < kChunk; u++)
                if (p0 + u < nparts) t[k] += v[k][u];
    }
    // t[0] = r^T D^-1 r (Jacobi part), t[1] = r.r, t[2] = y . Z^T r (coarse part);
    // rr: square of the stop-test norm
    const double rz = t[0] + t[2], rr = sc.nat ? t[0] : t[1];
    int st = st0;
    if (MODE == kFoldInit) {
        sc.bnorm2[slot] = rr;
        sc.tol2[slot] = fmax(rtol2 * rr, atol2);
        sc.rz[slot] = rz;
        sc.rz_prev[slot] = 1.0;
        sc.rr[slot] = rr;
        sc.beta[slot] = 0.0;
        sc.alpha[slot] = 0.0;
        sc.it[slot] = 0;
        sc.n_ru[slot] = 0;
        sc.it_ru[slot] = 0;
        sc.par[slot] = 0;
        sc.rr_hist[slot] = INFINITY;
        sc.rr_ru[slot] = rr;
        st = rr <= sc.tol2[slot] ? kConverged : kRun;
    } else if (MODE == kFoldBeta) {
        const int it = sc.it[slot] + 1;
        sc.it[slot] = it;
        sc.rz_prev[slot] = sc.rz[slot];
        sc.rz[slot] = rz;
        // Fletcher-Reeves: the two-level preconditioner is one fixed SPD operator
        sc.beta[slot] = rz / sc.rz_prev[slot];
        sc.rr[slot] = rr;
        if (!(rr == rr)) st = kBreakdown;
        else if (rr <= sc.tol2[slot]) st = kCandidate;  // confirmed on the true residual at the chunk end
        else if (it >= max_it) st = kMaxIt;
        else if (mixed && it - sc.it_ru[slot] > 1000 + sc.it_ru[slot]) st = kFp64;
    } else {  // reliable update
        sc.n_ru[slot] += 1;
        sc.rr_ru[slot] = rr;
        sc.it_ru[slot] = sc.it[slot];
        if (st0 != kRun) sc.beta[slot] = rz / sc.rz_prev[slot];  // re-formed direction below
        sc.rz[slot] = rz;
        sc.rr[slot] = rr;
        if (!(rr == rr)) st = kBreakdown;
        else if (st0 == kCandidate) {
            if (rr <= sc.tol2[slot]) st = kConverged;
            else {
                st = (rr > 0.01 * sc.rr_hist[slot]) ? kFp64 : kRun;
                sc.rr_hist[slot] = rr;
            }
        } else if (rr <= sc.tol2[slot]) st = kConverged;
        else if (sc.it[slot] >= max_it) st = kMaxIt;
        else st = kRun;
    }
    sc.status[slot] = st;
    // a slot that kept iterating up to its reliable update keeps its direction
    // (residual replacement); paused and candidate slots re-form theirs
    const bool keep = MODE == kFoldRU && st0 == kRun;
    sc.zp[slot] = (st == kRun && !keep ? 1 : 0) | (MODE == kFoldBeta ? 2 : 0);
}

// slot-major arrays ([slot][n]): out[s] = in[src[s]]
template <class T>
__global__ void k_b_perm_major(const T *__restrict__ in, T *__restrict__ out, size_t n, const int *__restrict__ src) {
    const int s = blockIdx.y;
    const T *a = in + (size_t)src[s] * n;
    T *b = out + (size_t)s * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// Slot permutation of a realization-interleaved array: out[n][s] = in[n][src[s]]
template <class T>
__global__ void k_b_perm(const T *__restrict__ in, T *__restrict__ out, size_t N, size_t Rp,
                         const int *__restrict__ src) {
    const int s = blockIdx.y * 64 + (threadIdx.x & 63);
    const int from = src[s];
    for (size_t n = blockIdx.x * (size_t)kWaves + (threadIdx.x >> 6); n < N; n += (size_t)gridDim.x * kWaves)
        out[n * Rp + s] = in[n * Rp + from];
}

// the same for the search direction, taken from the buffer each slot's parity names
template <class T>
__global__ void k_b_perm_p(const T *__restrict__ p0, const T *__restrict__ p1, T *__restrict__ out, size_t N,
                           size_t Rp, const int *__restrict__ src, const int *__restrict__ par) {
    const int s = blockIdx.y * 64 + (threadIdx.x & 63);
    const int from = src[s];
    const T *in = par[from] ? p1 : p0;
    for (size_t n = blockIdx.x * (size_t)kWaves + (threadIdx.x >> 6); n < N; n += (size_t)gridDim.x * kWaves)
        out[n * Rp + s] = in[n * Rp + from];
}

__global__ void k_b_perm_scal(BatchScal in, BatchScal out, const int *__restrict__ src, int Rp) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Rp) return;
    const int f = src[s];
    out.alpha[s] = in.alpha[f];
    out.beta[s] = in.beta[f];
    out.rz[s] = in.rz[f];
    out.rz_prev[s] = in.rz_prev[f];
    out.rr[s] = in.rr[f];
    out.tol2[s] = in.tol2[f];
    out.bnorm2[s] = in.bnorm2[f];
    out.rr_hist[s] = in.rr_hist[f];
    out.rr_ru[s] = in.rr_ru[f];
    out.status[s] = in.status[f];
    out.it[s] = in.it[f];
    out.n_ru[s] = in.n_ru[f];
    out.it_ru[s] = in.it_ru[f];
    out.zp[s] = in.zp[f];
    out.par[s] = 0;  // directions are moved into buffer 0
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <class T>
static T *b_up(std::vector<void *> &own, const std::vector<T> &v, hipStream_t s) {
    T *p = dev_upload(v.data(), v.size(), s);
    own.push_back(p);
    return p;
}

template <class V>
static BArgs<V> make_args(int F, int Rp, int G, int xcd, const int *rec, const int *row_ptr, const int *row_c,
                          const int *meta,
                          const int *op_off, const int *conn_off, const int *conn, const V *opT, const V *K,
                          const double *opT64, const double *K64,
                          const typename Pre<V>::T *dinv, V *x, V *r, V *z, V *p0, V *p1, V *q, double *part,
                          const BatchScal &sc) {
    BArgs<V> a;
    a.F = F; a.Rp = Rp; a.G = G; a.xcd = xcd;
    a.row_ptr = row_ptr; a.row_c = row_c; a.meta = meta; a.op_off = op_off; a.conn_off = conn_off; a.conn = conn;
    a.rec = rec;
    a.frec = nullptr;
    a.opT = opT; a.K = K; a.dinv = dinv;
    a.opT64 = opT64; a.K64 = K64;
    a.x = x; a.r = r; a.z = z; a.p0 = p0; a.p1 = p1; a.q = q;
    a.part0 = part; a.part1 = part + (size_t)kMaxGrid * Rp; a.part2 = part + 2 * (size_t)kMaxGrid * Rp;
    a.sc = sc;
    return a;
}

DarcyBatchGPU::DarcyBatchGPU(DarcyGPU &base, int R, int slots) : base_(base), dv_(base.dev()) {
    if (R < 1) throw Error(-1, "batch needs at least one realization");
    if (slots < 0) throw Error(-1, "slot count must be >= 0");
    R_ = R;
    Rout_ = (R + 63) / 64 * 64;
    Rp_ = std::min(Rout_, slots > 0 ? (slots + 63) / 64 * 64 : 256);
    nchunk_ = Rp_ / 64;
    FL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    FL_HIP(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
    FL_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    FL_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    build();
}

DarcyBatchGPU::~DarcyBatchGPU() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    free_graphs();
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
    for (void *p : owned_) dev_free(p);
    if (h_count_) (void)hipHostFree(h_count_);
    if (stream2_) (void)hipStreamSynchronize(stream2_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (stream2_) (void)hipStreamDestroy(stream2_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void DarcyBatchGPU::free_graphs() {
    if (g64_) (void)hipGraphExecDestroy(g64_);
    if (g32_) (void)hipGraphExecDestroy(g32_);
    g64_ = g32_ = nullptr;
    g64_key_ = g32_key_ = -1;
}

// Tiles of the FP32 operator (k_c_opl): general rows, and per tile of T rows
// its halo rows, its elements and the local-index records of its fast rows.
// T = 128 rows, halved while a tile's staged segments would not fit the LDS
// budget; a multiple of 8 tiles.
// The row tiles of k_c_opl for at most tmax rows per tile (fewer while a
// tile's neighbourhood exceeds the staging budget), uploaded as a candidate
// tiling; returns false when the tiling equals an existing candidate.
bool DarcyBatchGPU::build_tiling(const std::vector<int> &rec, int tmax) {
    const int F = F_;
    constexpr int kSegBudget = kStageMax * kOplThreads / 8;  // 56 KiB of staged rows per 32-slot pass
    std::vector<int> hptr, halo, eptr, els;
    std::vector<unsigned short> lrec;
    int T = tmax, nseg_max = 0, ntiles = 0;
    for (;; T /= 2) {
        ntiles = ((F + T - 1) / T + 7) / 8 * 8;
        hptr.assign(1, 0);
        eptr.assign(1, 0);
        halo.clear();
        els.clear();
        lrec.assign((size_t)F * kLrec, 0xFFFF);
        nseg_max = 0;
        std::unordered_map<int, int> hloc, eloc;
        for (int t = 0; t < ntiles; t++) {
            const int r0 = t * T, r1 = std::min(F, r0 + T), nr = std::max(0, r1 - r0);
            hloc.clear();
            eloc.clear();
            std::vector<int> th, te;
            auto rloc = [&](int row) {
                if (row >= r0 && row < r1) return row - r0;
                auto it = hloc.find(row);
                if (it != hloc.end()) return it->second;
                const int l = nr + (int)th.size();
                th.push_back(row);
                hloc[row] = l;
                return l;
            };
            auto elo = [&](int e) {
                auto it = eloc.find(e);
                if (it != eloc.end()) return it->second;
                const int l = (int)te.size();
                te.push_back(e);
                eloc[e] = l;
                return l;
            };
            for (int f = r0; f < r1; f++) {
                const int *R = &rec[(size_t)f * kRecInts];
                if (!op_fast_row(R)) continue;
                const int *Q2 = R[8] < 0 ? R : R + 8;  // one tetrahedron: repeated (its block row is zero)
                unsigned short *L = &lrec[(size_t)f * kLrec];
                L[0] = (unsigned short)elo(R[0]);
                L[1] = (unsigned short)elo(Q2[0]);
                for (int j = 0; j < 4; j++) {
                    L[2 + j] = (unsigned short)rloc(R[4 + j]);
                    L[6 + j] = (unsigned short)rloc(Q2[4 + j]);
                }
            }
            // element segments follow the halo: the records hold element indices
            // relative to that start
            halo.insert(halo.end(), th.begin(), th.end());
            els.insert(els.end(), te.begin(), te.end());
            hptr.push_back((int)halo.size());
            eptr.push_back((int)els.size());
            nseg_max = std::max(nseg_max, nr + (int)th.size() + (int)te.size());
        }
        if (nseg_max <= kSegBudget || T <= 16) break;
    }
    static_assert(kSegBudget <= 512, "staging registers cover 512 segments");
    if (nseg_max > 65535 || opl_lds_bytes(nseg_max, T) > 160 * 1024)
        throw Error(-5, "batched operator: a row tile's neighbourhood does not fit the LDS");
    for (const OplTiling &c : tilings_)
        if (c.T == T) return false;
    if (halo.empty()) halo.push_back(0);
    if (els.empty()) els.push_back(0);
    hipStream_t s = stream_;
    OplTiling c;
    c.T = T;
    c.ntiles = ntiles;
    c.nseg_max = nseg_max;
    c.hptr = b_up(owned_, hptr, s);
    c.halo = b_up(owned_, halo, s);
    c.eptr = b_up(owned_, eptr, s);
    c.els = b_up(owned_, els, s);
    c.lrec = b_up(owned_, lrec, s);
    tilings_.push_back(c);
    return true;
}

// Candidate tilings of the FP32 operator (at most 128 and at most 64 rows per
// tile; tune() times them on the filled slots and keeps the faster), and the
// general rows (elements with vb neighbours) of k_c_opg.
void DarcyBatchGPU::build_op_tiles(const std::vector<int> &rec) {
    const int F = F_;
    std::vector<int> gen;
    for (int f = 0; f < F; f++)
        if (!op_fast_row(&rec[(size_t)f * kRecInts])) gen.push_back(f);
    ngen_ = (int)gen.size();
    if (gen.empty()) gen.push_back(0);
    tilings_.clear();
    for (int tmax : {128, 64}) build_tiling(rec, tmax);
    int ntiles_max = 0;
    size_t lds_max = 0;
    for (const OplTiling &c : tilings_) {
        ntiles_max = std::max(ntiles_max, c.ntiles);
        lds_max = std::max(lds_max, opl_lds_bytes(c.nseg_max, c.T));
    }
    // general rows are long dependent chains: about 3 per wave, as many
    // blocks as the partial arrays leave beside the largest tiling
    gen_blocks_ = ngen_ ? std::min(kMaxGrid - ntiles_max, ((ngen_ + 11) / 12 + 7) / 8 * 8) / 8 * 8 : 0;
    if (gen_blocks_ < 0 || (ngen_ && gen_blocks_ == 0))
        throw Error(-5, "batched operator: too many row tiles for the partial arrays");
    d_gen_ = b_up(owned_, gen, stream_);
    FL_HIP(hipFuncSetAttribute((const void *)k_c_opl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
    use_tiling(0);
}

void DarcyBatchGPU::use_tiling(int i) {
    const OplTiling &c = tilings_.at(i);
    d_thptr_ = c.hptr;
    d_thalo_ = c.halo;
    d_teptr_ = c.eptr;
    d_tels_ = c.els;
    d_lrec_ = c.lrec;
    tile_T_ = c.T;
    ntiles_ = c.ntiles;
    nseg_max_ = c.nseg_max;
    tiling_ = i;
    free_graphs();  // captured launches hold the tiling
}

void DarcyBatchGPU::build() {
    const HostMesh &m = base_.mesh();
    E_ = m.n_elements();
    S_ = m.n_sides();
    F_ = m.n_edges();
    const int E = E_, F = F_;
    std::vector<int> conn_off(E + 1, 0), op_off(E + 1, 0), rs_off(E + 1, 0);
    for (int e = 0; e < E; e++) {
        const int n0 = m.el_dim[e] + 1, nvb = m.el_vb_ptr[e + 1] - m.el_vb_ptr[e];
        if (n0 + nvb > kMaxConn) throw Error(-3, "element with more than 16 connected edges");
        conn_off[e + 1] = conn_off[e] + n0 + nvb;
        op_off[e + 1] = op_off[e] + op_size(n0, nvb);
        rs_off[e + 1] = rs_off[e] + rs_size(n0, nvb);
    }
    if (E >= (1 << 27)) throw Error(-3, "too many elements for the batched contribution encoding");
    std::vector<int> conn(conn_off[E]);
    for (int e = 0; e < E; e++) {
        int o = conn_off[e];
        for (int i = 0; i <= m.el_dim[e]; i++) conn[o++] = m.side_edge[m.side_off[e] + i];
        for (int k = m.el_vb_ptr[e]; k < m.el_vb_ptr[e + 1]; k++) conn[o++] = m.vb_edge[m.el_vb[k]];
    }
    // Rows in Morton (Z-curve) order of the edge positions: a row's neighbours
    // are then close in memory, so the gathers of rows processed together hit
    // the same L2 lines.  perm: row -> edge; conn is renumbered to rows.
    std::vector<int> perm(F);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<uint64_t> key;  // Z-curve key per edge
    {
        std::vector<double> ec(3 * (size_t)F, 0.0);
        std::vector<int> cnt(F, 0);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int e = 0; e < E; e++) {
            double c[3] = {0, 0, 0};
            const int nn = m.el_dim[e] + 1;
            for (int k = 0; k < nn; k++)
                for (int d = 0; d < 3; d++) c[d] += m.xyz[3 * (size_t)m.el_nodes[4 * (size_t)e + k] + d] / nn;
            for (int i = conn_off[e]; i < conn_off[e + 1]; i++) {
                for (int d = 0; d < 3; d++) ec[3 * (size_t)conn[i] + d] += c[d];
                cnt[conn[i]]++;
            }
        }
        for (int f = 0; f < F; f++)
            for (int d = 0; d < 3; d++) {
                double v = ec[3 * (size_t)f + d] / std::max(cnt[f], 1);
                ec[3 * (size_t)f + d] = v;
                lo[d] = std::min(lo[d], v);
                hi[d] = std::max(hi[d], v);
            }
        auto spread = [](uint64_t v) {  // 21 bits -> every third bit
            v &= 0x1fffff;
            v = (v | v << 32) & 0x1f00000000ffffULL;
            v = (v | v << 16) & 0x1f0000ff0000ffULL;
            v = (v | v << 8) & 0x100f00f00f00f00fULL;
            v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
            v = (v | v << 2) & 0x1249249249249249ULL;
            return v;
        };
        key.assign(F, 0);
        for (int f = 0; f < F; f++) {
            uint64_t k = 0;
            for (int d = 0; d < 3; d++) {
                double w = hi[d] > lo[d] ? (ec[3 * (size_t)f + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
                uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, w * 2097151.0));
                k |= spread(q) << d;
            }
            key[f] = k;
        }
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    {
        std::vector<int> rinv(F);
        for (int r = 0; r < F; r++) rinv[perm[r]] = r;
        for (auto &c : conn) c = rinv[c];
    }
    h_perm_ = perm;
    // Aggregates of the two-level preconditioner: the cells of one level of
    // the Z-curve (about 800 rows each), i.e. contiguous row ranges; without
    // the curve, runs of 800 consecutive rows.
    {
        std::vector<int> agg(F);
        int n = 0;
        if (!key.empty()) {
            int L = (int)std::lround(std::log(std::max(1.0, F / 800.0)) / std::log(8.0));
            L = std::max(1, std::min(L, 3));
            const int sh = 63 - 3 * L;
            uint64_t prev = ~0ULL;
            for (int r = 0; r < F; r++) {
                const uint64_t pfx = key[perm[r]] >> sh;
                if (pfx != prev) {
                    prev = pfx;
                    n++;
                }
                agg[r] = n - 1;
            }
        } else {
            for (int r = 0; r < F; r++) agg[r] = r / 800;
            n = (F + 799) / 800;
        }
        nagg_ = n;
        h_agg_ = agg;
        std::vector<int> aptr(n + 1, 0);
        for (int r = 0; r < F; r++) aptr[agg[r] + 1]++;
        for (int a = 0; a < n; a++) aptr[a + 1] += aptr[a];
        std::vector<std::vector<int>> nb(n);
        for (int e = 0; e < E; e++) {
            int as[kMaxConn];
            int na = 0;
            for (int i = conn_off[e]; i < conn_off[e + 1]; i++) as[na++] = agg[conn[i]];
            for (int x = 0; x < na; x++)
                for (int y = 0; y < na; y++) nb[as[x]].push_back(as[y]);
        }
        std::vector<int> nbptr(n + 1, 0), nbv;
        nbmax_ = 0;
        for (int a = 0; a < n; a++) {
            nb[a].push_back(a);
            std::sort(nb[a].begin(), nb[a].end());
            nb[a].erase(std::unique(nb[a].begin(), nb[a].end()), nb[a].end());
            nbmax_ = std::max(nbmax_, (int)nb[a].size());
            nbv.insert(nbv.end(), nb[a].begin(), nb[a].end());
            nbptr[a + 1] = (int)nbv.size();
        }
        coarse_ok_ = nbmax_ <= kNbMax && n <= 1024 && E < (1 << 23);  // term packing e << 8
        std::vector<int> lbmap;
        if (coarse_ok_) {
            lbmap.assign((size_t)n * n, -1);
            for (int a = 0; a < n; a++)
                for (int k = nbptr[a]; k < nbptr[a + 1]; k++) lbmap[(size_t)a * n + nbv[k]] = k - nbptr[a];
        } else {
            lbmap.assign(1, -1);
        }
        h_nbptr_ = nbptr;
        h_nbv_ = nbv;
        h_lbmap_ = lbmap;
        d_agg_ = b_up(owned_, agg, stream_);
        d_aptr_ = b_up(owned_, aptr, stream_);
        d_nbptr_ = b_up(owned_, nbptr, stream_);
        d_nb_ = b_up(owned_, nbv, stream_);
        d_lbmap_ = b_up(owned_, lbmap, stream_);
    }
    // contributions of each row in element order
    std::vector<int> row_ptr(F + 1, 0), row_c(conn_off[E]);
    for (int i = 0; i < conn_off[E]; i++) row_ptr[conn[i] + 1]++;
    for (int f = 0; f < F; f++) row_ptr[f + 1] += row_ptr[f];
    {
        std::vector<int> fp(row_ptr.begin(), row_ptr.end() - 1);
        for (int e = 0; e < E; e++)
            for (int i = conn_off[e]; i < conn_off[e + 1]; i++) row_c[fp[conn[i]]++] = e * kMaxConn + (i - conn_off[e]);
    }
    ncontrib_ = conn_off[E];
    if (coarse_ok_) {
        // terms of the coarse blocks (a, b), grouped by block, in row /
        // contribution / column order: (row, e << 8 | i << 4 | j)
        const int ntg = (int)h_nbv_.size();
        std::vector<int> cnt(ntg + 1, 0), tagg(ntg);
        for (int a = 0; a < nagg_; a++)
            for (int k = h_nbptr_[a]; k < h_nbptr_[a + 1]; k++) tagg[k] = a;
        auto target = [&](int row, int g) { return h_nbptr_[h_agg_[row]] + h_lbmap_[(size_t)h_agg_[row] * nagg_ + h_agg_[g]]; };
        for (int pass = 0; pass < 2; pass++) {
            std::vector<int> fill;
            std::vector<int2> terms;
            if (pass == 1) {
                for (int t = 0; t < ntg; t++) cnt[t + 1] += cnt[t];
                fill.assign(cnt.begin(), cnt.end() - 1);
                terms.resize(cnt[ntg]);
            }
            for (int row = 0; row < F; row++)
                for (int cc = row_ptr[row]; cc < row_ptr[row + 1]; cc++) {
                    const int e = row_c[cc] / kMaxConn, i = row_c[cc] % kMaxConn;
                    for (int j = 0; j < conn_off[e + 1] - conn_off[e]; j++) {
                        const int t = target(row, conn[conn_off[e] + j]);
                        if (pass == 0) cnt[t + 1]++;
                        else terms[fill[t]++] = int2{row, (e << 8) | (i << 4) | j};
                    }
                }
            if (pass == 1) {
                // chunks of at most kAsmChunk terms, in block order
                std::vector<int> cbeg, cblk, cfirst(ntg + 1);
                for (int t = 0; t < ntg; t++) {
                    cfirst[t] = (int)cbeg.size();
                    for (int b = cnt[t]; b < cnt[t + 1]; b += kAsmChunk) {
                        cbeg.push_back(b);
                        cblk.push_back(t);
                    }
                }
                cfirst[ntg] = (int)cbeg.size();
                nasm_chunks_ = (int)cbeg.size();
                d_cbeg_ = b_up(owned_, cbeg, stream_);
                d_cblk_ = b_up(owned_, cblk, stream_);
                d_cfirst_ = b_up(owned_, cfirst, stream_);
                d_tptr_ = b_up(owned_, cnt, stream_);
                d_tagg_ = b_up(owned_, tagg, stream_);
                int2 *dt = dev_alloc<int2>(terms.size());
                owned_.push_back(dt);
                FL_HIP(hipMemcpyAsync(dt, terms.data(), terms.size() * sizeof(int2), hipMemcpyHostToDevice, stream_));
                FL_HIP(hipStreamSynchronize(stream_));
                d_terms_ = dt;
            }
        }
    }
    std::vector<int> rec((size_t)F * kRecInts, -1);
    for (int f = 0; f < F; f++) {
        int *Rr = &rec[(size_t)f * kRecInts];
        const int nc = row_ptr[f + 1] - row_ptr[f];
        bool simple = nc <= 2;
        for (int c = row_ptr[f]; c < row_ptr[f + 1] && simple; c++) {
            const int e = row_c[c] / kMaxConn;
            simple = m.el_vb_ptr[e + 1] == m.el_vb_ptr[e];
        }
        if (!simple) {
            Rr[0] = -2;
            continue;
        }
        for (int k = 0; k < nc; k++) {
            const int code = row_c[row_ptr[f] + k];
            const int e = code / kMaxConn, i = code % kMaxConn, n0 = m.el_dim[e] + 1;
            int *Q = Rr + 8 * k;
            Q[0] = e;
            Q[1] = i;
            Q[2] = n0;
            Q[3] = op_off[e];
            for (int j = 0; j < 4; j++) Q[4 + j] = j < n0 ? conn[conn_off[e] + j] : 0;
        }
    }
    d_rec_ = b_up(owned_, rec, stream_);
    build_op_tiles(rec);
    {
        // flattened records of the general rows (k_c_opg): up to kGenC element
        // contributions in the row's contribution order, each with its element,
        // local index, size, table offset and connected rows; more contributions
        // or rows: nC = -1, the row is evaluated by op_row
        std::vector<int> grec((size_t)std::max(ngen_, 1) * kGenRec, 0);
        std::vector<int> gen_rows;
        for (int f = 0; f < F; f++)
            if (!op_fast_row(&rec[(size_t)f * kRecInts])) gen_rows.push_back(f);
        for (int k = 0; k < ngen_; k++) {
            const int f = gen_rows[k];
            int *G = &grec[(size_t)k * kGenRec];
            const int nc = row_ptr[f + 1] - row_ptr[f];
            bool ok = nc <= kGenC;
            for (int c = 0; c < nc && ok; c++) {
                const int e = row_c[row_ptr[f] + c] / kMaxConn;
                ok = conn_off[e + 1] - conn_off[e] <= kGenR;
            }
            G[0] = ok ? nc : -1;
            if (!ok) continue;
            for (int c = 0; c < nc; c++) {
                const int code = row_c[row_ptr[f] + c];
                const int e = code / kMaxConn, i = code % kMaxConn;
                int *Q = G + 1 + c * kGenStride;
                Q[0] = e;
                Q[1] = i;
                Q[2] = m.el_dim[e] + 1;
                Q[3] = m.el_vb_ptr[e + 1] - m.el_vb_ptr[e];
                Q[4] = op_off[e];
                for (int j = 0; j < kGenR; j++)
                    Q[5 + j] = conn_off[e] + j < conn_off[e + 1] ? conn[conn_off[e] + j] : f;
            }
        }
        d_grec_ = b_up(owned_, grec, stream_);
    }
    d_frec_ = dev_alloc<int>((size_t)F * kFastRec);
    owned_.push_back(d_frec_);
    std::vector<int> dperm;
    for (int d = 1; d <= 3; d++) {
        for (int e = 0; e < E; e++)
            if (m.el_dim[e] == d) dperm.push_back(e);
        dim_start_[d] = (int)dperm.size();
    }
    hipStream_t s = stream_;
    d_conn_off_ = b_up(owned_, conn_off, s);
    d_conn_ = b_up(owned_, conn, s);
    d_op_off_ = b_up(owned_, op_off, s);
    d_rs_off_ = b_up(owned_, rs_off, s);
    d_row_ptr_ = b_up(owned_, row_ptr, s);
    d_row_c_ = b_up(owned_, row_c, s);
    d_perm_ = b_up(owned_, perm, s);
    d_eperm_ = b_up(owned_, dperm, s);
    op_total_ = op_off[E];
    rs_total_ = rs_off[E];
    tab_bytes_ = (long)op_total_ * 8;
    auto ad = [&](size_t n) { double *p = dev_alloc<double>(n); owned_.push_back(p); return p; };
    auto af = [&](size_t n) { float *p = dev_alloc<float>(n); owned_.push_back(p); return p; };
    auto ai = [&](size_t n) { int *p = dev_alloc<int>(n); owned_.push_back(p); return p; };
    d_meta_ = ai(E);
    d_tab_ = ad(op_total_);
    d_tab32_ = af(op_total_);
    d_rsT_ = ad(rs_total_);
    const size_t nv = (size_t)F * Rp_;
    d_K_ = ad((size_t)E * Rp_);
    d_K32_ = af((size_t)E * Rp_);
    d_b_ = ad(nv);
    d_dinv_ = ad(nv);
    d_dinv32_ = dev_alloc<unsigned short>(nv);
    owned_.push_back(d_dinv32_);
    d_x_ = ad(nv); d_r_ = ad(nv); d_z_ = ad(nv); d_p0_ = ad(nv); d_p1_ = ad(nv); d_q_ = ad(nv);
    d_sol_ = ad(((size_t)S_ + E + F) * Rout_);
    d_Kall_ = ad((size_t)R_ * E);
    d_var_ = ad(R_);
    d_pairs_ = ai(2 * (size_t)Rp_);
    d_Ksave_ = ad(E);
    d_lsel_ = ai(Rp_);
    d_ext_ = ad((size_t)S_ + E + F);
    d_part_ = ad(3 * (size_t)kMaxGrid * Rp_);
    d_pre_ = ad(3 * (size_t)kPre * Rp_);
    d_cpart_ = ad(3 * (size_t)coarse::kMaxTiles * Rp_);
    alloc_scal(sc_);
    alloc_scal(sc2_);
    d_src_ = ai(Rp_);
    d_sp64_ = ad(nv);
    d_spK_ = ad((size_t)E * Rp_);
    FL_HIP(hipHostMalloc((void **)&h_count_, 2 * nchunk_ * sizeof(int), hipHostMallocDefault));
    FL_HIP(hipStreamSynchronize(s));
}

void DarcyBatchGPU::alloc32() {
    if (d_x32_) return;
    const size_t nv = (size_t)F_ * Rp_;
    auto af = [&](size_t n) { float *p = dev_alloc<float>(n); owned_.push_back(p); return p; };
    d_x32_ = af(nv); d_r32_ = af(nv); d_z32_ = af(nv); d_p32a_ = af(nv); d_p32b_ = af(nv); d_q32_ = af(nv);
    d_sp32_ = af(nv);
    d_sp16_ = dev_alloc<unsigned short>(nv);
    owned_.push_back(d_sp16_);
    d_spK32_ = af((size_t)E_ * Rp_);
}

void DarcyBatchGPU::alloc_scal(BatchScal &sc) {
    double *scd = dev_alloc<double>(9 * (size_t)Rp_);
    int *sci = dev_alloc<int>(7 * (size_t)Rp_ + 2 * nchunk_);
    owned_.push_back(scd);
    owned_.push_back(sci);
    sc.alpha = scd; sc.beta = scd + Rp_; sc.rz = scd + 2 * Rp_; sc.rz_prev = scd + 3 * Rp_;
    sc.rr = scd + 4 * Rp_; sc.tol2 = scd + 5 * Rp_; sc.bnorm2 = scd + 6 * Rp_; sc.rr_hist = scd + 7 * Rp_;
    sc.rr_ru = scd + 8 * Rp_;
    sc.status = sci; sc.it = sci + Rp_; sc.n_ru = sci + 2 * Rp_; sc.par = sci + 3 * Rp_;
    sc.it_ru = sci + 4 * Rp_;
    sc.zp = sci + 5 * Rp_;
    sc.active_count = sci + 7 * Rp_;
    sc.ru_drop2 = kRuDrop2;
}

void DarcyBatchGPU::permute_slots(const std::vector<int> &src, int phase) {
    FL_HIP(hipMemcpyAsync(d_src_, src.data(), Rp_ * sizeof(int), hipMemcpyHostToDevice, stream_));
    const dim3 blk(256);
    const dim3 gF(std::min((F_ + kWaves - 1) / kWaves, 2048), nchunk_), gE(std::min((E_ + kWaves - 1) / kWaves, 2048), nchunk_);
    auto perm64 = [&](double *&a) {
        hipLaunchKernelGGL(k_b_perm<double>, gF, blk, 0, stream_, (const double *)a, d_sp64_, (size_t)F_,
                           (size_t)Rp_, (const int *)d_src_);
        std::swap(a, d_sp64_);
    };
    auto perm32 = [&](float *&a) {
        hipLaunchKernelGGL(k_b_perm<float>, gF, blk, 0, stream_, (const float *)a, d_sp32_, (size_t)F_,
                           (size_t)Rp_, (const int *)d_src_);
        std::swap(a, d_sp32_);
    };
    perm64(d_x_);
    perm64(d_b_);
    perm64(d_dinv_);
    hipLaunchKernelGGL(k_b_perm<double>, gE, blk, 0, stream_, (const double *)d_K_, d_spK_, (size_t)E_, (size_t)Rp_,
                       (const int *)d_src_);
    std::swap(d_K_, d_spK_);
    if (phase == 0) {
        perm64(d_r_);
        perm64(d_z_);
        hipLaunchKernelGGL(k_b_perm_p<double>, gF, blk, 0, stream_, (const double *)d_p0_, (const double *)d_p1_,
                           d_sp64_, (size_t)F_, (size_t)Rp_, (const int *)d_src_, (const int *)sc_.par);
        std::swap(d_p0_, d_sp64_);
    } else {
        perm32(d_x32_);
        perm32(d_r32_);
        perm32(d_z32_);
        hipLaunchKernelGGL(k_b_perm<unsigned short>, gF, blk, 0, stream_, (const unsigned short *)d_dinv32_, d_sp16_,
                           (size_t)F_, (size_t)Rp_, (const int *)d_src_);
        std::swap(d_dinv32_, d_sp16_);
        hipLaunchKernelGGL(k_b_perm_p<float>, gF, blk, 0, stream_, (const float *)d_p32a_, (const float *)d_p32b_,
                           d_sp32_, (size_t)F_, (size_t)Rp_, (const int *)d_src_, (const int *)sc_.par);
        std::swap(d_p32a_, d_sp32_);
        hipLaunchKernelGGL(k_b_perm<float>, gE, blk, 0, stream_, (const float *)d_K32_, d_spK32_, (size_t)E_,
                           (size_t)Rp_, (const int *)d_src_);
        std::swap(d_K32_, d_spK32_);
    }
    if (ncol_) {
        hipLaunchKernelGGL(k_b_perm<double>, dim3(std::min(nc_, 2048), nchunk_), blk, 0, stream_, (const double *)d_y_,
                           d_spy_, (size_t)nc_, (size_t)Rp_, (const int *)d_src_);
        std::swap(d_y_, d_spy_);
        hipLaunchKernelGGL(k_b_perm_major<float>, dim3(1024, Rp_), blk, 0, stream_, (const float *)d_ainv_,
                           d_spainv_, (size_t)coarse::ntiles_lower(nt_) * coarse::kT * coarse::kT,
                           (const int *)d_src_);
        std::swap(d_ainv_, d_spainv_);
        if (ncol_ == 2) {
            hipLaunchKernelGGL(k_b_perm<unsigned short>, gF, blk, 0, stream_, (const unsigned short *)d_phi_, d_spphi_,
                               (size_t)F_,
                               (size_t)Rp_, (const int *)d_src_);
            std::swap(d_phi_, d_spphi_);
        }
    }
    hipLaunchKernelGGL(k_b_perm_scal, dim3((Rp_ + 255) / 256), blk, 0, stream_, sc_, sc2_, (const int *)d_src_, Rp_);
    std::swap(sc_, sc2_);
    FL_HIP(hipGetLastError());
    std::vector<int> nr(Rp_);
    for (int s = 0; s < Rp_; s++) nr[s] = rid_[src[s]];
    rid_.swap(nr);
    free_graphs();
    FL_HIP(hipStreamSynchronize(stream_));  // src staging buffer is reused by the next call
}

void DarcyBatchGPU::set_pairs(const std::vector<int> &pairs) {
    npairs_ = (int)pairs.size() / 2;
    std::vector<int> sel(Rp_, 0);
    for (int k = 0; k < npairs_; k++) sel[pairs[2 * k]] = 1;
    if (npairs_) FL_HIP(hipMemcpyAsync(d_pairs_, pairs.data(), pairs.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipMemcpyAsync(d_lsel_, sel.data(), Rp_ * sizeof(int), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));  // host staging vectors go out of scope
}

__global__ void k_c_phi_rows(const float *__restrict__ src, const int *__restrict__ perm, int F, int R,
                             float *__restrict__ dst) {
    const int r = blockIdx.y;
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < F; row += gridDim.x * blockDim.x)
        dst[(size_t)r * F + row] = src[(size_t)r * F + perm[row]];
}

void DarcyBatchGPU::set_coarse(int ncol) {
    if (ncol < 0 || ncol > 2) throw Error(-1, "coarse columns must be 0, 1 or 2");
    if (ncol > 0 && !coarse_ok_)
        throw Error(-3, "aggregation not usable on this mesh (more than 32 neighbour aggregates or 1024 aggregates)");
    if (ncol == 2 && !have_phi_) throw Error(-4, "coarse basis phi not set");
    if (ncol != ncol_) free_graphs();
    ncol_ = ncol;
    if (ncol == 0) return;
    const size_t nv = (size_t)F_ * Rp_;
    auto ad = [&](size_t n) { double *p = dev_alloc<double>(n); owned_.push_back(p); return p; };
    auto af = [&](size_t n) { float *p = dev_alloc<float>(n); owned_.push_back(p); return p; };
    // coarse order padded to 64-row tiles (two columns: the largest)
    const int ntmax = (2 * nagg_ + coarse::kT - 1) / coarse::kT, ncpmax = ntmax * coarse::kT;
    const size_t tiles = (size_t)coarse::ntiles_lower(ntmax) * coarse::kT * coarse::kT;
    if (!d_rc_) {
        d_rc_ = ad((size_t)ncpmax * Rp_);
        d_y_ = ad((size_t)ncpmax * Rp_);
        d_spy_ = ad((size_t)ncpmax * Rp_);
        d_Asp_ = ad((size_t)nagg_ * kNbMax * 4 * Rp_);
        d_Apart_ = ad((size_t)std::max(nasm_chunks_, 1) * 4 * Rp_);
        d_ainv_ = af((size_t)Rp_ * tiles);
        d_spainv_ = af((size_t)Rp_ * tiles);
        d_Apack_ = ad((size_t)Rp_ * ncpmax * ncpmax);
        d_Wpack_ = ad((size_t)Rp_ * ncpmax * ncpmax);
        d_py_ = ad((size_t)Rp_ * ntmax * ntmax * coarse::kT);
        int *inf = dev_alloc<int>((size_t)Rp_);
        owned_.push_back(inf);
        d_info_ = inf;
    }
    nc_ = ncol * nagg_;
    nt_ = (nc_ + coarse::kT - 1) / coarse::kT;
    // padding rows of Z^T r stay zero
    FL_HIP(hipMemsetAsync(d_rc_, 0, (size_t)ncpmax * Rp_ * sizeof(double), stream_));
    if (ncol == 2 && !d_phi_) {
        auto ah = [&](size_t n) { unsigned short *p = dev_alloc<unsigned short>(n); owned_.push_back(p); return p; };
        d_phi_ = ah(nv);
        d_spphi_ = ah(nv);
    }
}

void DarcyBatchGPU::set_coarse_basis(const float *d_phi_edges) {
    if (!coarse_ok_) throw Error(-3, "aggregation not usable on this mesh");
    if (!d_phiAll_) {
        d_phiAll_ = dev_alloc<float>((size_t)R_ * F_);
        owned_.push_back(d_phiAll_);
    }
    hipLaunchKernelGGL(k_c_phi_rows, dim3(std::min(grid_for(F_), 1024), R_), dim3(kBlock), 0, stream_, d_phi_edges,
                       (const int *)d_perm_, F_, R_, d_phiAll_);
    FL_HIP(hipGetLastError());
    FL_HIP(hipStreamSynchronize(stream_));
    have_phi_ = true;
}

// Coarse operators of the listed (slot, realization) pairs: phi columns,
// Z^T S Z by aggregate blocks, batched Cholesky + inverse (coarse.hh), packed
// FP32 inverse tiles.
void DarcyBatchGPU::coarse_setup(const std::vector<int> &pairs) {
    const int np = (int)pairs.size() / 2;
    if (!np) return;
    const dim3 gagg(nagg_, nch_), blk(256);
    CArgs c{ncol_, nagg_, nc_, nt_, d_agg_, d_aptr_, d_nbptr_, d_nb_, d_lbmap_, d_phi_, d_ainv_, d_rc_, d_y_};
    if (ncol_ == 2) {
        hipLaunchKernelGGL(k_b_slot_col, dim3((F_ + 63) / 64, (np + 63) / 64), blk, 0, stream_,
                           (const float *)d_phiAll_, F_, Rp_, (const int *)d_pairs_, np, d_phi_);
        hipLaunchKernelGGL(k_c_phi_norm, gagg, dim3(64), 0, stream_, c, (size_t)Rp_, (const int *)d_lsel_, d_phi_);
    }
    BArgs<double> a = make_args<double>(F_, Rp_, grid_, xcd_map_, d_rec_, d_row_ptr_, d_row_c_, d_meta_, d_op_off_,
                                        d_conn_off_, d_conn_, d_tab_, d_K_, d_tab_, d_K_, d_dinv_, d_x_, d_r_, d_z_,
                                        d_p0_, d_p1_, d_q_, d_part_, sc_);
    if (ncol_ == 1) {
        if (nasm_chunks_ > 0)
            hipLaunchKernelGGL(k_c_assemble<1>, dim3(nasm_chunks_, nch_), dim3(64), 0, stream_, a, c,
                               (const int *)d_lsel_, (const int *)d_cbeg_, (const int *)d_cblk_, (const int *)d_tptr_,
                               (const int2 *)d_terms_, d_Apart_);
        hipLaunchKernelGGL(k_c_assemble_sum<1>, dim3((int)h_nbv_.size(), nch_), dim3(64), 0, stream_, c, (size_t)Rp_,
                           (const int *)d_lsel_, (const int *)d_cfirst_, (const int *)d_tagg_, (const double *)d_Apart_,
                           d_Asp_);
        hipLaunchKernelGGL(k_c_pack<1>, dim3(nt_ * coarse::kT, np), blk, 0, stream_, c, (const double *)d_Asp_, (size_t)Rp_,
                           (const int *)d_pairs_, d_Apack_);
    } else {
        if (nasm_chunks_ > 0)
            hipLaunchKernelGGL(k_c_assemble<2>, dim3(nasm_chunks_, nch_), dim3(64), 0, stream_, a, c,
                               (const int *)d_lsel_, (const int *)d_cbeg_, (const int *)d_cblk_, (const int *)d_tptr_,
                               (const int2 *)d_terms_, d_Apart_);
        hipLaunchKernelGGL(k_c_assemble_sum<2>, dim3((int)h_nbv_.size(), nch_), dim3(64), 0, stream_, c, (size_t)Rp_,
                           (const int *)d_lsel_, (const int *)d_cfirst_, (const int *)d_tagg_, (const double *)d_Apart_,
                           d_Asp_);
        hipLaunchKernelGGL(k_c_pack<2>, dim3(nt_ * coarse::kT, np), blk, 0, stream_, c, (const double *)d_Asp_, (size_t)Rp_,
                           (const int *)d_pairs_, d_Apack_);
    }
    FL_HIP(hipGetLastError());
    coarse::invert_batched(d_Apack_, d_Wpack_, nt_ * coarse::kT, np, d_info_, (const int *)d_pairs_, d_ainv_,
                           stream_);
}

void coarse_invert_host(const double *a, int n, int np, float *inv, int *info) {
    if (n < 1 || n > coarse::kMaxTiles * coarse::kT || np < 1) throw Error(-1, "coarse order must be in [1, 2048]");
    const int nt = (n + coarse::kT - 1) / coarse::kT, ncp = nt * coarse::kT, ntl = coarse::ntiles_lower(nt);
    const size_t tile = (size_t)coarse::kT * coarse::kT;
    std::vector<double> A((size_t)np * ncp * ncp, 0.0);
    for (int k = 0; k < np; k++)
        for (int i = 0; i < ncp; i++)
            for (int j = 0; j < ncp; j++)
                A[((size_t)k * ncp + i) * ncp + j] =
                    i < n && j < n ? a[((size_t)k * n + i) * n + j] : (i == j ? 1.0 : 0.0);
    std::vector<int> pairs(2 * (size_t)np);
    for (int k = 0; k < np; k++) pairs[2 * k] = pairs[2 * k + 1] = k;
    hipStream_t s;
    FL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<void *> own;
    try {
        double *dA = dev_upload(A.data(), A.size(), s);
        own.push_back(dA);
        double *dW = dev_alloc<double>(A.size());
        own.push_back(dW);
        int *dp = dev_upload(pairs.data(), pairs.size(), s), *dinfo = dev_alloc<int>(np);
        own.push_back(dp);
        own.push_back(dinfo);
        float *dinv = dev_alloc<float>((size_t)np * ntl * tile);
        own.push_back(dinv);
        coarse::invert_batched(dA, dW, ncp, np, dinfo, dp, dinv, s);
        std::vector<float> t((size_t)np * ntl * tile);
        FL_HIP(hipMemcpyAsync(t.data(), dinv, t.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        FL_HIP(hipMemcpyAsync(info, dinfo, np * sizeof(int), hipMemcpyDeviceToHost, s));
        FL_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < np; k++)
            for (int i = 0; i < n; i++)
                for (int j = 0; j <= i; j++) {
                    const float v = t[((size_t)k * ntl + coarse::tile_index(i / 64, j / 64)) * tile + (i % 64) * 64 +
                                      j % 64];
                    inv[((size_t)k * n + i) * n + j] = v;
                    inv[((size_t)k * n + j) * n + i] = v;
                }
    } catch (...) {
        for (void *p : own) dev_free(p);
        (void)hipStreamDestroy(s);
        throw;
    }
    for (void *p : own) dev_free(p);
    FL_HIP(hipStreamDestroy(s));
}

void DarcyBatchGPU::set_launch(int grid, int xcd_map, int chunk) {
    if (grid < 8 || grid > kMaxLaunchGrid || grid % 8)
        throw Error(-1, "batch grid must be a multiple of 8 in [8, 2048]");
    if (chunk < 0 || chunk > 1024) throw Error(-1, "batch chunk must be in [0, 1024] (0: automatic)");
    if (grid != grid_ || xcd_map != xcd_map_ || chunk != chunk_) free_graphs();
    if (xcd_map < 0 || xcd_map > 2) throw Error(-1, "batch row mapping must be 0, 1 or 2");
    grid_ = grid;
    xcd_map_ = xcd_map;
    chunk_ = chunk;
    tuned_ = true;
}

void DarcyBatchGPU::set_conductivity(const double *K, bool device_ptr) {
    have_K_ = false;
    FL_HIP(hipMemcpyAsync(d_Kall_, K, (size_t)R_ * E_ * sizeof(double),
                          device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
    const long i = first_not_positive_device(d_Kall_, (long)R_ * E_, base_.check_flag(), stream_);
    if (i >= 0) {
        double x = 0.0;
        FL_HIP(hipMemcpy(&x, d_Kall_ + i, sizeof(double), hipMemcpyDeviceToHost));
        char buf[64];
        snprintf(buf, sizeof buf, "%.17g", x);
        throw Error(kErrBadField, "conductivity of realization " + std::to_string(i / E_) + ", element " +
                                      std::to_string(base_.mesh().el_id[i % E_]) + " is not finite and positive: " +
                                      buf);
    }
    have_K_ = true;
}

void DarcyBatchGPU::setup_system() {
    const DarcyDev &d = dv_;
    for (int D = 1; D <= 3; D++) {
        int n = d.dim_start[D] - d.dim_start[D - 1];
        if (n <= 0) continue;
        dim3 g(grid_for(n)), b(kBlock);
        if (D == 1) hipLaunchKernelGGL(k_b_table<1>, g, b, 0, stream_, d, d_meta_, d_op_off_, d_rs_off_, d_tab_, d_rsT_);
        if (D == 2) hipLaunchKernelGGL(k_b_table<2>, g, b, 0, stream_, d, d_meta_, d_op_off_, d_rs_off_, d_tab_, d_rsT_);
        if (D == 3) hipLaunchKernelGGL(k_b_table<3>, g, b, 0, stream_, d, d_meta_, d_op_off_, d_rs_off_, d_tab_, d_rsT_);
    }
    hipLaunchKernelGGL(k_b_to_float, dim3(std::min(grid_for(op_total_), 4096)), dim3(kBlock), 0, stream_,
                       (size_t)op_total_, (const double *)d_tab_, d_tab32_);
    hipLaunchKernelGGL(k_b_fastrec, dim3(grid_for(F_)), dim3(kBlock), 0, stream_, F_, (const int *)d_rec_,
                       (const float *)d_tab32_, d_frec_);
    FL_HIP(hipGetLastError());
    FL_HIP(hipStreamSynchronize(stream_));
    base_.check_singular();
}


BatchStats DarcyBatchGPU::solve(double rtol, double atol, int max_it, int precision, int norm) {
    if (!have_K_) throw Error(-4, "batch conductivities not set");
    if (precision != 0 && precision != 1) throw Error(-1, "batch precision must be 0 (FP64) or 1 (FP32 + FP64)");
    if (max_it < 1) throw Error(-1, "max_it must be positive");
    if (precision == 1) alloc32();
    BatchStats st;
    st.n_real = R_;
    auto t_start = std::chrono::steady_clock::now();
    auto since_start = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    };
    hipEvent_t e0, e1;
    FL_HIP(hipEventCreate(&e0));
    FL_HIP(hipEventCreate(&e1));
    FL_HIP(hipEventRecord(e0, stream_));
    setup_system();
    FL_HIP(hipEventRecord(e1, stream_));
    const dim3 blk(256), bfold(kFoldThreads), blkA(kAggWaves * 64);
    const double rtol2 = rtol * rtol, atol2 = atol * atol;
    nch_ = nchunk_;  // chunks of 64 slots that hold live realizations (compaction shrinks it)
    lw_ = 64;        // lanes per row of the two-level kernels (narrowed for the tail)
    // FP64 runs no residual replacement: only candidates are checked
    sc_.ru_drop2 = sc2_.ru_drop2 = precision == 1 ? kRuDrop2 : 0.0;
    if (norm != 0 && norm != 1) throw Error(-1, "stop norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
    sc_.nat = sc2_.nat = norm;
    BArgs<double> a64;
    BArgs<float> a32;
    CArgs ca{};
    int G = grid_, X = xcd_map_;
    const int NC = ncol_;  // two-level preconditioner columns (0: Jacobi)
    // iterations per captured chunk (automatic: 64 two-level, 16 one-level).
    // Every chunk ends with the true residual of its running realizations;
    // on the two-level path 64 beat 32 (the residual pass costs about one
    // operator's worth of FP64 gathers) and 96 gained nothing more.
    const int CH = chunk_ > 0 ? chunk_ : (NC ? 64 : 16);
    auto rebuild = [&]() {
        G = grid_;
        X = xcd_map_;
        ca = CArgs{ncol_, nagg_, nc_, nt_, d_agg_, d_aptr_, d_nbptr_, d_nb_, d_lbmap_, d_phi_, d_ainv_, d_rc_, d_y_};
        a64 = make_args<double>(F_, Rp_, G, X, d_rec_, d_row_ptr_, d_row_c_, d_meta_, d_op_off_, d_conn_off_, d_conn_,
                                d_tab_, d_K_, d_tab_, d_K_, d_dinv_, d_x_, d_r_, d_z_, d_p0_, d_p1_, d_q_, d_part_,
                                sc_);
        a32 = make_args<float>(F_, Rp_, G, X, d_rec_, d_row_ptr_, d_row_c_, d_meta_, d_op_off_, d_conn_off_, d_conn_,
                               d_tab32_, d_K32_, d_tab_, d_K_, d_dinv32_, d_x32_, d_r32_, d_z32_, d_p32a_, d_p32b_,
                               d_q32_, d_part_, sc_);
        a32.frec = d_frec_;
    };
    rebuild();
    auto grow = [&]() { return dim3(G, nch_); };
    auto fold = [&](int mode, int mixed, const int *lsel, int gparts = 0) {
        const double *q0 = a64.part0, *q1 = a64.part1, *q2 = a64.part2;
        const dim3 gf(nch_);
        int Gf = gparts > 0 ? gparts : G;
        if (Gf > 2 * kPre) {  // two-stage: kPre partial rows per slot for the fold proper
            double *o0 = d_pre_, *o1 = d_pre_ + (size_t)kPre * Rp_, *o2 = d_pre_ + 2 * (size_t)kPre * Rp_;
            hipLaunchKernelGGL(k_b_prefold, dim3(kPre, nch_), blk, 0, stream_, mode == kFoldAlpha ? q0 : nullptr,
                               mode != kFoldAlpha ? q1 : nullptr,
                               mode != kFoldAlpha ? q2 : nullptr, Gf, Rp_, o0, o1, o2);
            q0 = o0;
            q1 = o1;
            q2 = o2;
            Gf = kPre;
        }
#define FL_FOLD(M)                                                                                                  \
    hipLaunchKernelGGL(k_b_fold<M>, gf, bfold, 0, stream_, sc_, q0, q1, q2, Gf, Rp_, R_, rtol2, atol2, max_it, mixed, \
                       lsel)
        switch (mode) {
            case kFoldInit: FL_FOLD(kFoldInit); break;
            case kFoldAlpha: FL_FOLD(kFoldAlpha); break;
            case kFoldBeta: FL_FOLD(kFoldBeta); break;
            case kFoldRU: FL_FOLD(kFoldRU); break;
            case kFoldWarm: FL_FOLD(kFoldWarm); break;
            default: FL_FOLD(kFoldFinal); break;
        }
#undef FL_FOLD
    };
    const int *nosel = nullptr;
    auto gagg = [&]() { return dim3(nagg_, nch_); };
    auto cfold = [&](int mode, int mixed, const int *lsel) {
        const dim3 gt((coarse::ntiles_lower(nt_) + kSolveTiles - 1) / kSolveTiles, 64 * nch_), gy(nt_, 64 * nch_),
            gfin(nch_), b64(64);
        const double *p1 = a64.part1, *p2 = a64.part2;
#define FL_CFOLD(M)                                                                                                  \
    hipLaunchKernelGGL(k_c_solve_t<M>, gt, blk, 0, stream_, sc_, ca, (size_t)Rp_, lsel, d_py_);                      \
    hipLaunchKernelGGL(k_c_solve_y<M>, gy, b64, 0, stream_, sc_, ca, p1, p2, (size_t)Rp_, lsel,                     \
                       (const double *)d_py_, d_cpart_);                                                             \
    hipLaunchKernelGGL(k_c_fin<M>, gfin, b64, 0, stream_, sc_, (const double *)d_cpart_, nt_, (size_t)Rp_, rtol2,     \
                       atol2, max_it, mixed, lsel)
        if (mode == kFoldInit) {
            FL_CFOLD(kFoldInit);
        } else if (mode == kFoldBeta) {
            FL_CFOLD(kFoldBeta);
        } else {
            FL_CFOLD(kFoldRU);
        }
#undef FL_CFOLD
    };
    // two-level kernels at the current lanes per row
#define FL_LW(...)                      \
    switch (lw_) {                      \
        case 16: {                      \
            constexpr int LW = 16;      \
            __VA_ARGS__;                \
        } break;                        \
        case 4: {                       \
            constexpr int LW = 4;       \
            __VA_ARGS__;                \
        } break;                        \
        default: {                      \
            constexpr int LW = 64;      \
            __VA_ARGS__;                \
        }                               \
    }
    auto zp64 = [&]() {
        if (NC == 1) FL_LW(hipLaunchKernelGGL((k_c_zp<double, 1, LW>), grow(), blk, 0, stream_, a64, ca))
        else FL_LW(hipLaunchKernelGGL((k_c_zp<double, 2, LW>), grow(), blk, 0, stream_, a64, ca))
    };
    // the FP32 operator: k_c_opg (general rows) on the second stream beside
    // k_c_opl (row tiles); the fork and join are captured into the graphs
    auto op_tiled_grid = [&]() { return gen_blocks_ + ntiles_; };
    auto op_tiled = [&](const BArgs<float> &a) {
        const OpTiles ot{d_gen_, d_grec_, d_thptr_, d_thalo_, d_teptr_, d_tels_, d_lrec_, ngen_, gen_blocks_,
                         ntiles_, tile_T_, nseg_max_};
        if (ngen_) {
            FL_HIP(hipEventRecord(ev_fork_, stream_));
            FL_HIP(hipStreamWaitEvent(stream2_, ev_fork_, 0));
            if (nch_ == 1)
                hipLaunchKernelGGL(k_c_opg<1>, dim3(gen_blocks_), blk, 0, stream2_, a, ot, ot.gen, ot.grec, a.opT,
                                   a.opT64);
            else
                hipLaunchKernelGGL(k_c_opg<2>, dim3(gen_blocks_), blk, 0, stream2_, a, ot, ot.gen, ot.grec, a.opT,
                                   a.opT64);
        }
        hipLaunchKernelGGL(k_c_opl, dim3(ntiles_), dim3(kOplThreads), opl_lds_bytes(nseg_max_, tile_T_), stream_, a,
                           ot, nch_, lw_ == 64 ? 2 * nch_ : 1);
        if (ngen_) {
            FL_HIP(hipEventRecord(ev_join_, stream2_));
            FL_HIP(hipStreamWaitEvent(stream_, ev_join_, 0));
        }
    };
    auto zp32 = [&]() {
        if (NC == 1) FL_LW(hipLaunchKernelGGL((k_c_zp<float, 1, LW>), grow(), blk, 0, stream_, a32, ca))
        else FL_LW(hipLaunchKernelGGL((k_c_zp<float, 2, LW>), grow(), blk, 0, stream_, a32, ca))
    };
    // one CG iteration; ev: optional events around the operator, the update and
    // the fold (with the coarse solve and the direction kernel)
    auto iter64 = [&](hipEvent_t *ev) {
        if (ev) FL_HIP(hipEventRecord(ev[0], stream_));
        if (NC) FL_LW(hipLaunchKernelGGL((k_c_op<double, LW, 1>), grow(), blk, 0, stream_, a64, ca))
        else hipLaunchKernelGGL(k_b_op<double>, grow(), blk, 0, stream_, a64);
        if (ev) FL_HIP(hipEventRecord(ev[1], stream_));
        fold(kFoldAlpha, 0, nosel);
        if (ev) FL_HIP(hipEventRecord(ev[2], stream_));
        if (NC == 1) { if (lw_ == 64) hipLaunchKernelGGL((k_c_update<double, 1, 64, 2>), gagg(), blkA, 0, stream_, a64, ca); else FL_LW(hipLaunchKernelGGL((k_c_update<double, 1, LW, 1>), gagg(), blkA, 0, stream_, a64, ca)) }
        else if (NC == 2) { if (lw_ == 64) hipLaunchKernelGGL((k_c_update<double, 2, 64, 2>), gagg(), blkA, 0, stream_, a64, ca); else FL_LW(hipLaunchKernelGGL((k_c_update<double, 2, LW, 1>), gagg(), blkA, 0, stream_, a64, ca)) }
        else hipLaunchKernelGGL(k_b_update<double>, grow(), blk, 0, stream_, a64);
        if (ev) FL_HIP(hipEventRecord(ev[3], stream_));
        if (NC) {
            cfold(kFoldBeta, 0, nosel);
            zp64();
        } else {
            fold(kFoldBeta, 0, nosel);
        }
        if (ev) FL_HIP(hipEventRecord(ev[4], stream_));
    };
    auto iter32 = [&](hipEvent_t *ev) {
        if (ev) FL_HIP(hipEventRecord(ev[0], stream_));
        if (NC && lw_ == 4)
            hipLaunchKernelGGL(k_c_opn<4>, grow(), blk, 0, stream_, a32, ca);
        else if (NC)
            op_tiled(a32);  // a tail narrowed to 16 slots computes one 32-slot group
        else
            hipLaunchKernelGGL(k_b_op<float>, grow(), blk, 0, stream_, a32);
        if (ev) FL_HIP(hipEventRecord(ev[1], stream_));
        fold(kFoldAlpha, 1, nosel, NC && lw_ != 4 ? op_tiled_grid() : 0);
        if (ev) FL_HIP(hipEventRecord(ev[2], stream_));
        if (NC == 1) { if (lw_ == 64) hipLaunchKernelGGL((k_c_update<float, 1, 64, 2>), gagg(), blkA, 0, stream_, a32, ca); else FL_LW(hipLaunchKernelGGL((k_c_update<float, 1, LW, 1>), gagg(), blkA, 0, stream_, a32, ca)) }
        else if (NC == 2) { if (lw_ == 64) hipLaunchKernelGGL((k_c_update<float, 2, 64, 2>), gagg(), blkA, 0, stream_, a32, ca); else FL_LW(hipLaunchKernelGGL((k_c_update<float, 2, LW, 1>), gagg(), blkA, 0, stream_, a32, ca)) }
        else hipLaunchKernelGGL(k_b_update<float>, grow(), blk, 0, stream_, a32);
        if (ev) FL_HIP(hipEventRecord(ev[3], stream_));
        if (NC) {
            cfold(kFoldBeta, 1, nosel);
            zp32();
        } else {
            fold(kFoldBeta, 1, nosel);
        }
        if (ev) FL_HIP(hipEventRecord(ev[4], stream_));
    };
    auto reliable_update = [&]() {
        if (NC) {
            FL_LW(hipLaunchKernelGGL(k_b_accum<LW>, grow(), blk, 0, stream_, F_, G, X, (size_t)Rp_, sc_, nosel, d_x_,
                                     d_x32_))
            if (NC == 1)
                FL_LW(hipLaunchKernelGGL((k_c_resid<float, 1, LW>), gagg(), blkA, 0, stream_, a32, ca,
                                         (const double *)d_b_, (const double *)d_x_))
            else
                FL_LW(hipLaunchKernelGGL((k_c_resid<float, 2, LW>), gagg(), blkA, 0, stream_, a32, ca,
                                         (const double *)d_b_, (const double *)d_x_))
            cfold(kFoldRU, 1, nosel);
            zp32();
            return;
        }
        hipLaunchKernelGGL(k_b_accum<64>, grow(), blk, 0, stream_, F_, G, X, (size_t)Rp_, sc_, nosel, d_x_, d_x32_);
        hipLaunchKernelGGL((k_b_resid<float, 0>), grow(), blk, 0, stream_, a32, -1, nosel, (const int *)d_op_off_,
                           (const double *)d_tab_, (const double *)d_K_, (const double *)d_b_,
                           (const double *)d_x_);
        fold(kFoldRU, 1, nosel);
    };
    // two-level FP64: realizations whose recursive residual met the tolerance
    // are confirmed (or resumed) on the true residual b - S x
    auto candidate_check64 = [&]() {
        if (NC == 1)
            FL_LW(hipLaunchKernelGGL((k_c_resid<double, 1, LW>), gagg(), blkA, 0, stream_, a64, ca,
                                     (const double *)d_b_, (const double *)d_x_))
        else
            FL_LW(hipLaunchKernelGGL((k_c_resid<double, 2, LW>), gagg(), blkA, 0, stream_, a64, ca,
                                     (const double *)d_b_, (const double *)d_x_))
        cfold(kFoldRU, 0, nosel);
        zp64();
    };
    // graph of `chunk_` iterations (kind 0: FP64; kind 1: FP32 + reliable update),
    // re-captured when the launch width or the iteration limit changes
    auto graph = [&](int kind) -> hipGraphExec_t {
        hipGraphExec_t &ex = kind ? g32_ : g64_;
        long long &key = kind ? g32_key_ : g64_key_;
        const long long want = ((long long)max_it << 24) | ((long long)nch_ << 8) | lw_;
        if (ex && key == want) return ex;
        if (ex) (void)hipGraphExecDestroy(ex);
        ex = nullptr;
        hipGraph_t gr;
        FL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < CH; k++) {
            if (kind) iter32(nullptr);
            else iter64(nullptr);
        }
        if (kind) reliable_update();
        else if (NC) candidate_check64();
        FL_HIP(hipStreamEndCapture(stream_, &gr));
        FL_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        key = want;
        return ex;
    };
    bool ksaved = false;  // the base solver's K, borrowed by hand-overs, is restored at the end
    // per-realization results
    h_it_.assign(R_, 0);
    h_status_.assign(R_, kIdle);
    h_ru_.assign(R_, 0);
    h_rr_.assign(R_, 0.0);
    h_bb_.assign(R_, 0.0);
    h_tol_.assign(R_, 0.0);
    std::vector<int> stv(Rp_), itv(Rp_), ruv(Rp_);
    std::vector<double> rrv(Rp_), bbv(Rp_), tolv(Rp_);
    auto read_status = [&]() {
        FL_HIP(hipMemcpyAsync(stv.data(), sc_.status, Rp_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
    };
    // Load realizations into empty slots: K, right-hand side, Jacobi diagonal,
    // cold-start vectors and scalars of those slots only.
    auto refill = [&](const std::vector<int> &pairs, int mixed) {
        if (pairs.empty()) return;
        set_pairs(pairs);
        const int np = (int)pairs.size() / 2;
        hipLaunchKernelGGL(k_b_slot_K, dim3((E_ + 63) / 64, (np + 63) / 64), blk, 0, stream_, (const double *)d_Kall_,
                           E_, Rp_, (const int *)d_pairs_, np, d_K_, d_K32_);
        hipLaunchKernelGGL(k_b_rhs, dim3(std::min((F_ + kWaves - 1) / kWaves, 4096), nch_), blk, 0, stream_, F_,
                           (size_t)Rp_, (const int *)d_row_ptr_, (const int *)d_row_c_, (const int *)d_meta_,
                           (const int *)d_rs_off_, (const double *)d_rsT_, (const double *)d_K_, d_b_, d_dinv_,
                           d_dinv32_, (const int *)d_lsel_);
        if (NC) {
            const auto tc = std::chrono::steady_clock::now();
            coarse_setup(pairs);
            FL_HIP(hipStreamSynchronize(stream_));
            st.ms_coarse_setup +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc).count();
            if (mixed && NC == 1)
                hipLaunchKernelGGL((k_c_init<float, 1>), gagg(), blkA, 0, stream_, a32, ca, (const double *)d_b_,
                                   d_x_, (const int *)d_lsel_);
            else if (mixed)
                hipLaunchKernelGGL((k_c_init<float, 2>), gagg(), blkA, 0, stream_, a32, ca, (const double *)d_b_,
                                   d_x_, (const int *)d_lsel_);
            else if (NC == 1)
                hipLaunchKernelGGL((k_c_init<double, 1>), gagg(), blkA, 0, stream_, a64, ca, (const double *)d_b_,
                                   (double *)nullptr, (const int *)d_lsel_);
            else
                hipLaunchKernelGGL((k_c_init<double, 2>), gagg(), blkA, 0, stream_, a64, ca, (const double *)d_b_,
                                   (double *)nullptr, (const int *)d_lsel_);
            cfold(kFoldInit, mixed, d_lsel_);
            if (mixed) zp32();
            else zp64();
        } else {
            if (mixed)
                hipLaunchKernelGGL(k_b_init<float>, grow(), blk, 0, stream_, a32, (const double *)d_b_, d_x_,
                                   (const int *)d_lsel_);
            else
                hipLaunchKernelGGL(k_b_init<double>, grow(), blk, 0, stream_, a64, (const double *)d_b_,
                                   (double *)nullptr, (const int *)d_lsel_);
            fold(kFoldInit, mixed, d_lsel_);
        }
        FL_HIP(hipGetLastError());
        for (int k = 0; k < np; k++) rid_[pairs[2 * k]] = pairs[2 * k + 1];
    };
    // Finished slots: final FP64 true residual, reconstruction into the
    // realization's output column, per-realization statistics; slot emptied.
    auto harvest = [&](const std::vector<int> &pairs, int mixed) {
        if (pairs.empty()) return;
        set_pairs(pairs);
        const int np = (int)pairs.size() / 2;
        if (mixed)
            hipLaunchKernelGGL(k_b_accum<64>, grow(), blk, 0, stream_, F_, G, X, (size_t)Rp_, sc_, (const int *)d_lsel_,
                               d_x_, d_x32_);
        hipLaunchKernelGGL((k_b_resid<double, 2>), grow(), blk, 0, stream_, a64, -2, (const int *)d_lsel_,
                           (const int *)d_op_off_, (const double *)d_tab_, (const double *)d_K_,
                           (const double *)d_b_, (const double *)d_x_);
        fold(kFoldFinal, 0, d_lsel_);
        double *lam = d_sol_ + ((size_t)S_ + E_) * Rout_;
        const dim3 gl(std::min((F_ + kWaves - 1) / kWaves, 4096), (np + 63) / 64);
        hipLaunchKernelGGL(k_b_scatter_lambda, gl, blk, 0, stream_, F_, (size_t)Rp_, (const int *)d_perm_,
                           (const double *)d_x_, lam, (const int *)d_pairs_, np, (size_t)Rout_);
        for (int D = 1; D <= 3; D++) {
            int n = dim_start_[D] - dim_start_[D - 1];
            if (n <= 0) continue;
            dim3 g((n + kWaves - 1) / kWaves, (np + 63) / 64);
#define FL_RESOLVE(DD)                                                                                        \
    hipLaunchKernelGGL(k_b_resolve<DD>, g, blk, 0, stream_, dim_start_[D - 1], dim_start_[D],                 \
                       (const int *)d_eperm_, dv_.side_off, S_, (size_t)Rp_, (const int *)d_meta_,            \
                       (const int *)d_rs_off_, (const double *)d_rsT_, (const int *)d_conn_off_,              \
                       (const int *)d_conn_, (const double *)d_K_, (const double *)d_x_, d_sol_,              \
                       (const int *)d_pairs_, np, (size_t)Rout_)
            if (D == 1) FL_RESOLVE(1);
            else if (D == 2) FL_RESOLVE(2);
            else FL_RESOLVE(3);
#undef FL_RESOLVE
        }
        FL_HIP(hipGetLastError());
        FL_HIP(hipMemcpyAsync(stv.data(), sc_.status, Rp_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(itv.data(), sc_.it, Rp_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(ruv.data(), sc_.n_ru, Rp_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(rrv.data(), sc_.rr, Rp_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(bbv.data(), sc_.bnorm2, Rp_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(tolv.data(), sc_.tol2, Rp_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        for (int k = 0; k < np; k++) {
            const int sl = pairs[2 * k], r = pairs[2 * k + 1];
            h_status_[r] = stv[sl];
            h_it_[r] = itv[sl];
            h_ru_[r] = ruv[sl];
            h_rr_[r] = rrv[sl];
            h_bb_[r] = bbv[sl];
            h_tol_[r] = tolv[sl];
            rid_[sl] = -1;
        }
    };
    // Hand realizations to the single-solve FP64 path, warm-started from the
    // batch iterate: a batched iteration costs a sweep over all rows whatever
    // the number of live lanes, a single solve's iteration touches one
    // realization only.  The single solves run in up to kHandWorkers worker
    // threads, each with a solver and stream of its own (the base solver and
    // copies of it on the same mesh, made by the workers themselves once and
    // kept), so the batch keeps iterating its other slots meanwhile and the
    // handed-over solves run side by side; their results are merged when the
    // batch is done.
    struct HandJob {
        int real, it0, n_ru;
        double bb, tol;
        double *x0;
    };
    struct HandResult {
        int real, it0, n_ru;
        double bb, tol;
        SolveStats ss;
    };
    std::mutex hmu;
    std::condition_variable hcv;
    std::deque<HandJob> hjobs;
    std::vector<HandResult> hres;
    std::vector<double *> hfree;  // extraction buffers ready for reuse
    bool hclose = false;
    std::exception_ptr herr;
    std::vector<std::thread> hworkers;
    double hbusy_ms = 0.0;
    int hdev = 0;
    FL_HIP(hipGetDevice(&hdev));
    auto hand_join = [&]() {
        if (hworkers.empty()) return;
        {
            std::lock_guard<std::mutex> lk(hmu);
            hclose = true;
        }
        hcv.notify_all();
        for (std::thread &t : hworkers)
            if (t.joinable()) t.join();
        hworkers.clear();
    };
    struct JoinOnExit {  // the worker never outlives this call, also on an exception
        std::function<void()> f;
        ~JoinOnExit() { f(); }
    } join_on_exit{hand_join};
    if (hand_solvers_.size() < kHandWorkers - 1) hand_solvers_.resize(kHandWorkers - 1);
    auto hand_worker = [&](int w) {
        (void)hipSetDevice(hdev);
        const size_t n = (size_t)S_ + E_ + F_;
        DarcyGPU *solver = &base_;
        if (w > 0) {
            try {
                std::unique_ptr<DarcyGPU> &h = hand_solvers_[w - 1];  // this worker's slot only
                if (!h) {
                    h = std::make_unique<DarcyGPU>(base_.mesh());
                    h->copy_data_from(base_);
                }
                solver = h.get();
            } catch (...) {
                std::lock_guard<std::mutex> lk(hmu);
                if (!herr) herr = std::current_exception();
                return;
            }
        }
        while (true) {
            HandJob j;
            {
                std::unique_lock<std::mutex> lk(hmu);
                hcv.wait(lk, [&] { return hclose || !hjobs.empty(); });
                if (hjobs.empty()) return;
                j = hjobs.front();
                hjobs.pop_front();
            }
            try {
                const auto t0 = std::chrono::steady_clock::now();
                solver->set_conductivity_device(d_Kall_ + (size_t)j.real * E_);
                const int left = std::max(1, max_it - j.it0);
                SolveStats ss = solver->solve_warm(j.x0, rtol, atol, left, j.bb, norm);
                hipLaunchKernelGGL(k_b_put_col, dim3(std::min(grid_for((long)n), 4096)), dim3(kBlock), 0,
                                   solver->stream(), n, (size_t)Rout_, j.real, solver->device_solution(), d_sol_);
                FL_HIP(hipGetLastError());
                FL_HIP(hipStreamSynchronize(solver->stream()));
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                std::lock_guard<std::mutex> lk(hmu);
                hres.push_back(HandResult{j.real, j.it0, j.n_ru, j.bb, j.tol, ss});
                hfree.push_back(j.x0);
                hbusy_ms += ms;
            } catch (...) {
                std::lock_guard<std::mutex> lk(hmu);
                if (!herr) herr = std::current_exception();
                hfree.push_back(j.x0);
            }
        }
    };
    auto handover = [&](const std::vector<int> &pairs, int mixed_now) {
        if (pairs.empty()) return;
        set_pairs(pairs);
        if (mixed_now)
            hipLaunchKernelGGL(k_b_accum<64>, grow(), blk, 0, stream_, F_, G, X, (size_t)Rp_, sc_, (const int *)d_lsel_,
                               d_x_, d_x32_);
        FL_HIP(hipMemcpyAsync(itv.data(), sc_.it, Rp_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(ruv.data(), sc_.n_ru, Rp_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(bbv.data(), sc_.bnorm2, Rp_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipMemcpyAsync(tolv.data(), sc_.tol2, Rp_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        if (!ksaved) {  // before the worker first changes the base solver's K
            FL_HIP(hipMemcpyAsync(d_Ksave_, base_.device_K(), E_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            FL_HIP(hipStreamSynchronize(stream_));
            ksaved = true;
        }
        std::vector<HandJob> js;
        for (size_t k = 0; k < pairs.size(); k += 2) {
            const int sl = pairs[k], r = pairs[k + 1];
            double *buf = nullptr;
            {
                std::lock_guard<std::mutex> lk(hmu);
                if (!hfree.empty()) {
                    buf = hfree.back();
                    hfree.pop_back();
                }
            }
            if (!buf) {
                buf = dev_alloc<double>((size_t)F_);
                owned_.push_back(buf);
            }
            hipLaunchKernelGGL(k_b_extract_x, dim3(std::min(grid_for(F_), 4096)), dim3(kBlock), 0, stream_, F_,
                               (size_t)Rp_, sl, (const int *)d_perm_, (const double *)d_x_, buf);
            hipLaunchKernelGGL(k_b_status_fill, dim3(1), dim3(64), 0, stream_, sc_.status + sl, 1, (int)kIdle);
            js.push_back(HandJob{r, itv[sl], ruv[sl], bbv[sl], tolv[sl], buf});
            rid_[sl] = -1;
            st.handovers++;
        }
        FL_HIP(hipGetLastError());
        FL_HIP(hipStreamSynchronize(stream_));  // extractions complete before the worker reads them
        size_t queued = 0;
        {
            std::lock_guard<std::mutex> lk(hmu);
            for (const HandJob &j : js) hjobs.push_back(j);
            queued = hjobs.size();
        }
        // one more worker per waiting job, up to kHandWorkers
        while (hworkers.size() < (size_t)kHandWorkers && hworkers.size() < queued + 1)
            hworkers.emplace_back(hand_worker, (int)hworkers.size());
        hcv.notify_all();
    };
    // move live slots to the front when that frees whole 64-slot chunks
    auto live = [&](int s) {
        return rid_[s] >= 0 && (stv[s] == kRun || stv[s] == kCandidate || stv[s] == kFp64);
    };
    auto maybe_compact = [&](int phase) {
        if (nch_ <= 1) return;
        std::vector<int> src;
        src.reserve(Rp_);
        for (int s = 0; s < Rp_; s++)
            if (live(s)) src.push_back(s);
        const int nn = std::max(1, ((int)src.size() + 63) / 64);
        if (nn >= nch_) return;
        for (int s = 0; s < Rp_; s++)
            if (!live(s)) src.push_back(s);
        permute_slots(src, phase);
        nch_ = nn;
        st.compactions++;
        rebuild();
        read_status();
    };
    // two-level tail: live realizations into slots [0, n), then 16 or 4 lanes
    // per row once n fits
    auto maybe_narrow = [&](int phase) {
        std::vector<int> src;
        src.reserve(Rp_);
        for (int s = 0; s < Rp_; s++)
            if (live(s)) src.push_back(s);
        const int n = (int)src.size();
        const int w = n <= 4 ? 4 : (n <= 16 ? 16 : 64);
        if (w >= lw_) return;
        (w == 4 ? st.ms_narrow4 : st.ms_narrow16) = since_start();
        for (int s = 0; s < Rp_; s++)
            if (!live(s)) src.push_back(s);
        permute_slots(src, phase);
        nch_ = 1;
        lw_ = w;
        st.compactions++;
        rebuild();
        read_status();
    };
    // queue: hardest realizations first (largest ln K variance), so the
    // longest solves start early and the final tail is short
    std::vector<int> queue(R_);
    {
        std::vector<double> var(R_);
        hipLaunchKernelGGL(k_b_logk_var, dim3(R_), blk, 0, stream_, (const double *)d_Kall_, E_, d_var_);
        FL_HIP(hipMemcpyAsync(var.data(), d_var_, R_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        std::iota(queue.begin(), queue.end(), 0);
        std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) { return var[a] > var[b]; });
    }
    size_t next = 0;
    rid_.assign(Rp_, -1);
    hipLaunchKernelGGL(k_b_status_fill, dim3((Rp_ + 255) / 256), blk, 0, stream_, sc_.status, Rp_, (int)kIdle);
    const int mixed = precision == 1 ? 1 : 0;
    {
        std::vector<int> pairs;
        for (int sl = 0; sl < Rp_ && next < queue.size(); sl++) {
            pairs.push_back(sl);
            pairs.push_back(queue[next++]);
        }
        refill(pairs, mixed);
    }
    if (!tuned_) {
        tune();  // times the operator on the filled slots; the CG state is unchanged
        rebuild();
    }
    // live kernel timing: the first iterations run eagerly between events
    const int kSample = 4;
    while ((int)ev_.size() < 5 * kSample) {
        hipEvent_t e;
        FL_HIP(hipEventCreate(&e));
        ev_.push_back(e);
    }
    for (int k = 0; k < kSample; k++) {
        if (mixed) iter32(&ev_[5 * k]);
        else iter64(&ev_[5 * k]);
    }
    if (mixed) reliable_update();
    {
        double so = 0, su = 0, sf = 0;
        FL_HIP(hipStreamSynchronize(stream_));
        for (int k = 0; k < kSample; k++) {
            float t1 = 0, t2 = 0, t3 = 0;
            FL_HIP(hipEventElapsedTime(&t1, ev_[5 * k], ev_[5 * k + 1]));
            FL_HIP(hipEventElapsedTime(&t2, ev_[5 * k + 2], ev_[5 * k + 3]));
            FL_HIP(hipEventElapsedTime(&t3, ev_[5 * k + 3], ev_[5 * k + 4]));
            so += t1;
            su += t2;
            sf += t3;
        }
        st.ms_op_avg = so / kSample;
        st.ms_update_avg = su / kSample;
        st.ms_coarse_avg = sf / kSample;
        st.op_samples = kSample;
    }
    // main loop: graph chunks; harvest finished slots, refill them from the
    // queue, compact once the queue is empty
    auto terminal = [&](int v) { return v == kConverged || v == kMaxIt || v == kBreakdown; };
    std::vector<int> free_slots;  // harvested slots waiting for a refill
    long guard = 0;
    const long guard_max = ((long)max_it + CH) * ((R_ + Rp_ - 1) / Rp_ + 2) / CH + 16;
    while (true) {
        read_status();
        std::vector<int> done, stalled;
        int nrun = 0, nf64 = 0;
        for (int sl = 0; sl < Rp_; sl++) {
            if (rid_[sl] < 0) continue;
            if (terminal(stv[sl])) {
                done.push_back(sl);
                done.push_back(rid_[sl]);
            } else if (stv[sl] == kFp64) {
                stalled.push_back(sl);
                stalled.push_back(rid_[sl]);
                nf64++;
            } else {
                nrun++;
            }
        }
        if (!stalled.empty()) {
            // the FP32 recursion of these stagnated: the single-solve FP64 path
            // finishes them from their current iterate, their slots are refilled
            st.fp64_finish += nf64;
            handover(stalled, mixed);
            for (size_t k = 0; k < stalled.size(); k += 2) {
                done.push_back(stalled[k]);
                done.push_back(-1);  // already harvested
            }
            nf64 = 0;
        }
        if (!done.empty()) {
            std::vector<int> h;
            for (size_t k = 0; k < done.size(); k += 2) {
                if (done[k + 1] >= 0) {
                    h.push_back(done[k]);
                    h.push_back(done[k + 1]);
                }
                free_slots.push_back(done[k]);
            }
            harvest(h, mixed);
        }
        // refill: at once with the one-level preconditioner; in groups of 16 with
        // the two-level one (each refill factorizes the new coarse operators)
        const size_t left = queue.size() - next;
        if (!free_slots.empty() && left > 0 &&
            (!NC || nrun == 0 || free_slots.size() >= std::min<size_t>(16, left))) {
            std::vector<int> pairs;
            for (size_t k = 0; k < free_slots.size() && next < queue.size(); k++) {
                pairs.push_back(free_slots[k]);
                pairs.push_back(queue[next++]);
            }
            free_slots.erase(free_slots.begin(), free_slots.begin() + pairs.size() / 2);
            if (!pairs.empty()) {
                // refilled slots may lie beyond the live chunks after a compaction
                int hi = 0;
                for (size_t k = 0; k < pairs.size(); k += 2) hi = std::max(hi, pairs[k] / 64 + 1);
                if (hi > nch_) {
                    nch_ = hi;
                    rebuild();
                }
                refill(pairs, mixed);
                nrun += (int)pairs.size() / 2;
            }
            read_status();
        }
        if (next >= queue.size() && st.ms_queue_empty == 0) st.ms_queue_empty = since_start();
        if (nrun == 0) break;
        const bool narrow = narrow_ && NC > 0;
        const int hand = handover_ >= 0 ? handover_ : (narrow ? 8 : 16);
        if (next >= queue.size() && nrun + nf64 <= hand) {
            std::vector<int> pairs;
            for (int sl = 0; sl < Rp_; sl++)
                if (rid_[sl] >= 0 && !terminal(stv[sl])) {
                    pairs.push_back(sl);
                    pairs.push_back(rid_[sl]);
                }
            handover(pairs, mixed);
            break;
        }
        if (next >= queue.size() && nrun + nf64 <= 64 * (nch_ - 1)) maybe_compact(mixed);
        if (narrow && next >= queue.size()) maybe_narrow(mixed);
        if (++guard > guard_max) throw Error(-5, "batched CG did not terminate");
        FL_HIP(hipGraphLaunch(graph(mixed), stream_));
        st.chunks++;
    }
    st.ms_loop_end = since_start();
    hand_join();
    st.ms_handover_end = since_start();
    if (herr) std::rethrow_exception(herr);
    for (const HandResult &h : hres) {
        h_status_[h.real] = h.ss.converged ? kConverged : kMaxIt;
        h_it_[h.real] = h.it0 + h.ss.iterations;
        h_ru_[h.real] = h.n_ru;
        h_rr_[h.real] = h.ss.residual * h.ss.residual;
        h_bb_[h.real] = h.bb;
        h_tol_[h.real] = h.tol;
        st.handover_iterations += h.ss.iterations;
    }
    st.ms_handover = hbusy_ms;
    if (ksaved) base_.set_conductivity_device(d_Ksave_);
    FL_HIP(hipStreamSynchronize(stream_));
    float t01 = 0;
    FL_HIP(hipEventElapsedTime(&t01, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.ms_setup = t01;
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    st.ms_cg = st.ms_total - st.ms_setup;
    st.iterations.resize(R_);
    st.converged.resize(R_);
    st.residual.resize(R_);
    for (int r = 0; r < R_; r++) {
        st.iterations[r] = h_it_[r];
        st.total_iterations += h_it_[r];
        st.max_iterations = std::max(st.max_iterations, h_it_[r]);
        if (precision == 1) st.reliable_updates += h_ru_[r];
        // the stop test is re-checked on the final FP64 true residual
        st.converged[r] = (h_status_[r] == kConverged && h_rr_[r] <= 4.0 * h_tol_[r]) ? 1 : 0;
        st.residual[r] = h_bb_[r] > 0 ? std::sqrt(h_rr_[r] / h_bb_[r]) : std::sqrt(h_rr_[r]);
    }
    st.grid = grid_;
    st.xcd_map = xcd_map_;
    st.chunk = CH;
    st.tile_rows = ncol_ && precision == 1 ? tile_T_ : 0;
    return st;
}

// Time the operator kernel (idempotent: it rewrites q and the other search
// direction buffer from z, beta and the current direction of the filled
// slots) over launch shapes and keep the fastest.
void DarcyBatchGPU::autotune() {
    if (!have_K_) throw Error(-4, "batch conductivities not set");
    tuned_ = false;
    (void)solve(1e-30, 0.0, 1, 1);  // fills the slots, tunes, one iteration
}

void DarcyBatchGPU::tune() {
    alloc32();
    if (ncol_) {
        // two-level path: the FP32 operator's grid follows its tiles, so the
        // tiling is what is tuned (each candidate times three operator
        // launches on the filled slots; q and the operator partials are
        // rewritten by the next iteration); the launch grid serves the
        // direction kernel and the FP64 operator
        set_launch(1024, 0, chunk_);
        if (tilings_.size() > 1 && lw_ == 64) {
            hipEvent_t t0, t1;
            FL_HIP(hipEventCreate(&t0));
            FL_HIP(hipEventCreate(&t1));
            float best = 1e30f;
            int bi = 0;
            for (int i = 0; i < (int)tilings_.size(); i++) {
                use_tiling(i);
                BArgs<float> a = make_args<float>(F_, Rp_, grid_, xcd_map_, d_rec_, d_row_ptr_, d_row_c_, d_meta_,
                                                  d_op_off_, d_conn_off_, d_conn_, d_tab32_, d_K32_, d_tab_, d_K_,
                                                  d_dinv32_, d_x32_, d_r32_, d_z32_, d_p32a_, d_p32b_, d_q32_,
                                                  d_part_, sc_);
                a.frec = d_frec_;
                const OpTiles ot{d_gen_, d_grec_, d_thptr_, d_thalo_, d_teptr_, d_tels_, d_lrec_, ngen_, gen_blocks_,
                                 ntiles_, tile_T_, nseg_max_};
                auto op = [&]() {
                    if (ngen_) {
                        if (nch_ == 1)
                            hipLaunchKernelGGL(k_c_opg<1>, dim3(gen_blocks_), dim3(256), 0, stream_, a, ot, ot.gen,
                                               ot.grec, a.opT, a.opT64);
                        else
                            hipLaunchKernelGGL(k_c_opg<2>, dim3(gen_blocks_), dim3(256), 0, stream_, a, ot, ot.gen,
                                               ot.grec, a.opT, a.opT64);
                    }
                    hipLaunchKernelGGL(k_c_opl, dim3(ntiles_), dim3(kOplThreads), opl_lds_bytes(nseg_max_, tile_T_),
                                       stream_, a, ot, nch_, 2 * nch_);
                };
                op();  // warm
                FL_HIP(hipEventRecord(t0, stream_));
                for (int k = 0; k < 3; k++) op();
                FL_HIP(hipEventRecord(t1, stream_));
                FL_HIP(hipEventSynchronize(t1));
                FL_HIP(hipGetLastError());
                float ms = 0;
                FL_HIP(hipEventElapsedTime(&ms, t0, t1));
                if (ms < best) {
                    best = ms;
                    bi = i;
                }
            }
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            use_tiling(bi);
        }
        tuned_ = true;
        return;
    }
    const int cand_g[] = {512, 1024, 2048};
    float best = 1e30f;
    int bg = grid_, bx = xcd_map_;
    hipEvent_t t0, t1;
    FL_HIP(hipEventCreate(&t0));
    FL_HIP(hipEventCreate(&t1));
    for (int x = 0; x <= 2; x++)
        for (int G : cand_g) {
            BArgs<float> a = make_args<float>(F_, Rp_, G, x, d_rec_, d_row_ptr_, d_row_c_, d_meta_, d_op_off_,
                                              d_conn_off_, d_conn_, d_tab32_, d_K32_, d_tab_, d_K_, d_dinv32_,
                                              d_x32_, d_r32_, d_z32_, d_p32a_, d_p32b_, d_q32_, d_part_, sc_);
            const dim3 grow(G, nch_), blk(256);
            // the one-level operator only rewrites q and the spare direction buffer
            auto op = [&]() { hipLaunchKernelGGL(k_b_op<float>, grow, blk, 0, stream_, a); };
            op();  // warm
            FL_HIP(hipEventRecord(t0, stream_));
            for (int k = 0; k < 3; k++) op();
            FL_HIP(hipEventRecord(t1, stream_));
            FL_HIP(hipEventSynchronize(t1));
            float ms = 0;
            FL_HIP(hipEventElapsedTime(&ms, t0, t1));
            if (ms < best) {
                best = ms;
                bg = G;
                bx = x;
            }
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    set_launch(bg, bx, chunk_);
    tuned_ = true;
}

void DarcyBatchGPU::diagnostics(int *status, int *n_ru) const {
    for (int r = 0; r < R_; r++) {
        if (status) status[r] = r < (int)h_status_.size() ? h_status_[r] : kIdle;
        if (n_ru) n_ru[r] = r < (int)h_ru_.size() ? h_ru_[r] : 0;
    }
}

void DarcyBatchGPU::lambda_device(double *out) const {
    hipLaunchKernelGGL(k_b_lambda_out, dim3(std::min(grid_for(F_), 1024), R_), dim3(kBlock), 0, stream_, R_,
                       (size_t)Rout_, F_, (size_t)S_ + E_, (const double *)d_sol_, out);
    FL_HIP(hipGetLastError());
    FL_HIP(hipStreamSynchronize(stream_));
}

void DarcyBatchGPU::get_solution(int r, double *out) const {
    if (r < 0 || r >= R_) throw Error(-1, "realization index out of range");
    const size_t n = (size_t)S_ + E_ + F_;
    hipLaunchKernelGGL(k_b_extract, dim3(std::min(grid_for((long)n), 4096)), dim3(kBlock), 0, stream_, n,
                       (size_t)Rout_, r, (const double *)d_sol_, d_ext_);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(out, d_ext_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

}  // namespace fl
