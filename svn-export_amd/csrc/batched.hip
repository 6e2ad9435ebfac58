// Batched Monte-Carlo Darcy solver (matrix-free condensed operator,
// realization-interleaved CG).  The closed form per element is the one of
// darcy.hip (assembly_steady_mh_matrix darcy_flow_mh.cc:525-694, make_schur1/2
// :1009-1131, SchurComplement::resolve schur.cc:714-773), split into its
// K-independent factors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <numeric>

#include "batched.hh"
#include "common.hh"
#include "condense.hh"

namespace fl {

namespace {

constexpr int kWaves = 4;            // waves per block of the row kernels
constexpr int kFoldThreads = 1024;   // 16 waves per fold block
constexpr int kMaxGrid = 2048;       // partial arrays are sized for this many blocks
constexpr int kMaxConn = 16;         // element index packing e * 16 + local index

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
    const V *opT, *K, *dinv;
    const double *opT64, *K64;  // FP64 table and K (elements with vb neighbours)
    V *x, *r, *z, *p0, *p1, *q;
    double *part0, *part1, *part2;
    BatchScal sc;
};

// rows of this wave: start, stride, end
__device__ __forceinline__ void row_range(int F, int G, int xcd, int &start, int &step, int &end) {
    const int wid = threadIdx.x >> 6;
    if (!xcd) {
        start = blockIdx.x * kWaves + wid;
        step = G * kWaves;
        end = F;
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

template <class V>
__device__ __forceinline__ V gather(const V *__restrict__ za, const V *__restrict__ pb, V beta, size_t idx,
                                    bool fly) {
    return fly ? za[idx] + beta * pb[idx] : za[idx];
}

// Row i of S_e(K) x for an element without vb neighbours.  FLY: x = za + beta
// pb (the new search direction, formed on the fly from z and the previous one).
template <class V, int N0, bool FLY>
__device__ __forceinline__ V elem_row_plain(const V *__restrict__ T, V K, const int *__restrict__ cn, int i,
                                            const V *__restrict__ za, const V *__restrict__ pb, V beta, size_t Rp,
                                            int r) {
    V t = V(0), xi = V(0);
#pragma unroll
    for (int j = 0; j < N0; j++) {
        const V xj = gather(za, pb, beta, (size_t)cn[j] * Rp + r, FLY);
        t += T[i * N0 + j] * xj;
        if (j == i) xi = xj;
    }
    return K * t + T[N0 * N0 + i] * xi;
}

// Row i of S_e(K) x for an element with vb neighbours, in FP64.
template <class V, int N0, bool FLY>
__device__ __forceinline__ double elem_row_vb(const double *__restrict__ T, int nvb, double K,
                                              const int *__restrict__ cn, int i, const V *__restrict__ za,
                                              const V *__restrict__ pb, V beta, size_t Rp, int r) {
    double xv[N0];
#pragma unroll
    for (int j = 0; j < N0; j++) xv[j] = (double)gather(za, pb, beta, (size_t)cn[j] * Rp + r, FLY);
    const double *C0 = T, *w0 = T + N0 * N0, *dlt = w0 + N0, *sv = dlt + N0, *vv = sv + 2;
    double ux = 0.0;
#pragma unroll
    for (int j = 0; j < N0; j++) ux += w0[j] * xv[j];
    ux *= K;
    double xi = 0.0;
    for (int k = 0; k < nvb; k++) {
        const double xk = (double)gather(za, pb, beta, (size_t)cn[N0 + k] * Rp + r, FLY);
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

template <class V, bool FLY>
__device__ __forceinline__ V op_row(const int *__restrict__ row_ptr, const int *__restrict__ row_c,
                                    const int *__restrict__ meta, const int *__restrict__ op_off,
                                    const int *__restrict__ conn_off, const int *__restrict__ conn,
                                    const V *__restrict__ opT, const V *__restrict__ Kv,
                                    const double *__restrict__ opT64, const double *__restrict__ K64, int row,
                                    const V *__restrict__ za, const V *__restrict__ pb, V beta, size_t Rp, int r) {
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
            if (n0 == 4) y += elem_row_plain<V, 4, FLY>(T, K, cn, i, za, pb, beta, Rp, r);
            else if (n0 == 3) y += elem_row_plain<V, 3, FLY>(T, K, cn, i, za, pb, beta, Rp, r);
            else y += elem_row_plain<V, 2, FLY>(T, K, cn, i, za, pb, beta, Rp, r);
        } else {
            const double *T = opT64 + op_off[e];
            const double K = K64[(size_t)e * Rp + r];
            double t;
            if (n0 == 4) t = elem_row_vb<V, 4, FLY>(T, nvb, K, cn, i, za, pb, beta, Rp, r);
            else if (n0 == 3) t = elem_row_vb<V, 3, FLY>(T, nvb, K, cn, i, za, pb, beta, Rp, r);
            else t = elem_row_vb<V, 2, FLY>(T, nvb, K, cn, i, za, pb, beta, Rp, r);
            y += (V)t;
        }
    }
    return y;
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

// K [R][E] -> K64 / K32 [E][Rp]; padding realizations repeat realization 0
__global__ void k_b_transpose_K(const double *__restrict__ K, int R, int Rp, int E, double *__restrict__ K64,
                                float *__restrict__ K32) {
    __shared__ double t[64][65];
    const int e0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = wid; j < 64; j += 4) {
        int r = r0 + j, e = e0 + lane;
        if (r >= R) r = 0;
        t[j][lane] = e < E ? K[(size_t)r * E + e] : 0.0;
    }
    __syncthreads();
    for (int j = wid; j < 64; j += 4) {
        int e = e0 + j;
        if (e < E) {
            double v = t[lane][j];
            K64[(size_t)e * Rp + r0 + lane] = v;
            K32[(size_t)e * Rp + r0 + lane] = (float)v;
        }
    }
}

// right-hand side and Jacobi diagonal of every realization (FP64)
__global__ void __launch_bounds__(256) k_b_rhs(int F, size_t Rp, const int *__restrict__ row_ptr,
                                               const int *__restrict__ row_c, const int *__restrict__ meta,
                                               const int *__restrict__ rs_off, const double *__restrict__ rsT,
                                               const double *__restrict__ Kv, double *__restrict__ b,
                                               double *__restrict__ dinv, float *__restrict__ dinv32) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
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
        dinv32[o] = (float)di;
    }
}

// cold start: x = 0, r = b, z = D^-1 b, p = 0 (both buffers); partials r.z, r.r
template <class V>
__global__ void __launch_bounds__(256) k_b_init(BArgs<V> a, const double *__restrict__ b64, double *x64) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    int s0, st, en;
    row_range(a.F, a.G, a.xcd, s0, st, en);
    double v[2] = {0.0, 0.0};
    for (int row = s0; row < en; row += st) {
        const size_t o = (size_t)row * a.Rp + r;
        const double b = b64[o];
        const V bv = (V)b;
        const V z = a.dinv[o] * bv;
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
                const V q = op_row<V, true>(a.row_ptr, a.row_c, a.meta, a.op_off, a.conn_off, a.conn, a.opT, a.K,
                                            a.opT64, a.K64, row, a.z, pin, beta, a.Rp, r);
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
                const V z = a.dinv[o] * rv;
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
constexpr double kRuDrop2 = 0.01;
__device__ __forceinline__ bool ru_pick(const BatchScal &sc, int r, int st) {
    return st == kCandidate || (st == kRun && sc.rr[r] < kRuDrop2 * sc.rr_ru[r]);
}

// mixed precision: x64 += x32, x32 = 0 for the selected realizations
// (all: every realization, at the end of the FP32 phase)
__global__ void __launch_bounds__(256) k_b_accum(int F, int G, int xcd, size_t Rp, BatchScal sc, int all,
                                                 double *__restrict__ x64, float *__restrict__ x32) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    const bool act = all || ru_pick(sc, r, sc.status[r]);
    if (!__any(act)) return;
    int s0, st, en;
    row_range(F, G, xcd, s0, st, en);
    for (int row = s0; row < en; row += st) {
        if (act) {
            const size_t o = (size_t)row * Rp + r;
            x64[o] += (double)x32[o];
            x32[o] = 0.0f;
        }
    }
}

// True FP64 residual r64 = b - S64 x64 for the realizations whose status bit
// is set in sel (sel < 0: the reliable-update selection).  MODE 0: store r and
// z = D^-1 r in V; MODE 1: also zero both search-direction buffers (FP64 warm
// start); MODE 2: partial r.r only.
template <class V, int MODE>
__global__ void __launch_bounds__(256) k_b_resid(BArgs<V> a, int sel, const int *__restrict__ op_off64,
                                                 const double *__restrict__ opT64, const double *__restrict__ K64,
                                                 const double *__restrict__ b64, const double *__restrict__ x64) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    const int st0 = a.sc.status[r];
    const bool act = sel < 0 ? ru_pick(a.sc, r, st0) : ((sel >> st0) & 1) != 0;
    double v[2] = {0.0, 0.0};
    if (__any(act)) {
        int s0, st, en;
        row_range(a.F, a.G, a.xcd, s0, st, en);
        for (int row = s0; row < en; row += st) {
            if (act) {
                const double ax = op_row<double, false>(a.row_ptr, a.row_c, a.meta, op_off64, a.conn_off, a.conn,
                                                        opT64, K64, opT64, K64, row, x64, nullptr, 0.0, a.Rp, r);
                const size_t o = (size_t)row * a.Rp + r;
                const double rr = b64[o] - ax;
                v[1] += rr * rr;
                if (MODE != 2) {
                    const V rv = (V)rr;
                    const V z = a.dinv[o] * rv;
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
                                                         double rtol2, double atol2, int max_it, int mixed) {
    __shared__ double sh[3][kFoldThreads / 64][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    double s[3] = {0.0, 0.0, 0.0};
    for (int g = wid; g < G; g += kFoldThreads / 64) {
        const size_t o = (size_t)g * Rp + r;
        if (MODE == kFoldAlpha) s[0] += part0[o];
        else {
            if (MODE != kFoldFinal) s[1] += part1[o];
            s[2] += part2[o];
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
    if (MODE == kFoldInit) {
        // s[1] = r.z, s[2] = r.r = b.b
        sc.bnorm2[r] = s[2];
        sc.tol2[r] = fmax(rtol2 * s[2], atol2);
        sc.rz[r] = s[1];
        sc.rz_prev[r] = 1.0;
        sc.rr[r] = s[2];
        sc.beta[r] = 0.0;
        sc.alpha[r] = 0.0;
        sc.it[r] = 0;
        sc.n_ru[r] = 0;
        sc.it_ru[r] = 0;
        sc.par[r] = 0;
        sc.rr_hist[r] = INFINITY;
        sc.rr_ru[r] = s[2];
        st = (r < R) ? (s[2] <= sc.tol2[r] ? kConverged : kRun) : kConverged;
        sc.status[r] = st;
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
            const double rz = s[1], rr = s[2];
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
            const double rz = s[1], rr = s[2];
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
            sc.rr[r] = s[2];
            sc.beta[r] = 0.0;
            sc.par[r] = 0;
            st = s[2] <= sc.tol2[r] ? kConverged : kRun;
            sc.status[r] = st;
        }
    } else {  // final true residual
        sc.rr[r] = s[2];
    }
    const unsigned long long run = __ballot(st == kRun), f64 = __ballot(st == kFp64);
    if (lane == 0) {
        sc.active_count[2 * blockIdx.x] = __popcll(run);
        sc.active_count[2 * blockIdx.x + 1] = __popcll(f64);
    }
}

// Reconstruction of element pressure and side fluxes (schur.cc:714-773) for
// every realization; output interleaved sol[idx * Rp + r].
template <int D>
__global__ void __launch_bounds__(256) k_b_resolve(int e_begin, int e_end, const int *__restrict__ perm,
                                                   const int *__restrict__ side_off, int S, size_t Rp,
                                                   const int *__restrict__ meta, const int *__restrict__ rs_off,
                                                   const double *__restrict__ rsT, const int *__restrict__ conn_off,
                                                   const int *__restrict__ conn, const double *__restrict__ Kv,
                                                   const double *__restrict__ lam, double *__restrict__ sol) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    const int idx = e_begin + blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (idx >= e_end) return;
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
        sol[(size_t)(so + i) * Rp + r] = K * Q[n0 * n0 + n0 + i] + pe * (K * Q[n0 * n0 + i]) - K * t;
    }
    sol[(size_t)(S + e) * Rp + r] = pe;
}

__global__ void k_b_scatter_lambda(int F, size_t Rp, const int *__restrict__ perm, const double *__restrict__ x,
                                   double *__restrict__ lam_out) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 64 + lane;
    for (int row = blockIdx.x * kWaves + (threadIdx.x >> 6); row < F; row += gridDim.x * kWaves)
        lam_out[(size_t)perm[row] * Rp + r] = x[(size_t)row * Rp + r];
}

__global__ void k_b_extract(size_t n, size_t Rp, int r, const double *__restrict__ sol, double *__restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = sol[i * Rp + r];
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

DarcyBatchGPU::DarcyBatchGPU(DarcyGPU &base, int R) : base_(base), dv_(base.dev()) {
    if (R < 1) throw Error(-1, "batch needs at least one realization");
    R_ = R;
    Rp_ = (R + 63) / 64 * 64;
    nchunk_ = Rp_ / 64;
    FL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    build();
}

DarcyBatchGPU::~DarcyBatchGPU() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    free_graphs();
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
    for (void *p : owned_) dev_free(p);
    if (h_count_) (void)hipHostFree(h_count_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void DarcyBatchGPU::free_graphs() {
    if (g64_) (void)hipGraphExecDestroy(g64_);
    if (g32_) (void)hipGraphExecDestroy(g32_);
    g64_ = g32_ = nullptr;
    g64_key_ = g32_key_ = -1;
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
    // rows in edge order; contributions of each row in element order
    std::vector<int> row_ptr(F + 1, 0), row_c(conn_off[E]);
    for (int i = 0; i < conn_off[E]; i++) row_ptr[conn[i] + 1]++;
    for (int f = 0; f < F; f++) row_ptr[f + 1] += row_ptr[f];
    {
        std::vector<int> fp(row_ptr.begin(), row_ptr.end() - 1);
        for (int e = 0; e < E; e++)
            for (int i = conn_off[e]; i < conn_off[e + 1]; i++) row_c[fp[conn[i]]++] = e * kMaxConn + (i - conn_off[e]);
    }
    ncontrib_ = conn_off[E];
    std::vector<int> perm(F);
    std::iota(perm.begin(), perm.end(), 0);
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
    d_dinv32_ = af(nv);
    d_x_ = ad(nv); d_r_ = ad(nv); d_z_ = ad(nv); d_p0_ = ad(nv); d_p1_ = ad(nv); d_q_ = ad(nv);
    d_sol_ = ad(((size_t)S_ + E + F) * Rp_);
    d_ext_ = ad((size_t)S_ + E + F);
    d_part_ = ad(3 * (size_t)kMaxGrid * Rp_);
    double *scd = ad(9 * (size_t)Rp_);
    sc_.alpha = scd; sc_.beta = scd + Rp_; sc_.rz = scd + 2 * Rp_; sc_.rz_prev = scd + 3 * Rp_;
    sc_.rr = scd + 4 * Rp_; sc_.tol2 = scd + 5 * Rp_; sc_.bnorm2 = scd + 6 * Rp_; sc_.rr_hist = scd + 7 * Rp_;
    sc_.rr_ru = scd + 8 * Rp_;
    int *sci = ai(5 * (size_t)Rp_ + 2 * nchunk_);
    sc_.status = sci; sc_.it = sci + Rp_; sc_.n_ru = sci + 2 * Rp_; sc_.par = sci + 3 * Rp_;
    sc_.it_ru = sci + 4 * Rp_;
    sc_.active_count = sci + 5 * Rp_;
    FL_HIP(hipHostMalloc((void **)&h_count_, 2 * nchunk_ * sizeof(int), hipHostMallocDefault));
    if (const char *g = std::getenv("FL_BATCH_GRID")) {  // launch-shape override: grid[,xcd[,chunk]]
        int gv = 0, xv = 0, cv = 16;
        if (std::sscanf(g, "%d,%d,%d", &gv, &xv, &cv) >= 1) {
            set_launch(gv, xv, cv);
            tuned_ = true;
        }
    }
    FL_HIP(hipStreamSynchronize(s));
}

void DarcyBatchGPU::alloc32() {
    if (d_x32_) return;
    const size_t nv = (size_t)F_ * Rp_;
    auto af = [&](size_t n) { float *p = dev_alloc<float>(n); owned_.push_back(p); return p; };
    d_x32_ = af(nv); d_r32_ = af(nv); d_z32_ = af(nv); d_p32a_ = af(nv); d_p32b_ = af(nv); d_q32_ = af(nv);
}

void DarcyBatchGPU::set_launch(int grid, int xcd_map, int chunk) {
    if (grid < 8 || grid > kMaxGrid || grid % 8) throw Error(-1, "batch grid must be a multiple of 8 in [8, 2048]");
    if (chunk < 1 || chunk > 1024) throw Error(-1, "batch chunk must be in [1, 1024]");
    if (grid != grid_ || (xcd_map != 0) != (xcd_map_ != 0) || chunk != chunk_) free_graphs();
    grid_ = grid;
    xcd_map_ = xcd_map ? 1 : 0;
    chunk_ = chunk;
    tuned_ = true;
}

void DarcyBatchGPU::set_conductivity(const double *K, bool device_ptr) {
    const double *src = K;
    double *tmp = nullptr;
    if (!device_ptr) {
        tmp = dev_alloc<double>((size_t)R_ * E_);
        FL_HIP(hipMemcpyAsync(tmp, K, (size_t)R_ * E_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        src = tmp;
    }
    dim3 g((E_ + 63) / 64, nchunk_);
    hipLaunchKernelGGL(k_b_transpose_K, g, dim3(256), 0, stream_, src, R_, Rp_, E_, d_K_, d_K32_);
    FL_HIP(hipGetLastError());
    FL_HIP(hipStreamSynchronize(stream_));
    if (tmp) dev_free(tmp);
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
    dim3 g(std::min((F_ + kWaves - 1) / kWaves, 4096), nchunk_);
    hipLaunchKernelGGL(k_b_rhs, g, dim3(256), 0, stream_, F_, (size_t)Rp_, d_row_ptr_, d_row_c_, d_meta_, d_rs_off_,
                       d_rsT_, d_K_, d_b_, d_dinv_, d_dinv32_);
    FL_HIP(hipGetLastError());
}

template <class V>
static BArgs<V> make_args(int F, int Rp, int G, int xcd, const int *row_ptr, const int *row_c, const int *meta,
                          const int *op_off, const int *conn_off, const int *conn, const V *opT, const V *K,
                          const double *opT64, const double *K64,
                          const V *dinv, V *x, V *r, V *z, V *p0, V *p1, V *q, double *part, const BatchScal &sc) {
    BArgs<V> a;
    a.F = F; a.Rp = Rp; a.G = G; a.xcd = xcd;
    a.row_ptr = row_ptr; a.row_c = row_c; a.meta = meta; a.op_off = op_off; a.conn_off = conn_off; a.conn = conn;
    a.opT = opT; a.K = K; a.dinv = dinv;
    a.opT64 = opT64; a.K64 = K64;
    a.x = x; a.r = r; a.z = z; a.p0 = p0; a.p1 = p1; a.q = q;
    a.part0 = part; a.part1 = part + (size_t)kMaxGrid * Rp; a.part2 = part + 2 * (size_t)kMaxGrid * Rp;
    a.sc = sc;
    return a;
}

BatchStats DarcyBatchGPU::solve(double rtol, double atol, int max_it, int precision) {
    if (!have_K_) throw Error(-4, "batch conductivities not set");
    if (precision != 0 && precision != 1) throw Error(-1, "batch precision must be 0 (FP64) or 1 (FP32 + FP64)");
    if (precision == 1) alloc32();
    BatchStats st;
    st.n_real = R_;
    auto t_start = std::chrono::steady_clock::now();
    hipEvent_t e0, e1, e2, e3;
    FL_HIP(hipEventCreate(&e0)); FL_HIP(hipEventCreate(&e1)); FL_HIP(hipEventCreate(&e2)); FL_HIP(hipEventCreate(&e3));
    FL_HIP(hipEventRecord(e0, stream_));
    setup_system();
    if (!tuned_) tune();
    FL_HIP(hipEventRecord(e1, stream_));
    const int G = grid_, X = xcd_map_;
    const dim3 grow(G, nchunk_), blk(256), gfold(nchunk_), bfold(kFoldThreads);
    const double rtol2 = rtol * rtol, atol2 = atol * atol;
    BArgs<double> a64 = make_args<double>(F_, Rp_, G, X, d_row_ptr_, d_row_c_, d_meta_, d_op_off_, d_conn_off_,
                                          d_conn_, d_tab_, d_K_, d_tab_, d_K_, d_dinv_, d_x_, d_r_, d_z_, d_p0_, d_p1_, d_q_,
                                          d_part_, sc_);
    BArgs<float> a32 = make_args<float>(F_, Rp_, G, X, d_row_ptr_, d_row_c_, d_meta_, d_op_off_, d_conn_off_,
                                        d_conn_, d_tab32_, d_K32_, d_tab_, d_K_, d_dinv32_, d_x32_, d_r32_, d_z32_, d_p32a_,
                                        d_p32b_, d_q32_, d_part_, sc_);
    double *P0 = a64.part0, *P1 = a64.part1, *P2 = a64.part2;
    auto fold = [&](int mode, int mixed) {
        const double *q0 = P0, *q1 = P1, *q2 = P2;
#define FL_FOLD(M) \
    hipLaunchKernelGGL(k_b_fold<M>, gfold, bfold, 0, stream_, sc_, q0, q1, q2, G, Rp_, R_, rtol2, atol2, max_it, mixed)
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
    // one CG iteration; ev: optional events around the operator and the update
    auto iter64 = [&](hipEvent_t *ev) {
        if (ev) FL_HIP(hipEventRecord(ev[0], stream_));
        hipLaunchKernelGGL(k_b_op<double>, grow, blk, 0, stream_, a64);
        if (ev) FL_HIP(hipEventRecord(ev[1], stream_));
        fold(kFoldAlpha, 0);
        if (ev) FL_HIP(hipEventRecord(ev[2], stream_));
        hipLaunchKernelGGL(k_b_update<double>, grow, blk, 0, stream_, a64);
        if (ev) FL_HIP(hipEventRecord(ev[3], stream_));
        fold(kFoldBeta, 0);
    };
    auto iter32 = [&](hipEvent_t *ev) {
        if (ev) FL_HIP(hipEventRecord(ev[0], stream_));
        hipLaunchKernelGGL(k_b_op<float>, grow, blk, 0, stream_, a32);
        if (ev) FL_HIP(hipEventRecord(ev[1], stream_));
        fold(kFoldAlpha, 1);
        if (ev) FL_HIP(hipEventRecord(ev[2], stream_));
        hipLaunchKernelGGL(k_b_update<float>, grow, blk, 0, stream_, a32);
        if (ev) FL_HIP(hipEventRecord(ev[3], stream_));
        fold(kFoldBeta, 1);
    };
    auto reliable_update = [&]() {
        hipLaunchKernelGGL(k_b_accum, grow, blk, 0, stream_, F_, G, X, (size_t)Rp_, sc_, 0, d_x_, d_x32_);
        hipLaunchKernelGGL((k_b_resid<float, 0>), grow, blk, 0, stream_, a32, -1,
                           (const int *)d_op_off_, (const double *)d_tab_, (const double *)d_K_,
                           (const double *)d_b_, (const double *)d_x_);
        fold(kFoldRU, 1);
    };
    auto counts = [&](int &run, int &f64) {
        FL_HIP(hipMemcpyAsync(h_count_, sc_.active_count, 2 * nchunk_ * sizeof(int), hipMemcpyDeviceToHost,
                              stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        run = f64 = 0;
        for (int c = 0; c < nchunk_; c++) {
            run += h_count_[2 * c];
            f64 += h_count_[2 * c + 1];
        }
    };
    // live kernel timing: the first iterations run eagerly between events
    const int kSample = 4;
    while ((int)ev_.size() < 4 * kSample) {
        hipEvent_t e;
        FL_HIP(hipEventCreate(&e));
        ev_.push_back(e);
    }
    auto sample_times = [&]() {
        double so = 0, su = 0;
        for (int k = 0; k < kSample; k++) {
            float t1 = 0, t2 = 0;
            FL_HIP(hipEventElapsedTime(&t1, ev_[4 * k], ev_[4 * k + 1]));
            FL_HIP(hipEventElapsedTime(&t2, ev_[4 * k + 2], ev_[4 * k + 3]));
            so += t1;
            su += t2;
        }
        st.ms_op_avg = so / kSample;
        st.ms_update_avg = su / kSample;
        st.op_samples = kSample;
    };
    int run = 0, f64 = 0;
    long launched = 0;
    if (precision == 0) {
        hipLaunchKernelGGL(k_b_init<double>, grow, blk, 0, stream_, a64, (const double *)d_b_, (double *)nullptr);
        fold(kFoldInit, 0);
        for (int k = 0; k < kSample; k++) iter64(&ev_[4 * k]);
        launched = kSample;
        counts(run, f64);
        sample_times();
        if (g64_key_ != max_it) {  // max_it is baked into the captured folds
            if (g64_) (void)hipGraphExecDestroy(g64_);
            g64_ = nullptr;
            hipGraph_t gr;
            FL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < chunk_; k++) iter64(nullptr);
            FL_HIP(hipStreamEndCapture(stream_, &gr));
            FL_HIP(hipGraphInstantiate(&g64_, gr, nullptr, nullptr, 0));
            (void)hipGraphDestroy(gr);
            g64_key_ = max_it;
        }
        while (run > 0 && launched < (long)max_it + chunk_) {
            FL_HIP(hipGraphLaunch(g64_, stream_));
            launched += chunk_;
            st.chunks++;
            counts(run, f64);
        }
    } else {
        hipLaunchKernelGGL(k_b_init<float>, grow, blk, 0, stream_, a32, (const double *)d_b_, d_x_);
        fold(kFoldInit, 1);
        for (int k = 0; k < kSample; k++) iter32(&ev_[4 * k]);
        launched = kSample;
        reliable_update();
        st.reliable_updates++;
        counts(run, f64);
        sample_times();
        if (g32_key_ != max_it) {
            if (g32_) (void)hipGraphExecDestroy(g32_);
            g32_ = nullptr;
            hipGraph_t gr;
            FL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < chunk_; k++) iter32(nullptr);
            reliable_update();
            FL_HIP(hipStreamEndCapture(stream_, &gr));
            FL_HIP(hipGraphInstantiate(&g32_, gr, nullptr, nullptr, 0));
            (void)hipGraphDestroy(gr);
            g32_key_ = max_it;
        }
        while (run > 0 && launched < (long)max_it + chunk_) {
            FL_HIP(hipGraphLaunch(g32_, stream_));
            launched += chunk_;
            st.chunks++;
            st.reliable_updates++;
            counts(run, f64);
        }
        // corrections of realizations stopped at the iteration limit
        hipLaunchKernelGGL(k_b_accum, grow, blk, 0, stream_, F_, G, X, (size_t)Rp_, sc_, 1, d_x_, d_x32_);
        if (f64 > 0) {
            // FP64 finish for the realizations whose FP32 residual stagnated
            st.fp64_finish = f64;
            hipLaunchKernelGGL((k_b_resid<double, 1>), grow, blk, 0, stream_, a64, 1 << kFp64,
                               (const int *)d_op_off_, (const double *)d_tab_, (const double *)d_K_,
                               (const double *)d_b_, (const double *)d_x_);
            fold(kFoldWarm, 0);
            counts(run, f64);
            if (g64_key_ != max_it) {  // max_it is baked into the captured folds
                if (g64_) (void)hipGraphExecDestroy(g64_);
                g64_ = nullptr;
                hipGraph_t gr;
                FL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < chunk_; k++) iter64(nullptr);
                FL_HIP(hipStreamEndCapture(stream_, &gr));
                FL_HIP(hipGraphInstantiate(&g64_, gr, nullptr, nullptr, 0));
                (void)hipGraphDestroy(gr);
                g64_key_ = max_it;
            }
            long l64 = 0;
            while (run > 0 && l64 < (long)max_it + chunk_) {
                FL_HIP(hipGraphLaunch(g64_, stream_));
                l64 += chunk_;
                st.chunks++;
                counts(run, f64);
            }
        }
    }
    // final true FP64 residual of every realization
    hipLaunchKernelGGL((k_b_resid<double, 2>), grow, blk, 0, stream_, a64, 0xff, (const int *)d_op_off_,
                       (const double *)d_tab_, (const double *)d_K_, (const double *)d_b_, (const double *)d_x_);
    fold(kFoldFinal, 0);
    FL_HIP(hipEventRecord(e2, stream_));
    // reconstruction
    {
        double *lam = d_sol_ + ((size_t)S_ + E_) * Rp_;
        dim3 gl(std::min((F_ + kWaves - 1) / kWaves, 4096), nchunk_);
        hipLaunchKernelGGL(k_b_scatter_lambda, gl, blk, 0, stream_, F_, (size_t)Rp_, (const int *)d_perm_,
                           (const double *)d_x_, lam);
        for (int D = 1; D <= 3; D++) {
            int n = dim_start_[D] - dim_start_[D - 1];
            if (n <= 0) continue;
            dim3 g((n + kWaves - 1) / kWaves, nchunk_);
#define FL_RESOLVE(DD)                                                                                         \
    hipLaunchKernelGGL(k_b_resolve<DD>, g, blk, 0, stream_, dim_start_[D - 1], dim_start_[D],                  \
                       (const int *)d_eperm_, dv_.side_off, S_, (size_t)Rp_, (const int *)d_meta_,             \
                       (const int *)d_rs_off_, (const double *)d_rsT_, (const int *)d_conn_off_,               \
                       (const int *)d_conn_, (const double *)d_K_, (const double *)lam, d_sol_)
            if (D == 1) FL_RESOLVE(1);
            else if (D == 2) FL_RESOLVE(2);
            else FL_RESOLVE(3);
#undef FL_RESOLVE
        }
    }
    FL_HIP(hipGetLastError());
    FL_HIP(hipEventRecord(e3, stream_));
    FL_HIP(hipEventSynchronize(e3));
    float t01, t12, t23;
    FL_HIP(hipEventElapsedTime(&t01, e0, e1));
    FL_HIP(hipEventElapsedTime(&t12, e1, e2));
    FL_HIP(hipEventElapsedTime(&t23, e2, e3));
    st.ms_setup = t01; st.ms_cg = t12; st.ms_resolve = t23;
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2); (void)hipEventDestroy(e3);
    // per-realization results
    std::vector<int> h_status(Rp_), h_it(Rp_);
    std::vector<double> h_rr(Rp_), h_bb(Rp_), h_tol(Rp_);
    FL_HIP(hipMemcpy(h_status.data(), sc_.status, Rp_ * sizeof(int), hipMemcpyDeviceToHost));
    FL_HIP(hipMemcpy(h_it.data(), sc_.it, Rp_ * sizeof(int), hipMemcpyDeviceToHost));
    FL_HIP(hipMemcpy(h_rr.data(), sc_.rr, Rp_ * sizeof(double), hipMemcpyDeviceToHost));
    FL_HIP(hipMemcpy(h_bb.data(), sc_.bnorm2, Rp_ * sizeof(double), hipMemcpyDeviceToHost));
    FL_HIP(hipMemcpy(h_tol.data(), sc_.tol2, Rp_ * sizeof(double), hipMemcpyDeviceToHost));
    st.iterations.resize(R_);
    st.converged.resize(R_);
    st.residual.resize(R_);
    for (int r = 0; r < R_; r++) {
        st.iterations[r] = h_it[r];
        st.total_iterations += h_it[r];
        st.max_iterations = std::max(st.max_iterations, h_it[r]);
        // the stop test is re-checked on the final FP64 true residual
        st.converged[r] = (h_status[r] == kConverged && h_rr[r] <= 4.0 * h_tol[r]) ? 1 : 0;
        st.residual[r] = h_bb[r] > 0 ? std::sqrt(h_rr[r] / h_bb[r]) : std::sqrt(h_rr[r]);
    }
    st.grid = grid_;
    st.xcd_map = xcd_map_;
    st.chunk = chunk_;
    return st;
}

// Time the operator kernel (idempotent: it rewrites q and the other search
// direction buffer from z, beta and the current direction) over launch shapes
// and keep the fastest.  Runs on a cold-started FP32 state.
void DarcyBatchGPU::autotune() {
    if (!have_K_) throw Error(-4, "batch conductivities not set");
    setup_system();
    tune();
}

void DarcyBatchGPU::tune() {
    alloc32();
    const int cand_g[] = {256, 512, 1024, 2048};
    float best = 1e30f;
    int bg = grid_, bx = xcd_map_;
    hipEvent_t t0, t1;
    FL_HIP(hipEventCreate(&t0));
    FL_HIP(hipEventCreate(&t1));
    for (int x = 0; x <= 1; x++)
        for (int G : cand_g) {
            BArgs<float> a = make_args<float>(F_, Rp_, G, x, d_row_ptr_, d_row_c_, d_meta_, d_op_off_, d_conn_off_,
                                              d_conn_, d_tab32_, d_K32_, d_tab_, d_K_, d_dinv32_, d_x32_, d_r32_, d_z32_, d_p32a_,
                                              d_p32b_, d_q32_, d_part_, sc_);
            const dim3 grow(G, nchunk_), blk(256);
            hipLaunchKernelGGL(k_b_init<float>, grow, blk, 0, stream_, a, (const double *)d_b_, d_x_);
            hipLaunchKernelGGL(k_b_fold<kFoldInit>, dim3(nchunk_), dim3(kFoldThreads), 0, stream_, sc_,
                               (const double *)a.part0, (const double *)a.part1, (const double *)a.part2, G, Rp_, R_,
                               0.0, 0.0, 1, 1);
            hipLaunchKernelGGL(k_b_op<float>, grow, blk, 0, stream_, a);  // warm
            FL_HIP(hipEventRecord(t0, stream_));
            for (int k = 0; k < 3; k++) hipLaunchKernelGGL(k_b_op<float>, grow, blk, 0, stream_, a);
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
    std::vector<int> h(Rp_);
    if (status) {
        FL_HIP(hipMemcpy(h.data(), sc_.status, Rp_ * sizeof(int), hipMemcpyDeviceToHost));
        std::copy(h.begin(), h.begin() + R_, status);
    }
    if (n_ru) {
        FL_HIP(hipMemcpy(h.data(), sc_.n_ru, Rp_ * sizeof(int), hipMemcpyDeviceToHost));
        std::copy(h.begin(), h.begin() + R_, n_ru);
    }
}

void DarcyBatchGPU::get_solution(int r, double *out) const {
    if (r < 0 || r >= R_) throw Error(-1, "realization index out of range");
    const size_t n = (size_t)S_ + E_ + F_;
    hipLaunchKernelGGL(k_b_extract, dim3(std::min(grid_for((long)n), 4096)), dim3(kBlock), 0, stream_, n,
                       (size_t)Rp_, r, (const double *)d_sol_, d_ext_);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(out, d_ext_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

}  // namespace fl
