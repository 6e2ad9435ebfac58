// Learned coarse basis: graph network kernels (BF16 MFMA forward and
// backward, FP32 accumulation), coarse-fit loss, Adam.  See gnn.hh.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "common.hh"
#include "gnn.hh"

namespace fl {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;
constexpr int H = GnnGPU::kH;
constexpr int NF = GnnGPU::kNF;
constexpr int kWgradRows = 1024;  // rows per block of the weight-gradient GEMM

__device__ __forceinline__ bf16 f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// MFMA 16x16x16 operand layout: lane l holds A[l % 16][4 (l / 16) .. +3] and
// B[4 (l / 16) .. +3][l % 16]; the result D[4 (l / 16) + i][l % 16], i < 4.

// Y[N][64] = relu(X1[N][K1] W1 + X2[N][K2] W2 + bias), W given transposed
// (Wt[64][K]).  One wave per 16 rows, 4 output tiles of 16 columns.  N is a
// multiple of 64 (padded buffers).
template <int K1, int K2>
__global__ void __launch_bounds__(256) k_g_fwd(const bf16 *__restrict__ X1, const bf16 *__restrict__ W1t,
                                               const bf16 *__restrict__ X2, const bf16 *__restrict__ W2t,
                                               const float *__restrict__ bias, bf16 *__restrict__ Y) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = l & 15, kg = l >> 4;
    const size_t rb = (size_t)blockIdx.x * 64 + w * 16;
    f4 acc[4];
    for (int t = 0; t < 4; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K1; k0 += 16) {
        const s4 a = *(const s4 *)(X1 + (rb + r) * K1 + k0 + kg * 4);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const s4 b = *(const s4 *)(W1t + (size_t)(t * 16 + r) * K1 + k0 + kg * 4);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    if constexpr (K2 > 0) {
#pragma unroll
        for (int k0 = 0; k0 < K2; k0 += 16) {
            const s4 a = *(const s4 *)(X2 + (rb + r) * K2 + k0 + kg * 4);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const s4 b = *(const s4 *)(W2t + (size_t)(t * 16 + r) * K2 + k0 + kg * 4);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = t * 16 + r;
        const float bb = bias[col];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float v = fmaxf(acc[t][i] + bb, 0.f);
            Y[(rb + kg * 4 + i) * H + col] = f2bf(v);
        }
    }
}

// Inference form of a message-passing layer: Y = relu(X1 W1 + M W2 + bias)
// with M[i] = mean of S[j] over j in idx[ptr[i] .. ptr[i + 1]), formed in the
// A-operand loads with the summation order and rounding of k_g_mean (so the
// result equals k_g_mean followed by k_g_fwd<64, 64>), and the output tile
// written through LDS as 32-byte row segments.  n rows, buffers padded to 64.
__global__ void __launch_bounds__(256) k_g_fwd_mean(int n, const bf16 *__restrict__ X1, const bf16 *__restrict__ W1t,
                                                    const int *__restrict__ ptr, const int *__restrict__ idx,
                                                    const bf16 *__restrict__ S, const bf16 *__restrict__ W2t,
                                                    const float *__restrict__ bias, bf16 *__restrict__ Y) {
    // both weight matrices staged once per block (rows padded against bank
    // conflicts); the output tile reuses the same LDS after the products
    constexpr int WS = H + 8;
    __shared__ __attribute__((aligned(16))) bf16 lds[2 * H * WS];
    bf16 (*wl)[H][WS] = reinterpret_cast<bf16 (*)[H][WS]>(lds);
    bf16 (*tile)[16][WS] = reinterpret_cast<bf16 (*)[16][WS]>(lds);
    static_assert(4 * 16 <= 2 * H, "output tile fits the weight staging");
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = l & 15, kg = l >> 4;
    const size_t rb = (size_t)blockIdx.x * 64 + w * 16;
    const int row = (int)rb + r;
    // weight loads first: 2 x 64 rows x 8 segments of 16 bytes, 4 per thread
    uint4 wv[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = threadIdx.x + 256 * j, m = q >> 9, wr = (q >> 3) & 63, seg = q & 7;
        wv[j] = *(const uint4 *)((m ? W2t : W1t) + (size_t)wr * H + seg * 8);
    }
    int a0 = 0, a1 = 0;
    if (row < n) {
        a0 = ptr[row];
        a1 = ptr[row + 1];
    }
    float m[16];
#pragma unroll
    for (int c = 0; c < 16; c++) m[c] = 0.f;
    for (int k = a0; k < a1; k++) {
        const bf16 *src = S + (size_t)idx[k] * H + 4 * kg;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const s4 v = *(const s4 *)(src + 16 * q);
#pragma unroll
            for (int c = 0; c < 4; c++) m[4 * q + c] += bf2f((bf16)v[c]);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = threadIdx.x + 256 * j, mm = q >> 9, wr = (q >> 3) & 63, seg = q & 7;
        *(uint4 *)&wl[mm][wr][seg * 8] = wv[j];
    }
    __syncthreads();
    f4 acc[4];
    for (int t = 0; t < 4; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < H; k0 += 16) {
        const s4 a = *(const s4 *)(X1 + (rb + r) * H + k0 + kg * 4);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const s4 b = *(const s4 *)&wl[0][t * 16 + r][k0 + kg * 4];
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    const int cnt = a1 - a0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        s4 a;
#pragma unroll
        for (int c = 0; c < 4; c++) a[c] = (short)f2bf(cnt > 0 ? m[4 * q + c] / (float)cnt : 0.f);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const s4 b = *(const s4 *)&wl[1][t * 16 + r][16 * q + kg * 4];
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    __syncthreads();  // every wave is done with the weights: the LDS becomes the output tile
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = t * 16 + r;
        const float bb = bias[col];
#pragma unroll
        for (int i = 0; i < 4; i++) tile[w][kg * 4 + i][col] = f2bf(fmaxf(acc[t][i] + bb, 0.f));
    }
    __syncthreads();
    const int rr = l >> 2, seg = (l & 3) * 16;
    const uint4 *src = (const uint4 *)&tile[w][rr][seg];
    uint4 *dst = (uint4 *)(Y + (rb + rr) * H + seg);
    dst[0] = src[0];
    dst[1] = src[1];
}

// dX[N][K1] (= or +=) dZ[N][64] W^T with W[K1][64]
template <int K1>
__global__ void __launch_bounds__(256) k_g_dx(const bf16 *__restrict__ dZ, const bf16 *__restrict__ W,
                                              float *__restrict__ dX, int accumulate) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = l & 15, kg = l >> 4;
    const size_t rb = (size_t)blockIdx.x * 64 + w * 16;
    constexpr int T = K1 / 16;
    f4 acc[T];
    for (int t = 0; t < T; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < H; k0 += 16) {
        const s4 a = *(const s4 *)(dZ + (rb + r) * H + k0 + kg * 4);
#pragma unroll
        for (int t = 0; t < T; t++) {
            const s4 b = *(const s4 *)(W + (size_t)(t * 16 + r) * H + k0 + kg * 4);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float *o = dX + (rb + kg * 4 + i) * K1 + t * 16 + r;
            *o = accumulate ? *o + acc[t][i] : acc[t][i];
        }
}

// partial weight gradients X^T dZ over kWgradRows rows per block:
// part[block][K1][64]; the 4 waves split the rows, LDS sums them in order
template <int K1>
__global__ void __launch_bounds__(256) k_g_wgrad(int N, const bf16 *__restrict__ X, const bf16 *__restrict__ dZ,
                                                 float *__restrict__ part) {
    constexpr int MT = K1 / 16;
    __shared__ float sh[4][K1][H];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = l & 15, kg = l >> 4;
    f4 acc[MT][4];
    for (int m = 0; m < MT; m++)
        for (int t = 0; t < 4; t++) acc[m][t] = f4{0.f, 0.f, 0.f, 0.f};
    const int n_begin = blockIdx.x * kWgradRows, n_end = min(N, n_begin + kWgradRows);
    for (int n0 = n_begin + w * 16; n0 < n_end; n0 += 64) {
        s4 bfr[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            s4 b;
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = (short)dZ[(size_t)(n0 + kg * 4 + j) * H + t * 16 + r];
            bfr[t] = b;
        }
#pragma unroll
        for (int m = 0; m < MT; m++) {
            s4 a;
#pragma unroll
            for (int j = 0; j < 4; j++) a[j] = (short)X[(size_t)(n0 + kg * 4 + j) * K1 + m * 16 + r];
#pragma unroll
            for (int t = 0; t < 4; t++) acc[m][t] = mfma(a, bfr[t], acc[m][t]);
        }
    }
    for (int m = 0; m < MT; m++)
        for (int t = 0; t < 4; t++)
            for (int i = 0; i < 4; i++) sh[w][m * 16 + kg * 4 + i][t * 16 + r] = acc[m][t][i];
    __syncthreads();
    for (int q = threadIdx.x; q < K1 * H; q += blockDim.x) {
        const int c = q / H, o = q % H;
        part[(size_t)blockIdx.x * K1 * H + q] = ((sh[0][c][o] + sh[1][c][o]) + sh[2][c][o]) + sh[3][c][o];
    }
}

// grad[q] += sum over blocks of part[b][q] (fixed order)
__global__ void k_g_reduce(int nblk, int n, const float *__restrict__ part, float *__restrict__ grad) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * n + q];
        grad[q] += s;
    }
}

// dZ = dY * (Y > 0) (BF16 for the GEMMs); bias gradient partials per block
__global__ void __launch_bounds__(256) k_g_relu_bwd(int N, const float *__restrict__ dY, const bf16 *__restrict__ Y,
                                                    bf16 *__restrict__ dZ, float *__restrict__ part) {
    __shared__ float sh[4][H];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    float s = 0.f;
    for (int n = blockIdx.x * 4 + w; n < N; n += gridDim.x * 4) {
        const size_t o = (size_t)n * H + l;
        const float z = bf2f(Y[o]) > 0.f ? dY[o] : 0.f;
        dZ[o] = f2bf(z);
        s += z;
    }
    sh[w][l] = s;
    __syncthreads();
    if (w == 0) part[(size_t)blockIdx.x * H + l] = ((sh[0][l] + sh[1][l]) + sh[2][l]) + sh[3][l];
}

// mean aggregation, one wave per target row, lane = feature
__global__ void k_g_mean(int n, const int *__restrict__ ptr, const int *__restrict__ idx, const bf16 *__restrict__ src,
                         bf16 *__restrict__ dst) {
    const int l = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        const int a = ptr[i], b = ptr[i + 1];
        float s = 0.f;
        for (int k = a; k < b; k++) s += bf2f(src[(size_t)idx[k] * H + l]);
        dst[(size_t)i * H + l] = f2bf(b > a ? s / (float)(b - a) : 0.f);
    }
}

// adjoint of the mean aggregation into `dst` rows: dst[i] (+)= sum over the
// sources k of row i of dM[k] / deg(k), deg from the other incidence
__global__ void k_g_mean_T(int n, const int *__restrict__ ptr, const int *__restrict__ idx,
                           const int *__restrict__ optr, const float *__restrict__ dM, float *__restrict__ dst,
                           int accumulate) {
    const int l = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        float s = 0.f;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = idx[k];
            s += dM[(size_t)j * H + l] / (float)(optr[j + 1] - optr[j]);
        }
        float *o = dst + (size_t)i * H + l;
        *o = accumulate ? *o + s : s;
    }
}

// phi = H3 w + c, one wave per edge
__global__ void k_g_out(int n, const bf16 *__restrict__ H3, const float *__restrict__ P, int ow, float *__restrict__ phi) {
    const int l = threadIdx.x & 63;
    const float wl = P[ow + l], c = P[ow + H];
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        float v = bf2f(H3[(size_t)i * H + l]) * wl;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (l == 0) phi[i] = v + c;
    }
}

// dH3 = dphi w; partial gradients of w (64) and c (1) per block
__global__ void __launch_bounds__(256) k_g_out_bwd(int n, const bf16 *__restrict__ H3, const float *__restrict__ dphi,
                                                   const float *__restrict__ P, int ow, float *__restrict__ dH3,
                                                   float *__restrict__ part) {
    __shared__ float sh[4][H + 1];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const float wl = P[ow + l];
    float gw = 0.f, gc = 0.f;
    for (int i = blockIdx.x * 4 + w; i < n; i += gridDim.x * 4) {
        const float d = dphi[i];
        dH3[(size_t)i * H + l] = d * wl;
        gw += bf2f(H3[(size_t)i * H + l]) * d;
        gc += d;
    }
    sh[w][l] = gw;
    if (l == 0) sh[w][H] = gc;
    __syncthreads();
    if (w == 0) {
        part[(size_t)blockIdx.x * (H + 1) + l] = ((sh[0][l] + sh[1][l]) + sh[2][l]) + sh[3][l];
        if (l == 0) part[(size_t)blockIdx.x * (H + 1) + H] = ((sh[0][H] + sh[1][H]) + sh[2][H]) + sh[3][H];
    }
}

// ln K statistics (mean, standard deviation), one block
// Mean and standard deviation of ln K over the elements, in two stages:
// kLkParts blocks write partial sums, one wave folds them in a fixed order.
constexpr int kLkParts = 128;
__global__ void __launch_bounds__(256) k_g_logk_part(int E, const double *__restrict__ K, double *__restrict__ part) {
    __shared__ double sh[2][4];
    double s1 = 0.0, s2 = 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const double v = log(K[e]);
        s1 += v;
        s2 += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
    }
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (l == 0) {
        sh[0][w] = s1;
        sh[1][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        part[2 * blockIdx.x + 1] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

__global__ void k_g_logk_stats(int E, const double *__restrict__ part, double *__restrict__ out) {
    double s1 = 0.0, s2 = 0.0;
    for (int k = threadIdx.x; k < kLkParts; k += 64) {
        s1 += part[2 * k];
        s2 += part[2 * k + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
    }
    if (threadIdx.x == 0) {
        const double mu = s1 / E;
        out[0] = mu;
        out[1] = sqrt(fmax(s2 / E - mu * mu, 1e-24));
    }
}

__global__ void k_g_feat_e(int E, const double *__restrict__ K, const double *__restrict__ cs,
                           const int *__restrict__ dim, const double *__restrict__ st, bf16 *__restrict__ xe) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const float k = (float)((log(K[e]) - st[0]) / st[1]);
        float f[NF] = {};
        f[0] = k;
        f[1] = k * k;
        f[2] = dim[e] == 1;
        f[3] = dim[e] == 2;
        f[4] = dim[e] == 3;
        f[5] = (float)(log(cs[e]) / 10.0);
        f[6] = 1.f;
        for (int j = 0; j < NF; j++) xe[(size_t)e * NF + j] = f2bf(f[j]);
    }
}

__global__ void k_g_to_bf(size_t n, const float *__restrict__ src, bf16 *__restrict__ dst) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = f2bf(src[i]);
}

// BF16 copies of one [K][64] weight matrix, plain and transposed
__global__ void k_g_pack(int K, const float *__restrict__ P, bf16 *__restrict__ W, bf16 *__restrict__ Wt) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < K * H; q += gridDim.x * blockDim.x) {
        const int k = q / H, o = q % H;
        const bf16 v = f2bf(P[q]);
        W[q] = v;
        Wt[o * K + k] = v;
    }
}

__global__ void k_g_adam(int n, float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m,
                         float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float g = G[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// coarse-fit loss, per aggregate (block): means, then centred sums
// st[a] = {ll, cc, nn}
__global__ void __launch_bounds__(256) k_g_loss_agg(const int *__restrict__ aptr, const int *__restrict__ aedge,
                                                    const float *__restrict__ phi, const double *__restrict__ lam,
                                                    double *__restrict__ st) {
    __shared__ double sh[3][4];
    const int a = blockIdx.x, b0 = aptr[a], b1 = aptr[a + 1];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    auto bsum = [&](double (&v)[3]) {
        for (int k = 0; k < 3; k++) {
            double x = v[k];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
            if (l == 0) sh[k][w] = x;
        }
        __syncthreads();
        for (int k = 0; k < 3; k++) v[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
        __syncthreads();
    };
    double v[3] = {0.0, 0.0, 0.0};
    for (int k = b0 + threadIdx.x; k < b1; k += blockDim.x) {
        const int f = aedge[k];
        v[0] += lam[f];
        v[1] += (double)phi[f];
    }
    bsum(v);
    const double n = (double)(b1 - b0);
    const double lm = v[0] / n, pm = v[1] / n;
    double u[3] = {0.0, 0.0, 0.0};
    for (int k = b0 + threadIdx.x; k < b1; k += blockDim.x) {
        const int f = aedge[k];
        const double lt = lam[f] - lm, pt = (double)phi[f] - pm;
        u[0] += lt * lt;
        u[1] += lt * pt;
        u[2] += pt * pt;
    }
    bsum(u);
    if (threadIdx.x == 0) {
        st[5 * a + 0] = u[0];
        st[5 * a + 1] = u[1];
        st[5 * a + 2] = u[2];
        st[5 * a + 3] = lm;
        st[5 * a + 4] = pm;
    }
}

// total: out[0] = loss, out[1] = sum of |l_a|^2 (one block, fixed order)
__global__ void k_g_loss_tot(int nagg, const double *__restrict__ st, double *__restrict__ out) {
    __shared__ double sh[2][4];
    double tot = 0.0, res = 0.0;
    for (int a = threadIdx.x; a < nagg; a += blockDim.x) {
        const double ll = st[5 * a], cc = st[5 * a + 1], nn = st[5 * a + 2];
        tot += ll;
        res += ll - (nn > 1e-30 ? cc * cc / nn : 0.0);
    }
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_down(tot, o, 64);
        res += __shfl_down(res, o, 64);
    }
    if (l == 0) {
        sh[0][w] = tot;
        sh[1][w] = res;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
        const double r = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
        out[0] = t > 0.0 ? r / t : 0.0;
        out[1] = t;
    }
}

__global__ void k_g_loss_grad(const int *__restrict__ aptr, const int *__restrict__ aedge,
                              const float *__restrict__ phi, const double *__restrict__ lam,
                              const double *__restrict__ st, const double *__restrict__ tot,
                              float *__restrict__ dphi) {
    const int a = blockIdx.x;
    const double cc = st[5 * a + 1], nn = st[5 * a + 2], lm = st[5 * a + 3], pm = st[5 * a + 4];
    const double c = nn > 1e-30 ? cc / nn : 0.0;
    const double T = tot[1] > 0.0 ? tot[1] : 1.0;
    for (int k = aptr[a] + threadIdx.x; k < aptr[a + 1]; k += blockDim.x) {
        const int f = aedge[k];
        const double lt = lam[f] - lm, pt = (double)phi[f] - pm;
        dphi[f] = (float)((-2.0 * c * lt + 2.0 * c * c * pt) / T);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <class T>
static T *g_up(std::vector<void *> &own, const std::vector<T> &v, hipStream_t s) {
    T *p = dev_upload(v.data(), v.size(), s);
    own.push_back(p);
    return p;
}

GnnGPU::GnnGPU(DarcyBatchGPU &batch, unsigned seed) : b_(batch) {
    FL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const HostMesh &m = batch.base().mesh();
    E_ = m.n_elements();
    F_ = m.n_edges();
    Ep_ = (E_ + 63) / 64 * 64;
    Fp_ = (F_ + 63) / 64 * 64;
    // element -> edges (own sides, vb edges) and edge -> elements
    std::vector<int> eptr(E_ + 1, 0), eedge;
    for (int e = 0; e < E_; e++) {
        for (int i = 0; i <= m.el_dim[e]; i++) eedge.push_back(m.side_edge[m.side_off[e] + i]);
        for (int k = m.el_vb_ptr[e]; k < m.el_vb_ptr[e + 1]; k++) eedge.push_back(m.vb_edge[m.el_vb[k]]);
        eptr[e + 1] = (int)eedge.size();
    }
    std::vector<int> fptr(F_ + 1, 0), fel(eedge.size());
    for (int f : eedge) fptr[f + 1]++;
    for (int f = 0; f < F_; f++) fptr[f + 1] += fptr[f];
    {
        std::vector<int> fp(fptr.begin(), fptr.end() - 1);
        for (int e = 0; e < E_; e++)
            for (int k = eptr[e]; k < eptr[e + 1]; k++) fel[fp[eedge[k]]++] = e;
    }
    // aggregates (edge order) and their edge lists
    std::vector<int> agg(F_);
    const auto &pe = batch.row_edges();
    const auto &ra = batch.row_aggregate();
    for (size_t r = 0; r < pe.size(); r++) agg[pe[r]] = ra[r];
    nagg_ = batch.aggregates();
    std::vector<int> aptr(nagg_ + 1, 0), aedge(F_);
    for (int f = 0; f < F_; f++) aptr[agg[f] + 1]++;
    for (int a = 0; a < nagg_; a++) aptr[a + 1] += aptr[a];
    {
        std::vector<int> fp(aptr.begin(), aptr.end() - 1);
        for (int f = 0; f < F_; f++) aedge[fp[agg[f]]++] = f;
    }
    // edge positions: mean of the centroids of the edge's elements
    std::vector<double> pos(3 * (size_t)F_, 0.0);
    for (int f = 0; f < F_; f++) {
        for (int k = fptr[f]; k < fptr[f + 1]; k++) {
            const int e = fel[k], nn = m.el_dim[e] + 1;
            for (int j = 0; j < nn; j++)
                for (int d = 0; d < 3; d++) pos[3 * f + d] += m.xyz[3 * (size_t)m.el_nodes[4 * (size_t)e + j] + d] / nn;
        }
        const int dg = std::max(1, fptr[f + 1] - fptr[f]);
        for (int d = 0; d < 3; d++) pos[3 * f + d] /= dg;
    }
    // Dirichlet edges (boundary sides with bc type 1)
    std::vector<int> bct(std::max(m.n_bc_elements(), 1), 0);
    if (m.n_bc_elements())
        FL_HIP(hipMemcpy(bct.data(), batch.base().dev().bc_type, m.n_bc_elements() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<float> dir(F_, 0.f);
    for (int e = 0; e < E_; e++)
        for (int i = 0; i <= m.el_dim[e]; i++) {
            const int b = m.side_bnd[m.side_off[e] + i];
            if (b >= 0 && bct[m.bnd_bcele[b]] == 1) dir[m.side_edge[m.side_off[e] + i]] = 1.f;
        }
    // edge features (K independent), restated in oracle/gnn.py edge_features
    std::vector<float> xf((size_t)Fp_ * NF, 0.f);
    {
        std::vector<double> cen(3 * (size_t)nagg_, 0.0), rad(nagg_, 0.0);
        for (int a = 0; a < nagg_; a++) {
            const int n = aptr[a + 1] - aptr[a];
            for (int k = aptr[a]; k < aptr[a + 1]; k++)
                for (int d = 0; d < 3; d++) cen[3 * a + d] += pos[3 * aedge[k] + d];
            for (int d = 0; d < 3; d++) cen[3 * a + d] /= std::max(n, 1);
            for (int k = aptr[a]; k < aptr[a + 1]; k++)
                for (int d = 0; d < 3; d++) {
                    const double q = pos[3 * aedge[k] + d] - cen[3 * a + d];
                    rad[a] += q * q;
                }
            rad[a] = std::sqrt(rad[a] / std::max(n, 1));
        }
        for (int f = 0; f < F_; f++) {
            const int a = agg[f];
            for (int d = 0; d < 3; d++) {
                xf[(size_t)f * NF + d] = (float)((pos[3 * f + d] - cen[3 * a + d]) / std::max(rad[a], 1e-30));
                xf[(size_t)f * NF + 3 + d] = (float)pos[3 * f + d];
            }
            xf[(size_t)f * NF + 6] = dir[f];
            xf[(size_t)f * NF + 7] = 1.f;
        }
    }
    std::vector<int> dim(m.el_dim.begin(), m.el_dim.end());
    hipStream_t s = stream_;
    d_eptr_ = g_up(owned_, eptr, s);
    d_eedge_ = g_up(owned_, eedge, s);
    d_fptr_ = g_up(owned_, fptr, s);
    d_fel_ = g_up(owned_, fel, s);
    d_aptr_ = g_up(owned_, aptr, s);
    d_aedge_ = g_up(owned_, aedge, s);
    d_dim_ = g_up(owned_, dim, s);
    d_Xf_ = g_up(owned_, xf, s);
    auto abf = [&](size_t n) {
        bf16 *p = dev_alloc<bf16>(n);
        owned_.push_back(p);
        FL_HIP(hipMemsetAsync(p, 0, n * sizeof(bf16), s));
        return p;
    };
    auto af = [&](size_t n) {
        float *p = dev_alloc<float>(n);
        owned_.push_back(p);
        FL_HIP(hipMemsetAsync(p, 0, n * sizeof(float), s));
        return p;
    };
    d_xe_ = abf((size_t)Ep_ * NF);
    d_xf_ = abf((size_t)Fp_ * NF);
    hipLaunchKernelGGL(k_g_to_bf, dim3(1024), dim3(256), 0, s, (size_t)Fp_ * NF, (const float *)d_Xf_, d_xf_);
    for (int l = 0; l < 3; l++) {
        d_He_[l] = abf((size_t)Ep_ * H);
        d_Me_[l] = abf((size_t)Ep_ * H);
    }
    for (int l = 0; l < 4; l++) {
        d_Hf_[l] = abf((size_t)Fp_ * H);
        d_Mf_[l] = abf((size_t)Fp_ * H);
    }
    const size_t mx = (size_t)std::max(Ep_, Fp_);
    d_phi_ = af(Fp_);
    for (int k = 0; k < 2; k++) {
        d_dHf_[k] = af((size_t)Fp_ * H);
        d_dHe_[k] = af((size_t)Ep_ * H);
    }
    d_dM_ = af(mx * H);
    d_dZ_ = abf(mx * H);
    d_wpart_ = af(((mx + kWgradRows - 1) / kWgradRows + 1) * (size_t)H * H + 4096 * (H + 1));
    d_lossbuf_ = af(Fp_);
    d_stats_ = dev_alloc<double>(5 * (size_t)nagg_ + 8);
    owned_.push_back(d_stats_);
    d_lkpart_ = dev_alloc<double>(2 * (size_t)kLkParts);
    owned_.push_back(d_lkpart_);
    // parameters (oracle/gnn.py param_shapes order)
    std::vector<std::pair<int, int>> shp = {{NF, H}, {H, 0}, {NF, H}, {H, 0}};
    for (int l = 1; l <= 3; l++) {
        shp.push_back({H, H});
        shp.push_back({H, H});
        shp.push_back({H, 0});
        if (l < 3) {
            shp.push_back({H, H});
            shp.push_back({H, H});
            shp.push_back({H, 0});
        }
    }
    shp.push_back({H, 0});
    shp.push_back({1, 0});
    np_ = 0;
    off_.clear();
    for (auto &p : shp) {
        off_.push_back(np_);
        np_ += p.second ? p.first * p.second : p.first;
    }
    std::vector<float> P(np_, 0.f);
    std::mt19937 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (size_t k = 0; k < shp.size(); k++) {
        const int rows = shp[k].first, cols = shp[k].second;
        if (cols) {
            const float sc = std::sqrt(2.f / rows);
            for (int q = 0; q < rows * cols; q++) P[off_[k] + q] = nd(rng) * sc;
        }
    }
    const int ow = off_[off_.size() - 2];
    for (int q = 0; q < H; q++) P[ow + q] = nd(rng) * std::sqrt(1.f / H);
    d_P_ = g_up(owned_, P, s);
    d_G_ = af(np_);
    d_m_ = af(np_);
    d_v_ = af(np_);
    d_W_ = abf(np_);
    d_Wt_ = abf(np_);
    FL_HIP(hipStreamSynchronize(s));
    pack_weights();
}

GnnGPU::~GnnGPU() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void *p : owned_) dev_free(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void GnnGPU::pack_weights() {
    // matrices: We0, Wf0, then A1, A2 (and B1, B2) per layer
    std::vector<std::pair<int, int>> mats = {{off_[0], NF}, {off_[2], NF}};
    int k = 4;
    for (int l = 1; l <= 3; l++) {
        mats.push_back({off_[k], H});
        mats.push_back({off_[k + 1], H});
        k += 3;
        if (l < 3) {
            mats.push_back({off_[k], H});
            mats.push_back({off_[k + 1], H});
            k += 3;
        }
    }
    for (auto &mm : mats)
        hipLaunchKernelGGL(k_g_pack, dim3(16), dim3(256), 0, stream_, mm.second, (const float *)(d_P_ + mm.first),
                           d_W_ + mm.first, d_Wt_ + mm.first);
    FL_HIP(hipGetLastError());
    FL_HIP(hipStreamSynchronize(stream_));
}

void GnnGPU::get_params(float *out) const {
    FL_HIP(hipMemcpy(out, d_P_, np_ * sizeof(float), hipMemcpyDeviceToHost));
}

void GnnGPU::set_params(const float *in) {
    FL_HIP(hipMemcpy(d_P_, in, np_ * sizeof(float), hipMemcpyHostToDevice));
    pack_weights();
}

void GnnGPU::zero_grad() {
    FL_HIP(hipMemsetAsync(d_G_, 0, np_ * sizeof(float), stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

void GnnGPU::get_grad(float *out) const {
    FL_HIP(hipMemcpy(out, d_G_, np_ * sizeof(float), hipMemcpyDeviceToHost));
}

void GnnGPU::get_features(float *xe, float *xf) const {
    std::vector<bf16> t((size_t)std::max(Ep_, Fp_) * NF);
    auto cvt = [&](const bf16 *d, int n, float *o) {
        FL_HIP(hipMemcpy(t.data(), d, (size_t)n * NF * sizeof(bf16), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)n * NF; i++) {
            unsigned u = (unsigned)t[i] << 16;
            std::memcpy(&o[i], &u, 4);
        }
    };
    if (xe) cvt(d_xe_, E_, xe);
    if (xf) cvt(d_xf_, F_, xf);
}

void GnnGPU::forward(const double *dK, float *dphi_out) {
    hipStream_t s = stream_;
    const dim3 blk(256);
    hipLaunchKernelGGL(k_g_logk_part, dim3(kLkParts), dim3(256), 0, s, E_, dK, d_lkpart_);
    hipLaunchKernelGGL(k_g_logk_stats, dim3(1), dim3(64), 0, s, E_, (const double *)d_lkpart_,
                       d_stats_ + 5 * (size_t)nagg_);
    hipLaunchKernelGGL(k_g_feat_e, dim3(std::min(grid_for(E_), 2048)), blk, 0, s, E_, dK, b_.base().dev().cs,
                       (const int *)d_dim_, (const double *)(d_stats_ + 5 * (size_t)nagg_), d_xe_);
    const float *P = d_P_;
    // input layers
    hipLaunchKernelGGL((k_g_fwd<NF, 0>), dim3(Ep_ / 64), blk, 0, s, (const bf16 *)d_xe_, (const bf16 *)(d_Wt_ + off_[0]),
                       (const bf16 *)nullptr, (const bf16 *)nullptr, P + off_[1], d_He_[0]);
    hipLaunchKernelGGL((k_g_fwd<NF, 0>), dim3(Fp_ / 64), blk, 0, s, (const bf16 *)d_xf_, (const bf16 *)(d_Wt_ + off_[2]),
                       (const bf16 *)nullptr, (const bf16 *)nullptr, P + off_[3], d_Hf_[0]);
    const dim3 gF(std::min((F_ + 3) / 4, 8192)), gE(std::min((E_ + 3) / 4, 8192));
    int k = 4;
    for (int l = 1; l <= 3; l++) {
        hipLaunchKernelGGL(k_g_mean, gF, blk, 0, s, F_, (const int *)d_fptr_, (const int *)d_fel_,
                           (const bf16 *)d_He_[l - 1], d_Mf_[l]);
        hipLaunchKernelGGL((k_g_fwd<H, H>), dim3(Fp_ / 64), blk, 0, s, (const bf16 *)d_Hf_[l - 1],
                           (const bf16 *)(d_Wt_ + off_[k]), (const bf16 *)d_Mf_[l], (const bf16 *)(d_Wt_ + off_[k + 1]),
                           P + off_[k + 2], d_Hf_[l]);
        k += 3;
        if (l < 3) {
            hipLaunchKernelGGL(k_g_mean, gE, blk, 0, s, E_, (const int *)d_eptr_, (const int *)d_eedge_,
                               (const bf16 *)d_Hf_[l], d_Me_[l]);
            hipLaunchKernelGGL((k_g_fwd<H, H>), dim3(Ep_ / 64), blk, 0, s, (const bf16 *)d_He_[l - 1],
                               (const bf16 *)(d_Wt_ + off_[k]), (const bf16 *)d_Me_[l],
                               (const bf16 *)(d_Wt_ + off_[k + 1]), P + off_[k + 2], d_He_[l]);
            k += 3;
        }
    }
    hipLaunchKernelGGL(k_g_out, gF, blk, 0, s, F_, (const bf16 *)d_Hf_[3], P, off_[off_.size() - 2], d_phi_);
    FL_HIP(hipGetLastError());
    if (dphi_out) FL_HIP(hipMemcpyAsync(dphi_out, d_phi_, F_ * sizeof(float), hipMemcpyDeviceToDevice, s));
    FL_HIP(hipStreamSynchronize(s));
}

double GnnGPU::loss_backward(const double *dlambda) {
    hipStream_t s = stream_;
    const dim3 blk(256);
    double *st = d_stats_;
    double *tot = d_stats_ + 5 * (size_t)nagg_ + 2;
    float *dphi = d_lossbuf_;
    hipLaunchKernelGGL(k_g_loss_agg, dim3(nagg_), blk, 0, s, (const int *)d_aptr_, (const int *)d_aedge_,
                       (const float *)d_phi_, dlambda, st);
    hipLaunchKernelGGL(k_g_loss_tot, dim3(1), blk, 0, s, nagg_, (const double *)st, tot);
    hipLaunchKernelGGL(k_g_loss_grad, dim3(nagg_), blk, 0, s, (const int *)d_aptr_, (const int *)d_aedge_,
                       (const float *)d_phi_, dlambda, (const double *)st, (const double *)tot, dphi);
    const float *P = d_P_;
    const int ow = off_[off_.size() - 2];
    // output layer
    const int nbo = 1024;
    hipLaunchKernelGGL(k_g_out_bwd, dim3(nbo), blk, 0, s, F_, (const bf16 *)d_Hf_[3], (const float *)dphi, P, ow,
                       d_dHf_[1], d_wpart_);
    hipLaunchKernelGGL(k_g_reduce, dim3(1), dim3(128), 0, s, nbo, H + 1, (const float *)d_wpart_, d_G_ + ow);
    // pad rows of the gradient buffers stay zero (k_g_out_bwd writes rows < F only)
    auto relu_bwd = [&](int N, int Np, const float *dY, const bf16 *Y, int boff) {
        const int nb = 512;
        hipLaunchKernelGGL(k_g_relu_bwd, dim3(nb), blk, 0, s, Np, dY, Y, d_dZ_, d_wpart_);
        hipLaunchKernelGGL(k_g_reduce, dim3(1), dim3(64), 0, s, nb, H, (const float *)d_wpart_, d_G_ + boff);
        (void)N;
    };
    auto wgrad = [&](int Np, const bf16 *X, int K1, int woff) {
        const int nb = (Np + kWgradRows - 1) / kWgradRows;
        if (K1 == NF) hipLaunchKernelGGL(k_g_wgrad<NF>, dim3(nb), blk, 0, s, Np, X, (const bf16 *)d_dZ_, d_wpart_);
        else hipLaunchKernelGGL(k_g_wgrad<H>, dim3(nb), blk, 0, s, Np, X, (const bf16 *)d_dZ_, d_wpart_);
        hipLaunchKernelGGL(k_g_reduce, dim3((K1 * H + 255) / 256), blk, 0, s, nb, K1 * H, (const float *)d_wpart_,
                           d_G_ + woff);
    };
    // parameter offsets per layer
    int ka[4], kb[3];
    {
        int k = 4;
        for (int l = 1; l <= 3; l++) {
            ka[l] = k;
            k += 3;
            if (l < 3) {
                kb[l] = k;
                k += 3;
            }
        }
    }
    const dim3 gF(std::min((F_ + 3) / 4, 8192)), gE(std::min((E_ + 3) / 4, 8192));
    // dHf[cur] holds dL/dHf(l); dHe[cur] holds dL/dHe(l) (for l < 3)
    float *dHf_l = d_dHf_[1], *dHf_p = d_dHf_[0];
    float *dHe_l = d_dHe_[1], *dHe_p = d_dHe_[0];
    for (int l = 3; l >= 1; l--) {
        bool he_set = false;  // dHe(l-1) written already
        if (l < 3) {
            // element update of layer l
            relu_bwd(E_, Ep_, dHe_l, d_He_[l], off_[kb[l] + 2]);
            wgrad(Ep_, d_He_[l - 1], H, off_[kb[l]]);
            wgrad(Ep_, d_Me_[l], H, off_[kb[l] + 1]);
            hipLaunchKernelGGL(k_g_dx<H>, dim3(Ep_ / 64), blk, 0, s, (const bf16 *)d_dZ_,
                               (const bf16 *)(d_W_ + off_[kb[l]]), dHe_p, 0);
            he_set = true;
            hipLaunchKernelGGL(k_g_dx<H>, dim3(Ep_ / 64), blk, 0, s, (const bf16 *)d_dZ_,
                               (const bf16 *)(d_W_ + off_[kb[l] + 1]), d_dM_, 0);
            // dHf(l) += adjoint of Me = mean over the edges of an element
            hipLaunchKernelGGL(k_g_mean_T, gF, blk, 0, s, F_, (const int *)d_fptr_, (const int *)d_fel_,
                               (const int *)d_eptr_, (const float *)d_dM_, dHf_l, 1);
        }
        // edge update of layer l
        relu_bwd(F_, Fp_, dHf_l, d_Hf_[l], off_[ka[l] + 2]);
        wgrad(Fp_, d_Hf_[l - 1], H, off_[ka[l]]);
        wgrad(Fp_, d_Mf_[l], H, off_[ka[l] + 1]);
        hipLaunchKernelGGL(k_g_dx<H>, dim3(Fp_ / 64), blk, 0, s, (const bf16 *)d_dZ_, (const bf16 *)(d_W_ + off_[ka[l]]),
                           dHf_p, 0);
        hipLaunchKernelGGL(k_g_dx<H>, dim3(Fp_ / 64), blk, 0, s, (const bf16 *)d_dZ_,
                           (const bf16 *)(d_W_ + off_[ka[l] + 1]), d_dM_, 0);
        // dHe(l-1) (+)= adjoint of Mf = mean over the elements of an edge
        hipLaunchKernelGGL(k_g_mean_T, gE, blk, 0, s, E_, (const int *)d_eptr_, (const int *)d_eedge_,
                           (const int *)d_fptr_, (const float *)d_dM_, dHe_p, he_set ? 1 : 0);
        std::swap(dHf_l, dHf_p);
        std::swap(dHe_l, dHe_p);
    }
    // input layers: dHe_l = dL/dHe0, dHf_l = dL/dHf0
    relu_bwd(E_, Ep_, dHe_l, d_He_[0], off_[1]);
    wgrad(Ep_, d_xe_, NF, off_[0]);
    relu_bwd(F_, Fp_, dHf_l, d_Hf_[0], off_[3]);
    wgrad(Fp_, d_xf_, NF, off_[2]);
    FL_HIP(hipGetLastError());
    double out[2];
    FL_HIP(hipMemcpyAsync(out, tot, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return out[0];
}

void GnnGPU::adam_step(double lr) {
    adam_t_++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t_), c2 = 1.f - std::pow(b2, (float)adam_t_);
    hipLaunchKernelGGL(k_g_adam, dim3((np_ + 255) / 256), dim3(256), 0, stream_, np_, d_P_, (const float *)d_G_, d_m_,
                       d_v_, (float)lr, b1, b2, 1e-8f, c1, c2);
    FL_HIP(hipGetLastError());
    pack_weights();
}

// Inference for R realizations: the forward pass without the stored
// aggregation tensors (k_g_fwd_mean), phi written straight into dphi, one
// stream synchronization at the end.
void GnnGPU::infer(const double *dK, int R, float *dphi) {
    hipStream_t s = stream_;
    const dim3 blk(256);
    const float *P = d_P_;
    const dim3 gF(std::min((F_ + 3) / 4, 8192));
    for (int r = 0; r < R; r++) {
        const double *K = dK + (size_t)r * E_;
        hipLaunchKernelGGL(k_g_logk_part, dim3(kLkParts), dim3(256), 0, s, E_, K, d_lkpart_);
        hipLaunchKernelGGL(k_g_logk_stats, dim3(1), dim3(64), 0, s, E_, (const double *)d_lkpart_,
                           d_stats_ + 5 * (size_t)nagg_);
        hipLaunchKernelGGL(k_g_feat_e, dim3(std::min(grid_for(E_), 2048)), blk, 0, s, E_, K, b_.base().dev().cs,
                           (const int *)d_dim_, (const double *)(d_stats_ + 5 * (size_t)nagg_), d_xe_);
        hipLaunchKernelGGL((k_g_fwd<NF, 0>), dim3(Ep_ / 64), blk, 0, s, (const bf16 *)d_xe_,
                           (const bf16 *)(d_Wt_ + off_[0]), (const bf16 *)nullptr, (const bf16 *)nullptr,
                           P + off_[1], d_He_[0]);
        hipLaunchKernelGGL((k_g_fwd<NF, 0>), dim3(Fp_ / 64), blk, 0, s, (const bf16 *)d_xf_,
                           (const bf16 *)(d_Wt_ + off_[2]), (const bf16 *)nullptr, (const bf16 *)nullptr,
                           P + off_[3], d_Hf_[0]);
        int k = 4;
        for (int l = 1; l <= 3; l++) {
            hipLaunchKernelGGL(k_g_fwd_mean, dim3(Fp_ / 64), blk, 0, s, F_, (const bf16 *)d_Hf_[l - 1],
                               (const bf16 *)(d_Wt_ + off_[k]), (const int *)d_fptr_, (const int *)d_fel_,
                               (const bf16 *)d_He_[l - 1], (const bf16 *)(d_Wt_ + off_[k + 1]), P + off_[k + 2],
                               d_Hf_[l]);
            k += 3;
            if (l < 3) {
                hipLaunchKernelGGL(k_g_fwd_mean, dim3(Ep_ / 64), blk, 0, s, E_, (const bf16 *)d_He_[l - 1],
                                   (const bf16 *)(d_Wt_ + off_[k]), (const int *)d_eptr_, (const int *)d_eedge_,
                                   (const bf16 *)d_Hf_[l], (const bf16 *)(d_Wt_ + off_[k + 1]), P + off_[k + 2],
                                   d_He_[l]);
                k += 3;
            }
        }
        hipLaunchKernelGGL(k_g_out, gF, blk, 0, s, F_, (const bf16 *)d_Hf_[3], P, off_[off_.size() - 2],
                           dphi + (size_t)r * F_);
    }
    FL_HIP(hipGetLastError());
    FL_HIP(hipStreamSynchronize(s));
}

}  // namespace fl
