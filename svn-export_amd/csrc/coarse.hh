// Batched inverse of the two-level preconditioner's coarse operators
// A_c = Z^T S Z (one dense SPD matrix per realization, order nc <= 2048),
// written for the GPU: right-looking blocked Cholesky A = L L^T over 64 x 64
// FP64 tiles, W = L^-1 by tile columns, then A^-1 = W^T W tile by tile.  The
// inverse leaves as the packed lower tiles in FP32 that the coarse solve of
// every CG iteration streams (half the bytes of the full matrix).
//
// Reference: the coarse space has no counterpart in the reference (its
// solver is PETSc's, src/la/solve.cc:370); the factorization replaces a
// vendor library call (LAPACK potrf/potri semantics: lower triangle in,
// symmetric inverse out, per-matrix failure flag for a matrix that is not
// positive definite).
#pragma once
#include "common.hh"

namespace fl {
namespace coarse {

constexpr int kT = 64;        // tile order
constexpr int kTP = kT + 1;   // LDS row stride (doubles) of a staged tile
constexpr int kThreads = 256; // 16 x 16 threads, 4 x 4 outputs each
constexpr int kMaxTiles = 32; // coarse order <= 2048

// number of packed lower tiles of an nt x nt tile matrix, and the index of
// tile (i, j), i >= j
__host__ __device__ inline int ntiles_lower(int nt) { return nt * (nt + 1) / 2; }
__host__ __device__ inline int tile_index(int i, int j) { return i * (i + 1) / 2 + j; }
__device__ inline void tile_of(int t, int &i, int &j) {
    int r = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) r++;
    while (r * (r + 1) / 2 > t) r--;
    i = r;
    j = t - r * (r + 1) / 2;
}

// 64 x 64 tile at `src` (row stride ld) into LDS, as is or transposed
template <bool TRANS>
__device__ __forceinline__ void load_tile(double (*dst)[kTP], const double *__restrict__ src, size_t ld) {
#pragma unroll 4
    for (int idx = threadIdx.x; idx < kT * kT; idx += kThreads) {
        const int r = idx >> 6, c = idx & 63;
        const double v = src[(size_t)r * ld + c];
        if (TRANS) dst[c][r] = v;
        else dst[r][c] = v;
    }
}

// acc[a][b] += sum_k As[r0 + a][k] * Bs[k][c0 + b]
__device__ __forceinline__ void tile_mac(const double (*As)[kTP], const double (*Bs)[kTP], int r0, int c0,
                                         double (&acc)[4][4]) {
#pragma unroll 4
    for (int k = 0; k < kT; k++) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++) av[a] = As[r0 + a][k];
#pragma unroll
        for (int b = 0; b < 4; b++) bv[b] = Bs[k][c0 + b];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
}

__device__ __forceinline__ void acc_zero(double (&acc)[4][4]) {
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
}

// Diagonal tile k: unblocked Cholesky of A_kk in LDS (lower), then the
// triangular inverse W_kk = L_kk^-1 by columns.  Both are written back as full
// tiles (zero upper part).  A pivot that is not positive (or NaN) marks the
// matrix as failed; its factor is then garbage and discarded.
__global__ void __launch_bounds__(kThreads) k_ch_potrf(double *A, double *W, int ncp, int k, int *info) {
    __shared__ double L[kT][kTP], Wi[kT][kTP];
    __shared__ int bad;
    const int tid = threadIdx.x;
    const size_t off = (size_t)blockIdx.x * ncp * ncp + (size_t)k * kT * ncp + (size_t)k * kT;
    load_tile<false>(L, A + off, ncp);
    if (tid == 0) bad = 0;
    for (int j = 0; j < kT; j++) {
        __syncthreads();  // the previous step's trailing update is complete
        const double d = L[j][j];
        const double s = d > 0.0 ? sqrt(d) : 1.0;
        __syncthreads();  // every thread has read the pivot
        if (tid == 0) {
            if (!(d > 0.0)) bad = 1;
            L[j][j] = s;
        }
        if (tid > j && tid < kT) L[tid][j] /= s;
        __syncthreads();
        for (int idx = tid; idx < kT * kT; idx += kThreads) {
            const int i = idx >> 6, l = idx & 63;
            if (l > j && l <= i) L[i][l] -= L[i][j] * L[l][j];
        }
    }
    __syncthreads();
    if (tid < kT) {
        const int c = tid;
        for (int i = 0; i < c; i++) Wi[i][c] = 0.0;
        Wi[c][c] = 1.0 / L[c][c];
        for (int i = c + 1; i < kT; i++) {
            double s = 0.0;
            for (int m = c; m < i; m++) s += L[i][m] * Wi[m][c];
            Wi[i][c] = -s / L[i][i];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < kT * kT; idx += kThreads) {
        const int r = idx >> 6, c = idx & 63;
        A[off + (size_t)r * ncp + c] = c <= r ? L[r][c] : 0.0;
        W[off + (size_t)r * ncp + c] = Wi[r][c];
    }
    if (tid == 0 && bad) info[blockIdx.x] = 1;
}

// Panel below tile k: L_ik = A_ik L_kk^-T = A_ik W_kk^T, i = k + 1 + blockIdx.x.
__global__ void __launch_bounds__(kThreads) k_ch_trsm(double *A, const double *W, int ncp, int k) {
    __shared__ double As[kT][kTP], Bs[kT][kTP];
    const int i = k + 1 + blockIdx.x, tid = threadIdx.x;
    const size_t m0 = (size_t)blockIdx.y * ncp * ncp;
    double *Aik = A + m0 + (size_t)i * kT * ncp + (size_t)k * kT;
    load_tile<false>(As, Aik, ncp);
    load_tile<true>(Bs, W + m0 + (size_t)k * kT * ncp + (size_t)k * kT, ncp);
    __syncthreads();
    const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    double acc[4][4];
    acc_zero(acc);
    tile_mac(As, Bs, r0, c0, acc);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) Aik[(size_t)(r0 + a) * ncp + c0 + b] = acc[a][b];
}

// Trailing update after step k: A_ij -= L_ik L_jk^T for k < j <= i, one
// block per lower tile of the trailing matrix.
__global__ void __launch_bounds__(kThreads) k_ch_syrk(double *A, int ncp, int k) {
    __shared__ double As[kT][kTP], Bs[kT][kTP];
    int ii, jj;
    tile_of(blockIdx.x, ii, jj);
    const int i = k + 1 + ii, j = k + 1 + jj, tid = threadIdx.x;
    const size_t m0 = (size_t)blockIdx.y * ncp * ncp;
    load_tile<false>(As, A + m0 + (size_t)i * kT * ncp + (size_t)k * kT, ncp);
    load_tile<true>(Bs, A + m0 + (size_t)j * kT * ncp + (size_t)k * kT, ncp);
    __syncthreads();
    const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    double acc[4][4];
    acc_zero(acc);
    tile_mac(As, Bs, r0, c0, acc);
    double *Aij = A + m0 + (size_t)i * kT * ncp + (size_t)j * kT;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) Aij[(size_t)(r0 + a) * ncp + c0 + b] -= acc[a][b];
}

// W = L^-1 below the diagonal, tile column k = blockIdx.x:
// W_ik = -W_ii (sum_{m=k}^{i-1} L_im W_mk), i = k+1 .. nt-1 in order (each
// uses the column's tiles written before it by this block).
__global__ void __launch_bounds__(kThreads) k_ch_trtri(const double *A, double *W, int ncp, int nt) {
    __shared__ double As[kT][kTP], Bs[kT][kTP];
    const int k = blockIdx.x, tid = threadIdx.x;
    const size_t m0 = (size_t)blockIdx.y * ncp * ncp;
    const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    for (int i = k + 1; i < nt; i++) {
        double acc[4][4];
        acc_zero(acc);
        for (int m = k; m < i; m++) {
            __syncthreads();
            load_tile<false>(As, A + m0 + (size_t)i * kT * ncp + (size_t)m * kT, ncp);
            load_tile<false>(Bs, W + m0 + (size_t)m * kT * ncp + (size_t)k * kT, ncp);
            __syncthreads();
            tile_mac(As, Bs, r0, c0, acc);
        }
        __syncthreads();
        load_tile<false>(As, W + m0 + (size_t)i * kT * ncp + (size_t)i * kT, ncp);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) Bs[r0 + a][c0 + b] = acc[a][b];
        __syncthreads();
        double w[4][4];
        acc_zero(w);
        tile_mac(As, Bs, r0, c0, w);
        double *Wik = W + m0 + (size_t)i * kT * ncp + (size_t)k * kT;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) Wik[(size_t)(r0 + a) * ncp + c0 + b] = -w[a][b];
        __threadfence_block();  // the next tiles of this column read it back
    }
}

// Inverse tiles X_ij = sum_{m=i}^{nt-1} W_mi^T W_mj (i >= j), rounded to FP32
// into the packed lower tiles of the matrix's slot (pairs[2 * matrix]); a
// failed factorization leaves zeros (that slot's coarse correction vanishes,
// the preconditioner falls back to Jacobi).
__global__ void __launch_bounds__(kThreads) k_ch_lauum(const double *W, int ncp, int nt, const int *__restrict__ info,
                                                       const int *__restrict__ pairs, float *__restrict__ ainv) {
    __shared__ double As[kT][kTP], Bs[kT][kTP];
    int i, j;
    tile_of(blockIdx.x, i, j);
    const int tid = threadIdx.x, mat = blockIdx.y;
    const size_t m0 = (size_t)mat * ncp * ncp;
    const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
    double acc[4][4];
    acc_zero(acc);
    for (int m = i; m < nt; m++) {
        __syncthreads();
        load_tile<true>(As, W + m0 + (size_t)m * kT * ncp + (size_t)i * kT, ncp);
        load_tile<false>(Bs, W + m0 + (size_t)m * kT * ncp + (size_t)j * kT, ncp);
        __syncthreads();
        tile_mac(As, Bs, r0, c0, acc);
    }
    const bool ok = info[mat] == 0;
    const int slot = pairs[2 * mat];
    float *out = ainv + ((size_t)slot * ntiles_lower(nt) + blockIdx.x) * (kT * kT);
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const float4 v = ok ? make_float4((float)acc[a][0], (float)acc[a][1], (float)acc[a][2], (float)acc[a][3])
                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        *reinterpret_cast<float4 *>(out + (r0 + a) * kT + c0) = v;
    }
}

__global__ void k_ch_zero_info(int *info, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) info[i] = 0;
}

// Inverse of np SPD matrices A[np][ncp][ncp] (row-major, lower triangle
// read; ncp a multiple of 64, rows/columns past the true order hold an
// identity block).  A is overwritten by L, W ([np][ncp][ncp]) receives L^-1;
// info[k] != 0 marks a matrix that is not positive definite.
inline void invert_batched(double *A, double *W, int ncp, int np, int *info, const int *pairs, float *ainv,
                           hipStream_t s) {
    if (np <= 0) return;
    if (ncp % kT) throw Error(-1, "coarse order must be padded to a multiple of 64");
    const int nt = ncp / kT;
    const dim3 blk(kThreads);
    hipLaunchKernelGGL(k_ch_zero_info, dim3((np + 255) / 256), dim3(256), 0, s, info, np);
    for (int k = 0; k < nt; k++) {
        hipLaunchKernelGGL(k_ch_potrf, dim3(np), blk, 0, s, A, W, ncp, k, info);
        const int m = nt - 1 - k;
        if (m > 0) {
            hipLaunchKernelGGL(k_ch_trsm, dim3(m, np), blk, 0, s, A, (const double *)W, ncp, k);
            hipLaunchKernelGGL(k_ch_syrk, dim3(ntiles_lower(m), np), blk, 0, s, A, ncp, k);
        }
    }
    if (nt > 1) hipLaunchKernelGGL(k_ch_trtri, dim3(nt - 1, np), blk, 0, s, (const double *)A, W, ncp, nt);
    hipLaunchKernelGGL(k_ch_lauum, dim3(ntiles_lower(nt), np), blk, 0, s, (const double *)W, ncp, nt,
                       (const int *)info, pairs, ainv);
    FL_HIP(hipGetLastError());
}

}  // namespace coarse
}  // namespace fl
