// Hermitian contraction C = (X diag(sx)) Y^H in split BF16 on the matrix
// cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate) -- the precision form of
// the TCC build (calcTCC accumulate, litho_thread.cpp L490-520) and of the
// source-space Gram matrix.
//
// Every FP64 operand v is split into hi = bf16(v) and lo = bf16(v - hi); a real
// product a*b is formed as ah*bh + ah*bl + al*bh (the lo*lo term, 2^-16
// relative, is dropped).  bf16 keeps FP32's exponent range, so lo never
// underflows for the values met here.
//
// Work split: one block per 64 x 64 upper-triangle tile (tj >= ti) and per
// K-slice; each block writes its FP32 partial tile, and a reduction kernel
// sums the slices in FP64 and writes both triangles of the FP64 result.
#include <hip/hip_runtime.h>

#include "tcc_kernels.hpp"

namespace litho {

typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BT = 64;       // output tile
constexpr int BK = 32;       // K per LDS stage (one MFMA k-step)
constexpr int LROW = BK + 8; // padded LDS row (bf16 elements)

__device__ __forceinline__ void split_bf16(double v, __bf16& hi, __bf16& lo) {
    hi = (__bf16)(float)v;
    lo = (__bf16)(float)(v - (double)(float)hi);
}

// planes: 0 re hi, 1 re lo, 2 im hi, 3 im lo
template <bool IMAG>
__global__ __launch_bounds__(256) void herm_gemm_bf16s_kernel(const double2* __restrict__ X,
                                                              const double* __restrict__ sx,
                                                              const double2* __restrict__ Y,
                                                              int n, int m, int kslice,
                                                              float2* __restrict__ part) {
    const int nt = (n + BT - 1) / BT;
    long bid = blockIdx.x;
    int ti = 0;
    {
        long rowlen = nt;
        while (bid >= rowlen) {
            bid -= rowlen;
            ti++;
            rowlen--;
        }
    }
    const int tj = ti + (int)bid;
    const int ks = blockIdx.y;
    const int k0 = ks * kslice, k1 = min(m, k0 + kslice);
    constexpr int NP = IMAG ? 4 : 2;
    __shared__ __attribute__((aligned(16))) __bf16 Xs[NP][BT][LROW];
    __shared__ __attribute__((aligned(16))) __bf16 Ys[NP][BT][LROW];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const int wi = wave >> 1, wj = wave & 1;  // 32 x 32 quadrant of the tile
    const int i0 = ti * BT, j0 = tj * BT;
    f4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            cr[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
            ci[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
        }
    for (int s0 = k0; s0 < k1; s0 += BK) {
        for (int e = tid; e < BT * BK; e += 256) {
            const int r = e / BK, c = e - r * BK;
            const int s = s0 + c;
            double2 xv = make_double2(0.0, 0.0), yv = make_double2(0.0, 0.0);
            if (s < k1) {
                if (i0 + r < n) {
                    xv = X[(long)(i0 + r) * m + s];
                    if (sx) {
                        const double f = sx[s];
                        xv.x *= f;
                        xv.y *= f;
                    }
                }
                if (j0 + r < n) yv = Y[(long)(j0 + r) * m + s];
            }
            split_bf16(xv.x, Xs[0][r][c], Xs[1][r][c]);
            split_bf16(yv.x, Ys[0][r][c], Ys[1][r][c]);
            if (IMAG) {
                split_bf16(xv.y, Xs[2][r][c], Xs[3][r][c]);
                split_bf16(yv.y, Ys[2][r][c], Ys[3][r][c]);
            }
        }
        __syncthreads();
        b8 xa[2][NP], yb[2][NP];
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int p = 0; p < NP; p++) {
                xa[q][p] = *(const b8*)&Xs[p][wi * 32 + q * 16 + r16][8 * g];
                yb[q][p] = *(const b8*)&Ys[p][wj * 32 + q * 16 + r16][8 * g];
            }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                // Cr += Xr Yr (+ Xi Yi); Ci += Xi Yr - Xr Yi, each split 3 ways
                f4 c = cr[a][b];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][0], yb[b][0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][0], yb[b][1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][1], yb[b][0], c, 0, 0, 0);
                if (IMAG) {
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][2], yb[b][2], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][2], yb[b][3], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][3], yb[b][2], c, 0, 0, 0);
                    f4 d = ci[a][b];
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][2], yb[b][0], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][2], yb[b][1], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][3], yb[b][0], d, 0, 0, 0);
                    // - Xr Yi: negate the product through the accumulator
                    d = -d;
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][0], yb[b][2], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][0], yb[b][3], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a][1], yb[b][2], d, 0, 0, 0);
                    ci[a][b] = -d;
                }
                cr[a][b] = c;
            }
        __syncthreads();
    }
    // partial tile [ks][i][j] (full n x n layout, upper-triangle tiles only)
    float2* P = part + (long)ks * n * n;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + wi * 32 + a * 16 + 4 * g + v;
                const int j = j0 + wj * 32 + b * 16 + r16;
                if (i < n && j < n) P[(long)i * n + j] = make_float2(cr[a][b][v], ci[a][b][v]);
            }
}

// C[i][j] (j >= i) = sum of the K-slice partials in FP64; mirrors to C[j][i].
__global__ __launch_bounds__(256) void herm_reduce_kernel(const float2* __restrict__ part,
                                                          int n, int nks,
                                                          double2* __restrict__ C) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e / n), j = (int)(e - (long)i * n);
    if (j < i) return;
    // only upper-triangle tiles were written; (i, j) with j >= i lies in one
    if ((j / BT) < (i / BT)) return;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nks; k++) {
        const float2 v = part[(long)k * n * n + e];
        re += v.x;
        im += v.y;
    }
    if (i == j) {
        C[e] = make_double2(re, 0.0);
    } else {
        C[e] = make_double2(re, im);
        C[(long)j * n + i] = make_double2(re, -im);
    }
}

int herm_gemm_bf16s_slices(int n, int m) {
    const long nt = (n + BT - 1) / BT;
    const long tiles = nt * (nt + 1) / 2;
    // enough blocks to fill the chip, slices of at least 4 LDS stages
    long ks = (1024 + tiles - 1) / tiles;
    const long kmax = (m + 4 * BK - 1) / (4 * BK);
    if (ks > kmax) ks = kmax;
    if (ks < 1) ks = 1;
    return (int)ks;
}

int launch_herm_gemm_bf16s(const double2* X, const double* sx, const double2* Y, int n, int m,
                           int has_imag, float2* part, int nks, double2* C, hipStream_t st) {
    if (n <= 0 || m <= 0 || nks <= 0) return LITHO_EARG;
    const long nt = (n + BT - 1) / BT;
    const long tiles = nt * (nt + 1) / 2;
    int kslice = (m + nks - 1) / nks;
    kslice = (kslice + BK - 1) / BK * BK;
    const dim3 grid((unsigned)tiles, (unsigned)nks);
    if (has_imag)
        hipLaunchKernelGGL(herm_gemm_bf16s_kernel<true>, grid, dim3(256), 0, st, X, sx, Y, n, m,
                           kslice, part);
    else
        hipLaunchKernelGGL(herm_gemm_bf16s_kernel<false>, grid, dim3(256), 0, st, X, sx, Y, n, m,
                           kslice, part);
    if (hipGetLastError() != hipSuccess) return LITHO_EHIP;
    const long nn = (long)n * n;
    hipLaunchKernelGGL(herm_reduce_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, st, part, n, nks, C);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

}  // namespace litho
