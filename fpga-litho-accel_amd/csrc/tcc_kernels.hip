// TCC build kernels for gfx950 (Module 4, calcTCC, litho_thread.cpp L441-532).
//
//   pupil_kernel   : A[ID][s] = P(f_ID + s) for every live source point, with
//                    the reference's float arithmetic (L458-484), no FMA
//                    contraction, so the 0/1 support is bit-identical.
//   herm_gemm_mfma : C = (X diag(sx)) Y^H in FP64 on v_mfma_f64_16x16x4f64,
//                    upper block triangle only, lower mirrored as conj (the
//                    reference fills its lower triangle the same way, L527-532).
//                    TCC = A diag(S) A^H; Gram = Bt Bt^H with Bt = conj(A)^T
//                    diag(sqrt S).
//   conj_transpose_scale : Bt[s][i] = conj(A[i][s]) * sqrt(S[s]).
#include "common.hpp"
#include <algorithm>

#include "tcc_kernels.hpp"

namespace litho {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// pupil: one thread per (ID, s).  srcs[s] = {p, q}; geometry in `g`.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pupil_kernel(const int2* __restrict__ srcs, int nsrc,
                                                    PupilGeom g, double2* __restrict__ A,
                                                    int* __restrict__ range_err) {
#pragma clang fp contract(off)
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int W = 2 * g.Nx + 1;
    const long n = (long)W * (2 * g.Ny + 1);
    if (idx >= n * nsrc) return;
    const int s = (int)(idx % nsrc);
    const int ID = (int)(idx / nsrc);
    const int nx = ID % W - g.Nx;
    const int ny = ID / W - g.Ny;
    const int p = srcs[s].x, q = srcs[s].y;
    const float sx = (float)p / (float)g.sh;
    const float sy = (float)q / (float)g.sh;
    const int nxMin = (int)((-1.0f - sx) * g.Lx_norm);
    const int nxMax = (int)((1.0f - sx) * g.Lx_norm);
    const int nyMin = (int)((-1.0f - sy) * g.Ly_norm);
    const int nyMax = (int)((1.0f - sy) * g.Ly_norm);
    double2 v = make_double2(0.0, 0.0);
    if (nx >= nxMin && nx <= nxMax && ny >= nyMin && ny <= nyMax) {
        const float fy = (float)ny / g.Ly_norm + sy;
        if (fy * fy <= 1.0f) {
            const float fx = (float)nx / g.Lx_norm + sx;
            const float rho2 = fx * fx + fy * fy;
            if (rho2 <= 1.0f) {
                const float na2 = rho2 * g.NA * g.NA;
                const float ph = (float)((double)(g.dz * g.k) * sqrt(1.0 - (double)na2));
                v = make_double2((double)cosf(ph), (double)sinf(ph));
            }
        }
    }
    A[(long)ID * nsrc + s] = v;
    // The reference's loops can reach (nx,ny) outside the window; detect that
    // (only the edge IDs can see it, so test the box against the window).
    if (ID == 0 && (nxMin < -g.Nx || nxMax > g.Nx || nyMin < -g.Ny || nyMax > g.Ny)) {
        // out-of-window points matter only if they pass the disk test; check them
        for (int yy = nyMin; yy <= nyMax; yy++) {
            const float fy = (float)yy / g.Ly_norm + sy;
            if (!(fy * fy <= 1.0f)) continue;
            for (int xx = nxMin; xx <= nxMax; xx++) {
                if (xx >= -g.Nx && xx <= g.Nx && yy >= -g.Ny && yy <= g.Ny) continue;
                const float fx = (float)xx / g.Lx_norm + sx;
                if (fx * fx + fy * fy <= 1.0f) atomicOr(range_err, 1);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Bt[s][i] = conj(A[i][s]) * sqrt(S[s])  (A: [n][m], Bt: [m][n])
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conj_transpose_scale_kernel(const double2* __restrict__ A,
                                                                   const double* __restrict__ S,
                                                                   int n, int m,
                                                                   double2* __restrict__ Bt) {
    __shared__ double2 tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int i0 = blockIdx.y * 32, s0 = blockIdx.x * 32;
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, s = s0 + tx;
        tile[r][tx] = (i < n && s < m) ? A[(long)i * m + s] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int s = s0 + r, i = i0 + tx;
        if (s < m && i < n) {
            double2 a = tile[tx][r];
            double f = sqrt(S[s]);
            Bt[(long)s * n + i] = make_double2(a.x * f, -a.y * f);
        }
    }
}

// ---------------------------------------------------------------------------
// C = (X diag(sx)) Y^H, X: [n][m], Y: [n][m] complex double, C: [n][n].
// Block = 4 waves computing a 32 x 32 tile (each wave 16 x 16), K staged in
// LDS 16 at a time.  Only blocks with tj >= ti run; they also write the
// conjugate mirror.  MFMA f64 16x16x4 lane layout (measured on gfx950): A
// operand lane l holds A[l%16][l/16], B operand B[l/16][l%16], result
// register v of lane l is C[l/16 + 4v][l%16].
// ---------------------------------------------------------------------------
constexpr int TG_T = 32;  // tile
constexpr int TG_K = 16;  // K stage

// blockIdx.y splits K into slices of kslice: with more than one slice, slice
// z writes its upper-triangle tile sums to P[z] (n x n) and herm_sum_kernel
// adds them up and mirrors.
__global__ __launch_bounds__(256) void herm_gemm_mfma_kernel(const double2* __restrict__ X,
                                                             const double* __restrict__ sx,
                                                             const double2* __restrict__ Y, int n,
                                                             int m, int has_imag, int kslice,
                                                             double2* __restrict__ C,
                                                             double2* __restrict__ P) {
    const int nt = (n + TG_T - 1) / TG_T;
    // linear block index -> (ti, tj) with tj >= ti
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
    __shared__ double2 Xs[TG_T][TG_K + 1];
    __shared__ double2 Ys[TG_T][TG_K + 1];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1;
    const int i0 = ti * TG_T, j0 = tj * TG_T;
    d4 cr = {0.0, 0.0, 0.0, 0.0};
    d4 ci = {0.0, 0.0, 0.0, 0.0};
    const int sbeg = blockIdx.y * kslice, send = min(m, sbeg + kslice);
    for (int s0 = sbeg; s0 < send; s0 += TG_K) {
        for (int e = tid; e < TG_T * TG_K; e += 256) {
            int r = e / TG_K, c = e - r * TG_K;
            int s = s0 + c;
            double2 xv = make_double2(0.0, 0.0), yv = make_double2(0.0, 0.0);
            if (s < send) {
                if (i0 + r < n) {
                    xv = X[(long)(i0 + r) * m + s];
                    if (sx) {
                        double f = sx[s];
                        xv.x *= f;
                        xv.y *= f;
                    }
                }
                if (j0 + r < n) yv = Y[(long)(j0 + r) * m + s];
            }
            Xs[r][c] = xv;
            Ys[r][c] = yv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TG_K; kk += 4) {
            double2 a = Xs[wi * 16 + (lane & 15)][kk + (lane >> 4)];
            double2 b = Ys[wj * 16 + (lane & 15)][kk + (lane >> 4)];
            // C += X Y^H : Cr += Xr Yr + Xi Yi ; Ci += Xi Yr - Xr Yi
            cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, cr, 0, 0, 0);
            if (has_imag) {
                cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, cr, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, ci, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.x, b.y, ci, 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int i = i0 + wi * 16 + (lane >> 4) + 4 * v;
        int j = j0 + wj * 16 + (lane & 15);
        if (i >= n || j >= n) continue;
        if (ti == tj && j < i) continue;  // diagonal tile: upper part only
        double re = cr[v], im = ci[v];
        if (P) {
            P[(long)blockIdx.y * n * n + (long)i * n + j] = make_double2(re, im);
            continue;
        }
        if (i == j) {
            C[(long)i * n + j] = make_double2(re, 0.0);
        } else {
            C[(long)i * n + j] = make_double2(re, im);
            C[(long)j * n + i] = make_double2(re, -im);
        }
    }
}

// ---------------------------------------------------------------------------
// Real-pupil contraction (zero defocus: the pupil matrix is 0/1, so X and Y
// are real and C = (X diag(sx)) Y^T is real symmetric): the real parts only,
// 64 x 64 block tiles of four 32 x 32 wave tiles (2 x 2 MFMA f64 16x16x4
// each, A and B fragments reused twice), K staged 16 at a time through two
// LDS buffers with the next stage's global loads in flight during the MFMAs,
// one barrier per stage.  Against the 32 x 32 complex tiles above: a quarter
// of the L2 reads per output and four MFMAs per fragment load pair.
// ---------------------------------------------------------------------------
constexpr int TR_T = 64;  // block tile
constexpr int TR_K = 16;  // K stage

__global__ __launch_bounds__(256) void herm_gemm_real_kernel(const double2* __restrict__ X,
                                                             const double* __restrict__ sx,
                                                             const double2* __restrict__ Y, int n,
                                                             int m, int kslice,
                                                             double2* __restrict__ C,
                                                             double2* __restrict__ P) {
    const int nt = (n + TR_T - 1) / TR_T;
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
    __shared__ double Xs[2][TR_T][TR_K + 1];
    __shared__ double Ys[2][TR_T][TR_K + 1];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1;
    const int i0 = ti * TR_T, j0 = tj * TR_T;
    const int sbeg = blockIdx.y * kslice, send = min(m, sbeg + kslice);
    // staging: element e = tid + 256 q (q < 4) of a stage is row e / 16, term
    // e % 16: sixteen lanes read 256 contiguous bytes of one row
    const int lc = tid & 15, lr0 = tid >> 4;
    double xr[4], yr[4];
    auto gload = [&](int s0) {
        const int sidx = s0 + lc;
        const bool ok = sidx < send;
        const int sc = ok ? sidx : sbeg;  // clamped: every load in flight at once
        const double f = sx ? sx[sc] : 1.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = lr0 + 16 * q;
            const bool xin = i0 + r < n, yin = j0 + r < n;
            const double xv = X[(long)(xin ? i0 + r : 0) * m + sc].x * f;
            const double yv = Y[(long)(yin ? j0 + r : 0) * m + sc].x;
            xr[q] = ok && xin ? xv : 0.0;
            yr[q] = ok && yin ? yv : 0.0;
        }
    };
    auto sstore = [&](int bf) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Xs[bf][lr0 + 16 * q][lc] = xr[q];
            Ys[bf][lr0 + 16 * q][lc] = yr[q];
        }
    };
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    if (sbeg < send) {
        gload(sbeg);
        sstore(0);
    }
    __syncthreads();
    int bf = 0;
    for (int s0 = sbeg; s0 < send; s0 += TR_K) {
        const bool more = s0 + TR_K < send;
        if (more) gload(s0 + TR_K);
#pragma unroll
        for (int kk = 0; kk < TR_K; kk += 4) {
            const int k = kk + (lane >> 4), r = lane & 15;
            const double a0 = Xs[bf][wi * 32 + r][k], a1 = Xs[bf][wi * 32 + 16 + r][k];
            const double b0 = Ys[bf][wj * 32 + r][k], b1 = Ys[bf][wj * 32 + 16 + r][k];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (more) sstore(bf ^ 1);
        __syncthreads();
        bf ^= 1;
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + wi * 32 + 16 * a + (lane >> 4) + 4 * v;
                const int j = j0 + wj * 32 + 16 * b + (lane & 15);
                if (i >= n || j >= n) continue;
                if (ti == tj && j < i) continue;  // diagonal tile: upper part only
                const double re = acc[a][b][v];
                if (P) {
                    P[(long)blockIdx.y * n * n + (long)i * n + j] = make_double2(re, 0.0);
                    continue;
                }
                C[(long)i * n + j] = make_double2(re, 0.0);
                if (i != j) C[(long)j * n + i] = make_double2(re, 0.0);
            }
}

int launch_pupil(const int2* srcs, int nsrc, const PupilGeom& g, double2* A, int* range_err,
                 hipStream_t st) {
    long n = (long)(2 * g.Nx + 1) * (2 * g.Ny + 1);
    long total = n * nsrc;
    if (total <= 0) return LITHO_EARG;
    hipLaunchKernelGGL(pupil_kernel, dim3(cdiv(total, 256)), dim3(256), 0, st, srcs, nsrc, g, A,
                       range_err);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_conj_transpose_scale(const double2* A, const double* S, int n, int m, double2* Bt,
                                hipStream_t st) {
    dim3 grid(cdiv(m, 32), cdiv(n, 32));
    hipLaunchKernelGGL(conj_transpose_scale_kernel, grid, dim3(256), 0, st, A, S, n, m, Bt);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

// C[i][j] (i <= j) = sum_z P[z][i][j], mirrored below the diagonal
__global__ __launch_bounds__(256) void herm_sum_kernel(const double2* __restrict__ P, int ns,
                                                       int n, double2* __restrict__ C) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    if (j < i) return;
    double re = 0.0, im = 0.0;
    for (int z = 0; z < ns; z++) {
        const double2 v = P[(long)z * n * n + e];
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

int herm_gemm_slices(int n, int m, int has_imag) {
    const long tile = has_imag ? TG_T : TR_T;
    const long nt = (n + tile - 1) / tile;
    const long nblocks = nt * (nt + 1) / 2;
    // about 4096 blocks in flight (2048 of the larger real tiles), slices of
    // at least 256 terms
    const long want = has_imag ? 4096 : 2048;
    long ns = (want + nblocks - 1) / nblocks;
    ns = std::min(ns, (long)std::max(1, m / 256));
    ns = std::min(ns, 64L);
    return (int)std::max(1L, ns);
}

int launch_herm_gemm(const double2* X, const double* sx, const double2* Y, int n, int m,
                     int has_imag, double2* C, double2* P, int nslices, hipStream_t st) {
    if (n <= 0 || m <= 0 || nslices <= 0 || (nslices > 1 && !P)) return LITHO_EARG;
    const long tile = has_imag ? TG_T : TR_T;
    long nt = (n + tile - 1) / tile;
    long nblocks = nt * (nt + 1) / 2;
    const int kslice = (int)(((long)m + nslices - 1) / nslices + TG_K - 1) / TG_K * TG_K;
    const int ns = (m + kslice - 1) / kslice;
    if (has_imag)
        hipLaunchKernelGGL(herm_gemm_mfma_kernel, dim3((unsigned)nblocks, ns), dim3(256), 0, st, X,
                           sx, Y, n, m, has_imag, kslice, C, ns > 1 ? P : nullptr);
    else
        hipLaunchKernelGGL(herm_gemm_real_kernel, dim3((unsigned)nblocks, ns), dim3(256), 0, st, X,
                           sx, Y, n, m, kslice, C, ns > 1 ? P : nullptr);
    if (ns > 1)
        hipLaunchKernelGGL(herm_sum_kernel, dim3(cdiv((long)n * n, 256)), dim3(256), 0, st, P, ns,
                           n, C);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}


// ---------------------------------------------------------------------------
// TCC-direct image spectrum (calcImage, litho_thread.cpp L539-583), FP64:
//   C[nx2][ny2 + 2Ny] = sum_{n1} M(n2+n1) conj(M(n1)) TCC[n1][n2+n1]
// for ny2 in [-2Ny, 2Ny], nx2 in [0, 2Nx], n1 and n2+n1 inside the window.
// M is the (2Ny+1) x (2Nx+1) centred mask spectrum window (float2); one
// block per output, the n1 sum reduced across the block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void tcc_image_spectrum_kernel(const float2* __restrict__ M,
                                                                 const double2* __restrict__ T,
                                                                 int Nx, int Ny,
                                                                 float2* __restrict__ C) {
    __shared__ double sr[256], si[256];
    const int W = 2 * Nx + 1, H = 2 * Ny + 1, n = W * H;
    const int nx2 = blockIdx.x, ny2 = (int)blockIdx.y - 2 * Ny;
    double ar = 0.0, ai = 0.0;
    for (int id1 = threadIdx.x; id1 < n; id1 += 256) {
        const int nx1 = id1 % W - Nx, ny1 = id1 / W - Ny;
        const int mx = nx2 + nx1, my = ny2 + ny1;
        if (mx < -Nx || mx > Nx || my < -Ny || my > Ny) continue;
        const int id2 = (my + Ny) * W + (mx + Nx);
        const float2 a = M[id2], b = M[id1];
        // (a * conj(b)) * T[id1][id2]
        const double pr = (double)a.x * b.x + (double)a.y * b.y;
        const double pi = (double)a.y * b.x - (double)a.x * b.y;
        const double2 t = T[(long)id1 * n + id2];
        ar += pr * t.x - pi * t.y;
        ai += pr * t.y + pi * t.x;
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sr[threadIdx.x] += sr[threadIdx.x + st];
            si[threadIdx.x] += si[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) C[(long)nx2 * (4 * Ny + 1) + blockIdx.y] = make_float2((float)sr[0], (float)si[0]);
}

int launch_tcc_image_spectrum(const float2* M, const double2* T, int Nx, int Ny, float2* C,
                              hipStream_t st) {
    dim3 grid(2 * Nx + 1, 4 * Ny + 1);
    hipLaunchKernelGGL(tcc_image_spectrum_kernel, grid, dim3(256), 0, st, M, T, Nx, Ny, C);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

// The column pass of the TCC-direct image on any Ly: from the band spectrum
// C [2Nx+1][4Ny+1] (kx2 = 0..2Nx, ky2 = -2Ny..2Ny, as the spectrum kernel
// writes it) to V [Ly][2Nx+1] = s (-1)^(kx2 + ky2) sum_ky2 C e^{+2 pi i ky2 y / Ly}.
// The (-1) factors put the phase origin at the clip centre (the fftshift of
// FT_c2r_padded, litho_thread.cpp L371-425); the row pass is a plain c2r.
__global__ __launch_bounds__(256) void tcc_image_cols_kernel(const float2* __restrict__ C, int Nx,
                                                             int Ny, int Ly, float s,
                                                             float2* __restrict__ V) {
    const int nb = 2 * Nx + 1, nk = 4 * Ny + 1;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)Ly * nb) return;
    const int y = (int)(e / nb), kx2 = (int)(e - (long)y * nb);
    double ar = 0.0, ai = 0.0;
    for (int j = 0; j < nk; j++) {
        const int ky2 = j - 2 * Ny;
        const long ph = (((long)ky2 * y) % Ly + Ly) % Ly;
        double sn, cs;
        sincospi(2.0 * (double)ph / Ly, &sn, &cs);
        const float2 c = C[(long)kx2 * nk + j];
        const double sg = (ky2 & 1) ? -1.0 : 1.0;
        ar += sg * (c.x * cs - c.y * sn);
        ai += sg * (c.x * sn + c.y * cs);
    }
    const double f = ((kx2 & 1) ? -1.0 : 1.0) * s;
    V[e] = make_float2((float)(f * ar), (float)(f * ai));
}

int launch_tcc_image_cols(const float2* C, int Nx, int Ny, int Ly, float s, float2* V,
                          hipStream_t st) {
    const long n = (long)Ly * (2 * Nx + 1);
    hipLaunchKernelGGL(tcc_image_cols_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       C, Nx, Ny, Ly, s, V);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

}  // namespace litho
