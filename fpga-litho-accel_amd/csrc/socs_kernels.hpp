// Host launchers of the SOCS pipeline kernels (socs_kernels.hip).  Each returns
// a LITHO_* status; sizes outside the compiled set return LITHO_EARG.
#pragma once

#include "common.hpp"

namespace litho {

// real rows -> window of their DFT, out[r][j] = sgn * DFT(row r)[kmin + j]
int launch_rows_dft(int L, const float* in, float2* out, long nrows, int kmin, int count,
                    bool alt, const float2* tw, hipStream_t st);
// row-DFT window R [B][N][cx] -> centred window M [B][cy][cx], /N^2
int launch_mask_cols(int N, const float2* R, float2* M, int B, int cx, int kymin, int cy,
                     const float2* tw, hipStream_t st);
// conv-grid rows handled per block for a grid size / window
int socs_rows_per_block(int G, int Nx);
// window M [B][W][W] x bank [nk][W][W] -> intensity [B][G][G] (shifted if shift);
// bank_fmt as launch_socs_band (16-bit formats on grids of 512 points and up)
// With Yt (socs_pre_bytes bytes, grids of 512 points and up) the column pass
// runs once per window column into Yt and the row blocks read their rows from
// there; without it every row block repeats the column pass.
int launch_socs_accum(int G, const float2* M, const void* bank, const float2* kscale,
                      int bank_fmt, const float* scales, int nk, int Nx, int B, float* out,
                      int shift, const float2* tw, hipStream_t st, float2* Yt = nullptr);
// bytes of the column-pass buffer of launch_socs_accum (0 below 512 points)
size_t socs_pre_bytes(int G, int Nx, int nk, int B);
// FI second axis: P [B][G][band+1] -> U [B][N][band+1]; yalt: P rows came
// from a centred (fftshifted) grid
int launch_fi_cols(int G, int N, const float2* P, float2* U, int B, int band, int yalt,
                   const float2* tw, hipStream_t st);
// FI last axis: U [nrows][band+1] -> image [nrows][N]
int launch_image_rows(int N, const float2* U, float* img, long nrows, int band, const float2* tw,
                      hipStream_t st);
// 4096^2 clips with the row-DFT window and the FI column output stored
// transposed per clip: R [B][count][4096] (|kx| < 256), U [B][band+1][4096].
int launch_rows4096_tr(const float* in, float2* out, int B, int kmin, int count, bool alt,
                       hipStream_t st);
int launch_mask_cols4096_tr(const float2* R, float2* M, int B, int cx, int kymin, int cy,
                            hipStream_t st);
int launch_fi_cols4096_tr(int G, const float2* P, float2* U, int B, int band, int yalt,
                          const float2* tw, hipStream_t st);
int launch_image4096_tr(const float2* U, float* img, int B, int band, hipStream_t st);

// Band pipeline (image path, square conv grid G = P*P in {36, 81, 144}); h16
// ... lean: smaller blocks (less LDS) so SOCS and FFT blocks can share a CU;
// selects half2 storage of the R / U intermediates (mixed-precision form).
// smallest such grid >= 4Nx+1, or 0 when none applies.
int band_grid(int Nx);
// mask rows -> transposed row-DFT window R [B][cx][N] (window kx in [kmin, kmin+count))
// in: f32 masks, or uint8 levels (mask = level / 255) when u8
int launch_rows_dft_t(int L, const void* in, void* out, long nrows, int kmin, int count, bool h16,
                      bool lean, bool u8, hipStream_t st);
// R [B][cx][N] -> M [B][cy][cx]; mirror mode mw = 2h+1 > 0: R holds kx = 0..h
// of a real mask (cx = h+1, kymin = -h, cy = mw) and M [B][mw][mw] is filled
// through M(-ky, -kx) = conj(M(ky, kx))
int launch_mask_cols_t(int N, const void* R, float2* M, int B, int cx, int kymin, int cy, bool h16,
                       bool lean, hipStream_t st, int mw = 0);
// M x bank -> band spectrum C [B][2Nx+1][4Nx+1] of the intensity (kx-major: column kx
// holds ky = -2Nx..2Nx contiguously), /G^2.
// bank_fmt 0: float2; 1: bf16 (re, im) pairs; 2: fp16 pairs; kscale[k] scales the real (x)
// and imaginary (y) parts of entry k.
// shape 0: 48-slot blocks, 1: 24-slot (lean) blocks at G = 144, 2: wave-synced
// transforms (socs_wave_kernel); ksplit = 2 (shape 2 only): each image's
// entries over two blocks, the second half's band spectrum in a second plane
// of C, cplane float2 after the first
int launch_socs_band(int G, const float2* M, const void* bank, const float2* kscale, int bank_fmt,
                     const float* scales, int nk, int Nx, int B, float2* C, int shape,
                     hipStream_t st, int ksplit = 1, long cplane = 0);
// C -> U [B][2Nx+1][N]; nsum = 2 adds the plane cplane float2 after C
int launch_fi_cols_t(int N, const float2* C, void* U, int B, int band, bool h16, bool lean,
                     hipStream_t st, int nsum = 1, long cplane = 0);
// C -> U in the compact form of launch_image_rows_mfma (usplit): FP32 for
// the ns split columns, fp16 above; N = 1024 with 2 band + 1 <= 160 only
// (else LITHO_EARG)
int launch_fi_cols_split(int N, const float2* C, void* U, int B, int band, int ns,
                         hipStream_t st, int nsum = 1, long cplane = 0);
// U [B][band+1][N] -> image [B*N][N]
// img: f32, or fp16 when out16
int launch_image_rows_t(int N, const void* U, void* img, long nrows, int band, bool h16,
                        bool lean, bool out16, hipStream_t st);

// window [-m, m]^2 of a centred spectrum given as Lx x Ly planes -> M [W][W]
// float2 (zero outside the planes)
int launch_spectrum_window(const float* re, const float* im, int Lx, int Ly, int m, float2* M,
                           hipStream_t st);
// unshifted G x G intensity -> fftshifted Gy x Gx samples (G / Gx, G / Gy
// integers) written to o1 and o2 (either may be null)
int launch_grid_sample_shift(const float* T, int G, int Gx, int Gy, float* o1, float* o2,
                             hipStream_t st);

}  // namespace litho
