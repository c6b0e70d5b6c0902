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
// window M [B][W][W] x bank [nk][W][W] -> intensity [B][G][G] (shifted if shift)
int launch_socs_accum(int G, const float2* M, const float2* bank, const float* scales, int nk,
                      int Nx, int B, float* out, int shift, const float2* tw, hipStream_t st);
// FI second axis: P [B][G][band+1] -> U [B][N][band+1]; yalt: P rows came
// from a centred (fftshifted) grid
int launch_fi_cols(int G, int N, const float2* P, float2* U, int B, int band, int yalt,
                   const float2* tw, hipStream_t st);
// FI last axis: U [nrows][band+1] -> image [nrows][N]
int launch_image_rows(int N, const float2* U, float* img, long nrows, int band, const float2* tw,
                      hipStream_t st);

}  // namespace litho
