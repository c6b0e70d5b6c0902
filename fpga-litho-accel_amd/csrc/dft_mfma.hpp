// Band DFTs of mask rows and image rows on the fp16 matrix cores
// (dft_mfma.hip).  Launchers return a LITHO_* status.
#pragma once

#include <vector>

#include "common.hpp"

namespace litho {

// fp16 cos/sin tables in MFMA fragment order (uploaded once per shape)
std::vector<unsigned short> mask_rows_table(int N, int kmin, int count);
std::vector<unsigned short> image_rows_table(int N, int nb);

// mask [B][N][N] -> R [B][count][N] (half2 if h16, else float2): window kx in
// [kmin, kmin + count) of each row's DFT, (-1)^kx applied.  mfmt: 0 f32 masks,
// 1 u8 levels (mask = level / 255), 2 one bit per pixel (pixel 8j + i at bit i
// of byte j of the row, N / 8 bytes per row).  N a multiple of 256, count <= 80.
int launch_mask_rows_mfma(int N, const void* mask, int mfmt, const void* table, void* R, int B,
                          int count, bool h16, hipStream_t st);
// U [B][nb][N] float2 (kx = 0..nb-1 of each image row) -> image [B][N][N]
// (fp16 if out16, else f32), the c2r sum with the DC imaginary part dropped.
// usplit: U in a compact form, per image the first ns =
// image_rows_split_cols(nb) columns as [ns][N] float2 (split into hi + lo here),
// then [nb - ns][N] half2 (the fp16 values the product uses for them).
int launch_image_rows_mfma(int N, const void* U, int nb, const void* table, void* img, int B,
                           bool out16, bool usplit, hipStream_t st);
// number of U columns (kx = 0..ns-1) the image-side product keeps as hi + lo
int image_rows_split_cols(int nb);

}  // namespace litho
