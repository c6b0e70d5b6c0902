// Host launchers of the TCC build kernels (tcc_kernels.hip).
#pragma once

#include "common.hpp"

namespace litho {

// calcTCC scalars (litho_thread.cpp L443-448), all in the reference's types.
struct PupilGeom {
    int sh;         // (srcSize-1)/2
    int Nx, Ny;
    float dz;       // defocus / (NA*NA/lambda)
    float k;        // (float)(2*pi/lambda)
    float Lx_norm;  // Lx*NA/lambda
    float Ly_norm;
    float NA;
};

int launch_pupil(const int2* srcs, int nsrc, const PupilGeom& g, double2* A, int* range_err,
                 hipStream_t st);
int launch_conj_transpose_scale(const double2* A, const double* S, int n, int m, double2* Bt,
                                hipStream_t st);
// C = (X diag(sx)) Y^H (sx may be null), Hermitian output, n x n; the m terms
// split into nslices K slices (herm_gemm_slices) whose partial sums go
// through P (nslices * n * n complex doubles; unused for one slice); without
// imaginary parts (has_imag = 0) the real-pupil kernel reads X.x, Y.x only
int herm_gemm_slices(int n, int m, int has_imag);
int launch_herm_gemm(const double2* X, const double* sx, const double2* Y, int n, int m,
                     int has_imag, double2* C, double2* P, int nslices, hipStream_t st);
// Same contraction in split BF16 on the matrix cores (tcc_bf16.hip): part is
// an FP32 scratch of nks * n * n float2 (nks from herm_gemm_bf16s_slices).
int herm_gemm_bf16s_slices(int n, int m);
int launch_herm_gemm_bf16s(const double2* X, const double* sx, const double2* Y, int n, int m,
                           int has_imag, float2* part, int nks, double2* C, hipStream_t st);

// TCC-direct image band spectrum C [(2Nx+1)][(4Ny+1)] (kx-major) from the mask window M
// [(2Ny+1)][(2Nx+1)] and the n x n TCC (calcImage)
int launch_tcc_image_spectrum(const float2* M, const double2* T, int Nx, int Ny, float2* C,
                              hipStream_t st);
// band spectrum C (as launch_tcc_image_spectrum writes it) -> V [Ly][2Nx+1]:
// the inverse column DFT with the centring signs, times s
int launch_tcc_image_cols(const float2* C, int Nx, int Ny, int Ly, float s, float2* V,
                          hipStream_t st);

}  // namespace litho
