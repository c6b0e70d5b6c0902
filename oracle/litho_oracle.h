/*
 * litho_oracle.h -- FP64 CPU restatement of the SOCS aerial-image path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (and the timed CPU baseline).  Nothing in the product
 * library links or calls it.
 *
 * Every function restates one function of the reference's CPU oracle
 * (validation/golden/src/litho_thread.cpp and its common/ sources); the citation is
 * given on each declaration.  FFTs are this file's own mixed-radix FFT
 * (the reference uses FFTW); the Hermitian eigensolver is this file's own
 * cyclic Jacobi (the reference calls LAPACK zheevr).
 *
 * Conventions (identical to the reference):
 *   - arrays are row-major, complex arrays are split into re/im planes;
 *   - kernel index ID(nx,ny) = (ny+Ny)*(2Nx+1) + (nx+Nx);
 *   - kernels are stored index-flipped: K[ID] = v[n-1-ID];
 *   - mask spectrum is centred (DC at [N/2][N/2]) and divided by N^2;
 *   - SOCS IFFT is unnormalised; FI divides the R2C by G^2.
 *
 * All functions return 0 on success and a negative code on bad arguments.
 */
#ifndef LITHO_ORACLE_H
#define LITHO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EARG -1
#define ORC_ENOMEM -2
#define ORC_ERANGE -3
#define ORC_ENOCONV -4

/* ---- FFT (own mixed radix, FP64) ------------------------------------- */
/* In-place 1-D complex DFT of length n: X[k] = sum_j x[j] e^{sign*2*pi*i*j*k/n}
 * (sign = -1 forward as FFTW_FORWARD, +1 backward as FFTW_BACKWARD), unnormalised. */
int orc_fft1d(int n, double* re, double* im, int sign);
/* In-place 2-D complex DFT on a sy x sx row-major grid. */
int orc_fft2d(int sy, int sx, double* re, double* im, int sign);

/* ---- Source (common/source.cpp) --------------------------------------- */
/* createAnnular L5-23 followed by normalizeMatrix L89-103. */
int orc_source_annular(float* src, int srcSize, float inner, float outer);
/* createDipole L58-87 + normalize. */
int orc_source_dipole(float* src, int srcSize, float radius, float offset, int onXAxis);
/* createCrossQuadrupole L25-55 + normalize. */
int orc_source_quadrupole(float* src, int srcSize, float radius, float offset);
/* createSource "Point" branch L122-129 + normalize. */
int orc_source_point(float* src, int srcSize, float px, float py);
/* normalizeMatrix L89-103: float sum, divide each element by it. */
void orc_normalize(float* m, int count);
/* getOuterSigma, litho_thread.cpp L427-439 (caller clamps to 1 as main L1344). */
float orc_outer_sigma(const float* src, int srcSize);
/* Nx = floor(NA * period * (1 + sigma) / lambda), float arithmetic, litho_thread.cpp L1346. */
int orc_compute_nx(float NA, int period, float outerSigma, float lambda);
/* Count of source points the TCC loop visits (src!=0 and p^2+q^2<=sh^2
 * inside [-oSgm,oSgm]^2), calcTCC L458-466. */
int orc_count_live_sources(const float* src, int srcSize, float outerSigma);

/* ---- Mask (common/mask.cpp) -------------------------------------------- */
/* generateLineSpace L5-29 (vertical when isHorizontal == 0). */
int orc_linespace(float* msk, int sizeX, int sizeY, int lineWidth, int spaceWidth, int isHorizontal);
/* createMask placement L54-59: centre an inner mask into Lx x Ly and multiply by dose. */
int orc_place_mask(float* msk, int Lx, int Ly, const float* inMsk, int maskSizeX, int maskSizeY, float dose);

/* ---- Module 3: FT_r2c_padded, litho_thread.cpp L255-331 ----------------- */
/* Full centred spectrum of a sy x sx real image (sizes used as-is; the
 * reference pads to the next power of two, which is the identity for the
 * power-of-two sizes it runs).  out[(sy/2+ky)*sx + sx/2+kx] =
 *   (1/(sx*sy)) * sum mask[y][x] e^{-2 pi i (kx (x-sx/2)/sx + ky (y-sy/2)/sy)}. */
int orc_mask_spectrum(const float* mask, int sy, int sx, double* out_re, double* out_im);

/* ---- Module 4: calcTCC, litho_thread.cpp L441-532 ------------------------ */
/* Float pupil arithmetic verbatim; upper triangle accumulated in the
 * reference's order, lower triangle mirrored.  Lx/Ly are the periods (nm). */
int orc_calc_tcc(double* tcc_re, double* tcc_im, int tccSize,
                 const float* src, int srcSize, float outerSigma,
                 float lambda, float NA, float defocus,
                 int Lx, int Ly, int Nx, int Ny);
/* Pupil matrix of the live sources, calcTCC L457-484: A[ID*nsrc + s]
 * (complex) and the source weight S[s] of column s; returns the number of
 * live sources in *nsrc_out.  Pass A_re == NULL to only count.  TCC = A diag(S) A^H. */
int orc_pupil_matrix(double* A_re, double* A_im, double* S, int* nsrc_out, int nsrc_cap,
                     const float* src, int srcSize, float outerSigma,
                     float lambda, float NA, float defocus,
                     int Lx, int Ly, int Nx, int Ny);

/* ---- Module 6: calcKernels, litho_thread.cpp L597-663 -------------------- */
/* Full Hermitian eigendecomposition (cyclic Jacobi, FP64).  H (n x n, both
 * triangles read) is overwritten.  w ascending; V column-major: V[j*n + i] is
 * component i of eigenvector j. */
int orc_eigh(int n, double* H_re, double* H_im, double* w, double* V_re, double* V_im);
/* Top-nk kernels from a TCC: scales descending (FP64), kernels index-flipped
 * (krn[k*n + j] = v_k[n-1-j], L643-652), degenerate clusters put in the canonical
 * mirror basis and every vector phase-fixed (see orc_canonicalize).
 * *cluster_split is set to 1 when nk cuts through a degenerate cluster. */
int orc_calc_kernels(double* krn_re, double* krn_im, double* scales, int nk,
                     const double* tcc_re, const double* tcc_im, int n,
                     int Nx, int Ny, int* cluster_split);
/* Same result through the source-space Gram route: G = B^H B with
 * B = A diag(sqrt(S)) (nsrc x nsrc), eigenvectors lifted v = B u / sqrt(lambda). */
int orc_calc_kernels_gram(double* krn_re, double* krn_im, double* scales, int nk,
                          const double* A_re, const double* A_im, const double* S,
                          int n, int nsrc, int Nx, int Ny, int* cluster_split);
/* Canonical basis for eigenvector clusters (V column-major n x m, w
 * descending): inside each cluster of relative width <= rtol*w[0], diagonalise
 * the mirror nx -> -nx (then ny -> -ny, then the transpose nx <-> ny) and order
 * by mirror eigenvalue descending; then rotate every vector's phase so its
 * first component with |v_j| >= (1-1e-6) max|v| is real positive. */
int orc_canonicalize(double* V_re, double* V_im, const double* w, int n, int m,
                     int Nx, int Ny, double rtol);

/* ---- Module 7: calcSOCS, litho_thread.cpp L813-1040 ------------------------ */
/* Conv grid G = nextPow2(4N+1) per axis (L833-834). */
int orc_conv_grid(int Nx);
/* Accumulated intensity on the conv grid, fftshifted (tmpImgp, L897-970):
 * out has Gy x Gx doubles.  mskf is the full fy x fx centred spectrum;
 * kernels nk x (2Ny+1)(2Nx+1) flipped; scales are float as in the reference. */
int orc_socs_tmpimg(double* tmpImgp, int Gx, int Gy,
                    const double* mskf_re, const double* mskf_im, int fx, int fy,
                    const double* krn_re, const double* krn_im, const float* scales,
                    int nk, int Nx, int Ny, int nthreads);
/* FI, litho_thread.cpp L692-772: band-limited interpolation of in (iy x ix,
 * zero frequency at the centre) to out (oy x ox). */
int orc_fi(const double* in, int ix, int iy, double* out, int ox, int oy);
/* calcSOCS complete: tmpImgp -> FI -> crop (L1033-1037); image is float. */
int orc_calc_socs(float* image, const double* mskf_re, const double* mskf_im, int fx, int fy,
                  const double* krn_re, const double* krn_im, const float* scales,
                  int nk, int Nx, int Ny, int nthreads);

/* ---- Module 5 (verification): calcImage + FT_c2r_padded, litho_thread.cpp L539-583, L371-425 */
int orc_calc_image_tcc(float* image, const double* mskf_re, const double* mskf_im, int fx, int fy,
                       const double* tcc_re, const double* tcc_im, int tccSize, int Nx, int Ny);

#ifdef __cplusplus
}
#endif
#endif
