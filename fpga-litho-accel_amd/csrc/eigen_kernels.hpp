// Host launchers of the FP64 Jacobi eigensolver and kernel-extraction
// kernels (eigen_kernels.hip).
#pragma once

#include "common.hpp"

namespace litho {

constexpr int CANON_MAX = 16;  // largest degenerate cluster put in canonical form

struct JacobiRot {
    int p, q, active, pad;
    double c, s, er, ei, dpp, dqq;
};

// H <- Hermitian from its upper triangle, V <- I, *tol <- 1e-17 ||H||_F
int jacobi_init(double2* H, double2* V, int n, double* partial, int npart, double* tol,
                hipStream_t st);
// one cyclic sweep (npad-1 rounds); *count accumulates the rotations done
int jacobi_sweep(double2* H, double2* V, int n, const double* tol, JacobiRot* rots, int* count,
                 hipStream_t st);
// H <- Hermitian from its upper triangle (zheevr 'U'), real diagonal
int launch_herm_from_upper(double2* H, int n, hipStream_t st);
// w = diag(H); order[r] = index of the r-th largest eigenvalue
int launch_diag_rank(const double2* H, int n, double* w, int* order, hipStream_t st);
// Vt[r][:] = eigenvector order[r] (vector-major), wd[r] = w[order[r]]
int launch_gather_top(const double2* V, const double* w, const int* order, int n, int m,
                      double2* Vt, double* wd, hipStream_t st);
// Gram route, the lift's right factor: Xs[r][s] = sqrt(S_s) u_{order[r]}[s] /
// sqrt(lambda), wd[r] = lambda; Vt = Xs A^T follows as a matrix-core product
int launch_lift_gather(const double* S, const double2* U, const double* w, const int* order,
                       int m, int mtop, double2* Xs, double* wd, hipStream_t st);
// canonical mirror basis inside each cluster {start, size}
int launch_canon(double2* Vt, int n, int Nx, int Ny, const int2* clusters, int ncl,
                 hipStream_t st);
// phase fix of m vectors, then flipped kernels for the first nk (any output may be null)
int launch_phase_flip(double2* Vt, int n, int m, int nk, double2* kd, float* kr, float* ki,
                      hipStream_t st);

}  // namespace litho
