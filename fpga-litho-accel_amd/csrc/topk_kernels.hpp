// Host launchers of the top-K subspace eigensolver kernels (topk_kernels.hip).
#pragma once

#include "common.hpp"

namespace litho {

// C[i][j] = sum_l opA(A[i*lda+l]) opB(B[j*ldb+l]) (op = conj when set), FP64
// MFMA; nsplit > 1 splits K over blocks, summing partials through work
// (nsplit * p * q complex doubles, requires ldc == q)
int launch_zgemm_nt(const double2* A, long lda, int conjA, const double2* B, long ldb, int conjB,
                    int p, int q, int k, double2* C, long ldc, int nsplit, double2* work,
                    hipStream_t st);
// Y[i][:] = sum_m Cf[i][m] X[m][:] (vector-major b_in x n -> b_out x n; Y != X)
int launch_rotate_rows(const double2* Cf, int bout, int bin, const double2* X, int n, double2* Y,
                       hipStream_t st);
// res[i] = |HX[i] - theta[i] X[i]|, i < b
int launch_ritz_residual(const double2* HX, const double2* X, const double* theta, int b, int n,
                         double* res, hipStream_t st);
// S = R^H R (S overwritten), Cf = (R^-1)^T; *status = 1 on breakdown (sticky)
int launch_chol_coeff(double2* S, int b, double2* Rw, double2* Cf, int* status, hipStream_t st);
// one-block Jacobi of a b x b Hermitian matrix (b <= 256): diag(A) = eigenvalues,
// V[i][k] = component i of vector k; *status = 1 if not converged (sticky)
int launch_jacobi_small(double2* A, double2* V, int b, int* status, hipStream_t st);

// Gram lift operand Xs[r][s] = X[r][s] sqrt(S[s] / theta[r]), r < mt, s < m
int launch_lift_scale(const double2* X, const double* S, const double* theta, int mt, int m,
                      double2* Xs, hipStream_t st);

}  // namespace litho
