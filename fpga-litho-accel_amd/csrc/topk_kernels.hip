// Top-K Hermitian eigensolver for large explicit TCCs (Module 6, calcKernels,
// litho_thread.cpp L597-663, where the reference calls LAPACK zheevr for the
// top nk eigenpairs).  A full-spectrum Jacobi costs O(n^3) per sweep on an
// n x n matrix in HBM; the TCC's top nk eigenpairs are found instead by block
// subspace iteration (Rutishauser; Golub & Van Loan 8.2.4) on b >= 2 nk
// vectors, every product on the FP64 matrix cores:
//
//   Y = H Q                      zgemm_nt (b x n outputs, K = n)
//   Q = orth(Y)                  S = Y^H Y, S = R^H R (Cholesky), Q = Y R^-1,
//                                done twice (CholQR2)
//   every other step:            T = Q^H H Q, T = W Theta W^H (one-block Jacobi),
//                                Ritz vectors X = Q W, residuals |H x - theta x|
//
// Vectors are stored vector-major, Qt[i][l] = component l of vector i, the
// layout of the kernel set.  Everything here is FP64.
#include <mutex>

#include "common.hpp"
#include "eigen_kernels.hpp"
#include "topk_kernels.hpp"

namespace litho {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// C[i][j] = sum_l opA(A[i*lda + l]) opB(B[j*ldb + l]), i < p, j < q, l < k;
// op = conj when the flag is set.  One block = 4 waves = a 32 x 32 tile, each
// wave a 16 x 16 MFMA tile (f64 16x16x4: A operand lane l holds A[l%16][l/16],
// B operand B[l/16][l%16], result register v of lane l is C[l/16 + 4v][l%16]);
// the four real products of the complex product accumulate separately, so
// consecutive MFMAs do not depend on each other.  blockIdx.z splits K: slice
// z writes its partial sums to C + z * p * ldc.
// ---------------------------------------------------------------------------
constexpr int ZG_T = 32;
constexpr int ZG_K = 16;

__global__ __launch_bounds__(256) void zgemm_nt_kernel(const double2* __restrict__ A, long lda,
                                                       int conjA, const double2* __restrict__ B,
                                                       long ldb, int conjB, int p, int q, int k,
                                                       int kslice, double2* __restrict__ C,
                                                       long ldc) {
    __shared__ double2 As[ZG_T][ZG_K + 1];
    __shared__ double2 Bs[ZG_T][ZG_K + 1];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1;
    const int i0 = blockIdx.y * ZG_T, j0 = blockIdx.x * ZG_T;
    const int l0 = blockIdx.z * kslice;
    const int l1 = min(k, l0 + kslice);
    d4 crr = {0.0, 0.0, 0.0, 0.0}, cii = crr, cri = crr, cir = crr;
    for (int s0 = l0; s0 < l1; s0 += ZG_K) {
        for (int e = tid; e < ZG_T * ZG_K; e += 256) {
            const int r = e / ZG_K, c = e - r * ZG_K;
            const int l = s0 + c;
            double2 av = make_double2(0.0, 0.0), bv = make_double2(0.0, 0.0);
            if (l < l1) {
                if (i0 + r < p) av = A[(long)(i0 + r) * lda + l];
                if (j0 + r < q) bv = B[(long)(j0 + r) * ldb + l];
            }
            As[r][c] = av;
            Bs[r][c] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < ZG_K; kk += 4) {
            const double2 a = As[wi * 16 + (lane & 15)][kk + (lane >> 4)];
            const double2 b = Bs[wj * 16 + (lane & 15)][kk + (lane >> 4)];
            crr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, crr, 0, 0, 0);
            cii = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, cii, 0, 0, 0);
            cri = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, cri, 0, 0, 0);
            cir = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, cir, 0, 0, 0);
        }
        __syncthreads();
    }
    // (ar + i sa ai)(br + i sb bi) = ar br - sa sb ai bi + i (sb ar bi + sa ai br)
    const double sa = conjA ? -1.0 : 1.0, sb = conjB ? -1.0 : 1.0;
    double2* Cz = C + (long)blockIdx.z * p * ldc;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int i = i0 + wi * 16 + (lane >> 4) + 4 * v;
        const int j = j0 + wj * 16 + (lane & 15);
        if (i < p && j < q)
            Cz[(long)i * ldc + j] =
                make_double2(crr[v] - sa * sb * cii[v], sb * cri[v] + sa * cir[v]);
    }
}

// C[i][j] = sum_z P[z][i][j]
__global__ __launch_bounds__(256) void zsum_slices_kernel(const double2* __restrict__ P, int ns,
                                                          long count, double2* __restrict__ C) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= count) return;
    double sr = 0.0, si = 0.0;
    for (int z = 0; z < ns; z++) {
        const double2 v = P[(long)z * count + e];
        sr += v.x;
        si += v.y;
    }
    C[e] = make_double2(sr, si);
}

// Y[i][l] = sum_m Cf[i][m] X[m][l], Cf b_out x b_in, X b_in x n (vector-major).
// Each thread keeps RI output rows of one column l.
constexpr int RT_RI = 8;
__global__ __launch_bounds__(256) void rotate_rows_kernel(const double2* __restrict__ Cf,
                                                          int bout, int bin,
                                                          const double2* __restrict__ X, int n,
                                                          double2* __restrict__ Y) {
    extern __shared__ double2 cs[];  // [RT_RI][bin]
    const int i0 = blockIdx.y * RT_RI;
    for (int e = threadIdx.x; e < RT_RI * bin; e += 256) {
        const int r = e / bin, m = e - r * bin;
        cs[e] = i0 + r < bout ? Cf[(long)(i0 + r) * bin + m] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int l = blockIdx.x * 256 + threadIdx.x;
    if (l >= n) return;
    double ar[RT_RI], ai[RT_RI];
#pragma unroll
    for (int r = 0; r < RT_RI; r++) ar[r] = ai[r] = 0.0;
    for (int m = 0; m < bin; m++) {
        const double2 x = X[(long)m * n + l];
#pragma unroll
        for (int r = 0; r < RT_RI; r++) {
            const double2 c = cs[r * bin + m];
            ar[r] += c.x * x.x - c.y * x.y;
            ai[r] += c.x * x.y + c.y * x.x;
        }
    }
#pragma unroll
    for (int r = 0; r < RT_RI; r++)
        if (i0 + r < bout) Y[(long)(i0 + r) * n + l] = make_double2(ar[r], ai[r]);
}

// res[i] = || HX[i] - theta[i] X[i] ||_2, one block per vector
__global__ __launch_bounds__(256) void ritz_residual_kernel(const double2* __restrict__ HX,
                                                            const double2* __restrict__ X,
                                                            const double* __restrict__ theta,
                                                            int n, double* __restrict__ res) {
    __shared__ double sh[256];
    const int i = blockIdx.x;
    const double th = theta[i];
    double acc = 0.0;
    for (int l = threadIdx.x; l < n; l += 256) {
        const double2 h = HX[(long)i * n + l], x = X[(long)i * n + l];
        const double dr = h.x - th * x.x, di = h.y - th * x.y;
        acc += dr * dr + di * di;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) res[i] = sqrt(sh[0]);
}

// Cholesky S = R^H R of a Hermitian positive definite b x b matrix (S is
// overwritten by the elimination, R goes to Rw, U = R^-1 is built in Cf as)
// Cf = U^T so that Q = Y R^-1 is q_j = sum_m y_m U[m][j] = sum_m Cf[j][m] y_m.
// One block.  *status = 1 when a pivot falls below 1e-26 of the largest
// diagonal entry (Y numerically rank deficient).
__global__ __launch_bounds__(1024) void chol_coeff_kernel(double2* __restrict__ S, int b,
                                                          double2* __restrict__ Rw,
                                                          double2* __restrict__ Cf,
                                                          int* __restrict__ status) {
    __shared__ double dmax;
    __shared__ int fail;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        double m = 0.0;
        for (int i = 0; i < b; i++) m = fmax(m, S[(long)i * b + i].x);
        dmax = m;
        fail = !(m > 0.0);
    }
    for (long e = tid; e < (long)b * b; e += nt) Rw[e] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int kk = 0; kk < b && !fail; kk++) {
        const double d = S[(long)kk * b + kk].x;
        if (!(d > 1e-26 * dmax)) {
            if (tid == 0) fail = 1;
            __syncthreads();
            break;
        }
        const double rkk = sqrt(d);
        for (int j = kk + tid; j < b; j += nt) {
            const double2 s = S[(long)kk * b + j];
            Rw[(long)kk * b + j] = j == kk ? make_double2(rkk, 0.0) : make_double2(s.x / rkk, s.y / rkk);
        }
        __syncthreads();
        // S[i][j] -= conj(R[k][i]) R[k][j] for kk < i <= j
        const int m = b - kk - 1;
        for (long e = tid; e < (long)m * m; e += nt) {
            const int i = kk + 1 + (int)(e / m), j = kk + 1 + (int)(e % m);
            if (j < i) continue;
            const double2 ri = Rw[(long)kk * b + i], rj = Rw[(long)kk * b + j];
            double2& s = S[(long)i * b + j];
            s.x -= ri.x * rj.x + ri.y * rj.y;
            s.y -= ri.x * rj.y - ri.y * rj.x;
        }
        __syncthreads();
    }
    if (fail) {
        if (tid == 0) *status = 1;  // sticky: the host clears it
        return;
    }
    // U = R^-1 column by column (back substitution), Cf[j][m] = U[m][j]
    for (int j = tid; j < b; j += nt) {
        for (int i = b - 1; i >= 0; i--) {
            double2 u;
            if (i > j) {
                u = make_double2(0.0, 0.0);
            } else {
                double sr = i == j ? 1.0 : 0.0, si = 0.0;
                for (int mm = i + 1; mm <= j; mm++) {
                    const double2 r = Rw[(long)i * b + mm], v = Cf[(long)j * b + mm];
                    sr -= r.x * v.x - r.y * v.y;
                    si -= r.x * v.y + r.y * v.x;
                }
                const double rii = Rw[(long)i * b + i].x;
                u = make_double2(sr / rii, si / rii);
            }
            Cf[(long)j * b + i] = u;
        }
    }
}

// Cyclic Jacobi (round-robin pairs, the rotation of eigen_kernels.hip) of a
// small b x b Hermitian matrix in one block: A is overwritten (its diagonal
// holds the eigenvalues), V receives the eigenvectors (V[i][k] = component i
// of vector k).  The matrix and V are staged in LDS when they fit.  A is read
// as given and made exactly Hermitian first ((A + A^H) / 2).
// *status is set to 1 when 60 sweeps do not converge (left as is otherwise).
__global__ __launch_bounds__(1024) void jacobi_small_kernel(double2* __restrict__ Ag,
                                                            double2* __restrict__ Vg, int b,
                                                            int use_lds,
                                                            int* __restrict__ status) {
    extern __shared__ double2 lds[];
    __shared__ JacobiRot rots[128];
    __shared__ double red[1024];
    __shared__ int nrot;
    const int tid = threadIdx.x, nt = blockDim.x;
    const long bb = (long)b * b;
    double2* A = use_lds ? lds : Ag;
    double2* V = use_lds ? lds + bb : Vg;
    for (long e = tid; e < bb; e += nt) {
        const int i = (int)(e / b), j = (int)(e % b);
        const double2 x = Ag[e], y = Ag[(long)j * b + i];
        V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        if (use_lds) A[e] = make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
    }
    __syncthreads();
    if (!use_lds) {
        // symmetrise in place: upper from both triangles, lower mirrored
        for (long e = tid; e < bb; e += nt) {
            const int i = (int)(e / b), j = (int)(e % b);
            if (i < j) {
                const double2 x = A[e], y = A[(long)j * b + i];
                const double2 h = make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
                A[e] = h;
                A[(long)j * b + i] = make_double2(h.x, -h.y);
            } else if (i == j) {
                A[e].y = 0.0;
            }
        }
        __syncthreads();
    }
    double acc = 0.0;
    for (long e = tid; e < bb; e += nt) acc += A[e].x * A[e].x + A[e].y * A[e].y;
    red[tid] = acc;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double tol = 1e-17 * sqrt(red[0]);
    const int npad = b + (b & 1), npairs = npad / 2, mm = npad - 1;
    int sweep = 0;
    for (; sweep < 60; sweep++) {
        if (tid == 0) nrot = 0;
        __syncthreads();
        for (int r = 0; r < mm; r++) {
            if (tid < npairs) {
                int a, c;
                if (tid == 0) {
                    a = r;
                    c = mm;
                } else {
                    a = (r + tid) % mm;
                    c = (r - tid + mm) % mm;
                }
                JacobiRot R;
                R.p = a;
                R.q = c;
                R.active = 0;
                R.c = 1.0;
                R.s = R.ei = R.dpp = R.dqq = 0.0;
                R.er = 1.0;
                if (a < b && c < b) {
                    const double2 z = A[(long)a * b + c];
                    const double rr = sqrt(z.x * z.x + z.y * z.y);
                    if (rr > tol) {
                        const double app = A[(long)a * b + a].x, aqq = A[(long)c * b + c].x;
                        const double theta = (aqq - app) / (2.0 * rr);
                        const double t =
                            (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        const double cc = 1.0 / sqrt(t * t + 1.0);
                        R.c = cc;
                        R.s = t * cc;
                        R.er = z.x / rr;
                        R.ei = z.y / rr;
                        R.dpp = app - t * rr;
                        R.dqq = aqq + t * rr;
                        R.active = 1;
                        atomicAdd(&nrot, 1);
                    }
                }
                rots[tid] = R;
            }
            __syncthreads();
            // rows p, q:  p' = c p - s e q ;  q' = s p + c e q
            for (long e = tid; e < (long)npairs * b; e += nt) {
                const JacobiRot& R = rots[e / b];
                if (!R.active) continue;
                const int j = (int)(e % b);
                double2* hp = A + (long)R.p * b;
                double2* hq = A + (long)R.q * b;
                const double2 x = hp[j], y = hq[j];
                const double eyr = R.er * y.x - R.ei * y.y, eyi = R.er * y.y + R.ei * y.x;
                hp[j] = make_double2(R.c * x.x - R.s * eyr, R.c * x.y - R.s * eyi);
                hq[j] = make_double2(R.s * x.x + R.c * eyr, R.s * x.y + R.c * eyi);
            }
            __syncthreads();
            // columns p, q of A and V:  p' = c p - s conj(e) q ;  q' = s p + c conj(e) q
            for (long e = tid; e < (long)npairs * b; e += nt) {
                const JacobiRot& R = rots[e / b];
                if (!R.active) continue;
                const int i = (int)(e % b);
                {
                    double2* row = A + (long)i * b;
                    const double2 x = row[R.p], y = row[R.q];
                    const double eyr = R.er * y.x + R.ei * y.y, eyi = R.er * y.y - R.ei * y.x;
                    double2 np = make_double2(R.c * x.x - R.s * eyr, R.c * x.y - R.s * eyi);
                    double2 nq = make_double2(R.s * x.x + R.c * eyr, R.s * x.y + R.c * eyi);
                    if (i == R.p) {
                        np = make_double2(R.dpp, 0.0);
                        nq = make_double2(0.0, 0.0);
                    } else if (i == R.q) {
                        np = make_double2(0.0, 0.0);
                        nq = make_double2(R.dqq, 0.0);
                    }
                    row[R.p] = np;
                    row[R.q] = nq;
                }
                {
                    double2* row = V + (long)i * b;
                    const double2 x = row[R.p], y = row[R.q];
                    const double eyr = R.er * y.x + R.ei * y.y, eyi = R.er * y.y - R.ei * y.x;
                    row[R.p] = make_double2(R.c * x.x - R.s * eyr, R.c * x.y - R.s * eyi);
                    row[R.q] = make_double2(R.s * x.x + R.c * eyr, R.s * x.y + R.c * eyi);
                }
            }
            __syncthreads();
        }
        if (nrot == 0) break;
        __syncthreads();
    }
    if (use_lds)
        for (long e = tid; e < bb; e += nt) {
            Ag[e] = A[e];
            Vg[e] = V[e];
        }
    if (tid == 0 && sweep >= 60) *status = 1;
}

// Gram-route lift operand: Xs[r][s] = X[r][s] sqrt(S[s]) / sqrt(theta[r])
// (theta > 0; a non-positive weight gives a zero row)
__global__ __launch_bounds__(256) void lift_scale_kernel(const double2* __restrict__ X,
                                                         const double* __restrict__ S,
                                                         const double* __restrict__ theta, int mt,
                                                         int m, double2* __restrict__ Xs) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)mt * m) return;
    const int r = (int)(e / m), sidx = (int)(e % m);
    const double th = theta[r];
    const double f = th > 0.0 ? sqrt(S[sidx]) / sqrt(th) : 0.0;
    const double2 x = X[e];
    Xs[e] = make_double2(x.x * f, x.y * f);
}

// ---------------------------------------------------------------------------

int launch_lift_scale(const double2* X, const double* S, const double* theta, int mt, int m,
                      double2* Xs, hipStream_t st) {
    hipLaunchKernelGGL(lift_scale_kernel, dim3(cdiv((long)mt * m, 256)), dim3(256), 0, st, X, S,
                       theta, mt, m, Xs);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_zgemm_nt(const double2* A, long lda, int conjA, const double2* B, long ldb, int conjB,
                    int p, int q, int k, double2* C, long ldc, int nsplit, double2* work,
                    hipStream_t st) {
    if (p <= 0 || q <= 0 || k <= 0 || nsplit <= 0 || (nsplit > 1 && !work)) return LITHO_EARG;
    const int kslice = (int)(((long)k + nsplit - 1) / nsplit + ZG_K - 1) / ZG_K * ZG_K;
    const int ns = (k + kslice - 1) / kslice;
    dim3 grid(cdiv(q, ZG_T), cdiv(p, ZG_T), ns);
    double2* out = ns > 1 ? work : C;
    const long ld = ns > 1 ? q : ldc;
    hipLaunchKernelGGL(zgemm_nt_kernel, grid, dim3(256), 0, st, A, lda, conjA, B, ldb, conjB, p, q,
                       k, kslice, out, ld);
    if (ns > 1) {
        if (ldc != q) return LITHO_EARG;
        const long count = (long)p * q;
        hipLaunchKernelGGL(zsum_slices_kernel, dim3(cdiv(count, 256)), dim3(256), 0, st, work, ns,
                           count, C);
    }
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_rotate_rows(const double2* Cf, int bout, int bin, const double2* X, int n, double2* Y,
                       hipStream_t st) {
    if (bout <= 0 || bin <= 0 || bin > 512 || n <= 0 || X == Y) return LITHO_EARG;
    const size_t lds = (size_t)RT_RI * bin * sizeof(double2);
    hipLaunchKernelGGL(rotate_rows_kernel, dim3(cdiv(n, 256), cdiv(bout, RT_RI)), dim3(256), lds,
                       st, Cf, bout, bin, X, n, Y);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_ritz_residual(const double2* HX, const double2* X, const double* theta, int b, int n,
                         double* res, hipStream_t st) {
    hipLaunchKernelGGL(ritz_residual_kernel, dim3(b), dim3(256), 0, st, HX, X, theta, n, res);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_chol_coeff(double2* S, int b, double2* Rw, double2* Cf, int* status, hipStream_t st) {
    if (b <= 0 || b > 512) return LITHO_EARG;
    hipLaunchKernelGGL(chol_coeff_kernel, dim3(1), dim3(1024), 0, st, S, b, Rw, Cf, status);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_jacobi_small(double2* A, double2* V, int b, int* status, hipStream_t st) {
    if (b <= 0 || b > 256) return LITHO_EARG;
    const size_t need = 2 * (size_t)b * b * sizeof(double2);
    // 2 b^2 complex doubles beside 16 KB of static LDS (b <= 64)
    const int use_lds = need <= 128 * 1024 ? 1 : 0;
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [] {
        attr = hipFuncSetAttribute((const void*)jacobi_small_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
    });
    if (attr != hipSuccess) return LITHO_EHIP;
    hipLaunchKernelGGL(jacobi_small_kernel, dim3(1), dim3(1024), use_lds ? need : 0, st, A, V, b,
                       use_lds, status);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

}  // namespace litho
