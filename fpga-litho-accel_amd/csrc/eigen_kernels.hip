// FP64 Hermitian eigensolver and SOCS kernel extraction for gfx950
// (Module 6, calcKernels, litho_thread.cpp L597-663).
//
// Two-sided cyclic Jacobi with the round-robin (circle) ordering: every round
// rotates npad/2 disjoint index pairs at once, so a round is three launches
// (rotation parameters, row update J^H H, column update H J and V J) with no
// atomics on matrix data.  npad-1 rounds make one sweep, which visits every
// pair once.  The rotation for a pair (p, q) with H[p][q] = r e^{i phi}:
//   t = sgn(theta)/(|theta| + sqrt(theta^2+1)), theta = (H_qq - H_pp)/(2r),
//   c = 1/sqrt(1+t^2), s = t c,  J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]],
//   H_pp' = H_pp - t r, H_qq' = H_qq + t r, H_pq' = 0.
// The kernel set is then: top-m selection, canonical basis inside degenerate
// clusters (mirror eigenbasis, same rule as the CPU oracle), phase fix, index
// flip K[j] = v[n-1-j] (L643-652).
#include "common.hpp"
#include "eigen_kernels.hpp"

namespace litho {

// circle-method pairing for round r of npad (even) players
__device__ __forceinline__ void round_pair(int r, int k, int npad, int& a, int& b) {
    const int m = npad - 1;
    if (k == 0) {
        a = r;
        b = m;
    } else {
        a = (r + k) % m;
        b = (r - k + m) % m;
    }
}

__global__ __launch_bounds__(256) void jacobi_params_kernel(const double2* __restrict__ H, int n,
                                                            int npad, int round,
                                                            const double* __restrict__ tolp,
                                                            JacobiRot* __restrict__ rots,
                                                            int* __restrict__ count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npad / 2) return;
    int a, b;
    round_pair(round, k, npad, a, b);
    JacobiRot R;
    R.p = a;
    R.q = b;
    R.active = 0;
    R.c = 1.0;
    R.s = 0.0;
    R.er = 1.0;
    R.ei = 0.0;
    R.dpp = 0.0;
    R.dqq = 0.0;
    if (a < n && b < n) {
        double2 z = H[(long)a * n + b];
        double r = sqrt(z.x * z.x + z.y * z.y);
        if (r > *tolp) {
            double app = H[(long)a * n + a].x, aqq = H[(long)b * n + b].x;
            double theta = (aqq - app) / (2.0 * r);
            double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            double c = 1.0 / sqrt(t * t + 1.0);
            R.c = c;
            R.s = t * c;
            R.er = z.x / r;
            R.ei = z.y / r;
            R.dpp = app - t * r;
            R.dqq = aqq + t * r;
            R.active = 1;
            atomicAdd(count, 1);
        }
    }
    rots[k] = R;
}

// rows p, q:  p' = c p - s e q ;  q' = s p + c e q   (e = e^{i phi})
__global__ __launch_bounds__(256) void jacobi_rows_kernel(double2* __restrict__ H, int n,
                                                          const JacobiRot* __restrict__ rots) {
    const JacobiRot R = rots[blockIdx.y];
    if (!R.active) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double2* hp = H + (long)R.p * n;
    double2* hq = H + (long)R.q * n;
    double2 x = hp[j], y = hq[j];
    double eyr = R.er * y.x - R.ei * y.y, eyi = R.er * y.y + R.ei * y.x;
    hp[j] = make_double2(R.c * x.x - R.s * eyr, R.c * x.y - R.s * eyi);
    hq[j] = make_double2(R.s * x.x + R.c * eyr, R.s * x.y + R.c * eyi);
}

// columns p, q of H and V:  p' = c p - s conj(e) q ;  q' = s p + c conj(e) q
__global__ __launch_bounds__(256) void jacobi_cols_kernel(double2* __restrict__ H,
                                                          double2* __restrict__ V, int n,
                                                          int npairs,
                                                          const JacobiRot* __restrict__ rots) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (k >= npairs) return;
    const JacobiRot R = rots[k];
    if (!R.active) return;
    {
        double2* row = H + (long)i * n;
        double2 x = row[R.p], y = row[R.q];
        double eyr = R.er * y.x + R.ei * y.y, eyi = R.er * y.y - R.ei * y.x;
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
        double2* row = V + (long)i * n;
        double2 x = row[R.p], y = row[R.q];
        double eyr = R.er * y.x + R.ei * y.y, eyi = R.er * y.y - R.ei * y.x;
        row[R.p] = make_double2(R.c * x.x - R.s * eyr, R.c * x.y - R.s * eyi);
        row[R.q] = make_double2(R.s * x.x + R.c * eyr, R.s * x.y + R.c * eyi);
    }
}

// ||H||_F^2 partial sums (one double per block) and identity V
__global__ __launch_bounds__(256) void fro_init_kernel(const double2* __restrict__ H,
                                                       double2* __restrict__ V, long nn, int n,
                                                       double* __restrict__ partial) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < nn; e += (long)gridDim.x * 256) {
        double2 h = H[e];
        acc += h.x * h.x + h.y * h.y;
        long i = e / n, j = e - i * n;
        V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void tol_finish_kernel(const double* __restrict__ partial, int np, double rel,
                                  double* __restrict__ tol) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; i++) s += partial[i];
        *tol = rel * sqrt(s);
    }
}

// Hermitian input cleanup: H = upper triangle mirrored (LAPACK 'U' semantics),
// real diagonal.
__global__ __launch_bounds__(256) void herm_from_upper_kernel(double2* __restrict__ H, int n) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * n) return;
    long i = e / n, j = e - i * n;
    if (i > j) {
        double2 u = H[j * (long)n + i];
        H[e] = make_double2(u.x, -u.y);
    } else if (i == j) {
        H[e].y = 0.0;
    }
}

__global__ __launch_bounds__(256) void diag_kernel(const double2* __restrict__ H, int n,
                                                   double* __restrict__ w) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) w[i] = H[(long)i * n + i].x;
}

// rank[i] = position of eigenvalue i in descending order (stable by index)
__global__ __launch_bounds__(256) void rank_kernel(const double* __restrict__ w, int n,
                                                   int* __restrict__ order) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double wi = w[i];
    int r = 0;
    for (int j = 0; j < n; j++) {
        double wj = w[j];
        r += (wj > wi) || (wj == wi && j < i);
    }
    order[r] = i;
}

// Vt[r][i] = V[i][order[r]] for r < m (vector-major output), wd[r] = w[order[r]]
__global__ __launch_bounds__(256) void gather_top_kernel(const double2* __restrict__ V,
                                                         const double* __restrict__ w,
                                                         const int* __restrict__ order, int n,
                                                         int m, double2* __restrict__ Vt,
                                                         double* __restrict__ wd) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (r >= m || i >= n) return;
    const int c = order[r];
    Vt[(long)r * n + i] = V[(long)i * n + c];
    if (i == 0) wd[r] = w[c];
}

// Gram route lift, first half: Xs[r][s] = sqrt(S_s) U[s][ord[r]] / sqrt(lam_r)
// (U row-major [m][m], column = vector), wd[r] = lam_r; the lift itself,
// Vt[r][i] = sum_s A[i][s] Xs[r][s], is one FP64 matrix-core product.
__global__ __launch_bounds__(256) void lift_gather_kernel(const double* __restrict__ S,
                                                          const double2* __restrict__ U,
                                                          const double* __restrict__ w,
                                                          const int* __restrict__ order, int m,
                                                          int mtop, double2* __restrict__ Xs,
                                                          double* __restrict__ wd) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)mtop * m) return;
    const int r = (int)(e / m), sidx = (int)(e % m);
    const int c = order[r];
    const double lam = w[c];
    const double f = lam > 0.0 ? sqrt(S[sidx]) / sqrt(lam) : 0.0;
    const double2 u = U[(long)sidx * m + c];
    Xs[e] = make_double2(u.x * f, u.y * f);
    if (sidx == 0) wd[r] = lam;
}

__device__ __forceinline__ int mirror_index(int id, int Nx, int Ny, int which) {
    const int W = 2 * Nx + 1;
    int ix = id % W - Nx, iy = id / W - Ny;
    if (which == 0) ix = -ix;
    else if (which == 1) iy = -iy;
    else {
        int t = ix;
        ix = iy;
        iy = t;
    }
    return (iy + Ny) * W + (ix + Nx);
}

__device__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// serial cyclic Jacobi on a small Hermitian matrix (row-major c x c, c <= CANON_MAX)
__device__ void small_eigh(double* Mr, double* Mi, int c, double* w, double* Qr, double* Qi) {
    for (int i = 0; i < c * c; i++) {
        Qr[i] = 0.0;
        Qi[i] = 0.0;
    }
    for (int i = 0; i < c; i++) {
        Qr[i * c + i] = 1.0;
        Mi[i * c + i] = 0.0;
    }
    double fro = 0.0;
    for (int i = 0; i < c * c; i++) fro += Mr[i] * Mr[i] + Mi[i] * Mi[i];
    const double tol = 1e-17 * sqrt(fro);
    for (int sweep = 0; sweep < 60; sweep++) {
        int rot = 0;
        for (int p = 0; p < c - 1; p++)
            for (int q = p + 1; q < c; q++) {
                double zr = Mr[p * c + q], zi = Mi[p * c + q];
                double r = sqrt(zr * zr + zi * zi);
                if (r <= tol) continue;
                double er = zr / r, ei = zi / r;
                double a = Mr[p * c + p], b = Mr[q * c + q];
                double theta = (b - a) / (2.0 * r);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cc = 1.0 / sqrt(t * t + 1.0), s = t * cc;
                for (int k = 0; k < c; k++) {
                    if (k == p || k == q) continue;
                    double xr = Mr[p * c + k], xi = Mi[p * c + k], yr = Mr[q * c + k], yi = Mi[q * c + k];
                    double eyr = er * yr - ei * yi, eyi = er * yi + ei * yr;
                    double npr = cc * xr - s * eyr, npi = cc * xi - s * eyi;
                    double nqr = s * xr + cc * eyr, nqi = s * xi + cc * eyi;
                    Mr[p * c + k] = npr; Mi[p * c + k] = npi;
                    Mr[q * c + k] = nqr; Mi[q * c + k] = nqi;
                    Mr[k * c + p] = npr; Mi[k * c + p] = -npi;
                    Mr[k * c + q] = nqr; Mi[k * c + q] = -nqi;
                }
                Mr[p * c + p] = a - t * r; Mr[q * c + q] = b + t * r;
                Mr[p * c + q] = Mi[p * c + q] = Mr[q * c + p] = Mi[q * c + p] = 0.0;
                Mi[p * c + p] = Mi[q * c + q] = 0.0;
                for (int k = 0; k < c; k++) {
                    double xr = Qr[k * c + p], xi = Qi[k * c + p], yr = Qr[k * c + q], yi = Qi[k * c + q];
                    double eyr = er * yr + ei * yi, eyi = er * yi - ei * yr;
                    Qr[k * c + p] = cc * xr - s * eyr; Qi[k * c + p] = cc * xi - s * eyi;
                    Qr[k * c + q] = s * xr + cc * eyr; Qi[k * c + q] = s * xi + cc * eyi;
                }
                rot++;
            }
        if (!rot) break;
    }
    for (int i = 0; i < c; i++) w[i] = Mr[i * c + i];
}

// One block per cluster (start, size <= CANON_MAX) of vector-major Vt [m][n]:
// mirror x, then y inside equal-eigenvalue subgroups, then the transpose
// (square windows); subgroups ordered by mirror eigenvalue descending.
__global__ __launch_bounds__(256) void canon_kernel(double2* __restrict__ Vt, int n, int Nx,
                                                    int Ny, const int2* __restrict__ clusters) {
    __shared__ double red[256];
    __shared__ double Mr[CANON_MAX * CANON_MAX], Mi[CANON_MAX * CANON_MAX];
    __shared__ double Qr[CANON_MAX * CANON_MAX], Qi[CANON_MAX * CANON_MAX];
    __shared__ double mu[CANON_MAX];
    __shared__ int ord[CANON_MAX];
    __shared__ int gstart[CANON_MAX], gsize[CANON_MAX], ngroups;
    __shared__ int nstart[CANON_MAX], nsize[CANON_MAX], nnew;
    const int2 cl = clusters[blockIdx.x];
    if (threadIdx.x == 0) {
        ngroups = 1;
        gstart[0] = cl.x;
        gsize[0] = cl.y;
    }
    __syncthreads();
    for (int which = 0; which < 3; which++) {
        if (which == 2 && Nx != Ny) break;
        if (threadIdx.x == 0) nnew = 0;
        __syncthreads();
        const int ng = ngroups;
        for (int g = 0; g < ng; g++) {
            const int a = gstart[g], c = gsize[g];
            if (c < 2) {
                if (threadIdx.x == 0) {
                    nstart[nnew] = a;
                    nsize[nnew] = c;
                    nnew++;
                }
                __syncthreads();
                continue;
            }
            for (int i = 0; i < c; i++)
                for (int j = 0; j < c; j++) {
                    const double2* vi = Vt + (long)(a + i) * n;
                    const double2* vj = Vt + (long)(a + j) * n;
                    double sr = 0.0, si = 0.0;
                    for (int k = threadIdx.x; k < n; k += blockDim.x) {
                        double2 x = vi[k], y = vj[mirror_index(k, Nx, Ny, which)];
                        sr += x.x * y.x + x.y * y.y;
                        si += x.x * y.y - x.y * y.x;
                    }
                    sr = block_sum(sr, red);
                    si = block_sum(si, red);
                    if (threadIdx.x == 0) {
                        Mr[i * c + j] = sr;
                        Mi[i * c + j] = si;
                    }
                }
            __syncthreads();
            if (threadIdx.x == 0) {
                // symmetrise, diagonalise, order descending
                for (int i = 0; i < c; i++)
                    for (int j = i; j < c; j++) {
                        double r = 0.5 * (Mr[i * c + j] + Mr[j * c + i]);
                        double im = 0.5 * (Mi[i * c + j] - Mi[j * c + i]);
                        Mr[i * c + j] = r; Mr[j * c + i] = r;
                        Mi[i * c + j] = im; Mi[j * c + i] = -im;
                    }
                double wv[CANON_MAX];
                small_eigh(Mr, Mi, c, wv, Qr, Qi);
                // stable ascending sort by eigenvalue, then reverse (matches the oracle)
                for (int i = 0; i < c; i++) ord[i] = i;
                for (int i = 1; i < c; i++) {
                    int v = ord[i], j = i - 1;
                    while (j >= 0 && wv[ord[j]] > wv[v]) { ord[j + 1] = ord[j]; j--; }
                    ord[j + 1] = v;
                }
                for (int i = 0; i < c / 2; i++) {
                    int t = ord[i]; ord[i] = ord[c - 1 - i]; ord[c - 1 - i] = t;
                }
                for (int i = 0; i < c; i++) mu[i] = wv[ord[i]];
            }
            __syncthreads();
            // rotate: new vector j = sum_i v_i Q[i][ord[j]]
            for (int k = threadIdx.x; k < n; k += blockDim.x) {
                double2 old[CANON_MAX];
                for (int i = 0; i < c; i++) old[i] = Vt[(long)(a + i) * n + k];
                for (int j = 0; j < c; j++) {
                    const int qj = ord[j];
                    double sr = 0.0, si = 0.0;
                    for (int i = 0; i < c; i++) {
                        double qr = Qr[i * c + qj], qi = Qi[i * c + qj];
                        sr += old[i].x * qr - old[i].y * qi;
                        si += old[i].x * qi + old[i].y * qr;
                    }
                    Vt[(long)(a + j) * n + k] = make_double2(sr, si);
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int i = 0;
                while (i < c) {
                    int j = i + 1;
                    while (j < c && fabs(mu[j] - mu[i]) < 1e-6) j++;
                    nstart[nnew] = a + i;
                    nsize[nnew] = j - i;
                    nnew++;
                    i = j;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            ngroups = nnew;
            for (int g = 0; g < nnew; g++) {
                gstart[g] = nstart[g];
                gsize[g] = nsize[g];
            }
        }
        __syncthreads();
    }
}

// One block per vector: rotate the phase so the first component with
// |v_k| >= (1-1e-6) max|v| is real positive.
__global__ __launch_bounds__(256) void phase_fix_kernel(double2* __restrict__ Vt, int n) {
    __shared__ double red[256];
    __shared__ int redi[256];
    double2* v = Vt + (long)blockIdx.x * n;
    double mx = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        double2 x = v[k];
        mx = fmax(mx, sqrt(x.x * x.x + x.y * x.y));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    if (mx == 0.0) return;
    int first = n;
    for (int k = threadIdx.x; k < n; k += 256) {
        double2 x = v[k];
        if (sqrt(x.x * x.x + x.y * x.y) >= (1.0 - 1e-6) * mx) {
            first = k;
            break;
        }
    }
    redi[threadIdx.x] = first;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) redi[threadIdx.x] = min(redi[threadIdx.x], redi[threadIdx.x + s]);
        __syncthreads();
    }
    const int kk = redi[0];
    const double2 p = v[kk];
    const double a = sqrt(p.x * p.x + p.y * p.y);
    const double pr = p.x / a, pi = -p.y / a;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += 256) {
        double2 x = v[k];
        double2 y = make_double2(x.x * pr - x.y * pi, x.x * pi + x.y * pr);
        if (k == kk) y.y = 0.0;
        v[k] = y;
    }
}

// krn[r][j] = v_r[n-1-j] as double2 and as f32 planes
__global__ __launch_bounds__(256) void flip_kernel(const double2* __restrict__ Vt, int n, int nk,
                                                   double2* __restrict__ kd,
                                                   float* __restrict__ kr,
                                                   float* __restrict__ ki) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (j >= n || r >= nk) return;
    double2 v = Vt[(long)r * n + (n - 1 - j)];
    if (kd) kd[(long)r * n + j] = v;
    if (kr) kr[(long)r * n + j] = (float)v.x;
    if (ki) ki[(long)r * n + j] = (float)v.y;
}

// ---------------------------------------------------------------------------

int jacobi_init(double2* H, double2* V, int n, double* partial, int npart, double* tol,
                hipStream_t st) {
    long nn = (long)n * n;
    hipLaunchKernelGGL(herm_from_upper_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, st, H, n);
    hipLaunchKernelGGL(fro_init_kernel, dim3(npart), dim3(256), 0, st, H, V, nn, n, partial);
    hipLaunchKernelGGL(tol_finish_kernel, dim3(1), dim3(1), 0, st, partial, npart, 1e-17, tol);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int jacobi_sweep(double2* H, double2* V, int n, const double* tol, JacobiRot* rots, int* count,
                 hipStream_t st) {
    const int npad = n + (n & 1);
    const int npairs = npad / 2;
    for (int r = 0; r < npad - 1; r++) {
        hipLaunchKernelGGL(jacobi_params_kernel, dim3(cdiv(npairs, 256)), dim3(256), 0, st, H, n,
                           npad, r, tol, rots, count);
        hipLaunchKernelGGL(jacobi_rows_kernel, dim3(cdiv(n, 256), npairs), dim3(256), 0, st, H, n,
                           rots);
        hipLaunchKernelGGL(jacobi_cols_kernel, dim3(cdiv(npairs, 256), n), dim3(256), 0, st, H, V,
                           n, npairs, rots);
    }
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_herm_from_upper(double2* H, int n, hipStream_t st) {
    hipLaunchKernelGGL(herm_from_upper_kernel, dim3(cdiv((long)n * n, 256)), dim3(256), 0, st, H,
                       n);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_diag_rank(const double2* H, int n, double* w, int* order, hipStream_t st) {
    hipLaunchKernelGGL(diag_kernel, dim3(cdiv(n, 256)), dim3(256), 0, st, H, n, w);
    hipLaunchKernelGGL(rank_kernel, dim3(cdiv(n, 256)), dim3(256), 0, st, w, n, order);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_gather_top(const double2* V, const double* w, const int* order, int n, int m,
                      double2* Vt, double* wd, hipStream_t st) {
    hipLaunchKernelGGL(gather_top_kernel, dim3(cdiv(n, 256), m), dim3(256), 0, st, V, w, order, n,
                       m, Vt, wd);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_lift_gather(const double* S, const double2* U, const double* w, const int* order,
                       int m, int mtop, double2* Xs, double* wd, hipStream_t st) {
    hipLaunchKernelGGL(lift_gather_kernel, dim3(cdiv((long)mtop * m, 256)), dim3(256), 0, st, S, U,
                       w, order, m, mtop, Xs, wd);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_canon(double2* Vt, int n, int Nx, int Ny, const int2* clusters, int ncl,
                 hipStream_t st) {
    if (ncl <= 0) return LITHO_OK;
    hipLaunchKernelGGL(canon_kernel, dim3(ncl), dim3(256), 0, st, Vt, n, Nx, Ny, clusters);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_phase_flip(double2* Vt, int n, int m, int nk, double2* kd, float* kr, float* ki,
                      hipStream_t st) {
    hipLaunchKernelGGL(phase_fix_kernel, dim3(m), dim3(256), 0, st, Vt, n);
    hipLaunchKernelGGL(flip_kernel, dim3(cdiv(n, 256), nk), dim3(256), 0, st, Vt, n, nk, kd, kr,
                       ki);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

}  // namespace litho
