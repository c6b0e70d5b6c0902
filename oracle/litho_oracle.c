/*
 * litho_oracle.c -- FP64 CPU restatement of the SOCS aerial-image path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  The product library never links or calls this file.
 *
 * Built with -ffp-contract=off so that the float pupil arithmetic of calcTCC
 * rounds exactly as the reference's (g++ -O2 on x86-64, no FMA) does.
 * Reference citations are to validation/golden/src/litho_thread.cpp unless
 * another file is named.
 */
#include "litho_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    double re, im;
} cd;

static inline cd cmul(cd a, cd b) {
    cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cd cadd(cd a, cd b) {
    cd r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline cd csub(cd a, cd b) {
    cd r = {a.re - b.re, a.im - b.im};
    return r;
}
static inline cd cconj(cd a) {
    cd r = {a.re, -a.im};
    return r;
}

/* ======================================================================== */
/* FFT: recursive mixed-radix decimation in time with one twiddle table.     */
/* ======================================================================== */

typedef struct {
    int n;
    int sign;
    int nf;
    int f[40];
    cd* tw; /* tw[j] = e^{sign 2 pi i j / n} */
} fft_plan;

static int fft_plan_init(fft_plan* p, int n, int sign) {
    int m = n, nf = 0;
    p->n = n;
    p->sign = sign;
    while (m % 4 == 0) { p->f[nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[nf++] = 2; m /= 2; }
    for (int q = 3; q * q <= m; q += 2)
        while (m % q == 0) { p->f[nf++] = q; m /= q; }
    if (m > 1) p->f[nf++] = m;
    p->nf = nf;
    p->tw = (cd*)malloc(sizeof(cd) * (size_t)(n > 0 ? n : 1));
    if (!p->tw) return ORC_ENOMEM;
    for (int j = 0; j < n; j++) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        p->tw[j].re = cos(a);
        p->tw[j].im = sign * sin(a);
    }
    return ORC_OK;
}
static void fft_plan_free(fft_plan* p) { free(p->tw); p->tw = NULL; }

/* out[0..n) = DFT of in[0], in[is], ..., using factors f[0..nf), root stride ts */
static void fft_rec(const fft_plan* P, const cd* in, int is, cd* out, int n,
                    const int* f, int ts) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = f[0], m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec(P, in + (size_t)r * is, is * p, out + (size_t)r * m, m, f + 1, ts * p);
    const cd* tw = P->tw;
    if (p == 2) {
        for (int k = 0; k < m; k++) {
            cd a = out[k];
            cd b = cmul(out[m + k], tw[(size_t)k * ts]);
            out[k] = cadd(a, b);
            out[m + k] = csub(a, b);
        }
    } else if (p == 4) {
        double sg = (double)P->sign;
        for (int k = 0; k < m; k++) {
            cd t0 = out[k];
            cd t1 = cmul(out[m + k], tw[(size_t)k * ts]);
            cd t2 = cmul(out[2 * m + k], tw[(size_t)2 * k * ts]);
            cd t3 = cmul(out[3 * m + k], tw[(size_t)3 * k * ts]);
            cd s02 = cadd(t0, t2), d02 = csub(t0, t2);
            cd s13 = cadd(t1, t3), d13 = csub(t1, t3);
            /* w = e^{sign i pi/2} = sign*i ; (sign*i)*d13 = (-sign*d13.im, sign*d13.re) */
            cd wd = {-sg * d13.im, sg * d13.re};
            out[k] = cadd(s02, s13);
            out[m + k] = cadd(d02, wd);
            out[2 * m + k] = csub(s02, s13);
            out[3 * m + k] = csub(d02, wd);
        }
    } else {
        cd t[64];
        cd* tt = p <= 64 ? t : (cd*)malloc(sizeof(cd) * (size_t)p);
        for (int k = 0; k < m; k++) {
            for (int r = 0; r < p; r++)
                tt[r] = r == 0 ? out[k] : cmul(out[(size_t)r * m + k], tw[(size_t)r * k * ts]);
            for (int q = 0; q < p; q++) {
                cd acc = tt[0];
                for (int r = 1; r < p; r++) {
                    size_t e = (size_t)(((long long)r * q) % p) * (size_t)m * (size_t)ts;
                    acc = cadd(acc, cmul(tt[r], tw[e]));
                }
                out[(size_t)q * m + k] = acc; /* inputs of this k are already in tt[] */
            }
        }
        if (tt != t) free(tt);
    }
}

static int fft_exec(const fft_plan* P, cd* data, cd* scratch) {
    memcpy(scratch, data, sizeof(cd) * (size_t)P->n);
    fft_rec(P, scratch, 1, data, P->n, P->f, 1);
    return ORC_OK;
}

int orc_fft1d(int n, double* re, double* im, int sign) {
    if (n <= 0 || !re || !im || (sign != 1 && sign != -1)) return ORC_EARG;
    fft_plan P;
    if (fft_plan_init(&P, n, sign)) return ORC_ENOMEM;
    cd* d = (cd*)malloc(sizeof(cd) * (size_t)n * 2);
    if (!d) { fft_plan_free(&P); return ORC_ENOMEM; }
    for (int i = 0; i < n; i++) { d[i].re = re[i]; d[i].im = im[i]; }
    fft_exec(&P, d, d + n);
    for (int i = 0; i < n; i++) { re[i] = d[i].re; im[i] = d[i].im; }
    free(d);
    fft_plan_free(&P);
    return ORC_OK;
}

/* 2-D complex DFT on a cd grid (rows then columns). */
static int fft2d_cd(int sy, int sx, cd* a, int sign) {
    fft_plan px, py;
    if (fft_plan_init(&px, sx, sign)) return ORC_ENOMEM;
    if (fft_plan_init(&py, sy, sign)) { fft_plan_free(&px); return ORC_ENOMEM; }
    int mx = sx > sy ? sx : sy;
    cd* buf = (cd*)malloc(sizeof(cd) * (size_t)mx * 2);
    if (!buf) { fft_plan_free(&px); fft_plan_free(&py); return ORC_ENOMEM; }
    for (int y = 0; y < sy; y++) fft_exec(&px, a + (size_t)y * sx, buf);
    for (int x = 0; x < sx; x++) {
        cd* col = buf + mx;
        for (int y = 0; y < sy; y++) col[y] = a[(size_t)y * sx + x];
        fft_exec(&py, col, buf);
        for (int y = 0; y < sy; y++) a[(size_t)y * sx + x] = col[y];
    }
    free(buf);
    fft_plan_free(&px);
    fft_plan_free(&py);
    return ORC_OK;
}

int orc_fft2d(int sy, int sx, double* re, double* im, int sign) {
    if (sy <= 0 || sx <= 0 || !re || !im || (sign != 1 && sign != -1)) return ORC_EARG;
    size_t N = (size_t)sx * sy;
    cd* a = (cd*)malloc(sizeof(cd) * N);
    if (!a) return ORC_ENOMEM;
    for (size_t i = 0; i < N; i++) { a[i].re = re[i]; a[i].im = im[i]; }
    int rc = fft2d_cd(sy, sx, a, sign);
    for (size_t i = 0; i < N; i++) { re[i] = a[i].re; im[i] = a[i].im; }
    free(a);
    return rc;
}

/* myShift, L147-159: shiftType true -> X/2 (fftshift), false -> (X+1)/2 (ifftshift). */
static void shift_d(const double* in, double* out, int sx, int sy, int fwd) {
    int xh = fwd ? sx / 2 : (sx + 1) / 2;
    int yh = fwd ? sy / 2 : (sy + 1) / 2;
    for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++)
            out[(size_t)((y + yh) % sy) * sx + (x + xh) % sx] = in[(size_t)y * sx + x];
}

/* c2r of a sy x (sx/2+1) half spectrum with FFTW semantics (imaginary parts of
 * the x DC and Nyquist bins ignored), unnormalised; out is sy x sx. */
static int c2r_2d(const cd* half, int sy, int sx, double* out) {
    int wt = sx / 2 + 1;
    cd* h = (cd*)malloc(sizeof(cd) * (size_t)sy * wt);
    if (!h) return ORC_ENOMEM;
    memcpy(h, half, sizeof(cd) * (size_t)sy * wt);
    fft_plan py, px;
    if (fft_plan_init(&py, sy, 1)) { free(h); return ORC_ENOMEM; }
    if (fft_plan_init(&px, sx, 1)) { free(h); fft_plan_free(&py); return ORC_ENOMEM; }
    int mx = sx > sy ? sx : sy;
    cd* buf = (cd*)malloc(sizeof(cd) * (size_t)mx * 2);
    for (int x = 0; x < wt; x++) {
        cd* col = buf + mx;
        for (int y = 0; y < sy; y++) col[y] = h[(size_t)y * wt + x];
        fft_exec(&py, col, buf);
        for (int y = 0; y < sy; y++) h[(size_t)y * wt + x] = col[y];
    }
    for (int y = 0; y < sy; y++) {
        cd* row = buf + mx;
        const cd* hr = h + (size_t)y * wt;
        for (int k = 0; k < sx; k++) { row[k].re = 0; row[k].im = 0; }
        row[0].re = hr[0].re;
        int top = (sx % 2 == 0) ? sx / 2 : (sx - 1) / 2;
        for (int k = 1; k <= top; k++) {
            if (sx % 2 == 0 && k == sx / 2) {
                row[k].re = hr[k].re;
            } else {
                row[k] = hr[k];
                row[sx - k] = cconj(hr[k]);
            }
        }
        fft_exec(&px, row, buf);
        for (int x = 0; x < sx; x++) out[(size_t)y * sx + x] = row[x].re;
    }
    free(buf);
    free(h);
    fft_plan_free(&py);
    fft_plan_free(&px);
    return ORC_OK;
}

/* ======================================================================== */
/* Source and mask                                                           */
/* ======================================================================== */

void orc_normalize(float* m, int count) {
    float sum = 0.0f;
    for (int i = 0; i < count; ++i) sum += m[i];
    if (sum != 0.0f)
        for (int i = 0; i < count; ++i) m[i] /= sum;
}

int orc_source_annular(float* src, int srcSize, float inner, float outer) {
    if (!src || srcSize <= 0) return ORC_EARG;
    int half = (srcSize - 1) / 2;
    int innerSize = (int)(inner * (float)half);
    int outerSize = (int)(outer * (float)half);
    int cx = half, cy = half;
    for (int y = 0; y < srcSize; ++y)
        for (int x = 0; x < srcSize; ++x) {
            double d = sqrt(pow((double)(x - cx), 2) + pow((double)(y - cy), 2));
            src[y * srcSize + x] = (d >= innerSize && d <= outerSize) ? 1.0f : 0.0f;
        }
    orc_normalize(src, srcSize * srcSize);
    return ORC_OK;
}

int orc_source_quadrupole(float* src, int srcSize, float radius, float offset) {
    if (!src || srcSize <= 0) return ORC_EARG;
    int half = (srcSize - 1) / 2;
    int off = (int)(offset * (float)half);
    int rad = (int)(radius * (float)half);
    int cx = half, cy = half;
    int c[4][2] = {{cx, cy + off}, {cx, cy - off}, {cx + off, cy}, {cx - off, cy}};
    for (int y = 0; y < srcSize; ++y)
        for (int x = 0; x < srcSize; ++x) {
            int in = 0;
            for (int i = 0; i < 4 && !in; ++i) {
                double dx = x - c[i][0], dy = y - c[i][1];
                if (sqrt(dx * dx + dy * dy) <= rad) in = 1;
            }
            src[y * srcSize + x] = in ? 1.0f : 0.0f;
        }
    orc_normalize(src, srcSize * srcSize);
    return ORC_OK;
}

int orc_source_dipole(float* src, int srcSize, float radius, float offset, int onXAxis) {
    if (!src || srcSize <= 0) return ORC_EARG;
    int half = (srcSize - 1) / 2;
    int off = (int)(offset * (float)half);
    int rad = (int)(radius * (float)half);
    int cx = srcSize / 2, cy = srcSize / 2; /* source.cpp L61-62 uses size/2 here */
    int c[2][2] = {{cx + (onXAxis ? off : 0), cy + (onXAxis ? 0 : off)},
                   {cx - (onXAxis ? off : 0), cy - (onXAxis ? 0 : off)}};
    for (int y = 0; y < srcSize; ++y)
        for (int x = 0; x < srcSize; ++x) {
            int in = 0;
            for (int i = 0; i < 2 && !in; ++i) {
                double dx = x - c[i][0], dy = y - c[i][1];
                if (sqrt(dx * dx + dy * dy) <= rad) in = 1;
            }
            src[y * srcSize + x] = in ? 1.0f : 0.0f;
        }
    orc_normalize(src, srcSize * srcSize);
    return ORC_OK;
}

int orc_source_point(float* src, int srcSize, float px, float py) {
    if (!src || srcSize <= 0) return ORC_EARG;
    memset(src, 0, sizeof(float) * (size_t)srcSize * srcSize);
    int ix = (int)round((double)(px * (float)srcSize) / 2.0 + srcSize / 2.0);
    int iy = (int)round((double)(-1.0f * py * (float)srcSize) / 2.0 + srcSize / 2.0);
    if (px >= 0) ix -= 1;
    if (-1.0f * py >= 0) iy -= 1;
    if (ix < 0 || iy < 0 || ix >= srcSize || iy >= srcSize) return ORC_ERANGE;
    src[iy * srcSize + ix] = 1.0f;
    orc_normalize(src, srcSize * srcSize);
    return ORC_OK;
}

float orc_outer_sigma(const float* src, int srcSize) {
    int center = (srcSize - 1) / 2;
    float maxd = 0.0f;
    for (int y = 0; y < srcSize; y++)
        for (int x = 0; x < srcSize; x++)
            if (src[y * srcSize + x] != 0.0f) {
                float d = (float)sqrt((double)((y - center) * (y - center) + (x - center) * (x - center)));
                if (d > maxd) maxd = d;
            }
    return center == 0 ? 0.0f : maxd / (float)center;
}

int orc_compute_nx(float NA, int period, float outerSigma, float lambda) {
    float v = NA * (float)period * (1.0f + outerSigma) / lambda;
    return (int)floorf(v);
}

int orc_linespace(float* msk, int sizeX, int sizeY, int lineWidth, int spaceWidth, int isHorizontal) {
    if (!msk || sizeX <= 0 || sizeY <= 0 || lineWidth + spaceWidth <= 0) return ORC_EARG;
    int period = lineWidth + spaceWidth;
    for (int y = 0; y < sizeY; ++y)
        for (int x = 0; x < sizeX; ++x) {
            int c = isHorizontal ? y : x;
            msk[y * sizeX + x] = (c % period) < lineWidth ? 1.0f : 0.0f;
        }
    return ORC_OK;
}

int orc_place_mask(float* msk, int Lx, int Ly, const float* inMsk, int maskSizeX, int maskSizeY, float dose) {
    if (!msk || !inMsk || maskSizeX > Lx || maskSizeY > Ly) return ORC_EARG;
    int difYh = (Ly - maskSizeY) / 2, difXh = (Lx - maskSizeX) / 2;
    for (int y = 0; y < maskSizeY; y++)
        for (int x = 0; x < maskSizeX; x++)
            msk[(y + difYh) * Lx + x + difXh] = inMsk[y * maskSizeX + x] * dose;
    return ORC_OK;
}

/* ======================================================================== */
/* Module 3: mask spectrum                                                   */
/* ======================================================================== */

int orc_mask_spectrum(const float* mask, int sy, int sx, double* out_re, double* out_im) {
    if (!mask || !out_re || !out_im || sx <= 0 || sy <= 0 || (sx % 2) || (sy % 2)) return ORC_EARG;
    size_t N = (size_t)sx * sy;
    cd* a = (cd*)malloc(sizeof(cd) * N);
    if (!a) return ORC_ENOMEM;
    /* ifftshift (myShift false,false, L269) */
    int xh = (sx + 1) / 2, yh = (sy + 1) / 2;
    for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
            cd v = {(double)mask[(size_t)y * sx + x], 0.0};
            a[(size_t)((y + yh) % sy) * sx + (x + xh) % sx] = v;
        }
    int rc = fft2d_cd(sy, sx, a, -1);
    if (rc) { free(a); return rc; }
    /* centred layout: out[y][x] = X[(y - sy/2) mod sy][(x - sx/2) mod sx] / (sx*sy) */
    double inv = (double)(sx * sy);
    for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
            cd v = a[(size_t)((y + sy / 2) % sy) * sx + (x + sx / 2) % sx];
            out_re[(size_t)y * sx + x] = v.re / inv;
            out_im[(size_t)y * sx + x] = v.im / inv;
        }
    free(a);
    return ORC_OK;
}

/* ======================================================================== */
/* Module 4: pupil and TCC                                                   */
/* ======================================================================== */

typedef struct {
    int sh, oSgm;
    float dz, k, Lx_norm, Ly_norm, NA;
} tcc_par;

static void tcc_params(tcc_par* t, int srcSize, float outerSigma, float lambda, float NA,
                       float defocus, int Lx, int Ly) {
    t->sh = (int)((srcSize - 1) / 2.0);
    t->oSgm = (int)ceil((double)((float)t->sh * outerSigma));
    t->dz = defocus / (NA * NA / lambda);
    t->k = (float)(2 * M_PI / (double)lambda);
    t->Lx_norm = (float)Lx * NA / lambda;
    t->Ly_norm = (float)Ly * NA / lambda;
    t->NA = NA;
}

/* One source point's pupil support, calcTCC L458-484.  Calls back per (nx,ny). */
typedef struct {
    int nxMin, nxMax, nyMin, nyMax;
    float sx, sy;
} src_box;

static void src_box_of(const tcc_par* t, int p, int q, src_box* b) {
    b->sx = (float)p / (float)t->sh;
    b->sy = (float)q / (float)t->sh;
    b->nxMin = (int)((-1.0f - b->sx) * t->Lx_norm);
    b->nxMax = (int)((1.0f - b->sx) * t->Lx_norm);
    b->nyMin = (int)((-1.0f - b->sy) * t->Ly_norm);
    b->nyMax = (int)((1.0f - b->sy) * t->Ly_norm);
}

static inline int src_live(const float* src, int srcSize, const tcc_par* t, int p, int q) {
    int id = (q + t->sh) * srcSize + p + t->sh;
    return src[id] != 0.0f && (p * p + q * q) <= t->sh * t->sh;
}

/* Fill column s of a (n x ncol) pupil matrix for source (p,q).  Returns
 * ORC_ERANGE when the support leaves the (2Nx+1)x(2Ny+1) window. */
static int pupil_column(const tcc_par* t, int p, int q, int Nx, int Ny,
                        double* Pr, double* Pi, size_t stride) {
    src_box b;
    src_box_of(t, p, q, &b);
    for (int ny = b.nyMin; ny <= b.nyMax; ny++) {
        float fy = (float)ny / t->Ly_norm + b.sy;
        if (fy * fy <= 1.0f) {
            for (int nx = b.nxMin; nx <= b.nxMax; nx++) {
                float fx = (float)nx / t->Lx_norm + b.sx;
                float rho2 = fx * fx + fy * fy;
                if (rho2 <= 1.0f) {
                    if (nx < -Nx || nx > Nx || ny < -Ny || ny > Ny) return ORC_ERANGE;
                    float na2 = rho2 * t->NA * t->NA;
                    double arg = 1.0 - (double)na2;
                    float ph = (float)((double)(t->dz * t->k) * sqrt(arg));
                    size_t id = (size_t)(ny + Ny) * (2 * Nx + 1) + (nx + Nx);
                    Pr[id * stride] = (double)cosf(ph);
                    Pi[id * stride] = (double)sinf(ph);
                }
            }
        }
    }
    return ORC_OK;
}

int orc_count_live_sources(const float* src, int srcSize, float outerSigma) {
    tcc_par t;
    tcc_params(&t, srcSize, outerSigma, 193.0f, 1.0f, 0.0f, 1, 1);
    int c = 0;
    for (int q = -t.oSgm; q <= t.oSgm; q++)
        for (int p = -t.oSgm; p <= t.oSgm; p++)
            if (src_live(src, srcSize, &t, p, q)) c++;
    return c;
}

int orc_pupil_matrix(double* A_re, double* A_im, double* S, int* nsrc_out, int nsrc_cap,
                     const float* src, int srcSize, float outerSigma, float lambda, float NA,
                     float defocus, int Lx, int Ly, int Nx, int Ny) {
    if (!src || !nsrc_out || srcSize <= 0 || Nx < 0 || Ny < 0) return ORC_EARG;
    tcc_par t;
    tcc_params(&t, srcSize, outerSigma, lambda, NA, defocus, Lx, Ly);
    int ns = 0;
    for (int q = -t.oSgm; q <= t.oSgm; q++)
        for (int p = -t.oSgm; p <= t.oSgm; p++)
            if (src_live(src, srcSize, &t, p, q)) ns++;
    *nsrc_out = ns;
    if (!A_re) return ORC_OK;
    if (!A_im || !S || nsrc_cap < ns) return ORC_EARG;
    size_t n = (size_t)(2 * Nx + 1) * (2 * Ny + 1);
    memset(A_re, 0, sizeof(double) * n * ns);
    memset(A_im, 0, sizeof(double) * n * ns);
    int s = 0;
    for (int q = -t.oSgm; q <= t.oSgm; q++)
        for (int p = -t.oSgm; p <= t.oSgm; p++) {
            if (!src_live(src, srcSize, &t, p, q)) continue;
            int rc = pupil_column(&t, p, q, Nx, Ny, A_re + s, A_im + s, (size_t)ns);
            if (rc) return rc;
            S[s] = (double)src[(q + t.sh) * srcSize + p + t.sh];
            s++;
        }
    return ORC_OK;
}

int orc_calc_tcc(double* tcc_re, double* tcc_im, int tccSize, const float* src, int srcSize,
                 float outerSigma, float lambda, float NA, float defocus, int Lx, int Ly, int Nx,
                 int Ny) {
    int cols = 2 * Nx + 1, rows = 2 * Ny + 1;
    if (!tcc_re || !tcc_im || !src || tccSize != cols * rows) return ORC_EARG;
    int ns = 0;
    int rc = orc_pupil_matrix(NULL, NULL, NULL, &ns, 0, src, srcSize, outerSigma, lambda, NA,
                              defocus, Lx, Ly, Nx, Ny);
    if (rc) return rc;
    size_t n = (size_t)tccSize;
    double* Ar = (double*)calloc(n * (size_t)(ns ? ns : 1), sizeof(double));
    double* Ai = (double*)calloc(n * (size_t)(ns ? ns : 1), sizeof(double));
    double* S = (double*)calloc((size_t)(ns ? ns : 1), sizeof(double));
    int* bx0 = (int*)malloc(sizeof(int) * 3 * (size_t)(ns ? ns : 1));
    if (!Ar || !Ai || !S || !bx0) { free(Ar); free(Ai); free(S); free(bx0); return ORC_ENOMEM; }
    rc = orc_pupil_matrix(Ar, Ai, S, &ns, ns, src, srcSize, outerSigma, lambda, NA, defocus, Lx,
                          Ly, Nx, Ny);
    if (rc) { free(Ar); free(Ai); free(S); free(bx0); return rc; }
    tcc_par t;
    tcc_params(&t, srcSize, outerSigma, lambda, NA, defocus, Lx, Ly);
    {
        int s = 0;
        for (int q = -t.oSgm; q <= t.oSgm; q++)
            for (int p = -t.oSgm; p <= t.oSgm; p++) {
                if (!src_live(src, srcSize, &t, p, q)) continue;
                src_box b;
                src_box_of(&t, p, q, &b);
                bx0[3 * s + 0] = b.nxMin;
                bx0[3 * s + 1] = b.nxMax;
                bx0[3 * s + 2] = b.nyMax;
                s++;
            }
    }
    memset(tcc_re, 0, sizeof(double) * n * n);
    memset(tcc_im, 0, sizeof(double) * n * n);
    /* Row-parallel upper-triangle accumulation, L495-525. */
#pragma omp parallel for schedule(dynamic)
    for (int ID1 = 0; ID1 < tccSize; ++ID1) {
        int ny = ID1 / cols - Ny;
        int nx = ID1 % cols - Nx;
        double* tr = tcc_re + (size_t)ID1 * n;
        double* ti = tcc_im + (size_t)ID1 * n;
        for (int s = 0; s < ns; s++) {
            double p1r = Ar[(size_t)ID1 * ns + s], p1i = Ai[(size_t)ID1 * ns + s];
            if (p1r == 0.0 && p1i == 0.0) continue;
            double sv = S[s];
            /* (p1 * (sv,0)) as std::complex multiplication */
            double ar = p1r * sv - p1i * 0.0;
            double ai = p1r * 0.0 + p1i * sv;
            int nxMin = bx0[3 * s + 0], nxMax = bx0[3 * s + 1], nyMax = bx0[3 * s + 2];
            for (int my = ny; my <= nyMax; my++) {
                int startmx = (my == ny) ? nx : nxMin;
                for (int mx = startmx; mx <= nxMax; mx++) {
                    if (my < -Ny || my > Ny || mx < -Nx || mx > Nx) continue;
                    int ID2 = (my + Ny) * cols + (mx + Nx);
                    double br = Ar[(size_t)ID2 * ns + s], bi = -Ai[(size_t)ID2 * ns + s];
                    tr[ID2] += ar * br - ai * bi;
                    ti[ID2] += ar * bi + ai * br;
                }
            }
        }
    }
    /* Fill the lower triangle, L527-532. */
    for (size_t i = 0; i < n; i++)
        for (size_t j = i + 1; j < n; j++) {
            tcc_re[j * n + i] = tcc_re[i * n + j];
            tcc_im[j * n + i] = -tcc_im[i * n + j];
        }
    free(Ar); free(Ai); free(S); free(bx0);
    return ORC_OK;
}

/* ======================================================================== */
/* Module 6: eigen (cyclic Jacobi) and kernels                               */
/* ======================================================================== */

static int jacobi_real(int n, double* A, double* V) {
    double fro = 0.0;
    for (size_t i = 0; i < (size_t)n * n; i++) fro += A[i] * A[i];
    fro = sqrt(fro);
    double tol = 1e-17 * fro;
    for (size_t i = 0; i < (size_t)n * n; i++) V[i] = 0.0;
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    if (fro == 0.0) return ORC_OK;
    for (int sweep = 0; sweep < 60; sweep++) {
        int rot = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[(size_t)p * n + q];
                if (fabs(apq) <= tol) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                double* rp = A + (size_t)p * n;
                double* rq = A + (size_t)q * n;
                for (int k = 0; k < n; k++) {
                    if (k == p || k == q) continue;
                    double x = rp[k], y = rq[k];
                    double nxv = c * x - s * y, nyv = s * x + c * y;
                    rp[k] = nxv;
                    rq[k] = nyv;
                    A[(size_t)k * n + p] = nxv;
                    A[(size_t)k * n + q] = nyv;
                }
                rp[p] = app - t * apq;
                rq[q] = aqq + t * apq;
                rp[q] = 0.0;
                rq[p] = 0.0;
                double* vp = V + (size_t)p * n;
                double* vq = V + (size_t)q * n;
                for (int k = 0; k < n; k++) {
                    double x = vp[k], y = vq[k];
                    vp[k] = c * x - s * y;
                    vq[k] = s * x + c * y;
                }
                rot++;
            }
        if (!rot) return ORC_OK;
    }
    return ORC_ENOCONV;
}

static int jacobi_complex(int n, double* Ar, double* Ai, double* Vr, double* Vi) {
    double fro = 0.0;
    for (size_t i = 0; i < (size_t)n * n; i++) fro += Ar[i] * Ar[i] + Ai[i] * Ai[i];
    fro = sqrt(fro);
    double tol = 1e-17 * fro;
    for (size_t i = 0; i < (size_t)n * n; i++) { Vr[i] = 0.0; Vi[i] = 0.0; }
    for (int i = 0; i < n; i++) Vr[(size_t)i * n + i] = 1.0;
    for (int i = 0; i < n; i++) Ai[(size_t)i * n + i] = 0.0;
    if (fro == 0.0) return ORC_OK;
    for (int sweep = 0; sweep < 60; sweep++) {
        int rot = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double zr = Ar[(size_t)p * n + q], zi = Ai[(size_t)p * n + q];
                double r = hypot(zr, zi);
                if (r <= tol) continue;
                double er = zr / r, ei = zi / r; /* e^{i phi} */
                double a = Ar[(size_t)p * n + p], b = Ar[(size_t)q * n + q];
                double theta = (b - a) / (2.0 * r);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                /* rows: p' = c p - s e^{i phi} q ; q' = s p + c e^{i phi} q */
                double* rpr = Ar + (size_t)p * n; double* rpi = Ai + (size_t)p * n;
                double* rqr = Ar + (size_t)q * n; double* rqi = Ai + (size_t)q * n;
                for (int k = 0; k < n; k++) {
                    if (k == p || k == q) continue;
                    double xr = rpr[k], xi = rpi[k], yr = rqr[k], yi = rqi[k];
                    double eyr = er * yr - ei * yi, eyi = er * yi + ei * yr;
                    double npr = c * xr - s * eyr, npi = c * xi - s * eyi;
                    double nqr = s * xr + c * eyr, nqi = s * xi + c * eyi;
                    rpr[k] = npr; rpi[k] = npi; rqr[k] = nqr; rqi[k] = nqi;
                    Ar[(size_t)k * n + p] = npr; Ai[(size_t)k * n + p] = -npi;
                    Ar[(size_t)k * n + q] = nqr; Ai[(size_t)k * n + q] = -nqi;
                }
                rpr[p] = a - t * r; rpi[p] = 0.0;
                rqr[q] = b + t * r; rqi[q] = 0.0;
                rpr[q] = 0.0; rpi[q] = 0.0; rqr[p] = 0.0; rqi[p] = 0.0;
                /* V <- V J: col p' = c vp - s e^{-i phi} vq ; col q' = s vp + c e^{-i phi} vq */
                double* vpr = Vr + (size_t)p * n; double* vpi = Vi + (size_t)p * n;
                double* vqr = Vr + (size_t)q * n; double* vqi = Vi + (size_t)q * n;
                for (int k = 0; k < n; k++) {
                    double xr = vpr[k], xi = vpi[k], yr = vqr[k], yi = vqi[k];
                    double eyr = er * yr + ei * yi, eyi = er * yi - ei * yr;
                    vpr[k] = c * xr - s * eyr; vpi[k] = c * xi - s * eyi;
                    vqr[k] = s * xr + c * eyr; vqi[k] = s * xi + c * eyi;
                }
                rot++;
            }
        if (!rot) return ORC_OK;
    }
    return ORC_ENOCONV;
}

/* Sort eigenpairs ascending by w (selection sort on indices, stable). */
static int sort_ascending(int n, double* w, double* Vr, double* Vi) {
    int* idx = (int*)malloc(sizeof(int) * (size_t)n);
    double* w2 = (double*)malloc(sizeof(double) * (size_t)n);
    double* t = (double*)malloc(sizeof(double) * (size_t)n * n);
    if (!idx || !w2 || !t) { free(idx); free(w2); free(t); return ORC_ENOMEM; }
    for (int i = 0; i < n; i++) idx[i] = i;
    /* insertion sort: stable, n <= a few thousand */
    for (int i = 1; i < n; i++) {
        int v = idx[i], j = i - 1;
        while (j >= 0 && w[idx[j]] > w[v]) { idx[j + 1] = idx[j]; j--; }
        idx[j + 1] = v;
    }
    for (int i = 0; i < n; i++) w2[i] = w[idx[i]];
    memcpy(w, w2, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) memcpy(t + (size_t)i * n, Vr + (size_t)idx[i] * n, sizeof(double) * (size_t)n);
    memcpy(Vr, t, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; i++) memcpy(t + (size_t)i * n, Vi + (size_t)idx[i] * n, sizeof(double) * (size_t)n);
    memcpy(Vi, t, sizeof(double) * (size_t)n * n);
    free(idx); free(w2); free(t);
    return ORC_OK;
}

int orc_eigh(int n, double* H_re, double* H_im, double* w, double* V_re, double* V_im) {
    if (n <= 0 || !H_re || !H_im || !w || !V_re || !V_im) return ORC_EARG;
    int real = 1;
    for (size_t i = 0; i < (size_t)n * n && real; i++)
        if (H_im[i] != 0.0) real = 0;
    int rc;
    if (real) {
        rc = jacobi_real(n, H_re, V_re);
        memset(V_im, 0, sizeof(double) * (size_t)n * n);
    } else {
        rc = jacobi_complex(n, H_re, H_im, V_re, V_im);
    }
    if (rc) return rc;
    for (int i = 0; i < n; i++) w[i] = H_re[(size_t)i * n + i];
    return sort_ascending(n, w, V_re, V_im);
}

/* mirror operators on the (2Nx+1)x(2Ny+1) index grid */
static inline int mirror_index(int id, int Nx, int Ny, int which) {
    int W = 2 * Nx + 1;
    int ix = id % W - Nx, iy = id / W - Ny;
    if (which == 0) ix = -ix;
    else if (which == 1) iy = -iy;
    else { int tmp = ix; ix = iy; iy = tmp; }
    return (iy + Ny) * W + (ix + Nx);
}

/* Rotate the c columns [a, a+c) of V (column-major n x m) by the eigenvectors
 * of M = V_c^H Sigma V_c, ordered by eigenvalue descending; mu receives them. */
static int mirror_rotate(double* Vr, double* Vi, int n, int a, int c, int Nx, int Ny, int which,
                         double* mu) {
    size_t cc = (size_t)c * c;
    double* Mr = (double*)calloc(cc, sizeof(double));
    double* Mi = (double*)calloc(cc, sizeof(double));
    double* Qr = (double*)calloc(cc, sizeof(double));
    double* Qi = (double*)calloc(cc, sizeof(double));
    double* tr = (double*)calloc((size_t)n * c, sizeof(double));
    double* ti = (double*)calloc((size_t)n * c, sizeof(double));
    if (!Mr || !Mi || !Qr || !Qi || !tr || !ti) { free(Mr); free(Mi); free(Qr); free(Qi); free(tr); free(ti); return ORC_ENOMEM; }
    for (int i = 0; i < c; i++)
        for (int j = 0; j < c; j++) {
            const double* air = Vr + (size_t)(a + i) * n; const double* aii = Vi + (size_t)(a + i) * n;
            const double* ajr = Vr + (size_t)(a + j) * n; const double* aji = Vi + (size_t)(a + j) * n;
            double sr = 0, si = 0;
            for (int k = 0; k < n; k++) {
                int mk = mirror_index(k, Nx, Ny, which);
                /* conj(v_i[k]) * v_j[mk] */
                sr += air[k] * ajr[mk] + aii[k] * aji[mk];
                si += air[k] * aji[mk] - aii[k] * ajr[mk];
            }
            Mr[(size_t)i * c + j] = sr;
            Mi[(size_t)i * c + j] = si;
        }
    /* symmetrise against round-off */
    for (int i = 0; i < c; i++)
        for (int j = i; j < c; j++) {
            double r = 0.5 * (Mr[(size_t)i * c + j] + Mr[(size_t)j * c + i]);
            double im = 0.5 * (Mi[(size_t)i * c + j] - Mi[(size_t)j * c + i]);
            Mr[(size_t)i * c + j] = r; Mr[(size_t)j * c + i] = r;
            Mi[(size_t)i * c + j] = im; Mi[(size_t)j * c + i] = -im;
        }
    double* wv = (double*)malloc(sizeof(double) * (size_t)c);
    int rc = orc_eigh(c, Mr, Mi, wv, Qr, Qi);
    if (!rc) {
        /* descending order: column j of the result is Q column c-1-j */
        for (int j = 0; j < c; j++) {
            int qj = c - 1 - j;
            mu[j] = wv[qj];
            for (int k = 0; k < n; k++) {
                double sr = 0, si = 0;
                for (int i = 0; i < c; i++) {
                    double vr = Vr[(size_t)(a + i) * n + k], vi = Vi[(size_t)(a + i) * n + k];
                    double qr = Qr[(size_t)qj * c + i], qi = Qi[(size_t)qj * c + i];
                    sr += vr * qr - vi * qi;
                    si += vr * qi + vi * qr;
                }
                tr[(size_t)j * n + k] = sr;
                ti[(size_t)j * n + k] = si;
            }
        }
        memcpy(Vr + (size_t)a * n, tr, sizeof(double) * (size_t)n * c);
        memcpy(Vi + (size_t)a * n, ti, sizeof(double) * (size_t)n * c);
    }
    free(wv); free(Mr); free(Mi); free(Qr); free(Qi); free(tr); free(ti);
    return rc;
}

static int canon_group(double* Vr, double* Vi, int n, int a, int c, int Nx, int Ny, int which) {
    if (c < 2 || which > 2) return ORC_OK;
    if (which == 2 && Nx != Ny) return ORC_OK;
    double* mu = (double*)malloc(sizeof(double) * (size_t)c);
    int rc = mirror_rotate(Vr, Vi, n, a, c, Nx, Ny, which, mu);
    if (!rc) {
        int i = 0;
        while (i < c) {
            int j = i + 1;
            while (j < c && fabs(mu[j] - mu[i]) < 1e-6) j++;
            if (j - i > 1) {
                rc = canon_group(Vr, Vi, n, a + i, j - i, Nx, Ny, which + 1);
                if (rc) break;
            }
            i = j;
        }
    }
    free(mu);
    return rc;
}

int orc_canonicalize(double* V_re, double* V_im, const double* w, int n, int m, int Nx, int Ny,
                     double rtol) {
    if (!V_re || !V_im || !w || n <= 0 || m <= 0 || m > n * 4096) return ORC_EARG;
    double scale = fabs(w[0]);
    int i = 0;
    while (i < m) {
        int j = i + 1;
        while (j < m && fabs(w[j - 1] - w[j]) <= rtol * scale) j++;
        if (j - i > 1) {
            int rc = canon_group(V_re, V_im, n, i, j - i, Nx, Ny, 0);
            if (rc) return rc;
        }
        i = j;
    }
    /* phase: first component within (1-1e-6) of the max magnitude -> real positive */
    for (int c = 0; c < m; c++) {
        double* vr = V_re + (size_t)c * n;
        double* vi = V_im + (size_t)c * n;
        double mx = 0;
        for (int k = 0; k < n; k++) {
            double a = hypot(vr[k], vi[k]);
            if (a > mx) mx = a;
        }
        if (mx == 0) continue;
        int kk = 0;
        for (int k = 0; k < n; k++)
            if (hypot(vr[k], vi[k]) >= (1.0 - 1e-6) * mx) { kk = k; break; }
        double a = hypot(vr[kk], vi[kk]);
        double pr = vr[kk] / a, pi = -vi[kk] / a; /* conj(v_kk)/|v_kk| */
        for (int k = 0; k < n; k++) {
            double xr = vr[k], xi = vi[k];
            vr[k] = xr * pr - xi * pi;
            vi[k] = xr * pi + xi * pr;
        }
        vi[kk] = 0.0;
    }
    return ORC_OK;
}

/* Shared tail of both kernel routes: w ascending (length m), V column-major
 * n x m.  Takes the top nk, canonicalises and flips. */
static int kernels_from_eig(double* krn_re, double* krn_im, double* scales, int nk, double* w,
                            double* Vr, double* Vi, int n, int m, int Nx, int Ny,
                            int* cluster_split) {
    const double rtol = 1e-10;
    /* reorder to descending */
    double* wd = (double*)malloc(sizeof(double) * (size_t)m);
    double* Dr = (double*)malloc(sizeof(double) * (size_t)n * m);
    double* Di = (double*)malloc(sizeof(double) * (size_t)n * m);
    if (!wd || !Dr || !Di) { free(wd); free(Dr); free(Di); return ORC_ENOMEM; }
    for (int j = 0; j < m; j++) {
        wd[j] = w[m - 1 - j];
        memcpy(Dr + (size_t)j * n, Vr + (size_t)(m - 1 - j) * n, sizeof(double) * (size_t)n);
        memcpy(Di + (size_t)j * n, Vi + (size_t)(m - 1 - j) * n, sizeof(double) * (size_t)n);
    }
    int split = 0;
    if (nk < m && fabs(wd[nk - 1] - wd[nk]) <= rtol * fabs(wd[0])) split = 1;
    if (cluster_split) *cluster_split = split;
    /* canonicalise every cluster that starts inside the top nk */
    int mc = nk;
    while (mc < m && fabs(wd[mc - 1] - wd[mc]) <= rtol * fabs(wd[0])) mc++;
    int rc = orc_canonicalize(Dr, Di, wd, n, mc, Nx, Ny, rtol);
    if (!rc) {
        for (int k = 0; k < nk; k++) {
            scales[k] = wd[k];
            for (int j = 0; j < n; j++) {
                krn_re[(size_t)k * n + j] = Dr[(size_t)k * n + (n - 1 - j)];
                krn_im[(size_t)k * n + j] = Di[(size_t)k * n + (n - 1 - j)];
            }
        }
    }
    free(wd); free(Dr); free(Di);
    return rc;
}

int orc_calc_kernels(double* krn_re, double* krn_im, double* scales, int nk, const double* tcc_re,
                     const double* tcc_im, int n, int Nx, int Ny, int* cluster_split) {
    if (!krn_re || !krn_im || !scales || !tcc_re || !tcc_im || nk <= 0 || nk > n ||
        n != (2 * Nx + 1) * (2 * Ny + 1))
        return ORC_EARG;
    size_t nn = (size_t)n * n;
    double* Hr = (double*)malloc(sizeof(double) * nn);
    double* Hi = (double*)malloc(sizeof(double) * nn);
    double* Vr = (double*)malloc(sizeof(double) * nn);
    double* Vi = (double*)malloc(sizeof(double) * nn);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    if (!Hr || !Hi || !Vr || !Vi || !w) { free(Hr); free(Hi); free(Vr); free(Vi); free(w); return ORC_ENOMEM; }
    /* LAPACK 'U': the upper triangle defines the matrix (L621-640) */
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++) {
            Hr[(size_t)i * n + j] = tcc_re[(size_t)i * n + j];
            Hi[(size_t)i * n + j] = tcc_im[(size_t)i * n + j];
            Hr[(size_t)j * n + i] = tcc_re[(size_t)i * n + j];
            Hi[(size_t)j * n + i] = -tcc_im[(size_t)i * n + j];
        }
    int rc = orc_eigh(n, Hr, Hi, w, Vr, Vi);
    if (!rc) rc = kernels_from_eig(krn_re, krn_im, scales, nk, w, Vr, Vi, n, n, Nx, Ny, cluster_split);
    free(Hr); free(Hi); free(Vr); free(Vi); free(w);
    return rc;
}

int orc_calc_kernels_gram(double* krn_re, double* krn_im, double* scales, int nk,
                          const double* A_re, const double* A_im, const double* S, int n, int nsrc,
                          int Nx, int Ny, int* cluster_split) {
    if (!krn_re || !krn_im || !scales || !A_re || !A_im || !S || nk <= 0 || nk > nsrc ||
        nk > n || n != (2 * Nx + 1) * (2 * Ny + 1))
        return ORC_EARG;
    size_t m = (size_t)nsrc;
    double* sq = (double*)malloc(sizeof(double) * m);
    double* Gr = (double*)calloc(m * m, sizeof(double));
    double* Gi = (double*)calloc(m * m, sizeof(double));
    double* Ur = (double*)malloc(sizeof(double) * m * m);
    double* Ui = (double*)malloc(sizeof(double) * m * m);
    double* w = (double*)malloc(sizeof(double) * m);
    double* Lr = (double*)calloc((size_t)n * m, sizeof(double));
    double* Li = (double*)calloc((size_t)n * m, sizeof(double));
    if (!sq || !Gr || !Gi || !Ur || !Ui || !w || !Lr || !Li) {
        free(sq); free(Gr); free(Gi); free(Ur); free(Ui); free(w); free(Lr); free(Li);
        return ORC_ENOMEM;
    }
    for (size_t s = 0; s < m; s++) sq[s] = sqrt(S[s]);
    /* G[s][t] = sqrt(S_s S_t) sum_i conj(A[i][s]) A[i][t] */
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsrc; s++)
        for (int t = s; t < nsrc; t++) {
            double gr = 0, gi = 0;
            for (int i = 0; i < n; i++) {
                double ar = A_re[(size_t)i * m + s], ai = A_im[(size_t)i * m + s];
                double br = A_re[(size_t)i * m + t], bi = A_im[(size_t)i * m + t];
                gr += ar * br + ai * bi;
                gi += ar * bi - ai * br;
            }
            double f = sq[s] * sq[t];
            Gr[(size_t)s * m + t] = f * gr; Gi[(size_t)s * m + t] = f * gi;
            Gr[(size_t)t * m + s] = f * gr; Gi[(size_t)t * m + s] = -f * gi;
        }
    int rc = orc_eigh(nsrc, Gr, Gi, w, Ur, Ui);
    int keep = nsrc < n ? nsrc : n;
    if (!rc) {
        /* lift the top `keep` eigenvectors: v = B u / sqrt(lambda); columns stored ascending */
        for (int jj = 0; jj < keep; jj++) {
            int j = nsrc - keep + jj;
            double lam = w[j];
            double inv = lam > 0 ? 1.0 / sqrt(lam) : 0.0;
            const double* ur = Ur + (size_t)j * m;
            const double* ui = Ui + (size_t)j * m;
            for (int i = 0; i < n; i++) {
                double sr = 0, si = 0;
                for (size_t s = 0; s < m; s++) {
                    double br = A_re[(size_t)i * m + s] * sq[s], bi = A_im[(size_t)i * m + s] * sq[s];
                    sr += br * ur[s] - bi * ui[s];
                    si += br * ui[s] + bi * ur[s];
                }
                Lr[(size_t)jj * n + i] = sr * inv;
                Li[(size_t)jj * n + i] = si * inv;
            }
        }
        rc = kernels_from_eig(krn_re, krn_im, scales, nk, w + (nsrc - keep), Lr, Li, n, keep, Nx,
                              Ny, cluster_split);
    }
    free(sq); free(Gr); free(Gi); free(Ur); free(Ui); free(w); free(Lr); free(Li);
    return rc;
}

/* ======================================================================== */
/* Module 7: SOCS and FI                                                     */
/* ======================================================================== */

int orc_conv_grid(int Nx) {
    int n = 4 * Nx + 1;
    int g = 1;
    while (g < n) g <<= 1;
    return g;
}

int orc_socs_tmpimg(double* tmpImgp, int Gx, int Gy, const double* mskf_re,
                    const double* mskf_im, int fx, int fy, const double* krn_re,
                    const double* krn_im, const float* scales, int nk, int Nx, int Ny,
                    int nthreads) {
    int kerX = 2 * Nx + 1, kerY = 2 * Ny + 1;
    if (!tmpImgp || !mskf_re || !mskf_im || !krn_re || !krn_im || !scales || nk < 0 ||
        Gx < kerX || Gy < kerY)
        return ORC_EARG;
    size_t G = (size_t)Gx * Gy;
    int embedX = (Gx - kerX) / 2, embedY = (Gy - kerY) / 2; /* L847-848 */
    int fxh = fx / 2, fyh = fy / 2;
    double* tmpImg = (double*)calloc(G, sizeof(double));
    if (!tmpImg) return ORC_ENOMEM;
    int nt = nthreads > 0 ? nthreads : 1;
    double* priv = (double*)calloc(G * (size_t)nt, sizeof(double));
    if (!priv) { free(tmpImg); return ORC_ENOMEM; }
    int fail = 0;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* acc = priv + G * (size_t)tid;
        cd* buf = (cd*)malloc(sizeof(cd) * G);
        if (!buf) {
#pragma omp atomic write
            fail = 1;
        } else {
#pragma omp for schedule(static)
            for (int k = 0; k < nk; ++k) {
                memset(buf, 0, sizeof(cd) * G);
                /* embed K_k * M, L909-936 */
                for (int y = -Ny; y <= Ny; ++y)
                    for (int x = -Nx; x <= Nx; ++x) {
                        int ky = y + Ny, kx = x + Nx;
                        int by = embedY + ky, bx = embedX + kx;
                        int my = y + fyh, mx = x + fxh;
                        if (by >= 0 && by < Gy && bx >= 0 && bx < Gx && my >= 0 && my < fy &&
                            mx >= 0 && mx < fx) {
                            cd kv = {krn_re[(size_t)k * kerX * kerY + ky * kerX + kx],
                                     krn_im[(size_t)k * kerX * kerY + ky * kerX + kx]};
                            cd mv = {mskf_re[(size_t)my * fx + mx], mskf_im[(size_t)my * fx + mx]};
                            buf[(size_t)by * Gx + bx] = cmul(kv, mv);
                        }
                    }
                fft2d_cd(Gy, Gx, buf, +1); /* FFTW_BACKWARD, unnormalised, L942 */
                double sc = (double)scales[k];
                for (size_t i = 0; i < G; ++i) {
                    double re = buf[i].re, im = buf[i].im;
                    acc[i] += sc * (re * re + im * im); /* L951 */
                }
            }
            free(buf);
        }
    }
    /* reduction in thread order (the reference's omp critical, L955-960) */
    for (int t = 0; t < nt; t++)
        for (size_t i = 0; i < G; i++) tmpImg[i] += priv[G * (size_t)t + i];
    free(priv);
    if (fail) { free(tmpImg); return ORC_ENOMEM; }
    shift_d(tmpImg, tmpImgp, Gx, Gy, 1); /* fftshift, L970 */
    free(tmpImg);
    return ORC_OK;
}

int orc_fi(const double* in, int ix, int iy, double* out, int ox, int oy) {
    if (!in || !out || ox < ix || oy < iy || ix <= 0 || iy <= 0) return ORC_EARG;
    size_t Gi = (size_t)ix * iy, Go = (size_t)ox * oy;
    int in_wt = ix / 2 + 1, out_wt = ox / 2 + 1, difY = oy - iy;
    cd* a = (cd*)malloc(sizeof(cd) * Gi);
    double* sh = (double*)malloc(sizeof(double) * Gi);
    cd* half = (cd*)calloc((size_t)oy * out_wt, sizeof(cd));
    double* o = (double*)malloc(sizeof(double) * Go);
    if (!a || !sh || !half || !o) { free(a); free(sh); free(half); free(o); return ORC_ENOMEM; }
    shift_d(in, sh, ix, iy, 0); /* ifftshift, L724 */
    for (size_t i = 0; i < Gi; i++) { a[i].re = sh[i]; a[i].im = 0.0; }
    fft2d_cd(iy, ix, a, -1); /* r2c (full), L730 */
    double inv = (double)Gi;
    int pos = iy / 2 + 1; /* L739 */
    for (int y = 0; y < pos; y++)
        for (int x = 0; x < in_wt; x++)
            if (x < out_wt) {
                cd v = a[(size_t)y * ix + x];
                v.re /= inv; v.im /= inv;
                half[(size_t)y * out_wt + x] = v;
            }
    for (int y = pos; y < iy; y++) {
        int oyy = y + difY;
        for (int x = 0; x < in_wt; x++)
            if (x < out_wt && oyy < oy) {
                cd v = a[(size_t)y * ix + x];
                v.re /= inv; v.im /= inv;
                half[(size_t)oyy * out_wt + x] = v;
            }
    }
    int rc = c2r_2d(half, oy, ox, o); /* L762 */
    if (!rc) shift_d(o, out, ox, oy, 0); /* L765 */
    free(a); free(sh); free(half); free(o);
    return rc;
}

int orc_calc_socs(float* image, const double* mskf_re, const double* mskf_im, int fx, int fy,
                  const double* krn_re, const double* krn_im, const float* scales, int nk, int Nx,
                  int Ny, int nthreads) {
    if (!image) return ORC_EARG;
    int Gx = orc_conv_grid(Nx), Gy = orc_conv_grid(Ny);
    if (Gx > fx || Gy > fy) return ORC_EARG;
    double* t = (double*)malloc(sizeof(double) * (size_t)Gx * Gy);
    double* o = (double*)malloc(sizeof(double) * (size_t)fx * fy);
    if (!t || !o) { free(t); free(o); return ORC_ENOMEM; }
    int rc = orc_socs_tmpimg(t, Gx, Gy, mskf_re, mskf_im, fx, fy, krn_re, krn_im, scales, nk, Nx,
                             Ny, nthreads);
    if (!rc) rc = orc_fi(t, Gx, Gy, o, fx, fy);
    if (!rc)
        for (size_t i = 0; i < (size_t)fx * fy; i++) image[i] = (float)o[i];
    free(t); free(o);
    return rc;
}

/* ======================================================================== */
/* Module 5: TCC-direct image                                                */
/* ======================================================================== */

int orc_calc_image_tcc(float* image, const double* mskf_re, const double* mskf_im, int fx, int fy,
                       const double* tcc_re, const double* tcc_im, int tccSize, int Nx, int Ny) {
    int W = 2 * Nx + 1;
    if (!image || !mskf_re || !mskf_im || !tcc_re || !tcc_im || tccSize != W * (2 * Ny + 1) ||
        (fx % 2) || (fy % 2))
        return ORC_EARG;
    size_t F = (size_t)fx * fy;
    int tccSizeh = (tccSize - 1) / 2, fxh = fx / 2, fyh = fy / 2;
    cd* imgf = (cd*)calloc(F, sizeof(cd));
    if (!imgf) return ORC_ENOMEM;
    /* calcImage, L539-583 */
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int ny2 = -2 * Ny; ny2 <= 2 * Ny; ny2++)
        for (int nx2 = 0; nx2 <= 2 * Nx; nx2++) {
            cd val = {0, 0};
            for (int ny1 = -Ny; ny1 <= Ny; ny1++)
                for (int nx1 = -Nx; nx1 <= Nx; nx1++) {
                    if (abs(nx2 + nx1) <= Nx && abs(ny2 + ny1) <= Ny) {
                        int ay = ny2 + ny1 + fyh, ax = nx2 + nx1 + fxh;
                        int by = ny1 + fyh, bx = nx1 + fxh;
                        if (ay >= 0 && ay < fy && ax >= 0 && ax < fx && by >= 0 && by < fy &&
                            bx >= 0 && bx < fx) {
                            cd ma = {mskf_re[(size_t)ay * fx + ax], mskf_im[(size_t)ay * fx + ax]};
                            cd mb = {mskf_re[(size_t)by * fx + bx], -mskf_im[(size_t)by * fx + bx]};
                            size_t ti = (size_t)(ny1 * W + nx1 + tccSizeh) * tccSize +
                                        (size_t)((ny2 + ny1) * W + (nx2 + nx1) + tccSizeh);
                            cd tv = {tcc_re[ti], tcc_im[ti]};
                            val = cadd(val, cmul(cmul(ma, mb), tv));
                        }
                    }
                }
            int oy = ny2 + fyh, ox = nx2 + fxh;
            if (oy >= 0 && oy < fy && ox >= 0 && ox < fx) imgf[(size_t)oy * fx + ox] = val;
        }
    /* FT_c2r_padded, L371-425 */
    int wt = fx / 2 + 1, xh = (fx + 1) / 2, yh = (fy + 1) / 2;
    cd* datac = (cd*)calloc((size_t)fy * wt, sizeof(cd));
    double* o = (double*)malloc(sizeof(double) * F);
    double* o2 = (double*)malloc(sizeof(double) * F);
    if (!datac || !o || !o2) { free(imgf); free(datac); free(o); free(o2); return ORC_ENOMEM; }
    for (int y = 0; y < fy; y++)
        for (int x = xh - 1; x < fx; x++) {
            int sx = (x + xh) % fx, sy = (y + yh) % fy;
            if (sx < wt) datac[(size_t)sy * wt + sx] = imgf[(size_t)y * fx + x];
        }
    datac[wt - 1] = cconj(imgf[0]);
    for (int y = 1; y < fy; y++) datac[(size_t)y * wt + wt - 1] = cconj(imgf[(size_t)(fy - y) * fx]);
    int rc = c2r_2d(datac, fy, fx, o);
    if (!rc) {
        shift_d(o, o2, fx, fy, 0);
        for (size_t i = 0; i < F; i++) image[i] = (float)o2[i];
    }
    free(imgf); free(datac); free(o); free(o2);
    return rc;
}
