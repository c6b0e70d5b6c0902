// C-ABI of the MI355X SOCS engine (declared in include/litho_socs.h).
//
// Host-pointer entry points are synchronous conveniences that stage data
// through a per-thread default context; the device-pointer batch entry point
// is asynchronous on the caller's context stream.  No entry point falls back
// to the CPU: without a usable GPU every compute call returns LITHO_ENODEV.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "dft_mfma.hpp"
#include "eigen_kernels.hpp"
#include "fft_lds.hpp"
#include "socs_kernels.hpp"
#include "tcc_kernels.hpp"
#include "topk_kernels.hpp"

namespace litho {

static thread_local char g_last_err[256] = "";

void set_last_hip_error(hipError_t e, const char* what, const char* file, int line) {
    snprintf(g_last_err, sizeof(g_last_err), "%s: %s (%s:%d)", hipGetErrorString(e), what, file,
             line);
}

}  // namespace litho

using namespace litho;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

enum { NSTAGE = 6 };

struct litho_ctx {
    int device = 0;
    int ncu = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    float2* tw = nullptr;
    DevBuf ws[8];
    hipEvent_t ev[NSTAGE + 1] = {};
    bool timing = true;  // record per-stage events in litho_socs_image_batch
    bool timed = false;
    int nstage = 0;
    const char* stage_names[NSTAGE] = {};
    int path = LITHO_PATH_AUTO;
    // batched form: extra streams and their join events, cached graph
    std::vector<hipStream_t> xs;
    std::vector<hipEvent_t> xe;
    hipEvent_t evStart = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<long> gkey;
    // matrix-core DFT tables, keyed by (kind, N, a, b)
    std::map<std::tuple<int, int, int, int>, DevBuf> tables;
    // band-pipeline plans chosen by litho_autotune, keyed by (N, Nx, nk, precision)
    std::map<std::tuple<int, int, int, int>, std::array<int, 3>> plans;
    // mixed-size batches: four workspaces per bucket, per-bucket end events
    std::vector<DevBuf> bws;
    std::vector<hipEvent_t> bev;
    hipEvent_t bev0 = nullptr;
    int nbev = 0;
};

// One kernel set in HBM: n entries of W x W float2, the weight of each entry,
// and 16-bit copies with one scale per entry for the real and one for the
// imaginary parts.
struct KernelSet {
    float2* d_bank = nullptr;
    float* d_scales = nullptr;
    unsigned* d_bank16 = nullptr;  // bf16 (re | im << 16), scaled per entry
    unsigned* d_bankh = nullptr;   // fp16 (re, im) pairs, scaled per entry
    uint2* d_bank16x2 = nullptr;   // bf16 hi + lo pairs, scaled per entry
    float2* d_kscale = nullptr;
    int n = 0;
};

// The kernels as given (calcSOCS semantics: any mask spectrum), and the set the
// band pipeline runs, whose mask spectra come from real masks.  There the window
// spectrum M is conjugate-symmetric about the window centre, so a kernel with
// K(-f) = e^{ib} conj(K(f)) (real kernels of definite parity are such) gives a
// field IFFT(K M) that is real up to the phase ramp common to every field of
// the clip; two such kernels a, b then share one transform: the entry
// sqrt(la) Ka + i sqrt(lb) Kb, weight 1, has |IFFT|^2 = la |Ea|^2 + lb |Eb|^2.
// Kernels with Nx != Ny are stored zero-padded into the square window of
// half-width max(Nx, Ny) (a kernel is zero outside its window, so every
// product K M and every field is unchanged); Nx and Ny below are that square
// half-width, nx0 / ny0 the kernels' own.  serial is unique per bank for the
// life of the process: cached graphs are keyed by it, never by the address.
struct litho_bank {
    litho_ctx* ctx = nullptr;
    KernelSet full, band;  // band aliases full when no two kernels pair
    int nk = 0, Nx = 0, Ny = 0;
    int nx0 = 0, ny0 = 0;
    int pairs = 0;
    long serial = 0;
};

// context workspace slot of batch workspace i (slots 5 and 6 belong to the
// accelerator-shaped entry)
static int batch_slot(int i) { return i < 5 ? i : 7; }

// Blocks per image of the wave-synced SOCS at the 144 grid (one 130 KB block
// per CU): a launch of at most half as many images as the device has CUs
// splits each image's kernel entries over two blocks (1024^2, K = 24: one
// clip 0.109 -> 0.075 ms, 128 clips 0.313 -> 0.285 ms; 256 clips are 2 %
// faster unsplit)
static int socs_ksplit(const litho_ctx* c, int G, int B) {
    return G == 144 && 2 * B <= c->ncu ? 2 : 1;
}

static int ctx_ws(litho_ctx* c, int slot, size_t bytes, void** out) {
    DevBuf& b = c->ws[slot];
    if (b.bytes < bytes) {
        if (b.p) LITHO_HIP(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        if (hipMalloc(&b.p, bytes) != hipSuccess) {
            b.p = nullptr;
            return LITHO_ENOMEM;
        }
        b.bytes = bytes;
    }
    *out = b.p;
    return LITHO_OK;
}

namespace litho {
hipStream_t ctx_stream(litho_ctx* c) { return c->stream; }
int ctx_device(litho_ctx* c) { return c->device; }
}  // namespace litho

extern "C" {

const char* litho_status_string(int status) {
    switch (status) {
        case LITHO_OK: return "ok";
        case LITHO_EARG: return "invalid argument or unsupported shape";
        case LITHO_ENOMEM: return "allocation failed";
        case LITHO_EHIP: return g_last_err[0] ? g_last_err : "HIP runtime error";
        case LITHO_ENOCONV: return "eigensolver did not converge";
        case LITHO_ENODEV: return "no usable GPU";
        case LITHO_ERANGE: return "pupil support leaves the kernel window";
        case LITHO_ESCRATCH: return "batch exceeds the no-allocation scratch (LITHO_NOALLOC_SCRATCH_BYTES)";
        default: return "unknown status";
    }
}

int litho_version(void) { return 100; }

int litho_ctx_create(litho_ctx** out, int device) {
    if (!out) return LITHO_EARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LITHO_ENODEV;
    if (device < 0 || device >= ndev) return LITHO_EARG;
    std::unique_ptr<litho_ctx> c(new litho_ctx());
    c->device = device;
    LITHO_HIP(hipSetDevice(device));
    LITHO_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
    LITHO_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    std::vector<float2> tw(TW_MAX);
    for (int j = 0; j < TW_MAX; j++) {
        double a = -2.0 * M_PI * (double)j / (double)TW_MAX;
        tw[j] = make_float2((float)cos(a), (float)sin(a));
    }
    LITHO_HIP(hipMalloc(&c->tw, sizeof(float2) * TW_MAX));
    LITHO_HIP(hipMemcpy(c->tw, tw.data(), sizeof(float2) * TW_MAX, hipMemcpyHostToDevice));
    for (int i = 0; i <= NSTAGE; i++) LITHO_HIP(hipEventCreate(&c->ev[i]));
    // the no-allocation entry's scratch on this device; a failed lookup is
    // kept and reported by that entry point alone
    (void)noalloc_init(device);
    *out = c.release();
    return LITHO_OK;
}

int litho_ctx_destroy(litho_ctx* c) {
    if (!c) return LITHO_EARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& b : c->ws)
        if (b.p) (void)hipFree(b.p);
    for (auto& kv : c->tables)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto& b : c->bws)
        if (b.p) (void)hipFree(b.p);
    for (auto e : c->bev) (void)hipEventDestroy(e);
    if (c->bev0) (void)hipEventDestroy(c->bev0);
    if (c->tw) (void)hipFree(c->tw);
    for (int i = 0; i <= NSTAGE; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto e : c->xe) (void)hipEventDestroy(e);
    if (c->evStart) (void)hipEventDestroy(c->evStart);
    for (auto x : c->xs) (void)hipStreamSynchronize(x), (void)hipStreamDestroy(x);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LITHO_OK;
}

int litho_ctx_sync(litho_ctx* c) {
    if (!c) return LITHO_EARG;
    LITHO_HIP(hipStreamSynchronize(c->stream));
    return LITHO_OK;
}

int litho_dev_alloc(litho_ctx* c, void** dptr, size_t bytes) {
    if (!c || !dptr || !bytes) return LITHO_EARG;
    LITHO_HIP(hipSetDevice(c->device));
    if (hipMalloc(dptr, bytes) != hipSuccess) return LITHO_ENOMEM;
    return LITHO_OK;
}

int litho_dev_free(litho_ctx* c, void* dptr) {
    if (!c) return LITHO_EARG;
    LITHO_HIP(hipFree(dptr));
    return LITHO_OK;
}

int litho_memcpy_h2d(litho_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) return LITHO_EARG;
    LITHO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    LITHO_HIP(hipStreamSynchronize(c->stream));
    return LITHO_OK;
}

int litho_memcpy_d2h(litho_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) return LITHO_EARG;
    LITHO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    LITHO_HIP(hipStreamSynchronize(c->stream));
    return LITHO_OK;
}

int litho_memcpy_d2d(litho_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) return LITHO_EARG;
    LITHO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    LITHO_HIP(hipStreamSynchronize(c->stream));
    return LITHO_OK;
}

int litho_memset_dev(litho_ctx* c, void* dst, int value, size_t bytes) {
    if (!c || !dst) return LITHO_EARG;
    LITHO_HIP(hipMemsetAsync(dst, value, bytes, c->stream));
    return LITHO_OK;
}

}  // extern "C"

// per-thread default context for the host-pointer entry points
static int default_ctx(litho_ctx** out) {
    static thread_local litho_ctx* ctx = nullptr;
    if (!ctx) {
        int s = litho_ctx_create(&ctx, 0);
        if (s != LITHO_OK) {
            ctx = nullptr;
            return s;
        }
    }
    *out = ctx;
    return LITHO_OK;
}

// RAII device buffer for the synchronous entry points
struct Dev {
    void* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            p = nullptr;
            return LITHO_ENOMEM;
        }
        return LITHO_OK;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

// ---------------------------------------------------------------------------
// Module 1 helpers and the synthetic mask generator
// ---------------------------------------------------------------------------
namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 0;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    // uniform integer in [0, bound) without modulo bias
    uint32_t bounded(uint32_t bound) {
        uint32_t threshold = (uint32_t)(-bound) % bound;
        for (;;) {
            uint32_t r = next();
            if (r >= threshold) return r % bound;
        }
    }
};

}  // namespace

extern "C" {

static void normalize_source(float* m, int count) {
    float sum = 0.0f;
    for (int i = 0; i < count; ++i) sum += m[i];
    if (sum != 0.0f)
        for (int i = 0; i < count; ++i) m[i] /= sum;
}

int litho_source_annular(float* src, int srcSize, float inner, float outer) {
    if (!src || srcSize <= 0) return LITHO_EARG;
    const int half = (srcSize - 1) / 2;
    const int rin = (int)(inner * (float)half), rout = (int)(outer * (float)half);
    for (int y = 0; y < srcSize; ++y)
        for (int x = 0; x < srcSize; ++x) {
            double d = std::sqrt(std::pow((double)(x - half), 2) + std::pow((double)(y - half), 2));
            src[y * srcSize + x] = (d >= rin && d <= rout) ? 1.0f : 0.0f;
        }
    normalize_source(src, srcSize * srcSize);
    return LITHO_OK;
}

static int circles_source(float* src, int srcSize, const int (*c)[2], int nc, int rad) {
    for (int y = 0; y < srcSize; ++y)
        for (int x = 0; x < srcSize; ++x) {
            bool in = false;
            for (int i = 0; i < nc && !in; ++i) {
                double dx = x - c[i][0], dy = y - c[i][1];
                in = std::sqrt(dx * dx + dy * dy) <= rad;
            }
            src[y * srcSize + x] = in ? 1.0f : 0.0f;
        }
    normalize_source(src, srcSize * srcSize);
    return LITHO_OK;
}

int litho_source_quadrupole(float* src, int srcSize, float radius, float offset) {
    if (!src || srcSize <= 0) return LITHO_EARG;
    const int half = (srcSize - 1) / 2;
    const int off = (int)(offset * (float)half), rad = (int)(radius * (float)half);
    const int c[4][2] = {{half, half + off}, {half, half - off}, {half + off, half}, {half - off, half}};
    return circles_source(src, srcSize, c, 4, rad);
}

int litho_source_dipole(float* src, int srcSize, float radius, float offset, int onXAxis) {
    if (!src || srcSize <= 0) return LITHO_EARG;
    const int half = (srcSize - 1) / 2;
    const int off = (int)(offset * (float)half), rad = (int)(radius * (float)half);
    const int cx = srcSize / 2, cy = srcSize / 2;
    const int c[2][2] = {{cx + (onXAxis ? off : 0), cy + (onXAxis ? 0 : off)},
                         {cx - (onXAxis ? off : 0), cy - (onXAxis ? 0 : off)}};
    return circles_source(src, srcSize, c, 2, rad);
}

int litho_source_point(float* src, int srcSize, float px, float py) {
    if (!src || srcSize <= 0) return LITHO_EARG;
    std::memset(src, 0, sizeof(float) * (size_t)srcSize * srcSize);
    int ix = (int)std::round((double)(px * (float)srcSize) / 2.0 + srcSize / 2.0);
    int iy = (int)std::round((double)(-1.0f * py * (float)srcSize) / 2.0 + srcSize / 2.0);
    if (px >= 0) ix -= 1;
    if (-1.0f * py >= 0) iy -= 1;
    if (ix < 0 || iy < 0 || ix >= srcSize || iy >= srcSize) return LITHO_ERANGE;
    src[iy * srcSize + ix] = 1.0f;
    normalize_source(src, srcSize * srcSize);
    return LITHO_OK;
}

float litho_outer_sigma(const float* src, int srcSize) {
    if (!src || srcSize <= 0) return 0.0f;
    int center = (srcSize - 1) / 2;
    float maxd = 0.0f;
    for (int y = 0; y < srcSize; y++)
        for (int x = 0; x < srcSize; x++)
            if (src[y * srcSize + x] != 0.0f) {
                float d = (float)std::sqrt((double)((y - center) * (y - center) +
                                                    (x - center) * (x - center)));
                if (d > maxd) maxd = d;
            }
    float s = center == 0 ? 0.0f : maxd / (float)center;
    return s > 1.0f ? 1.0f : s;
}

int litho_compute_nx(float NA, int period, float outerSigma, float lambda) {
    volatile float v = NA * (float)period;
    v = v * (1.0f + outerSigma);
    v = v / lambda;
    return (int)std::floor((float)v);
}

int litho_manhattan_mask(float* mask, int N, unsigned long long seed, int wmin, int wmax,
                         float fill) {
    if (!mask || N <= 0 || wmin <= 0 || wmax < wmin || wmax > N || fill < 0.f || fill > 0.95f)
        return LITHO_EARG;
    std::memset(mask, 0, sizeof(float) * (size_t)N * N);
    Pcg32 rng;
    rng.seed(seed, 0x5eedULL);
    const double target = (double)fill * (double)N * (double)N;
    double filled = 0.0;
    long guard = 64L * N * N / ((long)wmin * wmin) + 1000;
    while (filled < target && guard-- > 0) {
        int w = wmin + (int)rng.bounded((uint32_t)(wmax - wmin + 1));
        int h = wmin + (int)rng.bounded((uint32_t)(wmax - wmin + 1));
        int x0 = (int)rng.bounded((uint32_t)(N - w + 1));
        int y0 = (int)rng.bounded((uint32_t)(N - h + 1));
        for (int y = y0; y < y0 + h; y++)
            for (int x = x0; x < x0 + w; x++) {
                float& m = mask[(size_t)y * N + x];
                if (m == 0.0f) {
                    m = 1.0f;
                    filled += 1.0;
                }
            }
    }
    return LITHO_OK;
}

// ---------------------------------------------------------------------------
// Module 3
// ---------------------------------------------------------------------------
int litho_mask_spectrum(const float* mask, int N, float* mskf_r, float* mskf_i) {
    if (!mask || !mskf_r || !mskf_i || !is_pow2(N) || N < 32 || N > 4096) return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    size_t NN = (size_t)N * N;
    Dev dm, dR, dM;
    LITHO_TRY(dm.alloc(NN * sizeof(float)));
    LITHO_TRY(dR.alloc(NN * sizeof(float2)));
    LITHO_TRY(dM.alloc(NN * sizeof(float2)));
    LITHO_HIP(hipMemcpyAsync(dm.p, mask, NN * sizeof(float), hipMemcpyHostToDevice, st));
    LITHO_TRY(launch_rows_dft(N, dm.as<float>(), dR.as<float2>(), N, -N / 2, N, true, c->tw, st));
    LITHO_TRY(launch_mask_cols(N, dR.as<float2>(), dM.as<float2>(), 1, N, -N / 2, N, c->tw, st));
    std::vector<float2> h(NN);
    LITHO_HIP(hipMemcpyAsync(h.data(), dM.p, NN * sizeof(float2), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < NN; i++) {
        mskf_r[i] = h[i].x;
        mskf_i[i] = h[i].y;
    }
    return LITHO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Module 4 / 6 internals
// ---------------------------------------------------------------------------
namespace {

struct LiveSources {
    std::vector<int2> pq;
    std::vector<double> S;
    PupilGeom g;
};

// Live source points in the reference's loop order (calcTCC L458-466) and the
// calcTCC scalars (L443-448).
int collect_sources(const LithoOptics* op, const float* src, int srcSize, LiveSources& L) {
    if (!op || !src || srcSize <= 0 || op->Nx < 0 || op->Ny < 0 || op->NA <= 0.f ||
        op->lambda <= 0.f)
        return LITHO_EARG;
    PupilGeom& g = L.g;
    g.sh = (int)((srcSize - 1) / 2.0);
    const int oSgm = (int)std::ceil((double)((float)g.sh * op->outerSigma));
    volatile float na2l = op->NA * op->NA;
    na2l = na2l / op->lambda;
    g.dz = op->defocus / (float)na2l;
    g.k = (float)(2 * M_PI / (double)op->lambda);
    volatile float lx = (float)op->Lx * op->NA;
    volatile float ly = (float)op->Ly * op->NA;
    g.Lx_norm = (float)lx / op->lambda;
    g.Ly_norm = (float)ly / op->lambda;
    g.NA = op->NA;
    g.Nx = op->Nx;
    g.Ny = op->Ny;
    if (g.sh <= 0) return LITHO_EARG;
    for (int q = -oSgm; q <= oSgm; q++)
        for (int p = -oSgm; p <= oSgm; p++) {
            if (q + g.sh < 0 || q + g.sh >= srcSize || p + g.sh < 0 || p + g.sh >= srcSize)
                continue;
            int id = (q + g.sh) * srcSize + p + g.sh;
            if (src[id] != 0.0f && (p * p + q * q) <= g.sh * g.sh) {
                L.pq.push_back(make_int2(p, q));
                L.S.push_back((double)src[id]);
            }
        }
    return L.pq.empty() ? LITHO_EARG : LITHO_OK;
}

struct Pupil {
    Dev A, S, srcs, err;
    int n = 0, m = 0;
    bool imag = false;
};

int build_pupil(litho_ctx* c, const LithoOptics* op, const float* src, int srcSize, Pupil& P) {
    LiveSources L;
    LITHO_TRY(collect_sources(op, src, srcSize, L));
    hipStream_t st = c->stream;
    P.n = (2 * op->Nx + 1) * (2 * op->Ny + 1);
    P.m = (int)L.pq.size();
    P.imag = op->defocus != 0.0f;
    LITHO_TRY(P.A.alloc((size_t)P.n * P.m * sizeof(double2)));
    LITHO_TRY(P.S.alloc((size_t)P.m * sizeof(double)));
    LITHO_TRY(P.srcs.alloc((size_t)P.m * sizeof(int2)));
    LITHO_TRY(P.err.alloc(sizeof(int)));
    LITHO_HIP(hipMemcpyAsync(P.S.p, L.S.data(), P.m * sizeof(double), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(P.srcs.p, L.pq.data(), P.m * sizeof(int2), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemsetAsync(P.err.p, 0, sizeof(int), st));
    LITHO_TRY(launch_pupil(P.srcs.as<int2>(), P.m, L.g, P.A.as<double2>(), P.err.as<int>(), st));
    int herr = 0;
    LITHO_HIP(hipMemcpyAsync(&herr, P.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    return herr ? LITHO_ERANGE : LITHO_OK;
}

// Jacobi to convergence on the device matrix H (n x n); V receives vectors.
int jacobi_solve(litho_ctx* c, double2* H, double2* V, int n) {
    hipStream_t st = c->stream;
    const int npart = 256;
    Dev part, tol, rots, cnt;
    LITHO_TRY(part.alloc(npart * sizeof(double)));
    LITHO_TRY(tol.alloc(sizeof(double)));
    LITHO_TRY(rots.alloc(((n + 1) / 2 + 1) * sizeof(JacobiRot)));
    LITHO_TRY(cnt.alloc(sizeof(int)));
    LITHO_TRY(jacobi_init(H, V, n, part.as<double>(), npart, tol.as<double>(), st));
    if (n == 1) return LITHO_OK;
    for (int sweep = 0; sweep < 60; sweep++) {
        LITHO_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), st));
        LITHO_TRY(jacobi_sweep(H, V, n, tol.as<double>(), rots.as<JacobiRot>(), cnt.as<int>(), st));
        int h = 0;
        LITHO_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipStreamSynchronize(st));
        if (h == 0) return LITHO_OK;
    }
    return LITHO_ENOCONV;
}

// From descending eigenvalues: how many vectors to canonicalise and the
// clusters among them (same rule as the oracle: consecutive gaps <= rtol*|w0|).
// rtol is 1e-10 after an FP64 contraction; after the split-BF16 contraction
// the FP32 K-slice sums split symmetry-degenerate pairs by ~1e-7 relative, so
// the rule widens to 1e-6 (still far below any non-symmetric gap met here).
void plan_clusters(const std::vector<double>& wd, int nk, int* mtop, int* split,
                   std::vector<int2>& clusters, double rtol = 1e-10) {
    const int m = (int)wd.size();
    const double scale = std::fabs(wd[0]);
    *split = (nk < m && std::fabs(wd[nk - 1] - wd[nk]) <= rtol * scale) ? 1 : 0;
    int mc = nk;
    while (mc < m && std::fabs(wd[mc - 1] - wd[mc]) <= rtol * scale) mc++;
    *mtop = mc;
    int i = 0;
    while (i < mc) {
        int j = i + 1;
        while (j < mc && std::fabs(wd[j - 1] - wd[j]) <= rtol * scale) j++;
        if (j - i > 1 && j - i <= CANON_MAX) clusters.push_back(make_int2(i, j - i));
        i = j;
    }
}

// Shared tail: Vt (mtop x n vector-major, descending) -> canonical, phase, flip.
int finish_kernels(litho_ctx* c, double2* Vt, int n, int mtop, int nk, int Nx, int Ny,
                   const std::vector<int2>& clusters, float* kr, float* ki, double2* kd_dev) {
    hipStream_t st = c->stream;
    Dev dcl, dkr, dki;
    if (!clusters.empty()) {
        LITHO_TRY(dcl.alloc(clusters.size() * sizeof(int2)));
        LITHO_HIP(hipMemcpyAsync(dcl.p, clusters.data(), clusters.size() * sizeof(int2),
                                 hipMemcpyHostToDevice, st));
        LITHO_TRY(launch_canon(Vt, n, Nx, Ny, dcl.as<int2>(), (int)clusters.size(), st));
    }
    float *pr = nullptr, *pi = nullptr;
    if (kr) {
        LITHO_TRY(dkr.alloc((size_t)nk * n * sizeof(float)));
        LITHO_TRY(dki.alloc((size_t)nk * n * sizeof(float)));
        pr = dkr.as<float>();
        pi = dki.as<float>();
    }
    LITHO_TRY(launch_phase_flip(Vt, n, mtop, nk, kd_dev, pr, pi, st));
    if (kr) {
        LITHO_HIP(hipMemcpyAsync(kr, pr, (size_t)nk * n * sizeof(float), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipMemcpyAsync(ki, pi, (size_t)nk * n * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    LITHO_HIP(hipStreamSynchronize(st));
    return LITHO_OK;
}

// Setup timing of the calling thread's last Module 4 / 6 call (host wall
// clock, ms): pupil, contraction, eigensolve; solver (0 Jacobi, 1 subspace);
// subspace iterations.
struct SetupTiming {
    double t[5] = {0, 0, 0, 0, 0};
};
thread_local SetupTiming g_setup;

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Top eigenpairs of device H (n x n, upper triangle valid) by block subspace
// iteration (topk_kernels.hip).  Returns LITHO_OK with the kernels written,
// 1 when the problem does not suit the block (H not PSD to working accuracy,
// a cluster that does not fit, a rank-deficient block): the caller then runs
// the full Jacobi.
constexpr int SUBSPACE_MIN_N = 1024;
// the source-space Gram eigenproblem (n_src x n_src) takes it from this size
constexpr int SUBSPACE_MIN_GRAM = 256;

// finish(X, theta, mtop, split, clusters) receives the converged Ritz
// vectors (vector-major, descending) and values.
template <class F>
int subspace_topk(litho_ctx* c, double2* H, int n, int nk, double rtol, F&& finish) {
    hipStream_t st = c->stream;
    const int b = std::min(n, std::max(64, (2 * nk + 16 + 15) / 16 * 16));
    if (b > 256 || b < nk + 8) return 1;
    const size_t bn = (size_t)b * n, bb = (size_t)b * b;
    const int nsplit = std::max(1, std::min(64, n / 128));
    Dev Q, Y, X, HX, T, V, S, Rw, Cf, work, th, res, stat;
    LITHO_TRY(Q.alloc(bn * sizeof(double2)));
    LITHO_TRY(Y.alloc(bn * sizeof(double2)));
    LITHO_TRY(X.alloc(bn * sizeof(double2)));
    LITHO_TRY(HX.alloc(bn * sizeof(double2)));
    LITHO_TRY(T.alloc(bb * sizeof(double2)));
    LITHO_TRY(V.alloc(bb * sizeof(double2)));
    LITHO_TRY(S.alloc(bb * sizeof(double2)));
    LITHO_TRY(Rw.alloc(bb * sizeof(double2)));
    LITHO_TRY(Cf.alloc(bb * sizeof(double2)));
    LITHO_TRY(work.alloc((size_t)nsplit * bb * sizeof(double2)));
    LITHO_TRY(th.alloc(b * sizeof(double)));
    LITHO_TRY(res.alloc(b * sizeof(double)));
    LITHO_TRY(stat.alloc(sizeof(int)));
    LITHO_TRY(launch_herm_from_upper(H, n, st));
    LITHO_HIP(hipMemsetAsync(stat.p, 0, sizeof(int), st));
    // deterministic start block (64-bit LCG, Knuth's MMIX constants)
    {
        std::vector<double2> h0(bn);
        uint64_t x = 0x9e3779b97f4a7c15ULL;
        auto u = [&]() {
            x = x * 6364136223846793005ULL + 1442695040888963407ULL;
            return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        };
        for (auto& v : h0) {
            const double re = u();
            v = make_double2(re, u());
        }
        LITHO_HIP(hipMemcpyAsync(Q.p, h0.data(), bn * sizeof(double2), hipMemcpyHostToDevice, st));
        LITHO_HIP(hipStreamSynchronize(st));  // h0 is a local
    }
    double2 *q = Q.as<double2>(), *y = Y.as<double2>();
    // CholQR2: z <- z R^-1 twice, through tmp
    auto orth = [&](double2* z, double2* tmp) -> int {
        for (int pass = 0; pass < 2; pass++) {
            double2* src = pass == 0 ? z : tmp;
            double2* dst = pass == 0 ? tmp : z;
            LITHO_TRY(launch_zgemm_nt(src, n, 1, src, n, 0, b, b, n, S.as<double2>(), b, nsplit,
                                      work.as<double2>(), st));
            LITHO_TRY(launch_chol_coeff(S.as<double2>(), b, Rw.as<double2>(), Cf.as<double2>(),
                                        stat.as<int>(), st));
            LITHO_TRY(launch_rotate_rows(Cf.as<double2>(), b, b, src, n, dst, st));
        }
        return LITHO_OK;
    };
    LITHO_TRY(orth(q, y));
    std::vector<double> hT(2 * bb), hV(2 * bb), theta(b), hres(b);
    std::vector<double2> hCf(bb);
    std::vector<int> ord(b);
    const int maxit = 600;
    double best = 1e300;
    int stall = 0, it = 0;
    for (; it < maxit; it++) {
        // y = H q (row i: (H q_i)^T)
        LITHO_TRY(launch_zgemm_nt(q, n, 0, H, n, 0, b, n, n, y, n, 1, nullptr, st));
        const bool rr = it >= 2 && it % 2 == 0;
        if (!rr) {
            LITHO_TRY(orth(y, X.as<double2>()));
            std::swap(q, y);
            continue;
        }
        // Rayleigh-Ritz: T = Q^H H Q
        LITHO_TRY(launch_zgemm_nt(q, n, 1, y, n, 0, b, b, n, T.as<double2>(), b, nsplit,
                                  work.as<double2>(), st));
        LITHO_TRY(launch_jacobi_small(T.as<double2>(), V.as<double2>(), b, stat.as<int>(), st));
        int hs = 0;
        LITHO_HIP(hipMemcpyAsync(hT.data(), T.p, bb * sizeof(double2), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipMemcpyAsync(hV.data(), V.p, bb * sizeof(double2), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipMemcpyAsync(&hs, stat.p, sizeof(int), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipStreamSynchronize(st));
        if (hs) return 1;
        for (int i = 0; i < b; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int d) {
            return hT[2 * ((size_t)a * b + a)] > hT[2 * ((size_t)d * b + d)];
        });
        for (int i = 0; i < b; i++) {
            theta[i] = hT[2 * ((size_t)ord[i] * b + ord[i])];
            if (!std::isfinite(theta[i])) return LITHO_ENOCONV;
            // Ritz vector i = sum_m q_m V[m][ord[i]]
            for (int m = 0; m < b; m++) {
                const size_t vi = (size_t)m * b + ord[i];
                hCf[(size_t)i * b + m] = make_double2(hV[2 * vi], hV[2 * vi + 1]);
            }
        }
        // a negative eigenvalue of weight could hide top ones from the block
        if (theta[b - 1] < -1e-8 * std::fabs(theta[0])) return 1;
        LITHO_HIP(hipMemcpyAsync(Cf.p, hCf.data(), bb * sizeof(double2), hipMemcpyHostToDevice, st));
        LITHO_HIP(hipMemcpyAsync(th.p, theta.data(), b * sizeof(double), hipMemcpyHostToDevice, st));
        LITHO_TRY(launch_rotate_rows(Cf.as<double2>(), b, b, q, n, X.as<double2>(), st));
        LITHO_TRY(launch_rotate_rows(Cf.as<double2>(), b, b, y, n, HX.as<double2>(), st));
        LITHO_TRY(launch_ritz_residual(HX.as<double2>(), X.as<double2>(), th.as<double>(), b, n,
                                       res.as<double>(), st));
        LITHO_HIP(hipMemcpyAsync(hres.data(), res.p, b * sizeof(double), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipMemcpyAsync(&hs, stat.p, sizeof(int), hipMemcpyDeviceToHost, st));
        LITHO_HIP(hipStreamSynchronize(st));  // hCf, theta are read by the copies above
        if (hs) return 1;
        int mtop, split;
        std::vector<int2> clusters;
        plan_clusters(theta, nk, &mtop, &split, clusters, rtol);
        if (mtop > b - 8) return 1;
        double worst = 0.0;
        for (int i = 0; i < mtop; i++) worst = std::max(worst, hres[i]);
        const double scale = std::fabs(theta[0]);
        // converged at 1e-12 of the norm; a residual that no longer falls is
        // accepted at 1e-10 (the FP64 floor of an n-term product)
        if (worst < 0.5 * best) {
            best = worst;
            stall = 0;
        } else {
            stall++;
        }
        const bool done = worst <= 1e-12 * scale || (stall >= 6 && worst <= 1e-10 * scale);
        if (done) {
            g_setup.t[3] = 1;
            g_setup.t[4] = it + 1;
            return finish(X.as<double2>(), theta, mtop, split, clusters);
        }
        // next block: the Ritz vectors' images, ordered by Ritz value
        LITHO_TRY(orth(HX.as<double2>(), y));
        LITHO_HIP(hipMemcpyAsync(q, HX.p, bn * sizeof(double2), hipMemcpyDeviceToDevice, st));
    }
    return LITHO_ENOCONV;
}

// Eigen of device H (n x n, upper triangle valid) -> top kernels.
int kernels_from_device_tcc(litho_ctx* c, double2* H, int n, int Nx, int Ny, int nk,
                            double* scales, float* kr, float* ki, int* flags,
                            double rtol = 1e-10) {
    hipStream_t st = c->stream;
    g_setup.t[3] = 0;
    g_setup.t[4] = 0;
    if (n >= SUBSPACE_MIN_N) {
        // the subspace route leaves H Hermitian from its upper triangle, as
        // the Jacobi route expects
        const int r = subspace_topk(
            c, H, n, nk, rtol,
            [&](double2* X, const std::vector<double>& theta, int mtop, int split,
                const std::vector<int2>& clusters) -> int {
                LITHO_TRY(finish_kernels(c, X, n, mtop, nk, Nx, Ny, clusters, kr, ki, nullptr));
                for (int k = 0; k < nk; k++) scales[k] = theta[k];
                if (flags) *flags = split ? LITHO_FLAG_CLUSTER_SPLIT : 0;
                return LITHO_OK;
            });
        if (r != 1) return r;
        g_setup.t[3] = 0;
        g_setup.t[4] = 0;
    }
    Dev V, w, order, Vt, wd;
    LITHO_TRY(V.alloc((size_t)n * n * sizeof(double2)));
    LITHO_TRY(jacobi_solve(c, H, V.as<double2>(), n));
    LITHO_TRY(w.alloc(n * sizeof(double)));
    LITHO_TRY(order.alloc(n * sizeof(int)));
    LITHO_TRY(launch_diag_rank(H, n, w.as<double>(), order.as<int>(), st));
    std::vector<double> hw(n);
    std::vector<int> ho(n);
    LITHO_HIP(hipMemcpyAsync(hw.data(), w.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipMemcpyAsync(ho.data(), order.p, n * sizeof(int), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    std::vector<double> wdesc(n);
    for (int r = 0; r < n; r++) {
        if (ho[r] < 0 || ho[r] >= n || !std::isfinite(hw[ho[r]])) return LITHO_ENOCONV;
        wdesc[r] = hw[ho[r]];
    }
    int mtop, split;
    std::vector<int2> clusters;
    plan_clusters(wdesc, nk, &mtop, &split, clusters, rtol);
    LITHO_TRY(Vt.alloc((size_t)mtop * n * sizeof(double2)));
    LITHO_TRY(wd.alloc(mtop * sizeof(double)));
    LITHO_TRY(launch_gather_top(V.as<double2>(), w.as<double>(), order.as<int>(), n, mtop,
                                Vt.as<double2>(), wd.as<double>(), st));
    LITHO_TRY(finish_kernels(c, Vt.as<double2>(), n, mtop, nk, Nx, Ny, clusters, kr, ki, nullptr));
    for (int k = 0; k < nk; k++) scales[k] = wdesc[k];
    if (flags) *flags = split ? LITHO_FLAG_CLUSTER_SPLIT : 0;
    return LITHO_OK;
}

// C = (X diag(sx)) Y^H in FP64 MFMA (prec LITHO_TCC_FP64) or split BF16 on the
// matrix cores with FP32 accumulation (LITHO_TCC_SPLIT_BF16).
static int contract_herm(litho_ctx* c, const double2* X, const double* sx, const double2* Y, int n,
                         int m, int imag, int prec, double2* C) {
    hipStream_t st = c->stream;
    if (prec == LITHO_TCC_FP64) {
        const int ns = herm_gemm_slices(n, m, imag);
        Dev part;
        if (ns > 1) LITHO_TRY(part.alloc((size_t)ns * n * n * sizeof(double2)));
        LITHO_TRY(launch_herm_gemm(X, sx, Y, n, m, imag, C, part.as<double2>(), ns, st));
        LITHO_HIP(hipStreamSynchronize(st));  // before the scratch is freed
        return LITHO_OK;
    }
    const int nks = herm_gemm_bf16s_slices(n, m);
    Dev part;
    LITHO_TRY(part.alloc((size_t)nks * n * n * sizeof(float2)));
    LITHO_TRY(launch_herm_gemm_bf16s(X, sx, Y, n, m, imag, part.as<float2>(), nks, C, st));
    LITHO_HIP(hipStreamSynchronize(st));  // before the scratch is freed
    return LITHO_OK;
}

static int tcc_build_impl(const LithoOptics* op, const float* src, int srcSize, int prec,
                          double* tcc_re, double* tcc_im) {
    if (!op || !src || !tcc_re || !tcc_im) return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    g_setup = SetupTiming();
    double t0 = now_ms();
    Pupil P;
    LITHO_TRY(build_pupil(c, op, src, srcSize, P));
    double t1 = now_ms();
    g_setup.t[0] = t1 - t0;
    const size_t nn = (size_t)P.n * P.n;
    Dev C;
    LITHO_TRY(C.alloc(nn * sizeof(double2)));
    t1 = now_ms();
    LITHO_TRY(contract_herm(c, P.A.as<double2>(), P.S.as<double>(), P.A.as<double2>(), P.n, P.m,
                            P.imag, prec, C.as<double2>()));
    LITHO_HIP(hipStreamSynchronize(st));
    g_setup.t[1] = now_ms() - t1;
    std::vector<double2> h(nn);
    LITHO_HIP(hipMemcpyAsync(h.data(), C.p, nn * sizeof(double2), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < nn; i++) {
        tcc_re[i] = h[i].x;
        tcc_im[i] = h[i].y;
    }
    return LITHO_OK;
}

static int kernels_from_source_impl(const LithoOptics* op, const float* src, int srcSize, int nk,
                                    int prec, double* scales, float* krn_r, float* krn_i,
                                    int* flags);

}  // namespace

extern "C" {

int litho_tcc_build(const LithoOptics* op, const float* src, int srcSize, double* tcc_re,
                    double* tcc_im) {
    return tcc_build_impl(op, src, srcSize, LITHO_TCC_FP64, tcc_re, tcc_im);
}

int litho_tcc_build_prec(const LithoOptics* op, const float* src, int srcSize, int tcc_prec,
                         double* tcc_re, double* tcc_im) {
    if (tcc_prec != LITHO_TCC_FP64 && tcc_prec != LITHO_TCC_SPLIT_BF16) return LITHO_EARG;
    return tcc_build_impl(op, src, srcSize, tcc_prec, tcc_re, tcc_im);
}

int litho_kernels_from_source(const LithoOptics* op, const float* src, int srcSize, int nk,
                              double* scales, float* krn_r, float* krn_i, int* flags) {
    return kernels_from_source_impl(op, src, srcSize, nk, LITHO_TCC_FP64, scales, krn_r, krn_i,
                                    flags);
}

int litho_kernels_from_source_prec(const LithoOptics* op, const float* src, int srcSize, int nk,
                                   int tcc_prec, double* scales, float* krn_r, float* krn_i,
                                   int* flags) {
    if (tcc_prec != LITHO_TCC_FP64 && tcc_prec != LITHO_TCC_SPLIT_BF16) return LITHO_EARG;
    return kernels_from_source_impl(op, src, srcSize, nk, tcc_prec, scales, krn_r, krn_i, flags);
}

int litho_kernels_topk(const double* tcc_re, const double* tcc_im, int n, int Nx, int Ny, int nk,
                       double* scales, float* krn_r, float* krn_i, int* flags) {
    if (!tcc_re || !tcc_im || !scales || !krn_r || !krn_i || nk <= 0 || nk > n ||
        n != (2 * Nx + 1) * (2 * Ny + 1))
        return LITHO_EARG;
    const size_t nn = (size_t)n * n;
    std::vector<double2> h(nn);
    for (size_t i = 0; i < nn; i++) {
        // a NaN or infinity would leave the eigensolver's ranking undefined
        if (!std::isfinite(tcc_re[i]) || !std::isfinite(tcc_im[i])) return LITHO_EARG;
        h[i] = make_double2(tcc_re[i], tcc_im[i]);
    }
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    g_setup = SetupTiming();
    Dev H;
    LITHO_TRY(H.alloc(nn * sizeof(double2)));
    LITHO_HIP(hipMemcpyAsync(H.p, h.data(), nn * sizeof(double2), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipStreamSynchronize(st));  // h is released on every return path below
    const double t0 = now_ms();
    const int r =
        kernels_from_device_tcc(c, H.as<double2>(), n, Nx, Ny, nk, scales, krn_r, krn_i, flags);
    g_setup.t[2] = now_ms() - t0;
    return r;
}

int litho_setup_timing(double* out, int n) {
    if (!out || n < 0) return LITHO_EARG;
    for (int i = 0; i < n && i < 5; i++) out[i] = g_setup.t[i];
    return 5;
}

}  // extern "C"

namespace {
static int kernels_from_source_impl(const LithoOptics* op, const float* src, int srcSize, int nk,
                                    int prec, double* scales, float* krn_r, float* krn_i,
                                    int* flags) {
    const double rtol = prec == LITHO_TCC_FP64 ? 1e-10 : 1e-6;
    if (!op || !src || !scales || !krn_r || !krn_i || nk <= 0) return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    g_setup = SetupTiming();
    double t0 = now_ms();
    Pupil P;
    LITHO_TRY(build_pupil(c, op, src, srcSize, P));
    double t1 = now_ms();
    g_setup.t[0] = t1 - t0;
    const int n = P.n, m = P.m;
    if (nk > n || nk > m) return LITHO_EARG;
    if (n <= m) {
        Dev H;
        LITHO_TRY(H.alloc((size_t)n * n * sizeof(double2)));
        t1 = now_ms();
        LITHO_TRY(contract_herm(c, P.A.as<double2>(), P.S.as<double>(), P.A.as<double2>(), n, m,
                                P.imag, prec, H.as<double2>()));
        LITHO_HIP(hipStreamSynchronize(st));
        const double t2 = now_ms();
        g_setup.t[1] = t2 - t1;
        const int r = kernels_from_device_tcc(c, H.as<double2>(), n, op->Nx, op->Ny, nk, scales,
                                              krn_r, krn_i, flags, rtol);
        g_setup.t[2] = now_ms() - t2;
        return r;
    }
    // Gram route: G = B^H B (m x m), B = A diag(sqrt S)
    Dev Bt, G, U, w, order, Vt, wd;
    LITHO_TRY(Bt.alloc((size_t)m * n * sizeof(double2)));
    LITHO_TRY(G.alloc((size_t)m * m * sizeof(double2)));
    LITHO_TRY(U.alloc((size_t)m * m * sizeof(double2)));
    t1 = now_ms();
    LITHO_TRY(launch_conj_transpose_scale(P.A.as<double2>(), P.S.as<double>(), n, m,
                                          Bt.as<double2>(), st));
    LITHO_TRY(contract_herm(c, Bt.as<double2>(), nullptr, Bt.as<double2>(), m, n, P.imag, prec,
                            G.as<double2>()));
    LITHO_HIP(hipStreamSynchronize(st));
    const double t2 = now_ms();
    g_setup.t[1] = t2 - t1;
    // below about nine blocks of 64 vectors the Gram's own Jacobi is the
    // faster of the two (572^2 with K = 64 at 4096^2: 85 against 290 ms)
    if (m >= SUBSPACE_MIN_GRAM && 2 * nk + 16 <= 64) {
        // top eigenvectors u of the Gram by subspace iteration, lifted as
        // v = A diag(sqrt S) u / sqrt(lambda) by one matrix-core product
        const int r = subspace_topk(
            c, G.as<double2>(), m, nk, rtol,
            [&](double2* X, const std::vector<double>& theta, int mtop, int split,
                const std::vector<int2>& clusters) -> int {
                Dev Xs, Vt2, th;
                LITHO_TRY(Xs.alloc((size_t)mtop * m * sizeof(double2)));
                LITHO_TRY(Vt2.alloc((size_t)mtop * n * sizeof(double2)));
                LITHO_TRY(th.alloc(mtop * sizeof(double)));
                LITHO_HIP(hipMemcpyAsync(th.p, theta.data(), mtop * sizeof(double),
                                         hipMemcpyHostToDevice, st));
                LITHO_TRY(launch_lift_scale(X, P.S.as<double>(), th.as<double>(), mtop, m,
                                            Xs.as<double2>(), st));
                LITHO_TRY(launch_zgemm_nt(Xs.as<double2>(), m, 0, P.A.as<double2>(), m, 0, mtop,
                                          n, m, Vt2.as<double2>(), n, 1, nullptr, st));
                LITHO_TRY(finish_kernels(c, Vt2.as<double2>(), n, mtop, nk, op->Nx, op->Ny,
                                         clusters, krn_r, krn_i, nullptr));
                for (int k = 0; k < nk; k++) scales[k] = theta[k];
                if (flags)
                    *flags = (split ? LITHO_FLAG_CLUSTER_SPLIT : 0) | LITHO_FLAG_GRAM_ROUTE;
                return LITHO_OK;
            });
        if (r != 1) {
            g_setup.t[2] = now_ms() - t2;
            return r;
        }
        // the subspace route leaves G as it was (Hermitian from its upper
        // triangle): the Jacobi below starts from it
        g_setup.t[3] = 0;
        g_setup.t[4] = 0;
    }
    LITHO_TRY(jacobi_solve(c, G.as<double2>(), U.as<double2>(), m));
    LITHO_TRY(w.alloc(m * sizeof(double)));
    LITHO_TRY(order.alloc(m * sizeof(int)));
    LITHO_TRY(launch_diag_rank(G.as<double2>(), m, w.as<double>(), order.as<int>(), st));
    std::vector<double> hw(m);
    std::vector<int> ho(m);
    LITHO_HIP(hipMemcpyAsync(hw.data(), w.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipMemcpyAsync(ho.data(), order.p, m * sizeof(int), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    std::vector<double> wdesc(m);
    for (int r = 0; r < m; r++) {
        if (ho[r] < 0 || ho[r] >= m || !std::isfinite(hw[ho[r]])) return LITHO_ENOCONV;
        wdesc[r] = hw[ho[r]];
    }
    int mtop, split;
    std::vector<int2> clusters;
    plan_clusters(wdesc, nk, &mtop, &split, clusters, rtol);
    LITHO_TRY(Vt.alloc((size_t)mtop * n * sizeof(double2)));
    LITHO_TRY(wd.alloc(mtop * sizeof(double)));
    {
        // v = A diag(sqrt S) u / sqrt(lambda) for the top mtop vectors: one
        // FP64 matrix-core product (mtop x m) x (m x n)
        Dev Xs;
        LITHO_TRY(Xs.alloc((size_t)mtop * m * sizeof(double2)));
        LITHO_TRY(launch_lift_gather(P.S.as<double>(), U.as<double2>(), w.as<double>(),
                                     order.as<int>(), m, mtop, Xs.as<double2>(), wd.as<double>(),
                                     st));
        LITHO_TRY(launch_zgemm_nt(Xs.as<double2>(), m, 0, P.A.as<double2>(), m, 0, mtop, n, m,
                                  Vt.as<double2>(), n, 1, nullptr, st));
        LITHO_HIP(hipStreamSynchronize(st));
    }
    LITHO_TRY(finish_kernels(c, Vt.as<double2>(), n, mtop, nk, op->Nx, op->Ny, clusters, krn_r,
                             krn_i, nullptr));
    for (int k = 0; k < nk; k++) scales[k] = wdesc[k];
    if (flags) *flags = (split ? LITHO_FLAG_CLUSTER_SPLIT : 0) | LITHO_FLAG_GRAM_ROUTE;
    g_setup.t[2] = now_ms() - t2;
    return LITHO_OK;
}
}  // namespace

extern "C" {

}  // extern "C"

// Upload one kernel set: float2 entries, weights, and the 16-bit copies (round
// to nearest even) scaled per entry by max |re| and max |im|.
static int upload_set(litho_ctx* c, const std::vector<float2>& h, const std::vector<float>& w,
                      size_t per, KernelSet* s) {
    const int n = (int)w.size();
    s->n = n;
    std::vector<unsigned> h16(h.size());
    std::vector<uint2> h16x2(h.size());
    std::vector<__half2> hh(h.size());
    std::vector<float2> ks(n);
    auto bf16 = [](float f) -> unsigned {
        unsigned u;
        std::memcpy(&u, &f, 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        return u >> 16;
    };
    for (int k = 0; k < n; k++) {
        float mx = 0.f, my = 0.f;
        for (size_t j = 0; j < per; j++) {
            mx = std::max(mx, std::fabs(h[k * per + j].x));
            my = std::max(my, std::fabs(h[k * per + j].y));
        }
        ks[k] = make_float2(mx > 0.f ? mx : 1.0f, my > 0.f ? my : 1.0f);
        auto bf16f = [&](unsigned v) {
            const unsigned u = v << 16;
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        for (size_t j = 0; j < per; j++) {
            const float re = h[k * per + j].x / ks[k].x, im = h[k * per + j].y / ks[k].y;
            h16[k * per + j] = bf16(re) | (bf16(im) << 16);
            const unsigned rlo = bf16(re - bf16f(bf16(re))), ilo = bf16(im - bf16f(bf16(im)));
            h16x2[k * per + j] = make_uint2(h16[k * per + j], rlo | (ilo << 16));
            hh[k * per + j] = __floats2half2_rn(re, im);
        }
    }
    if (hipMalloc(&s->d_bank, h.size() * sizeof(float2)) != hipSuccess ||
        hipMalloc(&s->d_scales, n * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->d_bank16, h16.size() * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&s->d_bankh, hh.size() * sizeof(__half2)) != hipSuccess ||
        hipMalloc(&s->d_bank16x2, h16x2.size() * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&s->d_kscale, n * sizeof(float2)) != hipSuccess)
        return LITHO_ENOMEM;
    hipStream_t st = c->stream;
    LITHO_HIP(hipMemcpyAsync(s->d_bank, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(s->d_scales, w.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(s->d_bank16, h16.data(), h16.size() * sizeof(unsigned),
                             hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(s->d_bankh, hh.data(), hh.size() * sizeof(__half2),
                             hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(s->d_bank16x2, h16x2.data(), h16x2.size() * sizeof(uint2),
                             hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(s->d_kscale, ks.data(), n * sizeof(float2), hipMemcpyHostToDevice, st));
    // the host copies are locals of this call: the copies finish before they go
    LITHO_HIP(hipStreamSynchronize(st));
    return LITHO_OK;
}

// The band pipeline's set (see litho_bank).  Entry j of a W x W kernel and
// entry per - 1 - j are the frequencies f and -f about the window centre.  A
// kernel pairs when its weight is positive and, after the phase a that
// maximises its conjugate-symmetric part, the rest carries at most 1e-7 of
// its energy: with s = sum_f K(f) K(-f), n = sum_f |K(f)|^2, that rest is
// (1 - |s|/n) / 2 and a = -arg(s) / 2.  The kernel is replaced by its
// conjugate-symmetric part (the dropped part changes its intensity by that
// fraction, second order in the residual field).  A real kernel of even
// parity gives a real symmetric part, one of odd parity an imaginary one; a
// kernel pairs with the next unpaired one of its class (more energy in the
// real or in the imaginary part), so each kernel of a pair keeps a component
// of the entry, and with it its own 16-bit scale.  The other kernels keep
// their own entries and weights.  Returns the number of pairs.
static int pack_band_set(const std::vector<float2>& h, const float* scales, int nk, size_t per,
                         std::vector<float2>* hp, std::vector<float>* sp) {
    using cd = std::complex<double>;
    std::vector<std::vector<cd>> sym(nk);
    std::vector<int> cls(nk, -1);  // -1: keeps its own entry
    for (int k = 0; k < nk; k++) {
        if (!(scales[k] > 0.f)) continue;
        const float2* K = h.data() + k * per;
        double n = 0.0;
        cd s = 0.0;
        for (size_t j = 0; j < per; j++) {
            const cd a(K[j].x, K[j].y), m(K[per - 1 - j].x, K[per - 1 - j].y);
            n += std::norm(a);
            s += a * m;
        }
        if (!(n > 0.0) || (1.0 - std::abs(s) / n) / 2.0 > 1e-7) continue;
        const cd ph = std::polar(1.0, -std::arg(s) / 2.0);
        std::vector<cd>& v = sym[k];
        v.resize(per);
        double er = 0.0, ei = 0.0;
        for (size_t j = 0; j < per; j++) {
            const cd a = ph * cd(K[j].x, K[j].y), m = ph * cd(K[per - 1 - j].x, K[per - 1 - j].y);
            v[j] = 0.5 * (a + std::conj(m));
            er += v[j].real() * v[j].real();
            ei += v[j].imag() * v[j].imag();
        }
        cls[k] = ei > er ? 1 : 0;
    }
    hp->clear();
    sp->clear();
    auto keep = [&](int k) {
        hp->insert(hp->end(), h.begin() + k * per, h.begin() + (k + 1) * per);
        sp->push_back(scales[k]);
    };
    std::vector<char> used(nk, 0);
    int pairs = 0;
    for (int k = 0; k < nk; k++) {
        if (used[k]) continue;
        used[k] = 1;
        int q = -1;
        if (cls[k] >= 0)
            for (int r = k + 1; r < nk && q < 0; r++)
                if (!used[r] && cls[r] == cls[k]) q = r;
        if (q < 0) {
            keep(k);
            continue;
        }
        used[q] = 1;
        const double ra = std::sqrt((double)scales[k]), rb = std::sqrt((double)scales[q]);
        for (size_t j = 0; j < per; j++) {
            const cd e = ra * sym[k][j] + cd(0.0, 1.0) * rb * sym[q][j];
            hp->push_back(make_float2((float)e.real(), (float)e.imag()));
        }
        sp->push_back(1.0f);
        pairs++;
    }
    return pairs;
}

extern "C" {

// ---------------------------------------------------------------------------
// Module 7
// ---------------------------------------------------------------------------
int litho_bank_create(litho_ctx* c, const float* krn_r, const float* krn_i, const float* scales,
                      int nk, int Nx, int Ny, litho_bank** out) {
    if (!c || !krn_r || !krn_i || !scales || !out || nk <= 0 || Nx < 0 || Ny < 0 || Nx > 1023 ||
        Ny > 1023)
        return LITHO_EARG;
    *out = nullptr;
    static std::atomic<long> next_serial{1};
    const int m = std::max(Nx, Ny);
    const int W = 2 * m + 1, Wx = 2 * Nx + 1, Wy = 2 * Ny + 1;
    const size_t per = (size_t)W * W;
    std::vector<float2> h((size_t)nk * per, make_float2(0.f, 0.f));
    for (int k = 0; k < nk; k++)
        for (int y = 0; y < Wy; y++)
            for (int x = 0; x < Wx; x++) {
                const size_t i = (size_t)k * Wx * Wy + (size_t)y * Wx + x;
                h[(size_t)k * per + (size_t)(y + m - Ny) * W + (x + m - Nx)] =
                    make_float2(krn_r[i], krn_i[i]);
            }
    std::unique_ptr<litho_bank, int (*)(litho_bank*)> b(new litho_bank(), litho_bank_destroy);
    b->ctx = c;
    b->nk = nk;
    b->Nx = m;
    b->Ny = m;
    b->nx0 = Nx;
    b->ny0 = Ny;
    b->serial = next_serial.fetch_add(1);
    LITHO_HIP(hipSetDevice(c->device));
    LITHO_TRY(upload_set(c, h, std::vector<float>(scales, scales + nk), per, &b->full));
    std::vector<float2> hp;
    std::vector<float> sp;
    b->pairs = pack_band_set(h, scales, nk, per, &hp, &sp);
    if (b->pairs)
        LITHO_TRY(upload_set(c, hp, sp, per, &b->band));
    else
        b->band = b->full;
    LITHO_HIP(hipStreamSynchronize(c->stream));
    *out = b.release();
    return LITHO_OK;
}

static void free_set(KernelSet& s) {
    if (s.d_bank) (void)hipFree(s.d_bank);
    if (s.d_scales) (void)hipFree(s.d_scales);
    if (s.d_bank16) (void)hipFree(s.d_bank16);
    if (s.d_bankh) (void)hipFree(s.d_bankh);
    if (s.d_bank16x2) (void)hipFree(s.d_bank16x2);
    if (s.d_kscale) (void)hipFree(s.d_kscale);
    s = KernelSet();
}

int litho_bank_destroy(litho_bank* b) {
    if (!b) return LITHO_EARG;
    if (b->band.d_bank != b->full.d_bank) free_set(b->band);
    free_set(b->full);
    delete b;
    return LITHO_OK;
}

int litho_bank_info(const litho_bank* b, int* entries, int* pairs) {
    if (!b) return LITHO_EARG;
    if (entries) *entries = b->band.n;
    if (pairs) *pairs = b->pairs;
    return LITHO_OK;
}

int litho_ctx_set_path(litho_ctx* c, int path) {
    if (!c || (path != LITHO_PATH_AUTO && path != LITHO_PATH_POW2)) return LITHO_EARG;
    c->path = path;
    return LITHO_OK;
}

}  // extern "C"

namespace litho {

// Workspaces of litho_socs_image_batch for B clips of N^2 under the context's
// path choice.  Band pipeline: row-DFT window R, mask window M, intensity
// band C, FI column output U.  nextPow2 path: R, M, conv-grid intensity T,
// its row-DFT window P, U, and on grids of 512 points and up the SOCS column
// pass Yt.  Every kernel overwrites its workspace in full before the next one
// reads it, so none has to start at zero.
// Conv grid of the pipeline that is not the band one: the reference's
// nextPow2(4 Nx + 1) (litho_thread.cpp calcSOCS, L833), except 4096-pixel
// clips on the automatic path, which take 576 = 24^2 when the intensity band
// fits (4 Nx + 1 <= 576): the image does not depend on the grid once it holds
// the band (exact band-limited interpolation), and 576 points do 3.4x fewer
// row transforms of about half the length.
static int image_grid(int path, int N, int Nx) {
    const int G = next_pow2(4 * Nx + 1);
    if (path == LITHO_PATH_AUTO && N == 4096 && 4 * Nx + 1 <= 576 && G > 576) return 576;
    return G;
}

int image_batch_sizes(int path, const litho_bank* bank, int B, int N, size_t sz[6], int* band_path,
                      int* grid) {
    if (!bank || B <= 0 || B > 65535) return LITHO_EARG;
    const int Nx = bank->Nx;
    const size_t W = 2 * Nx + 1, band = 2 * Nx;
    if ((!is_pow2(N) && N != 768) || N < 32 || N > 4096 || (int)W > N) return LITHO_EARG;
    const int Gb = band_grid(Nx);
    if (path == LITHO_PATH_AUTO && Gb && Gb <= N && N >= 256 && N <= 1024) {
        *band_path = 1;
        *grid = Gb;
        sz[0] = (size_t)B * (Nx + 1) * N * sizeof(float2);
        sz[1] = (size_t)B * W * W * sizeof(float2);
        sz[2] = 2 * (size_t)B * (2 * band + 1) * (band + 1) * sizeof(float2);  // two planes
        sz[3] = 0;
        sz[4] = (size_t)B * (band + 1) * N * sizeof(float2);
        sz[5] = 0;
        return LITHO_OK;
    }
    const int G = image_grid(path, N, Nx);
    if (!is_pow2(N) || G > N || G > 1024) return LITHO_EARG;
    *band_path = 0;
    *grid = G;
    sz[0] = (size_t)B * N * W * sizeof(float2);
    sz[1] = (size_t)B * W * W * sizeof(float2);
    sz[2] = (size_t)B * G * G * sizeof(float);
    sz[3] = (size_t)B * G * (band + 1) * sizeof(float2);
    sz[4] = (size_t)B * N * (band + 1) * sizeof(float2);
    sz[5] = socs_pre_bytes(G, Nx, bank->band.n, B);  // column pass of the large grids
    return LITHO_OK;
}

// The launches of one batch on the context stream, into the workspaces ws
// (sized by image_batch_sizes); per-stage events when timing.
int image_batch_enqueue(litho_ctx* c, const litho_bank* bank, const float* d_mask, int B, int N,
                        float* d_image, int band_path, int grid, void* const ws[6], bool timing,
                        int bank_fmt) {
    const int Nx = bank->Nx;
    const KernelSet& ks = bank->band;
    const void* kb = bank_fmt == 1   ? (const void*)ks.d_bank16
                     : bank_fmt == 2 ? (const void*)ks.d_bankh
                     : bank_fmt == 3 ? (const void*)ks.d_bank16x2
                                     : (const void*)ks.d_bank;
    const int W = 2 * Nx + 1;
    const int band = 2 * Nx;
    hipStream_t st = c->stream;
    c->timed = false;
    c->nstage = 0;
    if (band_path) {
        // band pipeline: 5 launches
        const int Gb = band_grid(Nx);
        void *R = ws[0], *M = ws[1], *C = ws[2], *U = ws[4];
        static const char* names[] = {"mask_rows", "mask_cols", "socs_band", "fi_cols",
                                      "image_rows"};
        if (timing) LITHO_HIP(hipEventRecord(c->ev[0], st));
        // a real mask: kx = 0..Nx of the row spectra, the rest by symmetry
        LITHO_TRY(launch_rows_dft_t(N, d_mask, R, (long)B * N, 0, Nx + 1, false, false, false, st));
        if (timing) LITHO_HIP(hipEventRecord(c->ev[1], st));
        LITHO_TRY(launch_mask_cols_t(N, R, (float2*)M, B, Nx + 1, -Nx, W, false, false, st, W));
        if (timing) LITHO_HIP(hipEventRecord(c->ev[2], st));
        const int ksplit = socs_ksplit(c, Gb, B);
        const long cplane = (long)B * (2 * band + 1) * (band + 1);
        LITHO_TRY(launch_socs_band(Gb, (float2*)M, kb, ks.d_kscale, bank_fmt, ks.d_scales, ks.n,
                                   Nx, B, (float2*)C, 2, st, ksplit, cplane));
        if (timing) LITHO_HIP(hipEventRecord(c->ev[3], st));
        LITHO_TRY(launch_fi_cols_t(N, (float2*)C, U, B, band, false, false, st, ksplit, cplane));
        if (timing) LITHO_HIP(hipEventRecord(c->ev[4], st));
        LITHO_TRY(launch_image_rows_t(N, U, d_image, (long)B * N, band, false, false, false, st));
        if (timing) LITHO_HIP(hipEventRecord(c->ev[5], st));
        c->nstage = 5;
        for (int i = 0; i < 5; i++) c->stage_names[i] = names[i];
        c->timed = timing;
        return LITHO_OK;
    }
    const int G = grid;
    void *R = ws[0], *M = ws[1], *T = ws[2], *P = ws[3], *U = ws[4];
    static const char* names[] = {"mask_rows", "mask_cols", "socs_accum", "fi_rows", "fi_cols",
                                  "image_rows"};
    // 4096^2 clips: the row-DFT window R and the FI column output U stored
    // transposed per clip, so the column passes read and write whole columns
    // (per 32 clips at K = 64: mask columns 237 -> 68 us, FI columns 256 ->
    // 104 us; the mask rows 410 -> 467 us, the image rows, which load their
    // rows' values of U in tiles of two row pairs, 433 -> 506 us)
    const bool tr = N == 4096 && Nx < 256;
    if (timing) LITHO_HIP(hipEventRecord(c->ev[0], st));
    if (tr)
        LITHO_TRY(launch_rows4096_tr(d_mask, (float2*)R, B, -Nx, W, true, st));
    else
        LITHO_TRY(launch_rows_dft(N, d_mask, (float2*)R, (long)B * N, -Nx, W, true, c->tw, st));
    if (timing) LITHO_HIP(hipEventRecord(c->ev[1], st));
    if (tr)
        LITHO_TRY(launch_mask_cols4096_tr((float2*)R, (float2*)M, B, W, -Nx, W, st));
    else
        LITHO_TRY(launch_mask_cols(N, (float2*)R, (float2*)M, B, W, -Nx, W, c->tw, st));
    if (timing) LITHO_HIP(hipEventRecord(c->ev[2], st));
    // a real mask's window spectrum, conjugate-symmetric to FP32 round-off: the
    // paired set (cross-talk of that order, far below the FP32 image error)
    LITHO_TRY(launch_socs_accum(G, (float2*)M, kb, ks.d_kscale, bank_fmt, ks.d_scales, ks.n, Nx,
                                B, (float*)T, 0, c->tw, st, (float2*)ws[5]));
    if (timing) LITHO_HIP(hipEventRecord(c->ev[3], st));
    LITHO_TRY(launch_rows_dft(G, (float*)T, (float2*)P, (long)B * G, 0, band + 1, false, c->tw, st));
    if (timing) LITHO_HIP(hipEventRecord(c->ev[4], st));
    if (tr)
        LITHO_TRY(launch_fi_cols4096_tr(G, (float2*)P, (float2*)U, B, band, 0, c->tw, st));
    else
        LITHO_TRY(launch_fi_cols(G, N, (float2*)P, (float2*)U, B, band, 0, c->tw, st));
    if (timing) LITHO_HIP(hipEventRecord(c->ev[5], st));
    if (tr)
        LITHO_TRY(launch_image4096_tr((float2*)U, d_image, B, band, st));
    else
        LITHO_TRY(launch_image_rows(N, (float2*)U, d_image, (long)B * N, band, c->tw, st));
    if (timing) LITHO_HIP(hipEventRecord(c->ev[6], st));
    c->nstage = 6;
    for (int i = 0; i < 6; i++) c->stage_names[i] = names[i];
    c->timed = timing;
    return LITHO_OK;
}

int ctx_path(litho_ctx* c) { return c->path; }

}  // namespace litho

extern "C" {

int litho_socs_image_batch(litho_ctx* c, const litho_bank* bank, const float* d_mask, int B,
                           int N, float* d_image) {
    if (!c || !bank || !d_mask || !d_image || B <= 0 || B > 65535) return LITHO_EARG;
    size_t sz[6];
    int band_path = 0, grid = 0;
    LITHO_TRY(image_batch_sizes(c->path, bank, B, N, sz, &band_path, &grid));
    void* ws[6] = {};
    for (int i = 0; i < 6; i++)
        if (sz[i]) LITHO_TRY(ctx_ws(c, batch_slot(i), sz[i], &ws[i]));
    return image_batch_enqueue(c, bank, d_mask, B, N, d_image, band_path, grid, ws, c->timing, 0);
}

// fp16 DFT table of the matrix-core kernels: kind 0 mask rows (N, kmin, count),
// kind 1 image rows (N, nb).  Built and uploaded on first use, outside any
// stream capture.
static int dft_table(litho_ctx* c, int kind, int N, int a, int b, const void** out) {
    const auto key = std::make_tuple(kind, N, a, b);
    auto it = c->tables.find(key);
    if (it == c->tables.end()) {
        const std::vector<unsigned short> t =
            kind == 0 ? mask_rows_table(N, a, b) : image_rows_table(N, a);
        DevBuf d;
        d.bytes = t.size() * sizeof(unsigned short);
        if (hipMalloc(&d.p, d.bytes) != hipSuccess) return LITHO_ENOMEM;
        if (hipMemcpy(d.p, t.data(), d.bytes, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d.p);
            return LITHO_EHIP;
        }
        it = c->tables.emplace(key, d).first;
    }
    *out = it->second.p;
    return LITHO_OK;
}

static int ensure_pipeline(litho_ctx* c, int nstreams) {
    while ((int)c->xs.size() < nstreams - 1) {
        hipStream_t s;
        LITHO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        c->xs.push_back(s);
        hipEvent_t e;
        LITHO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->xe.push_back(e);
    }
    if (!c->evStart) LITHO_HIP(hipEventCreateWithFlags(&c->evStart, hipEventDisableTiming));
    return LITHO_OK;
}

// ---------------------------------------------------------------------------
// Band pipeline of one batch of same-size clips under a plan: the conv grid
// G (36 = 6^2, 81 = 9^2 or 144 = 12^2: the radix of the SOCS FFTs), the SOCS
// block shape (lean: 24 kernel slots per block instead of 48) and the block
// shape of the FFT row / column kernels (lean: 128 threads).
// ---------------------------------------------------------------------------
struct BandPlan {
    bool tuned = false;  // set by litho_autotune or litho_plan_set
    int G = 0;
    int socs_shape = 2;  // 0: 48-slot blocks, 1: 24-slot blocks, 2: wave-synced transforms,
                         // 3: as 2 with each image's entries over two blocks
    bool lean_mem = false;
};

static BandPlan plan_of(litho_ctx* c, int N, int Nx, int nk, int prec) {
    BandPlan p;
    auto it = c->plans.find(std::make_tuple(N, Nx, nk, prec));
    if (it != c->plans.end()) {
        p.tuned = true;
        p.G = it->second[0];
        p.socs_shape = it->second[1];
        p.lean_mem = it->second[2] != 0;
    } else {
        p.G = band_grid(Nx);
    }
    return p;
}

static bool prec_ok(int prec) {
    const int known = LITHO_PREC_INTER_FP16 | LITHO_PREC_BANK_BF16 | LITHO_PREC_BANK_FP16 |
                      LITHO_PREC_MASK_U8 | LITHO_PREC_IMAGE_FP16 | LITHO_PREC_DFT_MFMA |
                      LITHO_PREC_BANK_BF16X2 | LITHO_PREC_MASK_1BIT;
    if (prec & ~known) return false;
    // one input format; the bit-packed masks go through the matrix-core rows only
    if ((prec & LITHO_PREC_MASK_1BIT) &&
        ((prec & LITHO_PREC_MASK_U8) || !(prec & LITHO_PREC_DFT_MFMA)))
        return false;
    const int banks = prec & (LITHO_PREC_BANK_BF16 | LITHO_PREC_BANK_FP16 | LITHO_PREC_BANK_BF16X2);
    if (banks & (banks - 1)) return false;  // at most one bank format
    return true;
}

static bool band_shape_ok(int N, int Nx) {
    const int Gb = band_grid(Nx);
    return (N == 256 || N == 512 || N == 768 || N == 1024) && Gb && Gb <= N;
}

static int grow(DevBuf& b, size_t bytes, void** out) {
    if (b.bytes < bytes) {
        if (b.p) LITHO_HIP(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        if (hipMalloc(&b.p, bytes ? bytes : 1) != hipSuccess) {
            b.p = nullptr;
            return LITHO_ENOMEM;
        }
        b.bytes = bytes;
    }
    *out = b.p;
    return LITHO_OK;
}

struct BandWs {
    void *R = nullptr, *U = nullptr, *M = nullptr, *C = nullptr;
    long cplane = 0;  // float2 between the two planes of C (the entry-split SOCS)
};

// four workspaces (row window R: kx = 0..Nx of each row, mask window M,
// intensity band C, FI column output U) for B clips in the given slots
static int band_ws(DevBuf* slots, int B, int N, int Nx, int prec, BandWs* w) {
    const bool mixed = (prec & LITHO_PREC_INTER_FP16) != 0;
    const bool mfma = (prec & LITHO_PREC_DFT_MFMA) != 0;
    const size_t W = 2 * Nx + 1, band = 2 * Nx;
    const size_t es = mixed ? 4 : sizeof(float2), ues = (mixed && !mfma) ? 4 : sizeof(float2);
    LITHO_TRY(grow(slots[0], (size_t)B * (Nx + 1) * N * es, &w->R));
    LITHO_TRY(grow(slots[1], (size_t)B * W * W * sizeof(float2), &w->M));
    // C in two planes: the SOCS of plan shape 3 splits each image's entries
    // over two blocks, one plane each
    const size_t cper = (size_t)B * (2 * band + 1) * (band + 1);
    LITHO_TRY(grow(slots[2], 2 * cper * sizeof(float2), &w->C));
    w->cplane = (long)cper;
    LITHO_TRY(grow(slots[3], (size_t)B * (band + 1) * N * ues, &w->U));
    return LITHO_OK;
}

// matrix-core tables of one clip shape (uploaded outside any capture)
// the matrix-core tables of a shape: mask rows kx = 0..Nx (the other half
// follows from the real mask), image rows kx = 0..2Nx
static std::tuple<int, int, int, int> mask_table_key(int N, int Nx) { return {0, N, 0, Nx + 1}; }
static std::tuple<int, int, int, int> image_table_key(int N, int Nx) {
    return {1, N, 2 * Nx + 1, 0};
}

static int mfma_tables(litho_ctx* c, int N, int Nx, const void** tm, const void** ti) {
    const auto km = mask_table_key(N, Nx), ki = image_table_key(N, Nx);
    LITHO_TRY(dft_table(c, 0, N, std::get<2>(km), std::get<3>(km), tm));
    return dft_table(c, 1, N, std::get<2>(ki), std::get<3>(ki), ti);
}

// the five launches of one same-size batch on stream st (bit s of stages:
// launch s = mask rows, mask columns, SOCS, FI columns, image rows)
static int enqueue_band(litho_ctx* c, const litho_bank* bank, const void* d_mask, int B, int N,
                        void* d_image, int prec, const BandPlan& pl, const BandWs& w,
                        hipStream_t st, int stages = 31) {
    const bool mixed = (prec & LITHO_PREC_INTER_FP16) != 0;
    const bool mfma = (prec & LITHO_PREC_DFT_MFMA) != 0;
    const bool u8 = (prec & LITHO_PREC_MASK_U8) != 0;
    const int mfmt = (prec & LITHO_PREC_MASK_1BIT) ? 2 : (u8 ? 1 : 0);
    const bool out16 = (prec & LITHO_PREC_IMAGE_FP16) != 0;
    const int Nx = bank->Nx, W = 2 * Nx + 1, band = 2 * Nx;
    const int bfmt = (prec & LITHO_PREC_BANK_BF16)     ? 1
                     : (prec & LITHO_PREC_BANK_FP16)   ? 2
                     : (prec & LITHO_PREC_BANK_BF16X2) ? 3
                                                       : 0;
    const void* kb = bfmt == 1   ? (const void*)bank->band.d_bank16
                     : bfmt == 2 ? (const void*)bank->band.d_bankh
                     : bfmt == 3 ? (const void*)bank->band.d_bank16x2
                                 : (const void*)bank->band.d_bank;
    const float2* ks = bfmt ? bank->band.d_kscale : nullptr;
    const void *tm = nullptr, *ti = nullptr;
    if (mfma) {
        auto a = c->tables.find(mask_table_key(N, Nx));
        auto b = c->tables.find(image_table_key(N, Nx));
        if (a == c->tables.end() || b == c->tables.end()) return LITHO_EARG;
        tm = a->second.p;
        ti = b->second.p;
    }
    if (stages & 1) {
        if (mfma)
            LITHO_TRY(launch_mask_rows_mfma(N, d_mask, mfmt, tm, w.R, B, Nx + 1, mixed, st));
        else
            LITHO_TRY(launch_rows_dft_t(N, d_mask, w.R, (long)B * N, 0, Nx + 1, mixed,
                                        pl.lean_mem, u8, st));
    }
    if (stages & 2)
        LITHO_TRY(launch_mask_cols_t(N, w.R, (float2*)w.M, B, Nx + 1, -Nx, W, mixed, pl.lean_mem,
                                     st, W));
    // plan shape 3: the wave-synced SOCS with each image's entries over two
    // blocks, their band spectra added by the FI column pass
    const int ksplit = pl.socs_shape == 3                 ? 2
                       : pl.socs_shape == 2 && !pl.tuned ? socs_ksplit(c, pl.G, B)
                                                          : 1;
    if (ksplit > 1 && w.cplane <= 0) return LITHO_EARG;
    if (stages & 4)
        LITHO_TRY(launch_socs_band(pl.G, (float2*)w.M, kb, ks, bfmt, bank->band.d_scales,
                                   bank->band.n, Nx, B, (float2*)w.C,
                                   ksplit > 1 ? 2 : pl.socs_shape, st, ksplit, w.cplane));
    if (mfma) {
        // U for the image-side product: at 1024 points the column pass writes
        // it already split into fp16 hi + lo parts (its own form, fewer bytes);
        // elsewhere FP32, split by the product on the fly
        const bool usplit = N == 1024 && 2 * band + 1 <= 160;
        if ((stages & 8) && usplit)
            LITHO_TRY(launch_fi_cols_split(N, (float2*)w.C, w.U, B, band,
                                           image_rows_split_cols(band + 1), st, ksplit,
                                           w.cplane));
        else if (stages & 8)
            LITHO_TRY(launch_fi_cols_t(N, (float2*)w.C, w.U, B, band, false, pl.lean_mem, st,
                                       ksplit, w.cplane));
        if (stages & 16)
            LITHO_TRY(launch_image_rows_mfma(N, w.U, band + 1, ti, d_image, B, out16, usplit, st));
    } else {
        if (stages & 8)
            LITHO_TRY(launch_fi_cols_t(N, (float2*)w.C, w.U, B, band, mixed, pl.lean_mem, st,
                                       ksplit, w.cplane));
        if (stages & 16)
            LITHO_TRY(launch_image_rows_t(N, w.U, d_image, (long)B * N, band, mixed, pl.lean_mem,
                                          out16, st));
    }
    return LITHO_OK;
}

extern "C++" {
// Capture `enqueue` on the context stream once per key and replay it.
template <class F>
static int run_graph(litho_ctx* c, const std::vector<long>& key, F&& enqueue) {
    if (!c->gexec || key != c->gkey) {
        if (c->gexec) LITHO_HIP(hipGraphExecDestroy(c->gexec));
        if (c->graph) LITHO_HIP(hipGraphDestroy(c->graph));
        c->gexec = nullptr;
        c->graph = nullptr;
        c->gkey.clear();
        LITHO_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (st != LITHO_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        LITHO_HIP(e);
        c->graph = g;
        LITHO_HIP(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        c->gkey = key;
    }
    LITHO_HIP(hipGraphLaunch(c->gexec, c->stream));
    return LITHO_OK;
}
}  // extern "C++"

int litho_socs_image_batch_pipelined(litho_ctx* c, const litho_bank* bank, const void* d_mask,
                                     int B, int N, void* d_image, int precision, int chunks,
                                     int use_graph) {
    if (!c || !bank || !d_mask || !d_image || B <= 0 || B > 65535 || chunks <= 0 ||
        chunks > 64 || !prec_ok(precision))
        return LITHO_EARG;
    const int Nx = bank->Nx;
    if (!band_shape_ok(N, Nx)) {
        // clips the band grids do not cover (above 1024^2, or 4Nx+1 > 144):
        // the nextPow2 pipeline in one sub-batch, the kernel bank format the
        // only storage choice
        const int bankflags =
            LITHO_PREC_BANK_BF16 | LITHO_PREC_BANK_FP16 | LITHO_PREC_BANK_BF16X2;
        if (precision & ~bankflags) return LITHO_EARG;
        const int bfmt = (precision & LITHO_PREC_BANK_BF16)   ? 1
                         : (precision & LITHO_PREC_BANK_FP16) ? 2
                         : (precision & LITHO_PREC_BANK_BF16X2) ? 3
                                                                : 0;
        size_t sz[6];
        int band_path = 0, grid = 0;
        LITHO_TRY(image_batch_sizes(c->path, bank, B, N, sz, &band_path, &grid));
        if (band_path) return LITHO_EARG;
        void* ws[6] = {};
        for (int i = 0; i < 6; i++)
            if (sz[i]) LITHO_TRY(ctx_ws(c, batch_slot(i), sz[i], &ws[i]));
        c->nstage = 0;
        c->timed = false;
        auto enqueue = [&]() -> int {
            return image_batch_enqueue(c, bank, (const float*)d_mask, B, N, (float*)d_image,
                                       band_path, grid, ws, false, bfmt);
        };
        if (!use_graph) return enqueue();
        std::vector<long> key = {3, bank->serial, (long)d_mask, (long)d_image, B, N, precision,
                                 grid};
        for (int i = 0; i < 6; i++) key.push_back((long)ws[i]);
        return run_graph(c, key, enqueue);
    }
    const bool mfma = (precision & LITHO_PREC_DFT_MFMA) != 0;
    chunks = std::min(chunks, B);
    const int ns = std::min(chunks, 4);  // GPU_MAX_HW_QUEUES is 4 by default
    const int cb = (B + chunks - 1) / chunks;
    if (mfma) {
        const void *tm, *ti;
        LITHO_TRY(mfma_tables(c, N, Nx, &tm, &ti));
    }
    // sub-batch k uses the R / U workspaces of stream k % ns; M and C hold
    // the whole batch
    BandWs w[4];
    {
        const bool mixed = (precision & LITHO_PREC_INTER_FP16) != 0;
        const size_t W = 2 * Nx + 1, band = 2 * Nx;
        const size_t es = mixed ? 4 : sizeof(float2);
        const size_t ues = (mixed && !mfma) ? 4 : sizeof(float2);
        const size_t rstr = (size_t)cb * W * N * es, ustr = (size_t)cb * (band + 1) * N * ues;
        void *R, *M, *C, *U;
        LITHO_TRY(ctx_ws(c, 0, ns * rstr, &R));
        LITHO_TRY(ctx_ws(c, 1, (size_t)B * W * W * sizeof(float2), &M));
        const size_t cper = (size_t)B * (2 * band + 1) * (band + 1);
        LITHO_TRY(ctx_ws(c, 2, 2 * cper * sizeof(float2), &C));  // two planes (shape 3)
        LITHO_TRY(ctx_ws(c, 4, ns * ustr, &U));
        for (int k = 0; k < ns; k++) {
            w[k].R = (char*)R + k * rstr;
            w[k].U = (char*)U + k * ustr;
            w[k].M = M;
            w[k].C = C;
            w[k].cplane = (long)cper;
        }
    }
    LITHO_TRY(ensure_pipeline(c, ns));
    c->nstage = 0;
    c->timed = false;
    const BandPlan tuned = plan_of(c, N, Nx, bank->nk, precision);
    auto enqueue = [&]() -> int {
        const size_t NN = (size_t)N * N;
        // mask bytes per 8 pixels
        const size_t mb8 = (precision & LITHO_PREC_MASK_1BIT) ? 1
                           : (precision & LITHO_PREC_MASK_U8) ? 8
                                                              : 8 * sizeof(float);
        const size_t ib = (precision & LITHO_PREC_IMAGE_FP16) ? 2 : sizeof(float);
        const size_t W = 2 * Nx + 1, band = 2 * Nx;
        hipStream_t s0 = c->stream;
        LITHO_HIP(hipEventRecord(c->evStart, s0));
        for (int i = 1; i < ns; i++) LITHO_HIP(hipStreamWaitEvent(c->xs[i - 1], c->evStart, 0));
        // stage-major over the sub-batches (launch s of each before launch
        // s + 1 of any): the graph issues its nodes in this order
        for (int sg = 0; sg < 5; sg++)
            for (int k = 0; k < chunks; k++) {
                const int b0 = k * cb;
                const int nb = std::min(cb, B - b0);
                if (nb <= 0) break;
                hipStream_t st = (k % ns) == 0 ? s0 : c->xs[(k % ns) - 1];
                BandWs wk = w[k % ns];
                wk.M = (float2*)w[0].M + (size_t)b0 * W * W;
                wk.C = (float2*)w[0].C + (size_t)b0 * (2 * band + 1) * (band + 1);
                BandPlan pl = tuned;
                if (chunks > 1) pl.lean_mem = true;
                LITHO_TRY(enqueue_band(c, bank, (const char*)d_mask + (size_t)b0 * NN / 8 * mb8, nb,
                                       N, (char*)d_image + (size_t)b0 * NN * ib, precision, pl, wk,
                                       st, 1 << sg));
            }
        for (int i = 1; i < ns; i++) {
            LITHO_HIP(hipEventRecord(c->xe[i - 1], c->xs[i - 1]));
            LITHO_HIP(hipStreamWaitEvent(s0, c->xe[i - 1], 0));
        }
        return LITHO_OK;
    };
    if (!use_graph) return enqueue();
    std::vector<long> key = {1, bank->serial, (long)d_mask, (long)d_image, B, N, precision, chunks,
                             (long)w[0].R, (long)w[0].U, (long)w[0].M, (long)w[0].C, tuned.G,
                             tuned.socs_shape, tuned.lean_mem};
    return run_graph(c, key, enqueue);
}

// ---------------------------------------------------------------------------
// Mixed clip sizes: one bucket per size (its own kernel bank), every bucket's
// five launches on one of up to four streams, the whole set as one graph.
// ---------------------------------------------------------------------------
static int check_buckets(const LithoBucket* bk, int nb, int prec, int* active) {
    if (!bk || nb <= 0 || nb > 64 || !prec_ok(prec)) return LITHO_EARG;
    *active = 0;
    for (int i = 0; i < nb; i++) {
        const LithoBucket& b = bk[i];
        if (b.count < 0 || b.count > 65535) return LITHO_EARG;
        if (b.count == 0) continue;
        if (!b.bank || !b.d_masks || !b.d_images || !band_shape_ok(b.N, b.bank->Nx))
            return LITHO_EARG;
        (*active)++;
    }
    return LITHO_OK;
}

static int prepare_buckets(litho_ctx* c, const LithoBucket* bk, int nb, int prec, int sample,
                           std::vector<BandWs>& w) {
    if ((int)c->bws.size() < 4 * nb) c->bws.resize(4 * nb);
    w.assign(nb, BandWs());
    for (int i = 0; i < nb; i++) {
        const LithoBucket& b = bk[i];
        if (b.count == 0) continue;
        const int B = sample > 0 ? std::min(sample, b.count) : b.count;
        LITHO_TRY(band_ws(&c->bws[4 * i], B, b.N, b.bank->Nx, prec, &w[i]));
        if (prec & LITHO_PREC_DFT_MFMA) {
            const void *tm, *ti;
            LITHO_TRY(mfma_tables(c, b.N, b.bank->Nx, &tm, &ti));
        }
    }
    return LITHO_OK;
}

int litho_socs_image_buckets(litho_ctx* c, const LithoBucket* bk, int nb, int precision,
                             int use_graph) {
    int active = 0;
    if (!c) return LITHO_EARG;
    LITHO_TRY(check_buckets(bk, nb, precision, &active));
    if (!active) return LITHO_OK;
    std::vector<BandWs> w;
    LITHO_TRY(prepare_buckets(c, bk, nb, precision, 0, w));
    const int ns = std::min(active, 4);
    LITHO_TRY(ensure_pipeline(c, ns));
    while ((int)c->bev.size() < nb) {
        hipEvent_t e;
        LITHO_HIP(hipEventCreate(&e));
        c->bev.push_back(e);
    }
    if (!c->bev0) LITHO_HIP(hipEventCreate(&c->bev0));
    c->nstage = 0;
    c->timed = false;
    std::vector<BandPlan> plans(nb);
    for (int i = 0; i < nb; i++)
        if (bk[i].count) plans[i] = plan_of(c, bk[i].N, bk[i].bank->Nx, bk[i].bank->nk, precision);
    // heaviest bucket first (pixels x clips): its chain is the longest, so it
    // starts at once and the lighter chains fill the CUs it leaves idle
    std::vector<int> order;
    for (int i = 0; i < nb; i++)
        if (bk[i].count) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long)bk[a].count * bk[a].N * bk[a].N > (long)bk[b].count * bk[b].N * bk[b].N;
    });
    auto enqueue = [&](bool timing) -> int {
        hipStream_t s0 = c->stream;
        if (timing) LITHO_HIP(hipEventRecord(c->bev0, s0));
        LITHO_HIP(hipEventRecord(c->evStart, s0));
        for (int i = 1; i < ns; i++) LITHO_HIP(hipStreamWaitEvent(c->xs[i - 1], c->evStart, 0));
        // stage-major: launch s of every bucket before launch s + 1 of any.
        // The graph keeps the buckets' chains independent either way, but it
        // issues its nodes in this order, so the chains advance together
        // (0.316 -> 0.282 ms for 512 mixed clips in the compact form).
        auto stream_of = [&](int k) { return (k % ns) == 0 ? s0 : c->xs[(k % ns) - 1]; };
        for (int sg = 0; sg < 5; sg++)
            for (int k = 0; k < (int)order.size(); k++) {
                const int i = order[k];
                LITHO_TRY(enqueue_band(c, bk[i].bank, bk[i].d_masks, bk[i].count, bk[i].N,
                                       bk[i].d_images, precision, plans[i], w[i], stream_of(k),
                                       1 << sg));
                if (timing && sg == 4) LITHO_HIP(hipEventRecord(c->bev[i], stream_of(k)));
            }
        for (int i = 1; i < ns; i++) {
            LITHO_HIP(hipEventRecord(c->xe[i - 1], c->xs[i - 1]));
            LITHO_HIP(hipStreamWaitEvent(s0, c->xe[i - 1], 0));
        }
        return LITHO_OK;
    };
    if (!use_graph) {
        c->nbev = 0;
        LITHO_TRY(enqueue(true));
        c->nbev = nb;
        return LITHO_OK;
    }
    c->nbev = 0;
    std::vector<long> key = {2, precision, nb};
    for (int i = 0; i < nb; i++) {
        const LithoBucket& b = bk[i];
        for (long v : {b.bank->serial, (long)b.N, (long)b.count, (long)b.d_masks, (long)b.d_images,
                       (long)w[i].R, (long)w[i].U, (long)w[i].M, (long)w[i].C, (long)plans[i].G,
                       (long)plans[i].socs_shape, (long)plans[i].lean_mem})
            key.push_back(v);
    }
    return run_graph(c, key, [&]() { return enqueue(false); });
}

int litho_bucket_times(litho_ctx* c, float* ms, int nb) {
    if (!c || !ms || nb <= 0 || c->nbev == 0) return LITHO_EARG;
    const int n = std::min(nb, c->nbev);
    LITHO_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; i++) {
        ms[i] = 0.f;
        if (hipEventElapsedTime(&ms[i], c->bev0, c->bev[i]) != hipSuccess) ms[i] = -1.f;
    }
    return n;
}

int litho_autotune(litho_ctx* c, const LithoBucket* bk, int nb, int precision, int reps) {
    int active = 0;
    if (!c || reps <= 0 || reps > 100) return LITHO_EARG;
    LITHO_TRY(check_buckets(bk, nb, precision, &active));
    const int sample = 32;
    std::vector<BandWs> w;
    LITHO_TRY(prepare_buckets(c, bk, nb, precision, sample, w));
    hipStream_t st = c->stream;
    c->timed = false;
    int tuned = 0;
    for (int i = 0; i < nb; i++) {
        const LithoBucket& b = bk[i];
        if (!b.count) continue;
        const int Nx = b.bank->Nx;
        const auto key = std::make_tuple(b.N, Nx, b.bank->nk, precision);
        if (c->plans.count(key)) continue;
        const int B = std::min(sample, b.count);
        float best = 1e30f;
        std::array<int, 3> pick = {band_grid(Nx), 0, 0};
        for (int G : {36, 81, 144}) {
            if (G < 4 * Nx + 1 || G > b.N) continue;
            // shape 3 (entries over two blocks) is not a candidate: it wins on
            // the 32-clip sample but loses in full mixed-size sets, whose
            // buckets already fill the chip together
            for (int ls = 0; ls < 3; ls++) {
                if (ls == 1 && G != 144) continue;
                for (int lm = 0; lm < 2; lm++) {
                    if (lm && b.N != 1024) continue;
                    BandPlan pl;
                    pl.tuned = true;  // each candidate as given (no automatic split)
                    pl.G = G;
                    pl.socs_shape = ls;
                    pl.lean_mem = lm;
                    int rc = enqueue_band(c, b.bank, b.d_masks, B, b.N, b.d_images, precision, pl,
                                          w[i], st);
                    if (rc == LITHO_EARG) continue;  // shape not instantiated
                    LITHO_TRY(rc);
                    LITHO_HIP(hipEventRecord(c->ev[0], st));
                    for (int r = 0; r < reps; r++)
                        LITHO_TRY(enqueue_band(c, b.bank, b.d_masks, B, b.N, b.d_images,
                                               precision, pl, w[i], st));
                    LITHO_HIP(hipEventRecord(c->ev[1], st));
                    LITHO_HIP(hipEventSynchronize(c->ev[1]));
                    float ms = 0.f;
                    LITHO_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                    if (ms < best) {
                        best = ms;
                        pick = {G, ls, lm};
                    }
                }
            }
        }
        c->plans[key] = pick;
        tuned++;
    }
    return tuned;
}

int litho_plan_set(litho_ctx* c, int N, int Nx, int nk, int precision, int G, int socs_shape,
                   int lean_mem) {
    if (!c || !band_shape_ok(N, Nx) || nk <= 0 || !prec_ok(precision) || G < 4 * Nx + 1 ||
        G > N || (G != 36 && G != 81 && G != 144) || socs_shape < 0 || socs_shape > 3 ||
        (socs_shape == 1 && G != 144) || lean_mem < 0 || lean_mem > 1 || (lean_mem && N != 1024))
        return LITHO_EARG;
    c->plans[std::make_tuple(N, Nx, nk, precision)] = {G, socs_shape, lean_mem};
    return LITHO_OK;
}

int litho_plan_get(litho_ctx* c, int N, int Nx, int nk, int precision, int* G, int* socs_shape,
                   int* lean_mem) {
    if (!c || !G || !socs_shape || !lean_mem) return LITHO_EARG;
    const BandPlan p = plan_of(c, N, Nx, nk, precision);
    *G = p.G;
    *socs_shape = p.socs_shape;
    *lean_mem = p.lean_mem;
    return c->plans.count(std::make_tuple(N, Nx, nk, precision)) ? 1 : 0;
}

int litho_ctx_set_timing(litho_ctx* c, int on) {
    if (!c) return LITHO_EARG;
    c->timing = on != 0;
    return LITHO_OK;
}

int litho_last_timing(litho_ctx* c, float* total_ms, float* stage_ms, int nstage) {
    if (!c || !c->timed) return LITHO_EARG;
    LITHO_HIP(hipEventSynchronize(c->ev[c->nstage]));
    if (total_ms) LITHO_HIP(hipEventElapsedTime(total_ms, c->ev[0], c->ev[c->nstage]));
    for (int i = 0; i < nstage && i < c->nstage; i++)
        LITHO_HIP(hipEventElapsedTime(&stage_ms[i], c->ev[i], c->ev[i + 1]));
    return c->nstage;
}

const char* litho_stage_name(litho_ctx* c, int i) {
    if (!c || i < 0 || i >= c->nstage || !c->stage_names[i]) return "";
    return c->stage_names[i];
}

int litho_socs_image(const float* mask, int N, const float* krn_r, const float* krn_i,
                     const float* scales, int nk, int Nx, int Ny, float* image) {
    if (!mask || !image) return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    litho_bank* bank = nullptr;
    LITHO_TRY(litho_bank_create(c, krn_r, krn_i, scales, nk, Nx, Ny, &bank));
    std::unique_ptr<litho_bank, int (*)(litho_bank*)> guard(bank, litho_bank_destroy);
    const size_t NN = (size_t)N * N;
    Dev dm, di;
    LITHO_TRY(dm.alloc(NN * sizeof(float)));
    LITHO_TRY(di.alloc(NN * sizeof(float)));
    LITHO_HIP(hipMemcpyAsync(dm.p, mask, NN * sizeof(float), hipMemcpyHostToDevice, c->stream));
    LITHO_TRY(litho_socs_image_batch(c, bank, dm.as<float>(), 1, N, di.as<float>()));
    LITHO_HIP(hipMemcpyAsync(image, di.p, NN * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LITHO_HIP(hipStreamSynchronize(c->stream));
    return LITHO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Accelerator-shaped SOCS (the v18 top): tmpImgp, the fftshifted intensity
// on a Gy x Gx conv grid, before FI.
// ---------------------------------------------------------------------------
namespace {

// the v18 top's fixed 128-point grid (v18 .h L34-35, good for n <= 8, its
// MAX_KERNEL_SIZE = 17 at L43), grown to the reference's nextPow2(4n+1)
// (litho_thread.cpp L833-834) where a wider window needs more
int tmpimg_grid(int n) { return std::max(128, next_pow2(4 * n + 1)); }

// square grid the SOCS loop runs on for a window of half-width m
int tmpimg_square(int m, int Gx, int Gy) {
    return std::max(std::max(Gx, Gy), next_pow2(4 * m + 1));
}

bool tmpimg_grid_ok(int nx, int ny, int m, int Gx, int Gy) {
    return is_pow2(Gx) && is_pow2(Gy) && Gx >= 2 * nx + 1 && Gy >= 2 * ny + 1 &&
           tmpimg_square(m, Gx, Gy) <= 1024;
}

// bytes of the unshifted square-grid intensity, needed unless the output grid
// is the square grid on both axes
size_t tmpimg_scratch_bytes(int m, int Gx, int Gy) {
    const int Gs = tmpimg_square(m, Gx, Gy);
    return (Gx == Gs && Gy == Gs) ? 0 : (size_t)Gs * Gs * sizeof(float);
}

// window M (W x W, W = 2m+1) x kernels as given -> tmpImgp (Gy x Gx) in o1
// and o2 (either may be null, they may alias).  The loop runs on the square
// grid Gs >= 4m+1; a coarser output axis samples it (the field of a window
// that fits the coarser grid takes the same values at the common points).
int tmpimg_enqueue(litho_ctx* c, const KernelSet& ks, int m, const float2* dM, int Gx, int Gy,
                   float* T, float* o1, float* o2, hipStream_t st) {
    const int Gs = tmpimg_square(m, Gx, Gy);
    if (o1 == o2) o2 = nullptr;
    if (Gx == Gs && Gy == Gs) {
        float* dst = o1 ? o1 : o2;
        if (!dst) return LITHO_OK;
        LITHO_TRY(launch_socs_accum(Gs, dM, ks.d_bank, nullptr, 0, ks.d_scales, ks.n, m, 1, dst, 1,
                                    c->tw, st));
        if (o1 && o2)
            LITHO_HIP(hipMemcpyAsync(o2, o1, (size_t)Gs * Gs * sizeof(float),
                                     hipMemcpyDeviceToDevice, st));
        return LITHO_OK;
    }
    if (!T) return LITHO_EARG;
    LITHO_TRY(launch_socs_accum(Gs, dM, ks.d_bank, nullptr, 0, ks.d_scales, ks.n, m, 1, T, 0,
                                c->tw, st));
    return launch_grid_sample_shift(T, Gs, Gx, Gy, o1, o2, st);
}

}  // namespace

extern "C" {

int litho_tmpimg_grid(int n) {
    if (n < 0 || n > 255) return LITHO_EARG;
    return tmpimg_grid(n);
}

int litho_socs_tmpimg_on_grid(const float* mskf_r, const float* mskf_i, const float* krn_r,
                              const float* krn_i, const float* scales, float* tmpImg,
                              float* output, int nk, int nx, int ny, int Lx, int Ly, int Gx,
                              int Gy) {
    if (!mskf_r || !mskf_i || !krn_r || !krn_i || !scales || nk <= 0 || nx < 0 || ny < 0 ||
        Lx <= 0 || Ly <= 0)
        return LITHO_EARG;
    const int m = std::max(nx, ny);
    if (!tmpimg_grid_ok(nx, ny, m, Gx, Gy)) return LITHO_EARG;
    const int W = 2 * m + 1;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    // window of the centred spectrum (the embed's bounds check, v18 .cpp
    // L114-116, litho_thread.cpp L909-936); the padded kernels are zero
    // outside their own (2ny+1) x (2nx+1) window
    std::vector<float2> win((size_t)W * W, make_float2(0.f, 0.f));
    for (int y = -m; y <= m; y++)
        for (int x = -m; x <= m; x++) {
            int my = Ly / 2 + y, mx = Lx / 2 + x;
            if (my >= 0 && my < Ly && mx >= 0 && mx < Lx)
                win[(size_t)(y + m) * W + x + m] =
                    make_float2(mskf_r[(size_t)my * Lx + mx], mskf_i[(size_t)my * Lx + mx]);
        }
    litho_bank* bank = nullptr;
    LITHO_TRY(litho_bank_create(c, krn_r, krn_i, scales, nk, nx, ny, &bank));
    std::unique_ptr<litho_bank, int (*)(litho_bank*)> guard(bank, litho_bank_destroy);
    Dev dM, dT, dO;
    LITHO_TRY(dM.alloc(win.size() * sizeof(float2)));
    const size_t tb = tmpimg_scratch_bytes(m, Gx, Gy);
    if (tb) LITHO_TRY(dT.alloc(tb));
    const size_t ob = (size_t)Gx * Gy * sizeof(float);
    LITHO_TRY(dO.alloc(ob));
    LITHO_HIP(hipMemcpyAsync(dM.p, win.data(), win.size() * sizeof(float2), hipMemcpyHostToDevice, st));
    LITHO_TRY(tmpimg_enqueue(c, bank->full, m, dM.as<float2>(), Gx, Gy, dT.as<float>(),
                             dO.as<float>(), nullptr, st));
    std::vector<float> h((size_t)Gx * Gy);
    LITHO_HIP(hipMemcpyAsync(h.data(), dO.p, ob, hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    if (tmpImg) std::memcpy(tmpImg, h.data(), ob);
    if (output && output != tmpImg) std::memcpy(output, h.data(), ob);
    return LITHO_OK;
}

int litho_socs_tmpimg(const float* mskf_r, const float* mskf_i, const float* krn_r,
                      const float* krn_i, const float* scales, float* tmpImg, float* output,
                      int nk, int nx, int ny, int Lx, int Ly) {
    if (nx < 0 || ny < 0 || nx > 255 || ny > 255) return LITHO_EARG;
    return litho_socs_tmpimg_on_grid(mskf_r, mskf_i, krn_r, krn_i, scales, tmpImg, output, nk, nx,
                                     ny, Lx, Ly, tmpimg_grid(nx), tmpimg_grid(ny));
}

int litho_socs_tmpimg_dev(litho_ctx* c, const litho_bank* bank, const float* d_mskf_r,
                          const float* d_mskf_i, int Lx, int Ly, int Gx, int Gy, float* d_tmpImg,
                          float* d_output) {
    if (!c || !bank || !d_mskf_r || !d_mskf_i || Lx <= 0 || Ly <= 0 || (!d_tmpImg && !d_output))
        return LITHO_EARG;
    const int m = bank->Nx;
    if (!tmpimg_grid_ok(bank->nx0, bank->ny0, m, Gx, Gy)) return LITHO_EARG;
    hipStream_t st = c->stream;
    void *M, *T = nullptr;
    LITHO_TRY(ctx_ws(c, 5, (size_t)(2 * m + 1) * (2 * m + 1) * sizeof(float2), &M));
    const size_t tb = tmpimg_scratch_bytes(m, Gx, Gy);
    if (tb) LITHO_TRY(ctx_ws(c, 6, tb, &T));
    LITHO_TRY(launch_spectrum_window(d_mskf_r, d_mskf_i, Lx, Ly, m, (float2*)M, st));
    return tmpimg_enqueue(c, bank->full, m, (const float2*)M, Gx, Gy, (float*)T, d_tmpImg,
                          d_output, st);
}

int litho_image_tcc_rect(const float* mask, int Lx, int Ly, const double* tcc_re,
                         const double* tcc_im, int Nx, int Ny, float* image) {
    if (!mask || !tcc_re || !tcc_im || !image || Nx < 0 || Ny < 0 || !is_pow2(Lx) ||
        !is_pow2(Ly) || Lx < 32 || Ly < 32 || Lx > 4096 || Ly > 4096 || 4 * Nx + 1 > Lx ||
        4 * Ny + 1 > Ly)
        return LITHO_EARG;
    const int W = 2 * Nx + 1, H = 2 * Ny + 1, n = W * H;
    for (size_t i = 0; i < (size_t)n * n; i++)
        if (!std::isfinite(tcc_re[i]) || !std::isfinite(tcc_im[i])) return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    const size_t LL = (size_t)Lx * Ly;
    std::vector<double2> ht((size_t)n * n);
    for (size_t i = 0; i < ht.size(); i++) ht[i] = make_double2(tcc_re[i], tcc_im[i]);
    Dev dm, dR, dM, dT, dC, dV, dI;
    LITHO_TRY(dm.alloc(LL * sizeof(float)));
    LITHO_TRY(dR.alloc((size_t)Ly * W * sizeof(float2)));
    LITHO_TRY(dM.alloc((size_t)n * sizeof(float2)));
    LITHO_TRY(dT.alloc(ht.size() * sizeof(double2)));
    LITHO_TRY(dC.alloc((size_t)W * (4 * Ny + 1) * sizeof(float2)));
    LITHO_TRY(dV.alloc((size_t)Ly * W * sizeof(float2)));
    LITHO_TRY(dI.alloc(LL * sizeof(float)));
    LITHO_HIP(hipMemcpyAsync(dm.p, mask, LL * sizeof(float), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(dT.p, ht.data(), ht.size() * sizeof(double2), hipMemcpyHostToDevice,
                             st));
    // mask window M [H][W] = centred spectrum / (Lx Ly) (FT_r2c_padded): the
    // column pass divides by Ly^2, so the band spectrum (quadratic in M) is
    // rescaled by (Ly / Lx)^2 in the column pass of the image
    LITHO_TRY(launch_rows_dft(Lx, dm.as<float>(), dR.as<float2>(), Ly, -Nx, W, true, c->tw, st));
    LITHO_TRY(launch_mask_cols(Ly, dR.as<float2>(), dM.as<float2>(), 1, W, -Ny, H, c->tw, st));
    LITHO_TRY(launch_tcc_image_spectrum(dM.as<float2>(), dT.as<double2>(), Nx, Ny,
                                        dC.as<float2>(), st));
    const float s = (float)(((double)Ly / Lx) * ((double)Ly / Lx));
    LITHO_TRY(launch_tcc_image_cols(dC.as<float2>(), Nx, Ny, Ly, s, dV.as<float2>(), st));
    LITHO_TRY(launch_image_rows(Lx, dV.as<float2>(), dI.as<float>(), Ly, 2 * Nx, c->tw, st));
    LITHO_HIP(hipMemcpyAsync(image, dI.p, LL * sizeof(float), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    return LITHO_OK;
}

int litho_mask_spectrum_rect(const float* mask, int Lx, int Ly, float* mskf_r, float* mskf_i) {
    if (!mask || !mskf_r || !mskf_i || !is_pow2(Lx) || !is_pow2(Ly) || Lx < 32 || Ly < 32 ||
        Lx > 4096 || Ly > 4096)
        return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    const size_t LL = (size_t)Lx * Ly;
    Dev dm, dR, dM;
    LITHO_TRY(dm.alloc(LL * sizeof(float)));
    LITHO_TRY(dR.alloc(LL * sizeof(float2)));
    LITHO_TRY(dM.alloc(LL * sizeof(float2)));
    LITHO_HIP(hipMemcpyAsync(dm.p, mask, LL * sizeof(float), hipMemcpyHostToDevice, st));
    LITHO_TRY(launch_rows_dft(Lx, dm.as<float>(), dR.as<float2>(), Ly, -Lx / 2, Lx, true, c->tw,
                              st));
    LITHO_TRY(launch_mask_cols(Ly, dR.as<float2>(), dM.as<float2>(), 1, Lx, -Ly / 2, Ly, c->tw,
                               st));
    std::vector<float2> h(LL);
    LITHO_HIP(hipMemcpyAsync(h.data(), dM.p, LL * sizeof(float2), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    // the column pass divides by Ly^2, FT_r2c_padded by Lx Ly (L327-329)
    const float s = (float)Ly / (float)Lx;
    for (size_t i = 0; i < LL; i++) {
        mskf_r[i] = h[i].x * s;
        mskf_i[i] = h[i].y * s;
    }
    return LITHO_OK;
}

int litho_socs_image_rect(const float* mask, int Lx, int Ly, const float* krn_r,
                          const float* krn_i, const float* scales, int nk, int Nx, int Ny,
                          float* image) {
    if (!mask || !krn_r || !krn_i || !scales || !image || nk <= 0 || Nx < 0 || Ny < 0 ||
        !is_pow2(Lx) || !is_pow2(Ly) || Lx < 32 || Ly < 32 || Lx > 4096 || Ly > 4096)
        return LITHO_EARG;
    // the SOCS loop on one square grid G >= 4 max(Nx, Ny) + 1 (the intensity is
    // band-limited to 2 Nx x 2 Ny, so any grid holding the band gives the same
    // image after FI as calcSOCS's nextPow2 grid per axis, L833-834)
    const int m = std::max(Nx, Ny), W = 2 * m + 1, band = 2 * m;
    const int G = std::max(32, next_pow2(4 * m + 1));
    if (G > 1024 || G > Lx || G > Ly) return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    // the window's column pass divides by Ly^2 where FT_r2c_padded divides by
    // Lx Ly: the intensity, quadratic in the window, takes (Ly / Lx)^2 in the
    // weights
    const double q = (double)Ly / Lx;
    std::vector<float> sc(scales, scales + nk);
    for (float& v : sc) v = (float)(v * q * q);
    litho_bank* bank = nullptr;
    LITHO_TRY(litho_bank_create(c, krn_r, krn_i, sc.data(), nk, Nx, Ny, &bank));
    std::unique_ptr<litho_bank, int (*)(litho_bank*)> guard(bank, litho_bank_destroy);
    const size_t LL = (size_t)Lx * Ly;
    Dev dm, dR, dM, dT, dP, dU, dI;
    LITHO_TRY(dm.alloc(LL * sizeof(float)));
    LITHO_TRY(dR.alloc((size_t)Ly * W * sizeof(float2)));
    LITHO_TRY(dM.alloc((size_t)W * W * sizeof(float2)));
    LITHO_TRY(dT.alloc((size_t)G * G * sizeof(float)));
    LITHO_TRY(dP.alloc((size_t)G * (band + 1) * sizeof(float2)));
    LITHO_TRY(dU.alloc((size_t)Ly * (band + 1) * sizeof(float2)));
    LITHO_TRY(dI.alloc(LL * sizeof(float)));
    LITHO_HIP(hipMemcpyAsync(dm.p, mask, LL * sizeof(float), hipMemcpyHostToDevice, st));
    // window (2m+1)^2 of the centred spectrum; the padded kernels are zero
    // outside their own (2Ny+1) x (2Nx+1)
    LITHO_TRY(launch_rows_dft(Lx, dm.as<float>(), dR.as<float2>(), Ly, -m, W, true, c->tw, st));
    LITHO_TRY(launch_mask_cols(Ly, dR.as<float2>(), dM.as<float2>(), 1, W, -m, W, c->tw, st));
    // tmpImgp (centred G x G), then FI to Ly x Lx as litho_fi
    LITHO_TRY(tmpimg_enqueue(c, bank->full, m, dM.as<float2>(), G, G, nullptr, dT.as<float>(),
                             nullptr, st));
    LITHO_TRY(launch_rows_dft(G, dT.as<float>(), dP.as<float2>(), G, 0, band + 1, true, c->tw, st));
    LITHO_TRY(launch_fi_cols(G, Ly, dP.as<float2>(), dU.as<float2>(), 1, band, 1, c->tw, st));
    LITHO_TRY(launch_image_rows(Lx, dU.as<float2>(), dI.as<float>(), Ly, band, c->tw, st));
    LITHO_HIP(hipMemcpyAsync(image, dI.p, LL * sizeof(float), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    return LITHO_OK;
}

int litho_image_tcc(const float* mask, int N, const double* tcc_re, const double* tcc_im, int Nx,
                    int Ny, float* image) {
    if (!mask || !tcc_re || !tcc_im || !image || Nx < 0 || Ny < 0) return LITHO_EARG;
    // the band-grid kernels for the square sizes they cover, the general
    // power-of-two path otherwise
    if (Ny != Nx || !is_pow2(N) || N < 256 || N > 1024 || 4 * Nx + 1 > N)
        return litho_image_tcc_rect(mask, N, N, tcc_re, tcc_im, Nx, Ny, image);
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    const int W = 2 * Nx + 1, n = W * W, band = 2 * Nx;
    const size_t NN = (size_t)N * N;
    std::vector<double2> ht((size_t)n * n);
    for (size_t i = 0; i < ht.size(); i++) ht[i] = make_double2(tcc_re[i], tcc_im[i]);
    Dev dm, dR, dM, dT, dC, dU, dI;
    LITHO_TRY(dm.alloc(NN * sizeof(float)));
    LITHO_TRY(dR.alloc((size_t)W * N * sizeof(float2)));
    LITHO_TRY(dM.alloc((size_t)n * sizeof(float2)));
    LITHO_TRY(dT.alloc(ht.size() * sizeof(double2)));
    LITHO_TRY(dC.alloc((size_t)(2 * band + 1) * W * sizeof(float2)));
    LITHO_TRY(dU.alloc((size_t)W * N * sizeof(float2)));
    LITHO_TRY(dI.alloc(NN * sizeof(float)));
    LITHO_HIP(hipMemcpyAsync(dm.p, mask, NN * sizeof(float), hipMemcpyHostToDevice, st));
    LITHO_HIP(hipMemcpyAsync(dT.p, ht.data(), ht.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    LITHO_TRY(launch_rows_dft_t(N, dm.as<float>(), dR.p, N, 0, Nx + 1, false, false, false, st));
    LITHO_TRY(launch_mask_cols_t(N, dR.p, dM.as<float2>(), 1, Nx + 1, -Nx, W, false, false, st, W));
    LITHO_TRY(launch_tcc_image_spectrum(dM.as<float2>(), dT.as<double2>(), Nx, Ny,
                                        dC.as<float2>(), st));
    LITHO_TRY(launch_fi_cols_t(N, dC.as<float2>(), dU.p, 1, band, false, false, st));
    LITHO_TRY(launch_image_rows_t(N, dU.p, dI.as<float>(), N, band, false, false, false, st));
    LITHO_HIP(hipMemcpyAsync(image, dI.p, NN * sizeof(float), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    return LITHO_OK;
}

int litho_fi(const float* tmpImgp, int G, int N, int band, float* image) {
    if (!tmpImgp || !image || !is_pow2(G) || !is_pow2(N) || G > N || G > 1024 || N > 4096 ||
        band < 0 || 2 * band + 1 > G || G < 32)
        return LITHO_EARG;
    litho_ctx* c;
    LITHO_TRY(default_ctx(&c));
    hipStream_t st = c->stream;
    Dev dT, dP, dU, dI;
    LITHO_TRY(dT.alloc((size_t)G * G * sizeof(float)));
    LITHO_TRY(dP.alloc((size_t)G * (band + 1) * sizeof(float2)));
    LITHO_TRY(dU.alloc((size_t)N * (band + 1) * sizeof(float2)));
    LITHO_TRY(dI.alloc((size_t)N * N * sizeof(float)));
    LITHO_HIP(hipMemcpyAsync(dT.p, tmpImgp, (size_t)G * G * sizeof(float), hipMemcpyHostToDevice, st));
    // input is centred (tmpImgp): its DFT carries (-1)^(kx+ky), undone by the
    // alternating-sign variants of both FI axes
    LITHO_TRY(launch_rows_dft(G, dT.as<float>(), dP.as<float2>(), G, 0, band + 1, true, c->tw, st));
    LITHO_TRY(launch_fi_cols(G, N, dP.as<float2>(), dU.as<float2>(), 1, band, 1, c->tw, st));
    LITHO_TRY(launch_image_rows(N, dU.as<float2>(), dI.as<float>(), N, band, c->tw, st));
    LITHO_HIP(hipMemcpyAsync(image, dI.p, (size_t)N * N * sizeof(float), hipMemcpyDeviceToHost, st));
    LITHO_HIP(hipStreamSynchronize(st));
    return LITHO_OK;
}

}  // extern "C"
