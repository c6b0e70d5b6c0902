// SOCS image pipeline kernels (plain FP32 path) for gfx950.
//
//   rows_dft     : real rows -> a window of their DFT (two rows per complex
//                  FFT).  Used for the mask rows (Module 3, first axis) and
//                  for the rows of the conv-grid intensity (FI R2C, first axis).
//   mask_cols    : mask spectrum along y, keeps the (2Ny+1) window, /N^2 and
//                  the centring signs (FT_r2c_padded, litho_thread.cpp L255-331).
//   socs_accum   : for every kernel k: K_k * M on the window, pruned 2-D
//                  inverse FFT on the G x G conv grid, lambda_k |.|^2
//                  accumulated in registers (calcSOCS L897-961).
//   fi_cols      : FI second axis: G-point forward FFT of each kept column,
//                  band-limited embed into N, N-point inverse FFT (FI L692-772).
//   image_rows   : FI last axis: C2R of two rows per complex FFT, writes the
//                  final image rows once.
//
// Every kernel takes its geometry as arguments and is launched with a grid
// that the host checks against the buffers it was given.
#include <hip/hip_fp16.h>

#include "fft_lds.hpp"
#include "fft_reg.hpp"
#include "socs_kernels.hpp"

#include <mutex>
#include <set>
#include <utility>

namespace litho {

// The dynamic-LDS limit (hipFuncAttributeMaxDynamicSharedMemorySize) belongs
// to a function on one device: raise it to the full 160 KB once per (kernel,
// device), from any host thread.
static int allow_full_lds(const void* kern) {
    static std::mutex mu;
    static std::set<std::pair<const void*, int>> done;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LITHO_EHIP;
    std::lock_guard<std::mutex> lock(mu);
    if (done.count({kern, dev})) return LITHO_OK;
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) !=
        hipSuccess)
        return LITHO_EHIP;
    done.insert({kern, dev});
    return LITHO_OK;
}

// ---------------------------------------------------------------------------
// Storage of the two HBM intermediates (row-DFT window R, FI column output U):
// float2, or half2 in the mixed-precision form (FP32 arithmetic either way).
// Kernel banks: float2, or bf16 (re, im) pairs with one FP32 scale per kernel.
// ---------------------------------------------------------------------------
template <bool H>
struct Store;
template <>
struct Store<false> {
    using T = float2;
    __device__ static __forceinline__ float2 ld(const T& v) { return v; }
    __device__ static __forceinline__ T st(float2 v) { return v; }
};
template <>
struct Store<true> {
    using T = __half2;
    __device__ static __forceinline__ float2 ld(const T& v) { return __half22float2(v); }
    __device__ static __forceinline__ T st(float2 v) { return __floats2half2_rn(v.x, v.y); }
};

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
    return __uint_as_float(((unsigned)h) << 16);
}

// BK = 0: float2; 1: bf16 pairs x kscale[k]; 2: fp16 pairs (half2) x kscale[k]; kscale[k]
// holds one scale for the real and one for the imaginary parts of entry k.
// ld = dec(raw(...)); raw / dec apart let a kernel prefetch the stored word.
template <int BK>
struct Bank;
template <>
struct Bank<0> {
    using R = float2;
    __device__ static __forceinline__ R raw(const void* b, long idx) {
        return ((const float2*)b)[idx];
    }
    __device__ static __forceinline__ float2 dec(R v, float2) { return v; }
    __device__ static __forceinline__ float2 ld(const void* b, const float2*, long idx, int) {
        return raw(b, idx);
    }
};
template <>
struct Bank<1> {
    using R = unsigned;
    __device__ static __forceinline__ R raw(const void* b, long idx) {
        return ((const unsigned*)b)[idx];
    }
    __device__ static __forceinline__ float2 dec(R v, float2 s) {
        return make_float2(bf16_to_f32((unsigned short)(v & 0xffffu)) * s.x,
                           bf16_to_f32((unsigned short)(v >> 16)) * s.y);
    }
    __device__ static __forceinline__ float2 ld(const void* b, const float2* ks, long idx, int k) {
        return dec(raw(b, idx), ks[k]);
    }
};
template <>
struct Bank<2> {
    using R = unsigned;
    __device__ static __forceinline__ R raw(const void* b, long idx) {
        return ((const unsigned*)b)[idx];
    }
    __device__ static __forceinline__ float2 dec(R v, float2 s) {
        const float2 f = __half22float2(__builtin_bit_cast(__half2, v));
        return make_float2(f.x * s.x, f.y * s.y);
    }
    __device__ static __forceinline__ float2 ld(const void* b, const float2* ks, long idx, int k) {
        return dec(raw(b, idx), ks[k]);
    }
};

// BK = 3: bf16 hi + lo pairs per component, (re_hi | im_hi << 16, re_lo | im_lo << 16),
// lo = bf16(x - hi): 16 significant bits, x kscale[k]
template <>
struct Bank<3> {
    using R = uint2;
    __device__ static __forceinline__ R raw(const void* b, long idx) {
        return ((const uint2*)b)[idx];
    }
    __device__ static __forceinline__ float2 dec(R v, float2 s) {
        const float re = bf16_to_f32((unsigned short)(v.x & 0xffffu)) +
                         bf16_to_f32((unsigned short)(v.y & 0xffffu));
        const float im = bf16_to_f32((unsigned short)(v.x >> 16)) +
                         bf16_to_f32((unsigned short)(v.y >> 16));
        return make_float2(re * s.x, im * s.y);
    }
    __device__ static __forceinline__ float2 ld(const void* b, const float2* ks, long idx, int k) {
        return dec(raw(b, idx), ks[k]);
    }
};

// Linear (column, image) index of this block in an order that keeps
// neighbouring columns on one XCD: the dispatcher hands linear block ids to
// the 8 XCDs in turn (block L runs on XCD L mod 8), so XCD x is given the x-th
// contiguous run of the column-major sequence, and blocks that share an L2
// read (or write) the same 128-byte lines of a [rows][columns] array instead
// of each XCD fetching every line.  A bijection of [0, gridDim.x gridDim.y).
__device__ __forceinline__ int xcd_block_order() {
    const int T = gridDim.x * gridDim.y;
    const int L = blockIdx.x + gridDim.x * blockIdx.y;
    const int q = T >> 3, r = T & 7, x = L & 7, i = L >> 3;
    return x * q + (x < r ? x : r) + i;
}

// Column transforms of 4096 points on the same register core, one column per
// block.  mask_cols4096: R [B][4096][cx] -> M [B][cy][cx] as mask_cols_kernel
// (ky = kymin + j, |ky| < 256, times (-1)^ky / N^2).
// Linear block index with each pair of neighbouring blocks on one XCD and
// otherwise the dispatch order: in every run of 16 linear ids, XCD x (ids x
// and x + 8) is given blocks 2x and 2x + 1, which run at about the same time
// and share the 128-byte lines they each read half of.  A partial last run
// keeps the identity.
__device__ __forceinline__ int xcd_pair_order() {
    const int T = gridDim.x * gridDim.y * gridDim.z;
    const int L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int tile = L >> 4;
    if ((tile + 1) * 16 > T) return L;
    const int r = L & 15;
    return (tile << 4) + 2 * (r & 7) + (r >> 3);
}

// ---------------------------------------------------------------------------
// rows_dft: in  [nrows][L] f32 (nrows even), out [nrows][count] float2
// out[r][j] = sgn(k) * X_r[k mod L], k = kmin + j, X_r = DFT_L(row r) forward,
// sgn = (-1)^k when ALT.  FPB complex FFTs (2*FPB rows) per block.
// ---------------------------------------------------------------------------
template <int L, int FPB, bool ALT>
__global__ __launch_bounds__(256) void rows_dft_kernel(const float* __restrict__ in,
                                                       float2* __restrict__ out, int nrows,
                                                       int kmin, int count,
                                                       const float2* __restrict__ tw) {
    __shared__ float2 sa[FPB * L];
    __shared__ float2 sb[FPB * L];
    const int tid = threadIdx.x;
    const long pair0 = (long)blockIdx.x * FPB;  // first row pair of this block
    const long npairs = nrows / 2;
    for (int i = tid; i < FPB * L; i += 256) {
        int f = i / L, x = i - f * L;
        long p = pair0 + f;
        float2 v = make_float2(0.f, 0.f);
        if (p < npairs) {
            const float* r0 = in + (2 * p) * (long)L;
            v = make_float2(r0[x], r0[L + x]);
        }
        sa[i] = v;
    }
    __syncthreads();
    float2* res = fft_lds<L, false>(sa, sb, FPB, tw, tid, 256);
    for (int i = tid; i < FPB * count; i += 256) {
        int f = i / count, j = i - f * count;
        long p = pair0 + f;
        if (p >= npairs) continue;
        int k = kmin + j;
        int kp = ((k % L) + L) % L;
        int km = ((-k % L) + L) % L;
        float2 z1 = res[f * L + kp];
        float2 z2 = cconj(res[f * L + km]);
        // X1 = (z1 + z2)/2, X2 = -i (z1 - z2)/2
        float2 x1 = cscale(cadd(z1, z2), 0.5f);
        float2 d = csub(z1, z2);
        float2 x2 = make_float2(0.5f * d.y, -0.5f * d.x);
        if (ALT && (k & 1)) {
            x1 = cscale(x1, -1.f);
            x2 = cscale(x2, -1.f);
        }
        out[(2 * p) * (long)count + j] = x1;
        out[(2 * p + 1) * (long)count + j] = x2;
    }
}

// ---------------------------------------------------------------------------
// mask_cols: R [B][N][cx] float2 (row DFT window) -> M [B][cy][cx] float2
// M[b][j][c] = (1/N^2) (-1)^ky Y_c[ky mod N], ky = kymin + j, Y_c = DFT_N of
// column c of R[b].  CG columns per block; grid (ceil(cx/CG), B).
// ---------------------------------------------------------------------------
template <int N, int CG>
__global__ __launch_bounds__(256) void mask_cols_kernel(const float2* __restrict__ R,
                                                        float2* __restrict__ M, int cx,
                                                        int kymin, int cy,
                                                        const float2* __restrict__ tw) {
    __shared__ float2 sa[CG * N];
    __shared__ float2 sb[CG * N];
    const int tid = threadIdx.x;
    const int lin = xcd_block_order();
    const int c0 = (lin % gridDim.x) * CG;
    const long b = lin / gridDim.x;
    const float2* Rb = R + b * (long)N * cx;
    for (int i = tid; i < CG * N; i += 256) {
        int y = i / CG, c = i - y * CG;
        int col = c0 + c;
        sa[c * N + y] = col < cx ? Rb[(long)y * cx + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* res = fft_lds<N, false>(sa, sb, CG, tw, tid, 256);
    const float inv = 1.0f / ((float)N * (float)N);
    float2* Mb = M + b * (long)cy * cx;
    for (int i = tid; i < CG * cy; i += 256) {
        int j = i / CG, c = i - j * CG;
        int col = c0 + c;
        if (col >= cx) continue;
        int ky = kymin + j;
        int kp = ((ky % N) + N) % N;
        float s = (ky & 1) ? -inv : inv;
        Mb[(long)j * cx + col] = cscale(res[c * N + kp], s);
    }
}

// ---------------------------------------------------------------------------
// socs_accum: one block = one image x GR consecutive conv-grid rows.
//   M    [B][W][W] float2 (mask spectrum window, W = 2Nx+1, square)
//   bank [nk][W][W] float2 (flipped kernels), scales[nk] f32
//   out  [B][G][G] f32 intensity (fftshifted when SHIFT, i.e. tmpImgp)
// Kernel k: Z = K_k * M placed at natural frequency indices (kx mod G),
// column IFFTs over ky for every window column, keep rows [g0, g0+GR) in LDS,
// row IFFTs over kx, acc += lambda_k |.|^2.  |IFFT|^2 does not depend on where
// the window sits in the grid (a frequency shift is a phase ramp), so this
// equals the reference's centred embed (L847-848, L909-936).
// ---------------------------------------------------------------------------
template <int G, int GR>
__global__ __launch_bounds__(512) void socs_accum_kernel(const float2* __restrict__ M,
                                                         const float2* __restrict__ bank,
                                                         const float* __restrict__ scales,
                                                         int nk, int Nx, float* __restrict__ out,
                                                         int shift,
                                                         const float2* __restrict__ tw) {
    constexpr int NTH = 512;
    constexpr int NF = (2048 / G) < GR ? (2048 / G) : GR;  // FFTs in flight (<= 32 KB LDS)
    constexpr int PER = NF * G / NTH;       // outputs per thread per row batch
    constexpr int NB = GR / NF;             // row batches per kernel
    static_assert(NF >= 1 && PER >= 1 && NB >= 1, "geometry");
    extern __shared__ float2 lds[];
    float2* fa = lds;
    float2* fb = lds + NF * G;
    float2* Y = lds + 2 * NF * G;           // [GR][W]
    const int W = 2 * Nx + 1;
    const int tid = threadIdx.x;
    const int nblk = G / GR;
    const long b = blockIdx.x / nblk;
    const int g0 = (blockIdx.x - (int)b * nblk) * GR;
    const float2* Mb = M + b * (long)W * W;

    float acc[NB * PER];
#pragma unroll
    for (int i = 0; i < NB * PER; i++) acc[i] = 0.f;

    for (int k = 0; k < nk; k++) {
        const float2* Kk = bank + (long)k * W * W;
        const float lam = scales[k];
        // column pass: NF window columns at a time
        for (int c0 = 0; c0 < W; c0 += NF) {
            for (int i = tid; i < NF * G; i += NTH) {
                int c = i / G, g = i - c * G;
                int col = c0 + c;
                int ky = g < G / 2 ? g : g - G;
                float2 v = make_float2(0.f, 0.f);
                if (col < W && ky >= -Nx && ky <= Nx) {
                    int j = (ky + Nx) * W + col;
                    v = cmul(Kk[j], Mb[j]);
                }
                fa[i] = v;
            }
            __syncthreads();
            float2* res = fft_lds<G, true>(fa, fb, NF, tw, tid, NTH);
            for (int i = tid; i < NF * GR; i += NTH) {
                int c = i / GR, r = i - c * GR;
                int col = c0 + c;
                if (col < W) Y[r * W + col] = res[c * G + g0 + r];
            }
            __syncthreads();
        }
        // row pass: NF conv-grid rows at a time, accumulate intensity
#pragma unroll
        for (int rb = 0; rb < NB; rb++) {
            for (int i = tid; i < NF * G; i += NTH) {
                int rr = i / G, h = i - rr * G;
                int kx = h < G / 2 ? h : h - G;
                fa[i] = (kx >= -Nx && kx <= Nx) ? Y[(rb * NF + rr) * W + kx + Nx]
                                                : make_float2(0.f, 0.f);
            }
            __syncthreads();
            float2* res = fft_lds<G, true>(fa, fb, NF, tw, tid, NTH);
#pragma unroll
            for (int q = 0; q < PER; q++) {
                float2 a = res[tid + q * NTH];
                acc[rb * PER + q] += lam * (a.x * a.x + a.y * a.y);
            }
            __syncthreads();
        }
    }
    float* ob = out + b * (long)G * G;
#pragma unroll
    for (int rb = 0; rb < NB; rb++)
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int i = tid + q * NTH;
            int rr = i / G, h = i - rr * G;
            int row = g0 + rb * NF + rr;
            if (shift) ob[(long)((row + G / 2) % G) * G + (h + G / 2) % G] = acc[rb * PER + q];
            else ob[(long)row * G + h] = acc[rb * PER + q];
        }
}


// ---------------------------------------------------------------------------
// LDS written by some lanes of a wavefront and read by others of the same
// wavefront: a wavefront-scope fence orders it (no block barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// socs_accum2: same contract as socs_accum, with the conv-grid FFTs done as
// two-stage register FFTs (G = P*Q, T = min(P,Q) threads per transform, one
// LDS transpose).  The window is placed at indices [0, W) on both axes (a
// frequency shift is a phase ramp along the other axis, invisible in |.|^2),
// so stage-1 inputs with n1 >= NZ = ceil(W/Q) are compile-time zeros.
// LDS: W_G table | SLOTS transpose buffers | Y rows [GR][W].
// ---------------------------------------------------------------------------
// PRE: the column pass was done by socs_cols_kernel into Yt [B][nk][W][G]
// (column-major per entry); the block loads its GR rows of every entry from
// there instead of recomputing every column (grids of 512 points and up,
// where a block holds only GR of the G rows).
// With PRE the block is held to 128 VGPRs, four wavefronts per SIMD (one
// spilled register, outside the entry loop): 3 % faster at the 576 grid.
template <int P, int Q, int GR, int NZ, int NTH, int BK = 0, bool PRE = false>
__global__ __launch_bounds__(NTH, PRE ? 4 : 1) void socs_accum2_kernel(const float2* __restrict__ M,
                                                          const void* __restrict__ bank,
                                                          const float2* __restrict__ kscale,
                                                          const float* __restrict__ scales,
                                                          int nk, int Nx, float* __restrict__ out,
                                                          int shift,
                                                          const float2* __restrict__ tw,
                                                          const float2* __restrict__ Yt) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int G = Geo::G;
    constexpr int T = Geo::T;
    constexpr int SLOTS = NTH / T;
    constexpr int RB = GR / SLOTS;
    constexpr int PER = Geo::PER;
    static_assert(GR % SLOTS == 0 && RB >= 1, "rows per block must fill the slots");
    // a transform's lanes within one wavefront: its transpose needs only
    // wavefront-level syncs (else block barriers, T = 24 at the 576 grid)
    constexpr bool WSYNC = 64 % T == 0;
    extern __shared__ float2 lds[];
    float2* twG = lds;
    float2* xbuf = lds + G;
    float2* Y = xbuf + SLOTS * Geo::XBUF;
    const int W = 2 * Nx + 1;
    const int tid = threadIdx.x;
    const int slot = tid / T, t = tid - slot * T;
    const int nblk = G / GR;
    // the row blocks of an image in turn on one XCD: with the rows read in
    // place (PRE), neighbouring row blocks share the column pass's lines
    const int lin = PRE ? xcd_block_order() : (int)blockIdx.x;
    const long b = lin / nblk;
    const int g0 = (lin - (int)b * nblk) * GR;
    const float2* Mb = M + b * (long)W * W;
    float2* xb = xbuf + slot * Geo::XBUF;
    fill_twiddles<G>(twG, tid, NTH);
    float acc[RB * PER];
#pragma unroll
    for (int i = 0; i < RB * PER; i++) acc[i] = 0.f;
    __syncthreads();

    for (int k = 0; k < nk; k++) {
        const long kbase = (long)k * W * W;
        const float lam = scales[k];
        // PRE: each transform reads its row of entry k straight from the
        // column pass's output (the block's GR rows share each 128-byte line)
        const float2* yk = PRE ? Yt + ((b * nk + k) * (long)W) * G + g0 : nullptr;
        for (int c0 = 0; c0 < (PRE ? 0 : W); c0 += SLOTS) {
            const int col = c0 + slot;
            const bool act = col < W;
            fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
                float2 v = make_float2(0.f, 0.f);
                if (act && n < W) {
                    const int j = n * W + col;
                    v = cmul(Bank<BK>::ld(bank, kscale, kbase + j, k), Mb[j]);
                }
                return v;
            });
            __syncthreads();
            fft2_stage2<Geo, true>(t, xb, [&](int, int g, float2 val) {
                const int r = g - g0;
                if (act && r >= 0 && r < GR) Y[r * W + col] = val;
            });
            __syncthreads();
        }
        // each transform's T lanes lie in one wavefront, as does its transpose
        // buffer xb: with the rows read in place (PRE), wave-level syncs order it
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int r = rb * SLOTS + slot;
            const float2* yr = Y + r * W;
            fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
                if (n >= W) return make_float2(0.f, 0.f);
                return PRE ? yk[(long)n * G + r] : yr[n];
            });
            if (PRE && WSYNC) wave_sync();
            else __syncthreads();
            fft2_stage2<Geo, true>(t, xb, [&](int sidx, int, float2 val) {
                acc[rb * PER + sidx] += lam * (val.x * val.x + val.y * val.y);
            });
            if (PRE && WSYNC) wave_sync();
            else __syncthreads();
        }
    }
    float* ob = out + b * (long)G * G;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int row = g0 + rb * SLOTS + slot;
#pragma unroll
        for (int s = 0; s < Geo::S2; s++)
#pragma unroll
            for (int k2 = 0; k2 < Q; k2++) {
                const int h = t + T * s + P * k2;
                const float v = acc[rb * PER + s * Q + k2];
                if (shift) ob[(long)((row + G / 2) % G) * G + (h + G / 2) % G] = v;
                else ob[(long)row * G + h] = v;
            }
    }
}


// ---------------------------------------------------------------------------
// Register-FFT row kernels for N = P*Q (two real rows per complex transform,
// T = min(P,Q) threads per transform, NTH/T transforms per block).
//
// rows_fwd2: same contract as rows_dft (window [kmin, kmin+count) of the
// forward DFT of each real row, (-1)^k when ALT).
// ---------------------------------------------------------------------------
template <int P, int Q, int NTH, bool ALT, bool OUT_T, bool H16 = false, bool IN8 = false>
__global__ __launch_bounds__(NTH) void rows_fwd2_kernel(const void* __restrict__ in_,
                                                        void* __restrict__ out_, int nrows,
                                                        int kmin, int count,
                                                        const float2* __restrict__ tw) {
    using Geo = Fft2Geom<P, Q>;
    using S = Store<H16>;
    typename S::T* out = (typename S::T*)out_;
    constexpr int L = Geo::G;
    constexpr int T = Geo::T;
    constexpr int FPB = NTH / T;
    __shared__ float2 twG[L];
    __shared__ float2 xbuf[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const long npairs = nrows / 2;
    const long p = (long)blockIdx.x * FPB + f;
    const bool act = p < npairs;
    fill_twiddles<L>(twG, tid, NTH);
    __syncthreads();
    float2* xb = xbuf + f * Geo::XBUF;
    const long roff = (2 * (act ? p : 0)) * (long)L;
    // 8-bit masks: level / 255 (the PNG -> BIN convention), exact in FP32
    const unsigned char* r8 = (const unsigned char*)in_ + roff;
    const float* r0 = (const float*)in_ + roff;
    // an idle transform reads pair 0 (in bounds) and writes nothing
    fft2_stage1<Geo, P, false>(t, xb, twG, [&](int n) {
        // each mask row is read once: streaming (nontemporal) loads
        if (IN8)
            return make_float2((float)__builtin_nontemporal_load(r8 + n) * (1.0f / 255.0f),
                               (float)__builtin_nontemporal_load(r8 + L + n) * (1.0f / 255.0f));
        return make_float2(__builtin_nontemporal_load(r0 + n), __builtin_nontemporal_load(r0 + L + n));
    });
    __syncthreads();
    float2 res[Geo::PER];
    fft2_stage2<Geo, false>(t, xb, [&](int sidx, int, float2 v) { res[sidx] = v; });
    __syncthreads();
    // natural-order spectrum into the (now free) transpose buffer
#pragma unroll
    for (int s = 0; s < Geo::S2; s++)
#pragma unroll
        for (int k2 = 0; k2 < Q; k2++) xb[t + T * s + P * k2] = res[s * Q + k2];
    __syncthreads();
    if (!act) return;
    for (int j = t; j < count; j += T) {
        const int k = kmin + j;
        const int kp = ((k % L) + L) % L;
        const int km = ((-k % L) + L) % L;
        const float2 z1 = xb[kp];
        const float2 z2 = cconj(xb[km]);
        float2 x1 = cscale(cadd(z1, z2), 0.5f);
        const float2 d = csub(z1, z2);
        float2 x2 = make_float2(0.5f * d.y, -0.5f * d.x);
        if (ALT && (k & 1)) {
            x1 = cscale(x1, -1.f);
            x2 = cscale(x2, -1.f);
        }
        if (OUT_T) {
            // [img][j][row in img]: L rows per image (square clips)
            const long r = 2 * p, img = r / L, y = r - img * L;
            typename S::T* o = out + (img * count + j) * (long)L + y;
            o[0] = S::st(x1);
            o[1] = S::st(x2);
        } else {
            out[(2 * p) * (long)count + j] = S::st(x1);
            out[(2 * p + 1) * (long)count + j] = S::st(x2);
        }
    }
}

// Stage transpose of a 32 x 32 two-stage FFT through a float plane buffer xb
// (32 rows of pitch 33), one component at a time: lane t holds the twiddled
// stage-1 outputs v[k1] of n2 = t and receives the stage-2 inputs u[n2] of
// k1 = t.  Half the LDS of a float2 buffer; every thread of the block calls it
// (it holds block barriers).
__device__ __forceinline__ void transpose32_planes(int t, float* __restrict__ xb,
                                                   const float2 (&v)[32], float2 (&u)[32]) {
    constexpr int LDQ = 33;
#pragma unroll
    for (int k1 = 0; k1 < 32; k1++) xb[k1 * LDQ + t] = v[k1].x;
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 32; n2++) u[n2].x = xb[t * LDQ + n2];
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 32; k1++) xb[k1 * LDQ + t] = v[k1].y;
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 32; n2++) u[n2].y = xb[t * LDQ + n2];
}

// rows_band1024: rows_fwd2 for 1024-point rows and the band window kx = 0..count-1
// (count <= 64), transposed output with the centring sign.  Two changes keep
// more blocks on a CU: the stage-1 -> stage-2 transpose goes through LDS one
// component plane (re, then im) at a time, halving the buffer, and the output
// skips the natural-order spectrum in LDS: lane t of a transform holds
// X[t + 32 k2] in u[k2] after stage 2, and X[N - k] comes from lane (32 - t) mod 32
// by ds_bpermute (u[31] for k < 32, u[30] for 32 <= k < 64).
//
// Row pair p of lane t: stage-1 inputs x[32 n1 + t] of both rows as re / im.
template <bool IN8>
__device__ __forceinline__ void rows_band1024_load(const void* __restrict__ in_, long p, int t,
                                                   float2 (&v)[32]) {
    constexpr int L = 1024;
    const long roff = 2 * p * (long)L;
    const unsigned char* r8 = (const unsigned char*)in_ + roff + t;
    const float* r0 = (const float*)in_ + roff + t;
#pragma unroll
    for (int n1 = 0; n1 < 32; n1++) {
        const int n = 32 * n1;
        if (IN8) {
            v[n1] = make_float2((float)__builtin_nontemporal_load(r8 + n) * (1.0f / 255.0f),
                                (float)__builtin_nontemporal_load(r8 + L + n) * (1.0f / 255.0f));
        } else {
            v[n1] = make_float2(__builtin_nontemporal_load(r0 + n),
                                __builtin_nontemporal_load(r0 + L + n));
        }
    }
}

// The transform of one row pair from its loaded inputs v (clobbered) to the
// window columns of R.  Every thread of the block calls it (block barriers);
// act = false computes and writes nothing.
template <bool H16>
__device__ __forceinline__ void rows_band1024_fft(int tid, int t, const float2* __restrict__ twG,
                                                  float* __restrict__ xb, float2 (&v)[32],
                                                  bool act, long p, void* __restrict__ out_,
                                                  int count) {
    constexpr int L = 1024;
    using S = Store<H16>;
    typename S::T* out = (typename S::T*)out_;
    dft_reg<32, false>(v);
#pragma unroll
    for (int k1 = 1; k1 < 32; k1++) v[k1] = cmul_w<false>(v[k1], twG[t * k1]);
    float2 u[32];
    transpose32_planes(t, xb, v, u);
    dft_reg<32, false>(u);  // u[k2] = X[t + 32 k2]
    // X[N - k] from lane (32 - t) mod 32 of this transform (every lane takes part)
    const int lane = tid & 63;
    const int src = ((lane & 32) | ((32 - t) & 31)) << 2;
    const float2 m31 = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[31].x))),
                                   __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[31].y))));
    const float2 m30 = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[30].x))),
                                   __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[30].y))));
    if (!act) return;
    const long r = 2 * p, img = r / L, y = r - img * L;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int k = t + 32 * h;  // window column j = k (kmin = 0)
        if (k >= count) break;
        const float2 z1 = h ? u[1] : u[0];
        const float2 zm = h ? (t == 0 ? m31 : m30) : (t == 0 ? u[0] : m31);  // X[N - k]
        const float2 z2 = cconj(zm);
        float2 x1 = cscale(cadd(z1, z2), 0.5f);
        const float2 d = csub(z1, z2);
        float2 x2 = make_float2(0.5f * d.y, -0.5f * d.x);
        if (k & 1) {
            x1 = cscale(x1, -1.f);
            x2 = cscale(x2, -1.f);
        }
        typename S::T* o = out + (img * count + k) * (long)L + y;
        o[0] = S::st(x1);
        o[1] = S::st(x2);
    }
}

// One row pair per transform, FPB = NTH / 32 transforms per block.  An idle
// transform (past the last row pair) reads pair 0 and writes nothing, so the
// 64 loads are one base plus immediate offsets.
template <bool H16, bool IN8, int NTH = 256>
__global__ __launch_bounds__(NTH) void rows_band1024_kernel(const void* __restrict__ in_,
                                                            void* __restrict__ out_, int nrows,
                                                            int count) {
    using Geo = Fft2Geom<32, 32>;
    constexpr int L = 1024, T = 32, FPB = NTH / T;
    __shared__ float2 twG[L];
    __shared__ float xplane[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const long npairs = nrows / 2;
    const long p = (long)blockIdx.x * FPB + f;
    const bool act = p < npairs;
    // the row loads go out first: the twiddle table is computed while they fly
    float2 v[32];
    rows_band1024_load<IN8>(in_, act ? p : 0, t, v);
    fill_twiddles<L>(twG, tid, NTH);
    __syncthreads();
    rows_band1024_fft<H16>(tid, t, twG, xplane + f * Geo::XBUF, v, act, p, out_, count);
}

// image_rows2: same contract as image_rows.  The Hermitian row spectrum
// (|kx| <= band) is placed at indices kx + band in [0, 2 band] so stage-1
// inputs past ceil((2band+1)/Q) rows are compile-time zeros; the shift is
// undone on the output by e^{-2 pi i band x / N}.
template <int P, int Q, int NZ, int NTH, bool IN_T, bool H16 = false, bool O16 = false>
__global__ __launch_bounds__(NTH) void image_rows2_kernel(const void* __restrict__ U_,
                                                          void* __restrict__ img, int nrows,
                                                          int band,
                                                          const float2* __restrict__ tw) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int L = Geo::G;
    constexpr int T = Geo::T;
    constexpr int FPB = NTH / T;
    __shared__ float2 twG[L];
    __shared__ float2 xbuf[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const int nb = band + 1;
    const long npairs = nrows / 2;
    const long p = (long)blockIdx.x * FPB + f;
    const bool act = p < npairs;
    fill_twiddles<L>(twG, tid, NTH);
    __syncthreads();
    float2* xb = xbuf + f * Geo::XBUF;
    // IN_T: U is [img][kx][row in img] (L rows per image), else [row][kx]
    using S = Store<H16>;
    const typename S::T* U = (const typename S::T*)U_;
    const long r0 = 2 * (act ? p : 0);
    const long img0 = r0 / L, y0 = r0 - img0 * L;
    const typename S::T* u1 = IN_T ? U + img0 * nb * (long)L + y0 : U + r0 * (long)nb;
    const long ust = IN_T ? L : 1;       // stride between kx
    const long u2o = IN_T ? 1 : nb;      // offset of the second row
    fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
        float2 z = make_float2(0.f, 0.f);
        const int kx = n - band;
        if (kx >= -band && kx <= band) {  // idle transforms read row pair 0
            const int a = kx < 0 ? -kx : kx;
            float2 a1 = S::ld(u1[a * ust]), a2 = S::ld(u1[a * ust + u2o]);
            if (a == 0) {
                z = make_float2(a1.x, a2.x);
            } else if (kx > 0) {
                z = make_float2(a1.x - a2.y, a1.y + a2.x);  // u1 + i u2
            } else {
                z = make_float2(a1.x + a2.y, a2.x - a1.y);  // conj(u1) + i conj(u2)
            }
        }
        return z;
    });
    __syncthreads();
    if (!act) return;
    const long o = (2 * p) * (long)L;
    fft2_stage2<Geo, true>(t, xb, [&](int, int x, float2 v) {
        // undo the index shift: multiply by e^{-2 pi i band x / L}
        const float2 w = twG[(int)(((long)band * x) % L)];
        const float2 r = cmul(v, w);
        // the image is written once and not read here: streaming stores
        if (O16) {
            __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, __float2half_rn(r.x)),
                                        (unsigned short*)img + o + x);
            __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, __float2half_rn(r.y)),
                                        (unsigned short*)img + o + L + x);
        } else {
            __builtin_nontemporal_store(r.x, (float*)img + o + x);
            __builtin_nontemporal_store(r.y, (float*)img + o + L + x);
        }
    });
}

// image_band1024: image_rows2 for 1024-point rows of the band path (U
// transposed [img][kx][row], 2 band + 1 <= 160 so five stage-1 inputs per
// lane), with the stage transpose through LDS one component plane at a time
// (42 KB per block: three blocks per CU).
template <bool H16, bool O16, int NTH = 256>
__global__ __launch_bounds__(NTH) void image_band1024_kernel(const void* __restrict__ U_,
                                                             void* __restrict__ img, int nrows,
                                                             int band) {
    using Geo = Fft2Geom<32, 32>;
    using S = Store<H16>;
    constexpr int L = 1024, T = 32, FPB = NTH / T, NZ = 5;
    __shared__ float2 twG[L];
    __shared__ float xplane[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const int nb = band + 1;
    const long npairs = nrows / 2;
    const long p = (long)blockIdx.x * FPB + f;
    const bool act = p < npairs;
    float* xb = xplane + f * Geo::XBUF;
    const typename S::T* U = (const typename S::T*)U_;
    const long r0 = 2 * (act ? p : 0);
    const long img0 = r0 / L, y0 = r0 - img0 * L;
    const typename S::T* u1 = U + img0 * nb * (long)L + y0;
    // stage 1 for n2 = t: band row n = kx + band as the two rows' Hermitian
    // extensions packed as re / im (u1 + i u2 for kx > 0)
    float2 v[32];
#pragma unroll
    for (int n1 = 0; n1 < 32; n1++) {
        float2 z = make_float2(0.f, 0.f);
        const int n = 32 * n1 + t;
        const int kx = n - band;
        if (n1 < NZ && kx >= -band && kx <= band) {  // idle transforms read row pair 0
            const int a = kx < 0 ? -kx : kx;
            const float2 a1 = S::ld(u1[(long)a * L]), a2 = S::ld(u1[(long)a * L + 1]);
            if (a == 0)
                z = make_float2(a1.x, a2.x);
            else if (kx > 0)
                z = make_float2(a1.x - a2.y, a1.y + a2.x);
            else
                z = make_float2(a1.x + a2.y, a2.x - a1.y);
        }
        v[n1] = z;
    }
    // the loads above go out before the twiddle table is computed
    fill_twiddles<L>(twG, tid, NTH);
    __syncthreads();
    dft_reg<32, true>(v);
#pragma unroll
    for (int k1 = 1; k1 < 32; k1++) {
        float2 w = twG[t * k1];
        w.y = -w.y;
        v[k1] = cmul_w<false>(v[k1], w);
    }
    float2 u[32];
    transpose32_planes(t, xb, v, u);
    if (!act) return;
    dft_reg<32, true>(u);  // u[k2] = pixel x = t + 32 k2 of the two rows
    const long o = (2 * p) * (long)L;
#pragma unroll
    for (int k2 = 0; k2 < 32; k2++) {
        const int x = t + 32 * k2;
        // undo the index shift: multiply by e^{-2 pi i band x / L}; streaming stores
        const float2 r = cmul(u[k2], twG[(int)(((long)band * x) % L)]);
        if (O16) {
            __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, __float2half_rn(r.x)),
                                        (unsigned short*)img + o + x);
            __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, __float2half_rn(r.y)),
                                        (unsigned short*)img + o + L + x);
        } else {
            __builtin_nontemporal_store(r.x, (float*)img + o + x);
            __builtin_nontemporal_store(r.y, (float*)img + o + L + x);
        }
    }
}


// ---------------------------------------------------------------------------
// Band row transforms of 16 x Q points (N = 256, 512, 768 for Q = 16, 32, 48)
// in 16-lane transforms, four per wavefront.  Lane t of a transform is the
// stage-1 index n2 = t + 16 s (s < Q / 16) and holds x[Q n1 + n2], n1 < 16;
// after the stage transpose it is k1 = t and holds X[t + 16 k2], k2 < Q.  The
// transpose goes through LDS one component plane at a time (16 rows of pitch
// Q + 1 floats per transform) and stays inside the wavefront, so it needs no
// block barrier.
// ---------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

template <int Q>
__device__ __forceinline__ void transpose16_planes(int t, float* __restrict__ xb,
                                                   const float2 (&v)[Q / 16][16],
                                                   float2 (&u)[Q]) {
    constexpr int LDQ = Q + 1, S1 = Q / 16;
#pragma unroll
    for (int s = 0; s < S1; s++)
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) xb[k1 * LDQ + t + 16 * s] = v[s][k1].x;
    wave_sync();
#pragma unroll
    for (int n2 = 0; n2 < Q; n2++) u[n2].x = xb[t * LDQ + n2];
    wave_sync();
#pragma unroll
    for (int s = 0; s < S1; s++)
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) xb[k1 * LDQ + t + 16 * s] = v[s][k1].y;
    wave_sync();
#pragma unroll
    for (int n2 = 0; n2 < Q; n2++) u[n2].y = xb[t * LDQ + n2];
    wave_sync();
}

// rows_band16: rows_fwd2 for N = 16 Q point rows and the band window
// kx = 0..count-1 (count <= 32), transposed output with the centring sign.  Two
// real rows per complex transform; X[N - k] comes from lane (16 - t) mod 16 of
// the transform by ds_bpermute (u[Q-1] for k < 16, u[Q-2] for 16 <= k < 32).
template <int Q, bool H16, bool IN8, int NTH>
__global__ __launch_bounds__(NTH) void rows_band16_kernel(const void* __restrict__ in_,
                                                          void* __restrict__ out_, int nrows,
                                                          int count) {
    constexpr int L = 16 * Q, S1 = Q / 16, FPB = NTH / 16, LDQ = Q + 1;
    using S = Store<H16>;
    typename S::T* out = (typename S::T*)out_;
    __shared__ float2 twG[L];
    __shared__ float xplane[FPB * 16 * LDQ];
    const int tid = threadIdx.x;
    const int f = tid >> 4, t = tid & 15;
    const long npairs = nrows / 2;
    const long p = (long)blockIdx.x * FPB + f;
    const bool act = p < npairs;
    float* xb = xplane + f * 16 * LDQ;
    // an idle transform reads pair 0 (in bounds) and writes nothing
    const long roff = 2 * (act ? p : 0) * (long)L;
    const unsigned char* r8 = (const unsigned char*)in_ + roff + t;
    const float* r0 = (const float*)in_ + roff + t;
    // A transform reads 64 bytes of a row per load, so the S1 pieces of one
    // row segment are loaded back to back, and through the caches (streaming
    // loads of half lines measured 1.2-1.5x slower)
    float2 v[S1][16];
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++)
#pragma unroll
        for (int s = 0; s < S1; s++) {
            const int n = Q * n1 + 16 * s;
            if (IN8)
                v[s][n1] = make_float2((float)r8[n] * (1.0f / 255.0f), (float)r8[L + n] * (1.0f / 255.0f));
            else
                v[s][n1] = make_float2(r0[n], r0[L + n]);
        }
    // the loads above go out before the twiddle table is computed
    fill_twiddles<L>(twG, tid, NTH);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S1; s++) {
        dft_reg<16, false>(v[s]);
#pragma unroll
        for (int k1 = 1; k1 < 16; k1++) v[s][k1] = cmul_w<false>(v[s][k1], twG[(t + 16 * s) * k1]);
    }
    float2 u[Q];
    transpose16_planes<Q>(t, xb, v, u);
    dft_reg<Q, false>(u);  // u[k2] = X[t + 16 k2]
    const int lane = tid & 63;
    const int src = ((lane & 48) | ((16 - t) & 15)) << 2;
    const float2 m1 = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[Q - 1].x))),
                                  __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[Q - 1].y))));
    const float2 m2 = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[Q - 2].x))),
                                  __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(u[Q - 2].y))));
    if (!act) return;
    const long r = 2 * p, img = r / L, y = r - img * L;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int k = t + 16 * h;  // window column j = k (kmin = 0)
        if (k >= count) break;
        const float2 z1 = h ? u[1] : u[0];
        const float2 zm = h ? (t == 0 ? m1 : m2) : (t == 0 ? u[0] : m1);  // X[N - k]
        const float2 z2 = cconj(zm);
        float2 x1 = cscale(cadd(z1, z2), 0.5f);
        const float2 d = csub(z1, z2);
        float2 x2 = make_float2(0.5f * d.y, -0.5f * d.x);
        if (k & 1) {
            x1 = cscale(x1, -1.f);
            x2 = cscale(x2, -1.f);
        }
        typename S::T* o = out + (img * count + k) * (long)L + y;
        o[0] = S::st(x1);
        o[1] = S::st(x2);
    }
}

// image_band16: image_rows2 for N = 16 Q point rows of the band path (U
// transposed [img][kx][row], 2 band + 1 <= NZ Q).  The band row n = kx + band
// (two rows' Hermitian extensions as re / im) is stage-1 input n = Q n1 + n2,
// zero for n1 >= NZ; the shift is undone on the output by e^{-2 pi i band x / N}.
template <int Q, int NZ, bool H16, bool O16, int NTH>
__global__ __launch_bounds__(NTH) void image_band16_kernel(const void* __restrict__ U_,
                                                           void* __restrict__ img, int nrows,
                                                           int band) {
    constexpr int L = 16 * Q, S1 = Q / 16, FPB = NTH / 16, LDQ = Q + 1;
    using S = Store<H16>;
    __shared__ float2 twG[L];
    __shared__ float xplane[FPB * 16 * LDQ];
    const int tid = threadIdx.x;
    const int f = tid >> 4, t = tid & 15;
    const int nb = band + 1;
    const long npairs = nrows / 2;
    const long p = (long)blockIdx.x * FPB + f;
    const bool act = p < npairs;
    fill_twiddles<L>(twG, tid, NTH);
    __syncthreads();
    float* xb = xplane + f * 16 * LDQ;
    const typename S::T* U = (const typename S::T*)U_;
    const long r0 = 2 * (act ? p : 0);  // an idle transform reads row pair 0
    const long img0 = r0 / L, y0 = r0 - img0 * L;
    const typename S::T* u1 = U + img0 * nb * (long)L + y0;
    float2 v[S1][16];
#pragma unroll
    for (int s = 0; s < S1; s++)
#pragma unroll
        for (int n1 = 0; n1 < 16; n1++) {
            float2 z = make_float2(0.f, 0.f);
            const int n = Q * n1 + t + 16 * s;
            const int kx = n - band;
            if (n1 < NZ && kx >= -band && kx <= band) {
                const int a = kx < 0 ? -kx : kx;
                const float2 a1 = S::ld(u1[(long)a * L]), a2 = S::ld(u1[(long)a * L + 1]);
                if (a == 0)
                    z = make_float2(a1.x, a2.x);
                else if (kx > 0)
                    z = make_float2(a1.x - a2.y, a1.y + a2.x);  // u1 + i u2
                else
                    z = make_float2(a1.x + a2.y, a2.x - a1.y);  // conj(u1) + i conj(u2)
            }
            v[s][n1] = z;
        }
#pragma unroll
    for (int s = 0; s < S1; s++) {
        dft_reg<16, true>(v[s]);
#pragma unroll
        for (int k1 = 1; k1 < 16; k1++) {
            float2 w = twG[(t + 16 * s) * k1];
            w.y = -w.y;
            v[s][k1] = cmul_w<false>(v[s][k1], w);
        }
    }
    float2 u[Q];
    transpose16_planes<Q>(t, xb, v, u);
    if (!act) return;
    dft_reg<Q, true>(u);  // u[k2] = pixel x = t + 16 k2 of the two rows
    // undo the index shift: multiply by e^{-2 pi i band x / L}, x = t + 16 k2;
    // the table index band x mod L advances by 16 band mod L per k2 (no
    // division per pixel when L is not a power of two)
    {
        const int step = (16 * band) % L;
        int m = (band * t) % L;
#pragma unroll
        for (int k2 = 0; k2 < Q; k2++) {
            u[k2] = cmul(u[k2], twG[m]);
            m += step;
            m = m >= L ? m - L : m;
        }
    }
    // Each row passes through the transform's LDS plane so that a lane stores
    // 4 consecutive pixels and one store covers whole lines (16-lane stores of
    // 64 bytes measured 1.5x slower); streaming stores, the image is not read here.
    const long o = (2 * p) * (long)L;
#pragma unroll
    for (int row = 0; row < 2; row++) {
#pragma unroll
        for (int k2 = 0; k2 < Q; k2++) xb[t + 16 * k2] = row ? u[k2].y : u[k2].x;
        wave_sync();
#pragma unroll
        for (int c = 0; c < Q / 4; c++) {
            const int x0 = 64 * c + 4 * t;
            const float4 q = *(const float4*)(xb + x0);
            if (O16) {
                const u2v h = {__builtin_bit_cast(unsigned, __halves2half2(__float2half_rn(q.x),
                                                                           __float2half_rn(q.y))),
                               __builtin_bit_cast(unsigned, __halves2half2(__float2half_rn(q.z),
                                                                           __float2half_rn(q.w)))};
                __builtin_nontemporal_store(h, (u2v*)((unsigned short*)img + o + row * L + x0));
            } else {
                const f4v f = {q.x, q.y, q.z, q.w};
                __builtin_nontemporal_store(f, (f4v*)((float*)img + o + row * L + x0));
            }
        }
        wave_sync();
    }
}

// mask_cols16 / fi_cols16: mask_cols2 and fi_cols2 for N = 16 Q point columns
// in 16-lane transforms (per-transform LDS planes, wavefront fences only); same
// inputs, outputs and index conventions.  A transform reads and writes 16
// complex values (128 bytes) of a column per access.
template <int Q, bool H16, int NTH>
__global__ __launch_bounds__(NTH) void mask_cols16_kernel(const void* __restrict__ R_,
                                                          float2* __restrict__ M, int cx,
                                                          int kymin, int cy, int mw) {
    constexpr int N = 16 * Q, S1 = Q / 16, FPB = NTH / 16, LDQ = Q + 1;
    __shared__ float2 twG[N];
    __shared__ float xplane[FPB * 16 * LDQ];
    const int tid = threadIdx.x;
    const int f = tid >> 4, t = tid & 15;
    const long b = blockIdx.y;
    const int col = blockIdx.x * FPB + f;
    const bool act = col < cx;
    using S = Store<H16>;
    // an idle transform reads column 0 and writes nothing
    const typename S::T* Rc = (const typename S::T*)R_ + (b * cx + (act ? col : 0)) * (long)N + t;
    float2 v[S1][16];
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++)
#pragma unroll
        for (int s = 0; s < S1; s++) v[s][n1] = S::ld(Rc[Q * n1 + 16 * s]);
    // the loads above go out before the twiddle table is computed
    fill_twiddles<N>(twG, tid, NTH);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S1; s++) {
        dft_reg<16, false>(v[s]);
#pragma unroll
        for (int k1 = 1; k1 < 16; k1++) v[s][k1] = cmul_w<false>(v[s][k1], twG[(t + 16 * s) * k1]);
    }
    float2 u[Q];
    transpose16_planes<Q>(t, xplane + f * 16 * LDQ, v, u);
    if (!act) return;
    dft_reg<Q, false>(u);  // u[k2] = X[t + 16 k2]
    const float inv = 1.0f / ((float)N * (float)N);
    float2* Mb = M + b * (long)cy * (mw ? mw : cx);
    const int h = mw >> 1;
    // |ky| < 32 (checked by the launcher): only k2 = 0, 1, Q-2, Q-1 hold kept
    // rows, so the rest of the stage-2 DFT folds away
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k2 = i < 2 ? i : Q - 4 + i;
        const int k = t + 16 * k2;
        const int ky = k < N / 2 ? k : k - N;
        const int j = ky - kymin;
        if (j >= 0 && j < cy) {
            const float2 val = cscale(u[k2], (ky & 1) ? -inv : inv);
            if (mw) {
                // column kx = 0: ky >= 0 computed, ky < 0 written as the conjugate
                if (col > 0 || ky >= 0) Mb[(long)j * mw + h + col] = val;
                if (col > 0 || ky > 0) Mb[(long)(cy - 1 - j) * mw + h - col] = make_float2(val.x, -val.y);
            } else {
                Mb[(long)j * cx + col] = val;
            }
        }
    }
}

template <int Q, int NZ, bool H16, int NTH>
__global__ __launch_bounds__(NTH) void fi_cols16_kernel(const float2* __restrict__ C,
                                                        void* __restrict__ U_, int band, int nsum,
                                                        long cplane) {
    constexpr int N = 16 * Q, S1 = Q / 16, FPB = NTH / 16, LDQ = Q + 1;
    __shared__ float2 twG[N];
    __shared__ float xplane[FPB * 16 * LDQ];
    const int tid = threadIdx.x;
    const int f = tid >> 4, t = tid & 15;
    const int nb = band + 1;
    const long b = blockIdx.y;
    const int col = blockIdx.x * FPB + f;
    const bool act = col < nb;
    // an idle transform reads column 0 and writes nothing
    const float2* Cb = C + (b * (long)nb + (act ? col : 0)) * (2 * band + 1);
    float2 v[S1][16];
#pragma unroll
    for (int s = 0; s < S1; s++)
#pragma unroll
        for (int n1 = 0; n1 < 16; n1++) {
            const int n = Q * n1 + t + 16 * s;
            float2 z = make_float2(0.f, 0.f);
            if (n1 < NZ && n <= 2 * band) {
                z = Cb[n];
                if (nsum > 1) z = cadd(z, Cb[n + cplane]);
                if ((n - band) & 1) z = cscale(z, -1.f);
            }
            v[s][n1] = z;
        }
    // the loads above go out before the twiddle table is computed
    fill_twiddles<N>(twG, tid, NTH);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S1; s++) {
        dft_reg<16, true>(v[s]);
#pragma unroll
        for (int k1 = 1; k1 < 16; k1++) {
            float2 w = twG[(t + 16 * s) * k1];
            w.y = -w.y;
            v[s][k1] = cmul_w<false>(v[s][k1], w);
        }
    }
    float2 u[Q];
    transpose16_planes<Q>(t, xplane + f * 16 * LDQ, v, u);
    if (!act) return;
    dft_reg<Q, true>(u);  // u[k2] = row y = t + 16 k2
    using S = Store<H16>;
    typename S::T* Uc = (typename S::T*)U_ + (b * nb + col) * (long)N;
    const float sx = (col & 1) ? -1.f : 1.f;
    // e^{-2 pi i band y / N}: the table index advances by 16 band mod N per k2
    const int step = (16 * band) % N;
    int m = (band * t) % N;
#pragma unroll
    for (int k2 = 0; k2 < Q; k2++) {
        Uc[t + 16 * k2] = S::st(cscale(cmul(u[k2], twG[m]), sx));
        m += step;
        m = m >= N ? m - N : m;
    }
}

// ---------------------------------------------------------------------------
// Band pipeline (image path): square conv grid G = P*P of any size with
// factors 2 and 3 (e.g. 36, 81, 144), whole image per block.
//
// socs_band: SOCS accumulate as socs_accum2, then the forward R2C of the
// accumulated intensity done in place: the row-pass outputs sit in exactly
// the stage-1 input layout of a forward row FFT, so rows are transformed from
// registers, kept for kx <= band, staged in LDS and transformed down the
// columns; Cout[b][ky + band][kx] = (1/G^2) DFT(intensity)[ky][kx] for
// |ky| <= band, 0 <= kx <= band (band = 2 Nx).  This is FI's R2C (L724-735)
// restricted to the band the intensity occupies.
// ---------------------------------------------------------------------------
template <int P, int NZ, int SLOTS, int BK = 0>
__global__ __launch_bounds__(P * SLOTS) void socs_band_kernel(const float2* __restrict__ M,
                                                              const void* __restrict__ bank,
                                                              const float2* __restrict__ kscale,
                                                              const float* __restrict__ scales,
                                                              int nk, int Nx,
                                                              float2* __restrict__ Cout) {
    using Geo = Fft2Geom<P, P>;
    constexpr int G = Geo::G;
    constexpr int NTH = P * SLOTS;
    constexpr int RB = G / SLOTS;
    static_assert(G % SLOTS == 0, "slots must divide the grid");
    extern __shared__ float2 lds[];
    float2* twG = lds;
    float2* xbuf = lds + G;
    float2* Y = xbuf + SLOTS * Geo::XBUF;  // [G][W]
    const int W = 2 * Nx + 1;
    const int band = 2 * Nx;
    const int tid = threadIdx.x;
    const int slot = tid / P, t = tid - slot * P;
    const long b = blockIdx.x;
    const float2* Mb = M + b * (long)W * W;
    float2* xb = xbuf + slot * Geo::XBUF;
    fill_twiddles<G>(twG, tid, NTH);
    float acc[RB * P];
#pragma unroll
    for (int i = 0; i < RB * P; i++) acc[i] = 0.f;
    __syncthreads();

    for (int k = 0; k < nk; k++) {
        const long kbase = (long)k * W * W;
        const float lam = scales[k];
        for (int c0 = 0; c0 < W; c0 += SLOTS) {
            const int col = c0 + slot;
            const bool act = col < W;
            fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
                float2 v = make_float2(0.f, 0.f);
                if (act && n < W) {
                    const int j = n * W + col;
                    v = cmul(Bank<BK>::ld(bank, kscale, kbase + j, k), Mb[j]);
                }
                return v;
            });
            __syncthreads();
            fft2_stage2<Geo, true>(t, xb, [&](int, int g, float2 val) {
                if (act) Y[g * W + col] = val;
            });
            __syncthreads();
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const float2* yr = Y + (rb * SLOTS + slot) * W;
            fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
                return n < W ? yr[n] : make_float2(0.f, 0.f);
            });
            __syncthreads();
            fft2_stage2<Geo, true>(t, xb, [&](int sidx, int, float2 val) {
                acc[rb * P + sidx] += lam * (val.x * val.x + val.y * val.y);
            });
            __syncthreads();
        }
    }
    // forward R2C of the intensity rows straight from registers
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int r = rb * SLOTS + slot;
        fft2_stage1_ij<Geo, P, false>(t, xb, twG, [&](int n1, int) {
            return make_float2(acc[rb * P + n1], 0.f);
        });
        __syncthreads();
        fft2_stage2<Geo, false>(t, xb, [&](int, int kx, float2 val) {
            if (kx <= band) Y[r * W + kx] = val;
        });
        __syncthreads();
    }
    // columns kx = 0..band: forward FFT over the rows, keep |ky| <= band
    const float inv = 1.0f / ((float)G * (float)G);
    float2* Cb = Cout + b * (long)(2 * band + 1) * W;
    for (int c0 = 0; c0 < W; c0 += SLOTS) {
        const int col = c0 + slot;
        const bool act = col < W;
        fft2_stage1<Geo, P, false>(t, xb, twG, [&](int n) {
            return act ? Y[n * W + col] : make_float2(0.f, 0.f);
        });
        __syncthreads();
        fft2_stage2<Geo, false>(t, xb, [&](int, int kk, float2 val) {
            const int ky = kk <= G / 2 ? kk : kk - G;
            if (act && ky >= -band && ky <= band) Cb[(long)col * (2 * band + 1) + ky + band] = cscale(val, inv);
        });
        __syncthreads();
    }
}

// mask_cols2: R [B][cx][N] (row-DFT window, transposed) -> M [B][cy][cx],
// M[b][j][c] = (1/N^2) (-1)^ky DFT_N(R[b][c][:])[ky mod N], ky = kymin + j.
// Mirror mode (mw = 2h + 1 > 0, cx = h + 1, kymin = -h, cy = mw): R holds only
// kx = 0..h of a real mask's spectrum, and M [B][mw][mw] gets both halves
// through M(-ky, -kx) = conj(M(ky, kx)) (the (-1)^(kx+ky) centring is real);
// in the kx = 0 column the ky < 0 half is the conjugate of the ky > 0 half too,
// so M is exactly conjugate-symmetric about the window centre.
template <int P, int Q, int NTH, bool H16 = false>
__global__ __launch_bounds__(NTH) void mask_cols2_kernel(const void* __restrict__ R_,
                                                         float2* __restrict__ M, int cx,
                                                         int kymin, int cy, int mw) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int N = Geo::G;
    constexpr int T = Geo::T;
    constexpr int FPB = NTH / T;
    __shared__ float2 twG[N];
    __shared__ float2 xbuf[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const long b = blockIdx.y;
    const int col = blockIdx.x * FPB + f;
    const bool act = col < cx;
    fill_twiddles<N>(twG, tid, NTH);
    __syncthreads();
    float2* xb = xbuf + f * Geo::XBUF;
    using S = Store<H16>;
    const typename S::T* Rc = (const typename S::T*)R_ + (b * cx + (act ? col : 0)) * (long)N;
    fft2_stage1<Geo, P, false>(t, xb, twG, [&](int n) { return S::ld(Rc[n]); });
    __syncthreads();
    if (!act) return;
    const float inv = 1.0f / ((float)N * (float)N);
    float2* Mb = M + b * (long)cy * (mw ? mw : cx);
    const int h = mw >> 1;
    fft2_stage2<Geo, false>(t, xb, [&](int, int k, float2 val) {
        const int ky = k < N / 2 ? k : k - N;
        const int j = ky - kymin;
        if (j >= 0 && j < cy) {
            const float2 v = cscale(val, (ky & 1) ? -inv : inv);
            if (mw) {
                // column kx = 0: ky >= 0 computed, ky < 0 written as the conjugate,
                // so M(-f) = conj(M(f)) holds exactly
                if (col > 0 || ky >= 0) Mb[(long)j * mw + h + col] = v;
                if (col > 0 || ky > 0) Mb[(long)(cy - 1 - j) * mw + h - col] = make_float2(v.x, -v.y);
            } else {
                Mb[(long)j * cx + col] = v;
            }
        }
    });
}

// fi_cols2: C [B][2band+1][nb] (band spectrum, kx = 0..band) -> U [B][nb][N]:
// U[b][kx][y] = (-1)^kx sum_ky (-1)^ky C[ky][kx] e^{+2 pi i ky y / N}, the
// band placed at indices ky + band and the shift undone on the output.
template <int P, int Q, int NZ, int NTH, bool H16 = false>
__global__ __launch_bounds__(NTH) void fi_cols2_kernel(const float2* __restrict__ C,
                                                       void* __restrict__ U_, int band, int nsum,
                                                       long cplane) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int N = Geo::G;
    constexpr int T = Geo::T;
    constexpr int FPB = NTH / T;
    __shared__ float2 twG[N];
    __shared__ float2 xbuf[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const int nb = band + 1;
    const long b = blockIdx.y;
    const int col = blockIdx.x * FPB + f;
    const bool act = col < nb;
    fill_twiddles<N>(twG, tid, NTH);
    __syncthreads();
    float2* xb = xbuf + f * Geo::XBUF;
    // an idle transform reads column 0 and writes nothing
    const float2* Cb = C + (b * (long)nb + (act ? col : 0)) * (2 * band + 1);
    fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
        float2 v = make_float2(0.f, 0.f);
        if (n <= 2 * band) {
            v = Cb[n];
            if (nsum > 1) v = cadd(v, Cb[n + cplane]);
            if ((n - band) & 1) v = cscale(v, -1.f);
        }
        return v;
    });
    __syncthreads();
    if (!act) return;
    using S = Store<H16>;
    typename S::T* Uc = (typename S::T*)U_ + (b * nb + col) * (long)N;
    const float sx = (col & 1) ? -1.f : 1.f;
    fft2_stage2<Geo, true>(t, xb, [&](int, int y, float2 val) {
        const float2 w = twG[(int)(((long)band * y) % N)];  // e^{-2 pi i band y / N}
        Uc[y] = S::st(cscale(cmul(val, w), sx));
    });
}

// mask_cols_band1024 / fi_cols_band1024: mask_cols2 and fi_cols2 for 1024-point
// columns with the plane-split stage transpose (42 KB per block: three blocks
// per CU); same inputs, outputs and index conventions.
template <bool H16, int NTH = 256>
__global__ __launch_bounds__(NTH) void mask_cols_band1024_kernel(const void* __restrict__ R_,
                                                                 float2* __restrict__ M, int cx,
                                                                 int kymin, int cy, int mw) {
    using Geo = Fft2Geom<32, 32>;
    constexpr int N = 1024, T = 32, FPB = NTH / T;
    __shared__ float2 twG[N];
    __shared__ float xplane[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const long b = blockIdx.y;
    const int col = blockIdx.x * FPB + f;
    const bool act = col < cx;
    using S = Store<H16>;
    const typename S::T* Rc = (const typename S::T*)R_ + (b * cx + (act ? col : 0)) * (long)N;
    // the column loads go out first: the twiddle table is computed while they fly
    float2 v[32];
#pragma unroll
    for (int n1 = 0; n1 < 32; n1++) v[n1] = S::ld(Rc[32 * n1 + t]);  // idle: column 0
    fill_twiddles<N>(twG, tid, NTH);
    __syncthreads();
    dft_reg<32, false>(v);
#pragma unroll
    for (int k1 = 1; k1 < 32; k1++) v[k1] = cmul_w<false>(v[k1], twG[t * k1]);
    float2 u[32];
    transpose32_planes(t, xplane + f * Geo::XBUF, v, u);
    if (!act) return;
    dft_reg<32, false>(u);  // u[k2] = X[t + 32 k2]
    const float inv = 1.0f / ((float)N * (float)N);
    float2* Mb = M + b * (long)cy * (mw ? mw : cx);
    const int h = mw >> 1;
    // |ky| < 64 (checked by the launcher): only k2 = 0, 1, 30, 31 hold kept
    // rows, so the rest of the stage-2 DFT folds away
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k2 = i < 2 ? i : 28 + i;
        const int k = t + 32 * k2;
        const int ky = k < N / 2 ? k : k - N;
        const int j = ky - kymin;
        if (j >= 0 && j < cy) {
            const float2 val = cscale(u[k2], (ky & 1) ? -inv : inv);
            if (mw) {
                // column kx = 0: ky >= 0 computed, ky < 0 written as the conjugate,
                // so M(-f) = conj(M(f)) holds exactly
                if (col > 0 || ky >= 0) Mb[(long)j * mw + h + col] = val;
                if (col > 0 || ky > 0) Mb[(long)(cy - 1 - j) * mw + h - col] = make_float2(val.x, -val.y);
            } else {
                Mb[(long)j * cx + col] = val;
            }
        }
    }
}

// SPLIT: U in the compact form of the image-side matrix-core product: per
// image [ns][N] FP32 for the columns kx < ns that the product splits into fp16
// hi + lo parts itself, then [nb - ns][N] half2 for the rest, which it uses
// as fp16 values alone (rounded here as it would round them).
template <bool H16, int NTH = 256, bool SPLIT = false>
__global__ __launch_bounds__(NTH) void fi_cols_band1024_kernel(const float2* __restrict__ C,
                                                               void* __restrict__ U_, int band,
                                                               int ns, int nsum, long cplane) {
    using Geo = Fft2Geom<32, 32>;
    constexpr int N = 1024, T = 32, FPB = NTH / T, NZ = 5;
    __shared__ float2 twG[N];
    __shared__ float xplane[FPB * Geo::XBUF];
    const int tid = threadIdx.x;
    const int f = tid / T, t = tid - f * T;
    const int nb = band + 1;
    const long b = blockIdx.y;
    const int col = blockIdx.x * FPB + f;
    const bool act = col < nb;
    // an idle transform reads column 0 and writes nothing; the band loads go
    // out before the twiddle table is computed
    const float2* Cb = C + (b * (long)nb + (act ? col : 0)) * (2 * band + 1);
    float2 v[32];
#pragma unroll
    for (int n1 = 0; n1 < 32; n1++) {
        const int n = 32 * n1 + t;
        float2 z = make_float2(0.f, 0.f);
        if (n1 < NZ && n <= 2 * band) {
            z = Cb[n];
            if (nsum > 1) z = cadd(z, Cb[n + cplane]);  // the second half of the entries
            if ((n - band) & 1) z = cscale(z, -1.f);
        }
        v[n1] = z;
    }
    fill_twiddles<N>(twG, tid, NTH);
    __syncthreads();
    dft_reg<32, true>(v);
#pragma unroll
    for (int k1 = 1; k1 < 32; k1++) {
        float2 w = twG[t * k1];
        w.y = -w.y;
        v[k1] = cmul_w<false>(v[k1], w);
    }
    float2 u[32];
    transpose32_planes(t, xplane + f * Geo::XBUF, v, u);
    if (!act) return;
    dft_reg<32, true>(u);  // u[k2] = row y = t + 32 k2
    const float sx = (col & 1) ? -1.f : 1.f;
    if (SPLIT) {
        unsigned char* Ui =
            (unsigned char*)U_ + b * ((long)ns * N * 8 + (long)(nb - ns) * N * 4);
        const bool lo = col < ns;
        float2* Uf = (float2*)Ui + (long)(lo ? col : 0) * N;
        __half2* Uh = (__half2*)(Ui + (long)ns * N * 8) + (long)(lo ? 0 : col - ns) * N;
#pragma unroll
        for (int k2 = 0; k2 < 32; k2++) {
            const int y = t + 32 * k2;
            const float2 w = twG[(int)(((long)band * y) % N)];  // e^{-2 pi i band y / N}
            const float2 z = cscale(cmul(u[k2], w), sx);
            if (lo)
                Uf[y] = z;
            else
                Uh[y] = __floats2half2_rn(z.x, z.y);
        }
        return;
    }
    using S = Store<H16>;
    typename S::T* Uc = (typename S::T*)U_ + (b * nb + col) * (long)N;
#pragma unroll
    for (int k2 = 0; k2 < 32; k2++) {
        const int y = t + 32 * k2;
        const float2 w = twG[(int)(((long)band * y) % N)];  // e^{-2 pi i band y / N}
        Uc[y] = S::st(cscale(cmul(u[k2], w), sx));
    }
}

// ---------------------------------------------------------------------------
// fi_cols: P [B][G][nb] float2 (row DFT of the unshifted conv-grid intensity,
// kx = 0..band) -> U [B][N][nb] float2.  For column kx:
//   C[ky] = (1/G^2) DFT_G(P[:,kx])[ky mod G], |ky| <= band
//   U[y][kx] = (-1)^kx sum_ky (-1)^ky C[ky] e^{+2 pi i ky y / N}
// CG columns per block, grid (ceil(nb/CG), B).  yalt = 1 when the intensity
// rows came from a centred (fftshifted) grid: its (-1)^ky cancels the
// centring sign, so no sign is applied along ky.
// ---------------------------------------------------------------------------
template <int G, int N, int CG>
__global__ __launch_bounds__(256) void fi_cols_kernel(const float2* __restrict__ P,
                                                      float2* __restrict__ U, int band, int yalt,
                                                      const float2* __restrict__ tw) {
    __shared__ float2 sa[CG * N];
    __shared__ float2 sb[CG * N];
    const int tid = threadIdx.x;
    const int nb = band + 1;
    const int lin = xcd_block_order();
    const int c0 = (lin % gridDim.x) * CG;
    const long b = lin / gridDim.x;
    const float2* Pb = P + b * (long)G * nb;
    for (int i = tid; i < CG * G; i += 256) {
        int g = i / CG, c = i - g * CG;
        int col = c0 + c;
        sa[c * G + g] = col < nb ? Pb[(long)g * nb + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* r1 = fft_lds<G, false>(sa, sb, CG, tw, tid, 256);
    float2* other = (r1 == sa) ? sb : sa;
    // embed the band into an N-point spectrum with the centring signs
    const float inv = 1.0f / ((float)G * (float)G);
    for (int i = tid; i < CG * N; i += 256) {
        int c = i / N, y = i - c * N;
        int ky = y < N / 2 ? y : y - N;
        float2 v = make_float2(0.f, 0.f);
        if (ky >= -band && ky <= band) {
            int src = ((ky % G) + G) % G;
            float s = (!yalt && (ky & 1)) ? -inv : inv;
            v = cscale(r1[c * G + src], s);
        }
        other[i] = v;
    }
    __syncthreads();
    float2* spare = (other == sa) ? sb : sa;
    float2* r2 = fft_lds<N, true>(other, spare, CG, tw, tid, 256);
    float2* Ub = U + b * (long)N * nb;
    for (int i = tid; i < CG * N; i += 256) {
        int y = i / CG, c = i - y * CG;
        int col = c0 + c;
        if (col >= nb) continue;
        float2 v = r2[c * N + y];
        if (col & 1) v = cscale(v, -1.f);
        Ub[(long)y * nb + col] = v;
    }
}

// ---------------------------------------------------------------------------
// image_rows: U [B*N][nb] float2 -> image [B*N][N] f32.  Two rows per complex
// FFT: Z[kx] = U1[kx] + i U2[kx], Z[-kx] = conj(U1[kx]) + i conj(U2[kx]),
// DC taken real (the C2R ignores its imaginary part); inverse FFT gives row 1
// in the real part and row 2 in the imaginary part.
// ---------------------------------------------------------------------------
template <int N, int FPB>
__global__ __launch_bounds__(256) void image_rows_kernel(const float2* __restrict__ U,
                                                         float* __restrict__ img, int nrows,
                                                         int band,
                                                         const float2* __restrict__ tw) {
    __shared__ float2 sa[FPB * N];
    __shared__ float2 sb[FPB * N];
    const int tid = threadIdx.x;
    const int nb = band + 1;
    const long pair0 = (long)blockIdx.x * FPB;
    const long npairs = nrows / 2;
    for (int i = tid; i < FPB * N; i += 256) sa[i] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int i = tid; i < FPB * nb; i += 256) {
        int f = i / nb, kx = i - f * nb;
        long p = pair0 + f;
        if (p >= npairs) continue;
        float2 u1 = U[(2 * p) * (long)nb + kx];
        float2 u2 = U[(2 * p + 1) * (long)nb + kx];
        if (kx == 0) {
            sa[f * N] = make_float2(u1.x, u2.x);
        } else {
            // u1 + i u2 and conj(u1) + i conj(u2)
            sa[f * N + kx] = make_float2(u1.x - u2.y, u1.y + u2.x);
            sa[f * N + N - kx] = make_float2(u1.x + u2.y, u2.x - u1.y);
        }
    }
    __syncthreads();
    float2* res = fft_lds<N, true>(sa, sb, FPB, tw, tid, 256);
    for (int i = tid; i < FPB * N; i += 256) {
        int f = i / N, x = i - f * N;
        long p = pair0 + f;
        if (p >= npairs) continue;
        float2 v = res[i];
        img[(2 * p) * (long)N + x] = v.x;
        img[(2 * p + 1) * (long)N + x] = v.y;
    }
}

// ---------------------------------------------------------------------------
// 4096-point row transforms of a row pair (two real rows as one complex
// sequence), three radix-16 stages in registers, one pair per 256-thread block:
//   n = t + 256 n3 (t = n1 + 16 n2), k = a + 16 c + 256 d
//   A: thread t, DFT16 over n3 -> a, times W_4096^(t a)
//   B: thread (a, n1) = (tid >> 4, tid & 15), DFT16 over n2 -> c, times W_256^(n1 c)
//   C: thread (a, c) = (tid & 15, tid >> 4), DFT16 over n1 -> d: X[a + 16 c + 256 d]
// with two transposes through LDS.  Stage C's thread order makes a + 16 c =
// tid, so thread tid ends with X[tid + 256 d]: a wavefront holds 64
// consecutive outputs of every d and stores them without a reordering pass.
// Each transpose moves one float component at a time through a single
// 17 KB plane (write x, read x, write y, read y), so a block takes 18 KB of
// LDS and eight blocks (32 wavefronts) fit a CU.  The inter-stage twiddles
// of a thread are the powers 1..15 of one root per stage (f4k_powers), the
// root taken from two 64-entry tables, W_4096^m = W_64^(m >> 6) W_4096^(m & 63).
// The inverse is the same algebra with conjugate roots (input index k,
// output index n).  Plane pitches (4-byte words): [a][t] at 272 (= 16 mod
// 64, so the stage-B reads of lanes n1 = 0..15 of four a fall on 64
// distinct banks), [n1][16 c + a] at 260 (= 4 mod 64: the stage-B writes
// likewise); the stage-A writes and stage-C reads are consecutive.
// ---------------------------------------------------------------------------
constexpr int F4K_PA = 272;  // pitch of the [a][t] plane (t < 256)
constexpr int F4K_PB = 260;  // pitch of the [n1][16 c + a] plane
constexpr int F4K_BUF = 16 * F4K_PA / 2;  // the plane, in float2 (also used as float2 after the FFT)
static_assert(16 * F4K_PA >= 16 * F4K_PB && F4K_BUF >= 2 * (256 + 256 / 32), "transpose plane");
constexpr int F4K_NEG = 256 + 256 / 32;  // the d = 15 band after the d = 0 band (padded)

__device__ __forceinline__ void f4k_tables(float2* __restrict__ lo, float2* __restrict__ hi,
                                           int tid) {
    if (tid < 128) {
        const int j = tid & 63;
        float sn, cs;
        // lo: W_4096^j; hi: W_64^j = W_4096^(64 j)
        sincospif(2.0f * (float)(tid < 64 ? j : 64 * j) / 4096.0f, &sn, &cs);
        (tid < 64 ? lo : hi)[j] = make_float2(cs, -sn);
    }
}

// v[j] *= w^j (j = 1..15) for w = W_4096^m (conjugate when INV), m < 4096:
// one table product for w, the other powers as products of depth <= 4
// (w^2, w^3 = w^2 w, w^4 = (w^2)^2, w^(4+j) = w^4 w^j, w^(8+j) = w^8 w^j).
template <bool INV>
__device__ __forceinline__ void f4k_powers(const float2* __restrict__ lo,
                                           const float2* __restrict__ hi, int m,
                                           float2 (&v)[16]) {
    float2 w1 = cmul(hi[m >> 6], lo[m & 63]);
    if (INV) w1.y = -w1.y;
    const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1), w4 = cmul(w2, w2);
    const float2 w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3), w8 = cmul(w4, w4);
    v[1] = cmul(v[1], w1);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], w5);
    v[6] = cmul(v[6], w6);
    v[7] = cmul(v[7], w7);
    v[8] = cmul(v[8], w8);
    v[9] = cmul(v[9], cmul(w8, w1));
    v[10] = cmul(v[10], cmul(w8, w2));
    v[11] = cmul(v[11], cmul(w8, w3));
    v[12] = cmul(v[12], cmul(w8, w4));
    v[13] = cmul(v[13], cmul(w8, w5));
    v[14] = cmul(v[14], cmul(w8, w6));
    v[15] = cmul(v[15], cmul(w8, w7));
}

// v (thread t: inputs t + 256 n3, n3 = 0..15) -> v (thread tid: outputs tid +
// 256 d, d = 0..15).  All 256 threads call it; buf is the F4K_BUF plane.  A
// component is overwritten only after its old value has been written out.
template <bool INV>
__device__ __forceinline__ void fft4096_pair(int tid, float2 (&v)[16], float2* __restrict__ buf,
                                             const float2* __restrict__ lo,
                                             const float2* __restrict__ hi) {
    float* pl = reinterpret_cast<float*>(buf);
    dft_reg<16, INV>(v);
    f4k_powers<INV>(lo, hi, tid, v);  // W_4096^(t a)
    const int a = tid >> 4, r = tid & 15;  // stage B: (a, n1 = r)
#pragma unroll
    for (int k = 0; k < 16; k++) pl[k * F4K_PA + tid] = v[k].x;
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 16; n2++) v[n2].x = pl[a * F4K_PA + r + 16 * n2];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) pl[k * F4K_PA + tid] = v[k].y;
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 16; n2++) v[n2].y = pl[a * F4K_PA + r + 16 * n2];
    dft_reg<16, INV>(v);
    f4k_powers<INV>(lo, hi, 16 * r, v);  // W_256^(n1 c)
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 16; c++) pl[r * F4K_PB + 16 * c + a] = v[c].x;
    __syncthreads();
    // stage C: thread (a, c) = (tid & 15, tid >> 4) reads its 16 n1 values
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) v[n1].x = pl[n1 * F4K_PB + tid];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 16; c++) pl[r * F4K_PB + 16 * c + a] = v[c].y;
    __syncthreads();
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) v[n1].y = pl[n1 * F4K_PB + tid];
    dft_reg<16, INV>(v);
}

// Mask rows of 4096 pixels -> window R [rows][count] (kx = kmin + j, |kx| < 256),
// as rows_dft_kernel: two real rows per transform, (-1)^kx when ALT.  TR: R is
// stored transposed per clip of ncl rows, R [clip][j][y], so that the column
// pass reads each column contiguously; a block writes its two rows' values as
// one 16-byte pair per j, and the row pairs are taken in XCD order, so the
// eight pairs that share a 128-byte line of R run on one XCD (one L2).
template <bool ALT, bool TR = false>
__global__ __launch_bounds__(256) void rows4096_kernel(const float* __restrict__ in,
                                                       float2* __restrict__ out, int nrows,
                                                       int kmin, int count, int ncl) {
    constexpr int L = 4096;
    __shared__ float2 buf[F4K_BUF];
    __shared__ float2 lo[64], hi[64];
    const int tid = threadIdx.x;
    const long p = TR ? xcd_block_order() : (long)blockIdx.x;  // row pair
    const float* r0 = in + 2 * p * (long)L + tid;
    float2 v[16];
#pragma unroll
    for (int n3 = 0; n3 < 16; n3++)
        v[n3] = make_float2(__builtin_nontemporal_load(r0 + 256 * n3),
                            __builtin_nontemporal_load(r0 + L + 256 * n3));
    f4k_tables(lo, hi, tid);
    __syncthreads();
    fft4096_pair<false>(tid, v, buf, lo, hi);
    // X[k] for |k| < 256 in natural order: d = 0 (k = a + 16 c) and d = 15
    // (k - 4096); entry m at m + m / 32 (lanes 16 apart on different banks)
    __syncthreads();
    buf[tid + (tid >> 5)] = v[0];             // X[tid]
    buf[F4K_NEG + tid + (tid >> 5)] = v[15];  // X[tid + 3840] = X[tid - 256]
    __syncthreads();
    auto at = [&](int k) {  // X[k mod 4096], |k| < 256
        const int m = k >= 0 ? k : k + 256;  // -k' -> 4096 - k' = 3840 + (256 - k')
        return buf[(k >= 0 ? 0 : F4K_NEG) + m + (m >> 5)];
    };
    for (int j = tid; j < count; j += 256) {
        const int k = kmin + j;
        const float2 z1 = at(k), z2 = cconj(at(-k));
        float2 x1 = cscale(cadd(z1, z2), 0.5f);
        const float2 dd = csub(z1, z2);
        float2 x2 = make_float2(0.5f * dd.y, -0.5f * dd.x);
        if (ALT && (k & 1)) {
            x1 = cscale(x1, -1.f);
            x2 = cscale(x2, -1.f);
        }
        if (TR) {
            const long b = (2 * p) / ncl, y = 2 * p - b * ncl;  // y even, ncl even
            *(float4*)(out + (b * count + j) * ncl + y) = make_float4(x1.x, x1.y, x2.x, x2.y);
        } else {
            out[(2 * p) * (long)count + j] = x1;
            out[(2 * p + 1) * (long)count + j] = x2;
        }
    }
}

// U [rows][band+1] (kx = 0..band) -> image rows of 4096 pixels, as
// image_rows_kernel: the c2r of two rows as one complex inverse transform.
__global__ __launch_bounds__(256) void image4096_kernel(const float2* __restrict__ U,
                                                        float* __restrict__ img, int nrows,
                                                        int band) {
    constexpr int L = 4096;
    __shared__ float2 buf[F4K_BUF];
    __shared__ float2 lo[64], hi[64];
    const int tid = threadIdx.x;
    const long p = blockIdx.x;
    const int nb = band + 1;
    const float2* u1 = U + (2 * p) * (long)nb;
    const float2* u2 = u1 + nb;
    // spectrum S[k], k = tid + 256 d: u1 + i u2 at k <= band, conj(u1) + i conj(u2) at
    // L - k (k = 1..band), zero elsewhere
    float2 v[16];
#pragma unroll
    for (int d = 0; d < 16; d++) {
        const int k = tid + 256 * d;
        float2 sv = make_float2(0.f, 0.f);
        if (k <= band) {
            const float2 a1 = u1[k], a2 = u2[k];
            sv = k == 0 ? make_float2(a1.x, a2.x) : make_float2(a1.x - a2.y, a1.y + a2.x);
        } else if (L - k <= band) {
            const float2 a1 = u1[L - k], a2 = u2[L - k];
            sv = make_float2(a1.x + a2.y, a2.x - a1.y);
        }
        v[d] = sv;
    }
    f4k_tables(lo, hi, tid);
    __syncthreads();
    fft4096_pair<true>(tid, v, buf, lo, hi);
    // thread tid holds pixels tid + 256 d: each store instruction of a
    // wavefront writes 256 contiguous bytes of a row
    float* o1 = img + (2 * p) * (long)L + tid;
#pragma unroll
    for (int d = 0; d < 16; d++) {
        __builtin_nontemporal_store(v[d].x, o1 + 256 * d);
        __builtin_nontemporal_store(v[d].y, o1 + L + 256 * d);
    }
}

// U transposed per clip, U [clip][nb][4096] -> image rows.  A block takes TPB
// consecutive tiles of NP row pairs (2 NP rows) of one clip.  For each tile it
// copies the rows' values of every kx into LDS (16-byte pieces, 16 NP bytes per
// kx, so the loads cover whole line segments) and runs its NP inverse
// transforms from there (the c2r of two rows as one complex transform, as
// image4096_kernel); the next tile's values are loaded into registers while
// the current tile's transforms run.  Dynamic LDS: nb * (2 NP + 1) float2
// (pitch padded by one).  band < 512 (PT pieces per thread).
template <int NP, int TPB>
__global__ __launch_bounds__(256) void image4096_tiles_kernel(const float2* __restrict__ U,
                                                              float* __restrict__ img, int band) {
    constexpr int L = 4096, R2 = 2 * NP, SP = R2 + 1, PK = NP;  // PK 16-byte pieces per kx
    constexpr int PT = (512 * PK + 255) / 256;                 // pieces per thread, nb <= 512
    constexpr int BPC = L / R2 / TPB;                           // blocks per clip
    __shared__ float2 buf[F4K_BUF];
    __shared__ float2 lo[64], hi[64];
    extern __shared__ float2 spec[];  // [nb][SP]
    const int tid = threadIdx.x;
    const int nb = band + 1;
    const long b = blockIdx.x / BPC;
    const int yb = (int)(blockIdx.x - b * BPC) * TPB * R2;
    const float2* Ub = U + b * (long)nb * L;
    float4 pre[PT];
    auto load = [&](int y0) {
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int idx = tid + 256 * i;
            if (idx < nb * PK) {
                const int kx = idx / PK, q = idx - kx * PK;
                pre[i] = *(const float4*)(Ub + (long)kx * L + y0 + 2 * q);
            }
        }
    };
    load(yb);
    f4k_tables(lo, hi, tid);
    for (int j = 0; j < TPB; j++) {
        const int y0 = yb + j * R2;
        // every thread is past the previous tile's transforms (they end in
        // block barriers after their last reads of spec), so spec is free
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int idx = tid + 256 * i;
            if (idx < nb * PK) {
                const int kx = idx / PK, q = idx - kx * PK;
                spec[kx * SP + 2 * q] = make_float2(pre[i].x, pre[i].y);
                spec[kx * SP + 2 * q + 1] = make_float2(pre[i].z, pre[i].w);
            }
        }
        __syncthreads();  // spec written; the previous transform's reads of buf done
        if (j + 1 < TPB) load(y0 + R2);
        for (int pp = 0; pp < NP; pp++) {
            float2 v[16];
#pragma unroll
            for (int d = 0; d < 16; d++) {
                const int k = tid + 256 * d;
                float2 sv = make_float2(0.f, 0.f);
                if (k <= band) {
                    const float2 a1 = spec[k * SP + 2 * pp], a2 = spec[k * SP + 2 * pp + 1];
                    sv = k == 0 ? make_float2(a1.x, a2.x) : make_float2(a1.x - a2.y, a1.y + a2.x);
                } else if (L - k <= band) {
                    const float2 a1 = spec[(L - k) * SP + 2 * pp],
                                 a2 = spec[(L - k) * SP + 2 * pp + 1];
                    sv = make_float2(a1.x + a2.y, a2.x - a1.y);
                }
                v[d] = sv;
            }
            if (pp > 0) __syncthreads();  // the previous transform's last reads of buf
            fft4096_pair<true>(tid, v, buf, lo, hi);
            float* o1 = img + (b * L + y0 + 2 * pp) * (long)L + tid;
#pragma unroll
            for (int d = 0; d < 16; d++) {
                __builtin_nontemporal_store(v[d].x, o1 + 256 * d);
                __builtin_nontemporal_store(v[d].y, o1 + L + 256 * d);
            }
        }
    }
}

// TR: R [clip][cx][4096] (the transposed form of rows4096_kernel<.., true>),
// each column read contiguously.
template <bool TR = false>
__global__ __launch_bounds__(256) void mask_cols4096_kernel(const float2* __restrict__ R,
                                                            float2* __restrict__ M, int cx,
                                                            int kymin, int cy) {
    constexpr int N = 4096;
    __shared__ float2 buf[F4K_BUF];
    __shared__ float2 lo[64], hi[64];
    const int tid = threadIdx.x;
    const int lin = xcd_block_order();
    const int col = lin % cx;
    const long b = lin / cx;
    float2 v[16];
    if (TR) {
        const float2* Rc = R + (b * cx + col) * (long)N + tid;
#pragma unroll
        for (int n3 = 0; n3 < 16; n3++) v[n3] = Rc[256 * n3];
    } else {
        const float2* Rc = R + b * (long)N * cx + col;
#pragma unroll
        for (int n3 = 0; n3 < 16; n3++) v[n3] = Rc[(long)(tid + 256 * n3) * cx];
    }
    f4k_tables(lo, hi, tid);
    __syncthreads();
    fft4096_pair<false>(tid, v, buf, lo, hi);
    // thread tid holds X[tid] (d = 0) and X[tid - 256] (d = 15); every ky in
    // [kymin, kymin + cy) (inside (-256, 256)) is one of them
    const float inv = 1.0f / ((float)N * (float)N);
    float2* Mb = M + b * (long)cy * cx + col;
    const int j0 = tid - kymin, j1 = tid - 256 - kymin;
    if (j0 >= 0 && j0 < cy) Mb[(long)j0 * cx] = cscale(v[0], (tid & 1) ? -inv : inv);
    if (j1 >= 0 && j1 < cy) Mb[(long)j1 * cx] = cscale(v[15], (tid & 1) ? -inv : inv);
}

// fi_cols4096: P [B][G][nb] -> U [B][4096][nb] as fi_cols_kernel: the G-point
// forward transform of a column in LDS, the band |ky| <= band embedded in a
// 4096-point spectrum with the centring signs, the inverse on the register core.
// TR: U is written transposed, U [clip][nb][4096], each column contiguously.
template <int G, bool TR = false>
__global__ __launch_bounds__(256) void fi_cols4096_kernel(const float2* __restrict__ P,
                                                          float2* __restrict__ U, int band,
                                                          int yalt,
                                                          const float2* __restrict__ tw) {
    constexpr int N = 4096;
    __shared__ float2 sa[G], sb[G];
    __shared__ float2 buf[F4K_BUF];
    __shared__ float2 lo[64], hi[64];
    const int tid = threadIdx.x;
    const int nb = band + 1;
    const int lin = xcd_block_order();
    const int col = lin % nb;
    const long b = lin / nb;
    const float2* Pb = P + b * (long)G * nb + col;
    for (int g = tid; g < G; g += 256) sa[g] = Pb[(long)g * nb];
    f4k_tables(lo, hi, tid);
    const float2* r1;
    if constexpr ((G & (G - 1)) == 0) {
        __syncthreads();
        r1 = fft_lds<G, false>(sa, sb, 1, tw, tid, 256);
    } else {
        // G = Pg^2 (576 = 24^2): the register two-stage FFT by Pg lanes,
        // twiddles W_G^m in sb, the transpose in buf (free until the inverse)
        constexpr int Pg = G == 576 ? 24 : 0;
        static_assert(Pg * Pg == G, "square non-power-of-two grids: 576");
        using Geo = Fft2Geom<Pg, Pg>;
        fill_twiddles<G>(sb, tid, 256);
        __syncthreads();
        if (tid < Pg)
            fft2_stage1<Geo, Pg, false>(tid, buf, sb, [&](int n) { return sa[n]; });
        __syncthreads();
        if (tid < Pg) fft2_stage2<Geo, false>(tid, buf, [&](int, int g, float2 v) { sa[g] = v; });
        __syncthreads();
        r1 = sa;
    }
    const float inv = 1.0f / ((float)G * (float)G);
    float2 v[16];
#pragma unroll
    for (int d = 0; d < 16; d++) {
        const int y = tid + 256 * d;
        const int ky = y < N / 2 ? y : y - N;
        float2 x = make_float2(0.f, 0.f);
        if (ky >= -band && ky <= band)
            x = cscale(r1[((ky % G) + G) % G], (!yalt && (ky & 1)) ? -inv : inv);
        v[d] = x;
    }
    fft4096_pair<true>(tid, v, buf, lo, hi);
    // thread tid: rows y = tid + 256 d
    const float sg = (col & 1) ? -1.f : 1.f;
    if (TR) {
        float2* Uc = U + (b * nb + col) * (long)N + tid;
#pragma unroll
        for (int d = 0; d < 16; d++) Uc[256 * d] = cscale(v[d], sg);
    } else {
        float2* Uc = U + b * (long)N * nb + col;
#pragma unroll
        for (int d = 0; d < 16; d++) Uc[(long)(tid + 256 * d) * nb] = cscale(v[d], sg);
    }
}

// ===========================================================================
// host-side launchers with size dispatch
// ===========================================================================

template <int L>
static int launch_rows_dft_L(const float* in, float2* out, long nrows, int kmin, int count,
                             bool alt, const float2* tw, hipStream_t st) {
    constexpr int FPB = (2048 / L) >= 1 ? (2048 / L) : 1;
    unsigned grid = cdiv((size_t)(nrows / 2), FPB);
    if (alt)
        hipLaunchKernelGGL((rows_dft_kernel<L, FPB, true>), dim3(grid), dim3(256), 0, st, in, out,
                           (int)nrows, kmin, count, tw);
    else
        hipLaunchKernelGGL((rows_dft_kernel<L, FPB, false>), dim3(grid), dim3(256), 0, st, in,
                           out, (int)nrows, kmin, count, tw);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_rows_dft(int L, const float* in, float2* out, long nrows, int kmin, int count,
                    bool alt, const float2* tw, hipStream_t st) {
    if (nrows % 2 || count <= 0 || nrows > 0x7fffffffL) return LITHO_EARG;
    if (L == 4096 && kmin > -256 && kmin + count - 1 < 256 && nrows / 2 <= 0x7fffffffL) {
        // the register transform: a window of |kx| < 256 (every Nx of a 4096 clip)
        const dim3 grid((unsigned)(nrows / 2));
        if (alt)
            hipLaunchKernelGGL((rows4096_kernel<true>), grid, dim3(256), 0, st, in, out,
                               (int)nrows, kmin, count, 0);
        else
            hipLaunchKernelGGL((rows4096_kernel<false>), grid, dim3(256), 0, st, in, out,
                               (int)nrows, kmin, count, 0);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
#define LITHO_RF2(p, q, nth)                                                                  \
    if (L == p * q) {                                                                         \
        constexpr int FPB = nth / Fft2Geom<p, q>::T;                                          \
        unsigned grid = cdiv((size_t)(nrows / 2), FPB);                                       \
        if (alt)                                                                              \
            hipLaunchKernelGGL((rows_fwd2_kernel<p, q, nth, true, false>), dim3(grid), dim3(nth), 0, \
                               st, in, out, (int)nrows, kmin, count, tw);                     \
        else                                                                                  \
            hipLaunchKernelGGL((rows_fwd2_kernel<p, q, nth, false, false>), dim3(grid), dim3(nth), 0,\
                               st, in, out, (int)nrows, kmin, count, tw);                     \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                       \
    }
    LITHO_RF2(16, 16, 256)
    LITHO_RF2(16, 32, 256)
    LITHO_RF2(32, 32, 256)
    LITHO_RF2(24, 24, 192)
#undef LITHO_RF2
    switch (L) {
        case 32: return launch_rows_dft_L<32>(in, out, nrows, kmin, count, alt, tw, st);
        case 64: return launch_rows_dft_L<64>(in, out, nrows, kmin, count, alt, tw, st);
        case 128: return launch_rows_dft_L<128>(in, out, nrows, kmin, count, alt, tw, st);
        case 256: return launch_rows_dft_L<256>(in, out, nrows, kmin, count, alt, tw, st);
        case 512: return launch_rows_dft_L<512>(in, out, nrows, kmin, count, alt, tw, st);
        case 1024: return launch_rows_dft_L<1024>(in, out, nrows, kmin, count, alt, tw, st);
        case 2048: return launch_rows_dft_L<2048>(in, out, nrows, kmin, count, alt, tw, st);
        case 4096: return launch_rows_dft_L<4096>(in, out, nrows, kmin, count, alt, tw, st);
        default: return LITHO_EARG;
    }
}

template <int N>
static int launch_mask_cols_N(const float2* R, float2* M, int B, int cx, int kymin, int cy,
                              const float2* tw, hipStream_t st) {
    constexpr int CG = (2048 / N) >= 1 ? (2048 / N) : 1;
    dim3 grid(cdiv(cx, CG), B);
    hipLaunchKernelGGL((mask_cols_kernel<N, CG>), grid, dim3(256), 0, st, R, M, cx, kymin, cy, tw);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_mask_cols(int N, const float2* R, float2* M, int B, int cx, int kymin, int cy,
                     const float2* tw, hipStream_t st) {
    if (B <= 0 || B > 65535 || cx <= 0 || cy <= 0) return LITHO_EARG;
    if (N == 4096 && kymin > -256 && kymin + cy - 1 < 256) {
        hipLaunchKernelGGL((mask_cols4096_kernel<false>), dim3(cx, B), dim3(256), 0, st, R, M, cx,
                           kymin, cy);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
    switch (N) {
        case 32: return launch_mask_cols_N<32>(R, M, B, cx, kymin, cy, tw, st);
        case 64: return launch_mask_cols_N<64>(R, M, B, cx, kymin, cy, tw, st);
        case 128: return launch_mask_cols_N<128>(R, M, B, cx, kymin, cy, tw, st);
        case 256: return launch_mask_cols_N<256>(R, M, B, cx, kymin, cy, tw, st);
        case 512: return launch_mask_cols_N<512>(R, M, B, cx, kymin, cy, tw, st);
        case 1024: return launch_mask_cols_N<1024>(R, M, B, cx, kymin, cy, tw, st);
        case 2048: return launch_mask_cols_N<2048>(R, M, B, cx, kymin, cy, tw, st);
        case 4096: return launch_mask_cols_N<4096>(R, M, B, cx, kymin, cy, tw, st);
        default: return LITHO_EARG;
    }
}

template <int G, int GR>
static int launch_socs_GR(const float2* M, const float2* bank, const float* scales, int nk,
                          int Nx, int B, float* out, int shift, const float2* tw,
                          hipStream_t st) {
    const int W = 2 * Nx + 1;
    constexpr int NF = (2048 / G) < GR ? (2048 / G) : GR;
    size_t lds = (size_t)(2 * NF * G + GR * W) * sizeof(float2);
    if (lds > 160 * 1024) return LITHO_EARG;
    auto kern = socs_accum_kernel<G, GR>;
    LITHO_TRY(allow_full_lds((const void*)kern));
    unsigned grid = (unsigned)((size_t)B * (G / GR));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(512), lds, st, M, bank, scales, nk, Nx, out, shift,
                       tw);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int socs_rows_per_block(int G, int Nx) {
    const int W = 2 * Nx + 1;
    int cands[3];
    int nc = 0;
    switch (G) {
        case 32: cands[nc++] = 32; break;
        case 64: cands[nc++] = 64; break;
        case 128: cands[nc++] = 128; cands[nc++] = 64; break;
        case 256: cands[nc++] = 128; cands[nc++] = 64; cands[nc++] = 32; break;
        case 512: cands[nc++] = 64; cands[nc++] = 32; cands[nc++] = 16; break;
        case 1024: cands[nc++] = 32; cands[nc++] = 16; break;
        default: return 0;
    }
    for (int i = 0; i < nc; i++)
        if ((size_t)(2 * 2048 + cands[i] * W) * sizeof(float2) <= 160 * 1024) return cands[i];
    return 0;
}

typedef float f2nt __attribute__((ext_vector_type(2)));  // for 8-byte streaming stores

// Column pass of the large-grid SOCS: one transform per (image b, entry k,
// window column col): the inverse G-point FFT over ky of K_k M (W nonzero
// inputs at indices 0..W-1, as socs_accum2), written to Yt[b][k][col][0..G).
// A block takes one column group of one entry for CB consecutive images
// (grid z = ceil(B / CB)).  CB > 1: the lane's bank values stay in registers
// for all CB images, and image c + 1's window values are loaded while image
// c is transformed.
template <int P, int Q, int NZ, int NTH, int BK, int CB>
__global__ __launch_bounds__(NTH) void socs_cols_kernel(const float2* __restrict__ M,
                                                        const void* __restrict__ bank,
                                                        const float2* __restrict__ kscale, int nk,
                                                        int Nx, int B, float2* __restrict__ Yt) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int G = Geo::G;
    constexpr int T = Geo::T;
    constexpr int SLOTS = NTH / T;
    constexpr int S1 = Geo::S1;
    extern __shared__ float2 lds[];
    float2* twG = lds;
    float2* xbuf = lds + G;
    const int W = 2 * Nx + 1;
    const int tid = threadIdx.x;
    const int slot = tid / T, t = tid - slot * T;
    // neighbouring column groups (each half of a bank line) on one XCD
    const int lin = xcd_pair_order();
    const int cgrp = lin % gridDim.x, kb = lin / gridDim.x;
    const int col = cgrp * SLOTS + slot;
    const int k = kb % gridDim.y;
    const long b0 = (long)(kb / gridDim.y) * CB;
    const bool act = col < W;
    const long kbase = (long)k * W * W;
    float2* xb = xbuf + slot * Geo::XBUF;
    fill_twiddles<G>(twG, tid, NTH);
    if constexpr (CB == 1) {
        const float2* Mb = M + b0 * (long)W * W;
        __syncthreads();
        fft2_stage1<Geo, NZ, true>(t, xb, twG, [&](int n) {
            float2 v = make_float2(0.f, 0.f);
            if (act && n < W) {
                const int j = n * W + col;
                v = cmul(Bank<BK>::ld(bank, kscale, kbase + j, k), Mb[j]);
            }
            return v;
        });
        __syncthreads();
        float2* yo = Yt + ((b0 * nk + k) * (long)W + (act ? col : 0)) * G;
        fft2_stage2<Geo, true>(t, xb, [&](int, int g, float2 val) {
            if (act) yo[g] = val;
        });
    } else {
        // the lane's inputs: rows n = Q n1 + t + T s (n1 < NZ) of column col
        float2 kv[S1][NZ], mv[S1][NZ];
        auto load_m = [&](long b) {
            const float2* Mb = M + b * (long)W * W;
#pragma unroll
            for (int s = 0; s < S1; s++)
#pragma unroll
                for (int n1 = 0; n1 < NZ; n1++) {
                    const int n = Q * n1 + t + T * s;
                    mv[s][n1] = act && n < W ? Mb[n * W + col] : make_float2(0.f, 0.f);
                }
        };
#pragma unroll
        for (int s = 0; s < S1; s++)
#pragma unroll
            for (int n1 = 0; n1 < NZ; n1++) {
                const int n = Q * n1 + t + T * s;
                kv[s][n1] = act && n < W ? Bank<BK>::ld(bank, kscale, kbase + n * W + col, k)
                                         : make_float2(0.f, 0.f);
            }
        load_m(b0);
        __syncthreads();
        for (int c = 0; c < CB; c++) {
            const long b = b0 + c;
            if (b >= B) break;  // the same for the whole block
            float2 xin[S1][NZ];
#pragma unroll
            for (int s = 0; s < S1; s++)
#pragma unroll
                for (int n1 = 0; n1 < NZ; n1++) xin[s][n1] = cmul(kv[s][n1], mv[s][n1]);
            if (c + 1 < CB && b + 1 < B) load_m(b + 1);
            if (c > 0) __syncthreads();  // the previous image's stage 2 has read xb
            fft2_stage1_ij<Geo, NZ, true>(t, xb, twG, [&](int n1, int s) { return xin[s][n1]; });
            __syncthreads();
            float2* yo = Yt + ((b * nk + k) * (long)W + (act ? col : 0)) * G;
            // Yt is read once, by the row pass: streaming stores
            fft2_stage2<Geo, true>(t, xb, [&](int, int g, float2 val) {
                f2nt w;
                w.x = val.x;
                w.y = val.y;
                if (act) __builtin_nontemporal_store(w, (f2nt*)(yo + g));
            });
        }
    }
}

// images per block of the column pass: 8 where a lane's stage-1 inputs fit in
// registers three times over (bank, window, product), else 1
template <int NZ, int S1>
constexpr int cols_images_per_block() {
    return NZ * S1 <= 12 ? 8 : 1;
}

template <int P, int Q, int GR, int NZ, int NTH, int BK>
static int launch_socs2_pre(const float2* M, const void* bank, const float2* kscale,
                            const float* scales, int nk, int Nx, int B, float* out, int shift,
                            const float2* tw, float2* Yt, hipStream_t st) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int G = P * Q;
    constexpr int SLOTS = NTH / Geo::T;
    // column pass: as many transforms as fit 256 threads (10 at T = 24: two
    // 4-wave blocks fill a CU's 8 wavefronts at its ~200 VGPRs, where 3-wave
    // blocks left two idle; 384 -> 355 us per 32 clips of 4096^2)
    constexpr int NTHC = (256 / Geo::T) * Geo::T;
    constexpr int SLOTSC = NTHC / Geo::T;
    const int W = 2 * Nx + 1;
    const size_t lds_c = (size_t)(G + SLOTSC * Geo::XBUF) * sizeof(float2);
    const size_t lds_r = (size_t)(G + SLOTS * Geo::XBUF) * sizeof(float2);  // rows read in place
    if (lds_c > 160 * 1024 || lds_r > 160 * 1024) return LITHO_EARG;
    constexpr int CB = cols_images_per_block<NZ, Geo::S1>();
    auto kc = socs_cols_kernel<P, Q, NZ, NTHC, BK, CB>;
    auto kr = socs_accum2_kernel<P, Q, GR, NZ, NTH, BK, true>;
    LITHO_TRY(allow_full_lds((const void*)kc));
    LITHO_TRY(allow_full_lds((const void*)kr));
    hipLaunchKernelGGL(kc, dim3(cdiv(W, SLOTSC), nk, cdiv(B, CB)), dim3(NTHC), lds_c, st, M, bank,
                       kscale, nk, Nx, B, Yt);
    hipLaunchKernelGGL(kr, dim3((unsigned)((size_t)B * (G / GR))), dim3(NTH), lds_r, st, M, bank,
                       kscale, scales, nk, Nx, out, shift, tw, (const float2*)Yt);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

template <int P, int Q, int GR, int NZ, int NTH, int BK>
static int launch_socs2_bk(const float2* M, const void* bank, const float2* kscale,
                           const float* scales, int nk, int Nx, int B, float* out, int shift,
                           const float2* tw, hipStream_t st) {
    using Geo = Fft2Geom<P, Q>;
    constexpr int G = P * Q;
    constexpr int SLOTS = NTH / Geo::T;
    const int W = 2 * Nx + 1;
    size_t lds = (size_t)(G + SLOTS * Geo::XBUF + GR * W) * sizeof(float2);
    if (lds > 160 * 1024) return LITHO_EARG;
    auto kern = socs_accum2_kernel<P, Q, GR, NZ, NTH, BK>;
    LITHO_TRY(allow_full_lds((const void*)kern));
    unsigned grid = (unsigned)((size_t)B * (G / GR));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NTH), lds, st, M, bank, kscale, scales, nk, Nx, out,
                       shift, tw, (const float2*)nullptr);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

// the 16-bit bank formats are built for the large grids (clips above 1024^2),
// the kernels as given everywhere
template <int P, int Q, int GR, int NZ, int NTH>
static int launch_socs2(const float2* M, const void* bank, const float2* kscale, int bank_fmt,
                        const float* scales, int nk, int Nx, int B, float* out, int shift,
                        const float2* tw, float2* Yt, hipStream_t st) {
    if constexpr (P * Q >= 512) {
        if (Yt) {
            switch (bank_fmt) {
                case 0:
                    return launch_socs2_pre<P, Q, GR, NZ, NTH, 0>(M, bank, kscale, scales, nk, Nx,
                                                                  B, out, shift, tw, Yt, st);
                case 1:
                    return launch_socs2_pre<P, Q, GR, NZ, NTH, 1>(M, bank, kscale, scales, nk, Nx,
                                                                  B, out, shift, tw, Yt, st);
                case 2:
                    return launch_socs2_pre<P, Q, GR, NZ, NTH, 2>(M, bank, kscale, scales, nk, Nx,
                                                                  B, out, shift, tw, Yt, st);
                case 3:
                    return launch_socs2_pre<P, Q, GR, NZ, NTH, 3>(M, bank, kscale, scales, nk, Nx,
                                                                  B, out, shift, tw, Yt, st);
                default:
                    return LITHO_EARG;
            }
        }
    }
    if (bank_fmt == 0)
        return launch_socs2_bk<P, Q, GR, NZ, NTH, 0>(M, bank, kscale, scales, nk, Nx, B, out,
                                                     shift, tw, st);
    if constexpr (P * Q >= 512) {
        if (bank_fmt == 1)
            return launch_socs2_bk<P, Q, GR, NZ, NTH, 1>(M, bank, kscale, scales, nk, Nx, B, out,
                                                         shift, tw, st);
        if (bank_fmt == 2)
            return launch_socs2_bk<P, Q, GR, NZ, NTH, 2>(M, bank, kscale, scales, nk, Nx, B, out,
                                                         shift, tw, st);
        if (bank_fmt == 3)
            return launch_socs2_bk<P, Q, GR, NZ, NTH, 3>(M, bank, kscale, scales, nk, Nx, B, out,
                                                         shift, tw, st);
    }
    return LITHO_EARG;
}

size_t socs_pre_bytes(int G, int Nx, int nk, int B) {
    return G >= 512 ? (size_t)B * nk * (2 * Nx + 1) * G * sizeof(float2) : 0;
}

int launch_socs_accum(int G, const float2* M, const void* bank, const float2* kscale,
                      int bank_fmt, const float* scales, int nk, int Nx, int B, float* out,
                      int shift, const float2* tw, hipStream_t st, float2* Yt) {
    if (B <= 0 || nk < 0 || 4 * Nx + 1 > G || bank_fmt < 0 || bank_fmt > 3 ||
        (bank_fmt && !kscale))
        return LITHO_EARG;
    const int W = 2 * Nx + 1;
    // register-FFT variant: (P, Q, rows per block, nonzero stage-1 rows, threads)
#define LITHO_S2(p, q, gr, nz, nth)                                                        \
    if (G == p * q && W <= nz * q)                                                         \
        return launch_socs2<p, q, gr, nz, nth>(M, bank, kscale, bank_fmt, scales, nk, Nx, B, out, \
                                               shift, tw, G >= 512 ? Yt : nullptr, st);
    LITHO_S2(4, 4, 16, 4, 64)
    LITHO_S2(4, 8, 32, 4, 128)
    LITHO_S2(8, 8, 64, 3, 512)
    LITHO_S2(8, 8, 64, 8, 512)
    LITHO_S2(8, 16, 128, 3, 512)
    LITHO_S2(8, 16, 128, 4, 512)
    LITHO_S2(8, 16, 128, 8, 512)
    LITHO_S2(16, 16, 128, 5, 512)
    LITHO_S2(16, 16, 128, 8, 512)
    LITHO_S2(16, 16, 64, 16, 512)
    LITHO_S2(16, 32, 32, 5, 256)
    LITHO_S2(16, 32, 32, 16, 256)
    LITHO_S2(32, 32, 16, 9, 256)
    LITHO_S2(32, 32, 16, 32, 256)
    LITHO_S2(24, 24, 8, 12, 192)
#undef LITHO_S2
    int GR = socs_rows_per_block(G, Nx);
    if (!GR || bank_fmt) return LITHO_EARG;
    const float2* bank2 = (const float2*)bank;
#define LITHO_SOCS_CASE(g, gr) \
    if (G == g && GR == gr) return launch_socs_GR<g, gr>(M, bank2, scales, nk, Nx, B, out, shift, tw, st);
    LITHO_SOCS_CASE(32, 32)
    LITHO_SOCS_CASE(64, 64)
    LITHO_SOCS_CASE(128, 128)
    LITHO_SOCS_CASE(128, 64)
    LITHO_SOCS_CASE(256, 128)
    LITHO_SOCS_CASE(256, 64)
    LITHO_SOCS_CASE(256, 32)
    LITHO_SOCS_CASE(512, 64)
    LITHO_SOCS_CASE(512, 32)
    LITHO_SOCS_CASE(512, 16)
    LITHO_SOCS_CASE(1024, 32)
    LITHO_SOCS_CASE(1024, 16)
#undef LITHO_SOCS_CASE
    return LITHO_EARG;
}

template <int G, int N>
static int launch_fi_cols_GN(const float2* P, float2* U, int B, int band, int yalt,
                             const float2* tw, hipStream_t st) {
    constexpr int CG = (2048 / N) >= 1 ? (2048 / N) : 1;
    dim3 grid(cdiv(band + 1, CG), B);
    hipLaunchKernelGGL((fi_cols_kernel<G, N, CG>), grid, dim3(256), 0, st, P, U, band, yalt, tw);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

template <int G>
static int launch_fi_cols_G(int N, const float2* P, float2* U, int B, int band, int yalt,
                            const float2* tw, hipStream_t st) {
    switch (N) {
        case 32: if constexpr (G <= 32) return launch_fi_cols_GN<G, 32>(P, U, B, band, yalt, tw, st); break;
        case 64: if constexpr (G <= 64) return launch_fi_cols_GN<G, 64>(P, U, B, band, yalt, tw, st); break;
        case 128: if constexpr (G <= 128) return launch_fi_cols_GN<G, 128>(P, U, B, band, yalt, tw, st); break;
        case 256: if constexpr (G <= 256) return launch_fi_cols_GN<G, 256>(P, U, B, band, yalt, tw, st); break;
        case 512: if constexpr (G <= 512) return launch_fi_cols_GN<G, 512>(P, U, B, band, yalt, tw, st); break;
        case 1024: if constexpr (G <= 1024) return launch_fi_cols_GN<G, 1024>(P, U, B, band, yalt, tw, st); break;
        case 2048: return launch_fi_cols_GN<G, 2048>(P, U, B, band, yalt, tw, st);
        case 4096:
            // the 4096-point inverse on the register core, one column per block
            hipLaunchKernelGGL((fi_cols4096_kernel<G>), dim3(band + 1, B), dim3(256), 0, st, P, U,
                               band, yalt, tw);
            return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
        default: break;
    }
    return LITHO_EARG;
}

int launch_fi_cols(int G, int N, const float2* P, float2* U, int B, int band, int yalt,
                   const float2* tw, hipStream_t st) {
    if (B <= 0 || B > 65535 || band < 0 || 2 * band + 1 > G || G > N) return LITHO_EARG;
    switch (G) {
        case 32: return launch_fi_cols_G<32>(N, P, U, B, band, yalt, tw, st);
        case 64: return launch_fi_cols_G<64>(N, P, U, B, band, yalt, tw, st);
        case 128: return launch_fi_cols_G<128>(N, P, U, B, band, yalt, tw, st);
        case 256: return launch_fi_cols_G<256>(N, P, U, B, band, yalt, tw, st);
        case 512: return launch_fi_cols_G<512>(N, P, U, B, band, yalt, tw, st);
        case 1024: return launch_fi_cols_G<1024>(N, P, U, B, band, yalt, tw, st);
        case 576:
            if (N != 4096) return LITHO_EARG;
            hipLaunchKernelGGL((fi_cols4096_kernel<576>), dim3(band + 1, B), dim3(256), 0, st, P,
                               U, band, yalt, tw);
            return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
        default: return LITHO_EARG;
    }
}

template <int N>
static int launch_image_rows_N(const float2* U, float* img, long nrows, int band,
                               const float2* tw, hipStream_t st) {
    constexpr int FPB = (2048 / N) >= 1 ? (2048 / N) : 1;
    unsigned grid = cdiv((size_t)(nrows / 2), FPB);
    hipLaunchKernelGGL((image_rows_kernel<N, FPB>), dim3(grid), dim3(256), 0, st, U, img,
                       (int)nrows, band, tw);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_image_rows(int N, const float2* U, float* img, long nrows, int band, const float2* tw,
                      hipStream_t st) {
    if (nrows % 2 || 2 * band + 1 > N || nrows > 0x7fffffffL) return LITHO_EARG;
    if (N == 4096) {
        hipLaunchKernelGGL(image4096_kernel, dim3((unsigned)(nrows / 2)), dim3(256), 0, st, U,
                           img, (int)nrows, band);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
#define LITHO_IR2(p, q, nz, nth)                                                          \
    if (N == p * q && 2 * band + 1 <= nz * q) {                                           \
        constexpr int FPB = nth / Fft2Geom<p, q>::T;                                      \
        unsigned grid = cdiv((size_t)(nrows / 2), FPB);                                   \
        hipLaunchKernelGGL((image_rows2_kernel<p, q, nz, nth, false>), dim3(grid), dim3(nth), 0, \
                           st, U, img, (int)nrows, band, tw);                             \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                   \
    }
    LITHO_IR2(16, 16, 5, 256)
    LITHO_IR2(16, 16, 16, 256)
    LITHO_IR2(16, 32, 3, 256)
    LITHO_IR2(16, 32, 5, 256)
    LITHO_IR2(16, 32, 16, 256)
    LITHO_IR2(32, 32, 3, 256)
    LITHO_IR2(32, 32, 9, 256)
    LITHO_IR2(32, 32, 32, 256)
#undef LITHO_IR2
    switch (N) {
        case 32: return launch_image_rows_N<32>(U, img, nrows, band, tw, st);
        case 64: return launch_image_rows_N<64>(U, img, nrows, band, tw, st);
        case 128: return launch_image_rows_N<128>(U, img, nrows, band, tw, st);
        case 256: return launch_image_rows_N<256>(U, img, nrows, band, tw, st);
        case 512: return launch_image_rows_N<512>(U, img, nrows, band, tw, st);
        case 1024: return launch_image_rows_N<1024>(U, img, nrows, band, tw, st);
        case 2048: return launch_image_rows_N<2048>(U, img, nrows, band, tw, st);
        case 4096: return launch_image_rows_N<4096>(U, img, nrows, band, tw, st);
        default: return LITHO_EARG;
    }
}


// The 4096^2 image path with R and U stored transposed per clip ([kx][y]):
// the column transforms read and write whole columns contiguously, the row
// transforms move one 16-byte pair per kept frequency (row pairs in XCD
// order).  Mask rows: window |kx| < 256 (kmin = -Nx, count = 2 Nx + 1).
int launch_rows4096_tr(const float* in, float2* out, int B, int kmin, int count, bool alt,
                       hipStream_t st) {
    if (B <= 0 || count <= 0 || kmin <= -256 || kmin + count - 1 >= 256) return LITHO_EARG;
    const dim3 grid((unsigned)((long)B * 2048));  // row pairs
    if (alt)
        hipLaunchKernelGGL((rows4096_kernel<true, true>), grid, dim3(256), 0, st, in, out,
                           B * 4096, kmin, count, 4096);
    else
        hipLaunchKernelGGL((rows4096_kernel<false, true>), grid, dim3(256), 0, st, in, out,
                           B * 4096, kmin, count, 4096);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_mask_cols4096_tr(const float2* R, float2* M, int B, int cx, int kymin, int cy,
                            hipStream_t st) {
    if (B <= 0 || B > 65535 || cx <= 0 || cy <= 0 || kymin <= -256 || kymin + cy - 1 >= 256)
        return LITHO_EARG;
    hipLaunchKernelGGL((mask_cols4096_kernel<true>), dim3(cx, B), dim3(256), 0, st, R, M, cx,
                       kymin, cy);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_fi_cols4096_tr(int G, const float2* P, float2* U, int B, int band, int yalt,
                          const float2* tw, hipStream_t st) {
    if (B <= 0 || B > 65535 || band < 0 || 2 * band + 1 > G) return LITHO_EARG;
    const dim3 grid(band + 1, B);
#define LITHO_FI4K(g)                                                                          \
    case g:                                                                                    \
        hipLaunchKernelGGL((fi_cols4096_kernel<g, true>), grid, dim3(256), 0, st, P, U, band,  \
                           yalt, tw);                                                          \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    switch (G) {
        LITHO_FI4K(32)
        LITHO_FI4K(64)
        LITHO_FI4K(128)
        LITHO_FI4K(256)
        LITHO_FI4K(512)
        LITHO_FI4K(576)
        LITHO_FI4K(1024)
        default: return LITHO_EARG;
    }
#undef LITHO_FI4K
}

template <int NP, int TPB>
static int launch_image4096_tiles(const float2* U, float* img, int B, int band, hipStream_t st) {
    // static (transpose plane, tables) + dynamic LDS within the default 64 KB
    // per block, so the kernel needs no raised limit (hipFuncSetAttribute
    // refuses a 160 KB dynamic limit beside static LDS)
    const size_t lds = (size_t)(band + 1) * (2 * NP + 1) * sizeof(float2);
    if (lds + (F4K_BUF + 128) * sizeof(float2) > 64 * 1024) return LITHO_EARG;
    auto kern = image4096_tiles_kernel<NP, TPB>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((long)B * (4096 / (2 * NP * TPB)))), dim3(256), lds,
                       st, U, img, band);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

// tiles of two row pairs, eight tiles per block (506 us per 32 clips at K =
// 64; one tile per block 538, four pairs per tile 550, eight 672; reading one
// row pair's values per block, one 16-byte piece from each of 275 lines, 641)
int launch_image4096_tr(const float2* U, float* img, int B, int band, hipStream_t st) {
    if (B <= 0 || band < 0 || band >= 512) return LITHO_EARG;
    return launch_image4096_tiles<2, 8>(U, img, B, band, st);
}

// ===========================================================================
// band pipeline launchers
// ===========================================================================
int band_grid(int Nx) {
    const int need = 4 * Nx + 1;
    const int cands[] = {36, 81, 144};
    for (int g : cands)
        if (g >= need) return g;
    return 0;
}

// ---------------------------------------------------------------------------
// socs_wave: socs_band with every transform inside one wavefront.  A P-point
// transform uses P lanes; a wave holds TPW = 64 / P of them (60 of 64 lanes
// at P = 12) and NW waves make the block.  The LDS transposes inside each
// two-stage FFT then need only a wavefront fence, so a kernel costs two block
// barriers (around the column -> row hand-off of Y) instead of two per round.
// Lanes past TPW * P sit out every FFT (they must not touch the slot buffers).
// What a lane needs in every kernel stays in registers: its inter-stage
// twiddles W_G^(t k1) (lane t is stage-1 index n2 = t), and its mask-spectrum
// inputs of the column pass; the next kernel's bank words are fetched while
// the current kernel runs.
// ---------------------------------------------------------------------------

// Stage 1 of a P x P two-stage FFT for the lane's n2 = t, inputs v[n1] in
// registers (zeros past NZ fold away), twiddles twr[k1] = W_G^(t k1) forward.
template <int P, bool INV>
__device__ __forceinline__ void stage1_regs(int t, float2 (&v)[P], const float2 (&twr)[P],
                                            float2* __restrict__ xb) {
    dft_reg<P, INV>(v);
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) {
        float2 w = twr[k1];
        if (INV) w.y = -w.y;
        xb[k1 * (P + 1) + t] = k1 == 0 ? v[0] : cmul_w<false>(v[k1], w);
    }
}

// ksplit > 1: the bank entries of an image are split over ksplit blocks
// (block b ksplit + part takes entries [part nk / ksplit, (part + 1) nk /
// ksplit)), each writing the band spectrum of its share to plane part of Cout
// (planes cplane float2 apart); the FI column pass adds the planes in order.
// For launches of fewer images than CUs: twice the blocks, each half the work.
template <int P, int NZ, int NW, int BK = 0>
__global__ __launch_bounds__(64 * NW) void socs_wave_kernel(const float2* __restrict__ M,
                                                            const void* __restrict__ bank,
                                                            const float2* __restrict__ kscale,
                                                            const float* __restrict__ scales,
                                                            int nk, int Nx,
                                                            float2* __restrict__ Cout, int ksplit,
                                                            long cplane) {
    using Geo = Fft2Geom<P, P>;
    static_assert(Geo::LDQ == P + 1 && Geo::S1 == 1 && Geo::S2 == 1, "square two-stage FFT");
    constexpr int G = Geo::G;
    constexpr int TPW = 64 / P;
    constexpr int SL = TPW * NW;
    constexpr int RB = (G + SL - 1) / SL;
    constexpr int WMAX = 2 * ((G - 1) / 4) + 1;  // widest band of a G grid (4 Nx + 1 <= G)
    constexpr int CR = (WMAX + SL - 1) / SL;     // column rounds at most
    using BR = typename Bank<BK>::R;
    extern __shared__ float2 lds[];
    float2* twG = lds;
    float2* xbuf = lds + G;
    float2* Y = xbuf + SL * Geo::XBUF;  // [G][W]
    const int W = 2 * Nx + 1;
    const int band = 2 * Nx;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int j = lane / P, t = lane - j * P;
    const bool on = j < TPW;
    const int slot = wv * TPW + (on ? j : 0);
    const long b = blockIdx.x / ksplit;
    const int part = (int)(blockIdx.x - b * ksplit);
    const int k0 = part * nk / ksplit, k1 = (part + 1) * nk / ksplit;
    const float2* Mb = M + b * (long)W * W;
    float2* xb = xbuf + slot * Geo::XBUF;
    fill_twiddles<G>(twG, tid, 64 * NW);
    float acc[RB * P];
#pragma unroll
    for (int i = 0; i < RB * P; i++) acc[i] = 0.f;
    // column-pass inputs of this lane: round cr, column cr * SL + slot, rows
    // n = P n1 + t (n1 < NZ); zero where the column or row is outside the band
    float2 mreg[CR][NZ];
    BR braw[CR][NZ];
#pragma unroll
    for (int cr = 0; cr < CR; cr++)
#pragma unroll
        for (int n1 = 0; n1 < NZ; n1++) {
            const int col = cr * SL + slot, n = P * n1 + t;
            const bool ok = on && col < W && n < W;
            mreg[cr][n1] = ok ? Mb[n * W + col] : make_float2(0.f, 0.f);
            braw[cr][n1] = BR{};
        }
    auto load_bank = [&](int k) {
        const long kbase = (long)k * W * W;
#pragma unroll
        for (int cr = 0; cr < CR; cr++)
#pragma unroll
            for (int n1 = 0; n1 < NZ; n1++) {
                const int col = cr * SL + slot, n = P * n1 + t;
                if (on && col < W && n < W) braw[cr][n1] = Bank<BK>::raw(bank, kbase + n * W + col);
            }
    };
    if (k1 > k0) load_bank(k0);
    __syncthreads();
    float2 twr[P];
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) twr[k1] = twG[t * k1];  // t, k1 < P: index < G

    for (int k = k0; k < k1; k++) {
        const float lam = scales[k];
        const float2 ksc = BK ? kscale[k] : make_float2(1.f, 1.f);
        float2 xin[CR][NZ];
#pragma unroll
        for (int cr = 0; cr < CR; cr++)
#pragma unroll
            for (int n1 = 0; n1 < NZ; n1++)
                xin[cr][n1] = cmul(Bank<BK>::dec(braw[cr][n1], ksc), mreg[cr][n1]);
        if (k + 1 < k1) load_bank(k + 1);
#pragma unroll
        for (int cr = 0; cr < CR; cr++) {
            if (cr * SL >= W) break;
            const int col = cr * SL + slot;
            const bool act = on && col < W;
            if (on) {
                float2 v[P];
#pragma unroll
                for (int n1 = 0; n1 < P; n1++) v[n1] = n1 < NZ ? xin[cr][n1] : make_float2(0.f, 0.f);
                stage1_regs<P, true>(t, v, twr, xb);
            }
            wave_sync();
            if (act)
                fft2_stage2<Geo, true>(t, xb, [&](int, int g, float2 val) { Y[g * W + col] = val; });
            wave_sync();
        }
        __syncthreads();
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int r = rb * SL + slot;
            const bool act = on && r < G;
            const float2* yr = Y + (act ? r : 0) * W;
            if (on) {
                float2 v[P];
#pragma unroll
                for (int n1 = 0; n1 < P; n1++) {
                    const int n = P * n1 + t;
                    v[n1] = (n1 < NZ && act && n < W) ? yr[n] : make_float2(0.f, 0.f);
                }
                stage1_regs<P, true>(t, v, twr, xb);
            }
            wave_sync();
            if (act)
                fft2_stage2<Geo, true>(t, xb, [&](int sidx, int, float2 val) {
                    acc[rb * P + sidx] += lam * (val.x * val.x + val.y * val.y);
                });
            wave_sync();
        }
        __syncthreads();
    }
    // forward R2C of the intensity rows straight from registers
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int r = rb * SL + slot;
        const bool act = on && r < G;
        if (on) {
            float2 v[P];
#pragma unroll
            for (int n1 = 0; n1 < P; n1++) v[n1] = make_float2(acc[rb * P + n1], 0.f);
            stage1_regs<P, false>(t, v, twr, xb);
        }
        wave_sync();
        if (act)
            fft2_stage2<Geo, false>(t, xb, [&](int, int kx, float2 val) {
                if (kx <= band) Y[r * W + kx] = val;
            });
        wave_sync();
    }
    __syncthreads();
    // columns kx = 0..band: forward FFT over the rows, keep |ky| <= band
    const float inv = 1.0f / ((float)G * (float)G);
    float2* Cb = Cout + part * cplane + b * (long)(2 * band + 1) * W;
    for (int c0 = 0; c0 < W; c0 += SL) {
        const int col = c0 + slot;
        const bool act = on && col < W;
        if (on) {
            float2 v[P];
#pragma unroll
            for (int n1 = 0; n1 < P; n1++)
                v[n1] = act ? Y[(P * n1 + t) * W + col] : make_float2(0.f, 0.f);
            stage1_regs<P, false>(t, v, twr, xb);
        }
        wave_sync();
        if (act)
            fft2_stage2<Geo, false>(t, xb, [&](int, int kk, float2 val) {
                const int ky = kk <= G / 2 ? kk : kk - G;
                if (ky >= -band && ky <= band) Cb[(long)col * (2 * band + 1) + ky + band] = cscale(val, inv);
            });
        wave_sync();
    }
}

template <int P, int NZ, int NW, int BK>
static int launch_wave(const float2* M, const void* bank, const float2* kscale, const float* scales,
                       int nk, int Nx, int B, float2* C, hipStream_t st, int ksplit, long cplane) {
    using Geo = Fft2Geom<P, P>;
    constexpr int SL = (64 / P) * NW;
    const int W = 2 * Nx + 1;
    size_t lds = (size_t)(Geo::G + SL * Geo::XBUF + Geo::G * W) * sizeof(float2);
    if (lds > 160 * 1024) return LITHO_EARG;
    auto kern = socs_wave_kernel<P, NZ, NW, BK>;
    LITHO_TRY(allow_full_lds((const void*)kern));
    hipLaunchKernelGGL(kern, dim3((unsigned)((long)B * ksplit)), dim3(64 * NW), lds, st, M, bank,
                       kscale, scales, nk, Nx, C, ksplit, cplane);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

template <int P, int NZ, int SLOTS, int BK>
static int launch_band(const float2* M, const void* bank, const float2* kscale, const float* scales,
                       int nk, int Nx, int B, float2* C, hipStream_t st) {
    using Geo = Fft2Geom<P, P>;
    const int W = 2 * Nx + 1;
    size_t lds = (size_t)(Geo::G + SLOTS * Geo::XBUF + Geo::G * W) * sizeof(float2);
    if (lds > 160 * 1024) return LITHO_EARG;
    auto kern = socs_band_kernel<P, NZ, SLOTS, BK>;
    LITHO_TRY(allow_full_lds((const void*)kern));
    hipLaunchKernelGGL(kern, dim3(B), dim3(P * SLOTS), lds, st, M, bank, kscale, scales, nk, Nx, C);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_socs_band(int G, const float2* M, const void* bank, const float2* kscale, int bank_fmt,
                     const float* scales, int nk, int Nx, int B, float2* C, int shape,
                     hipStream_t st, int ksplit, long cplane) {
    const int W = 2 * Nx + 1;
    if (B <= 0 || nk < 0 || 4 * Nx + 1 > G || bank_fmt < 0 || bank_fmt > 3 || shape < 0 ||
        shape > 2 || ksplit < 1 || ksplit > 2 || (ksplit > 1 && (shape != 2 || cplane <= 0)) ||
        (long)B * ksplit > 0x7fffffffL)
        return LITHO_EARG;
#define LITHO_BAND(p, nz, slots)                                                                  \
    if (G == p * p && W <= nz * p) {                                                              \
        if (bank_fmt == 3) return LITHO_EARG; /* hi + lo pairs: wave-synced blocks only */       \
        if (bank_fmt == 1)                                                                        \
            return launch_band<p, nz, slots, 1>(M, bank, kscale, scales, nk, Nx, B, C, st);       \
        if (bank_fmt == 2)                                                                        \
            return launch_band<p, nz, slots, 2>(M, bank, kscale, scales, nk, Nx, B, C, st);       \
        return launch_band<p, nz, slots, 0>(M, bank, kscale, scales, nk, Nx, B, C, st);           \
    }
#define LITHO_WAVE(p, nz, nw)                                                                     \
    if (G == p * p && W <= nz * p) {                                                              \
        if (bank_fmt == 1)                                                                        \
            return launch_wave<p, nz, nw, 1>(M, bank, kscale, scales, nk, Nx, B, C, st, ksplit,   \
                                             cplane);                                             \
        if (bank_fmt == 2)                                                                        \
            return launch_wave<p, nz, nw, 2>(M, bank, kscale, scales, nk, Nx, B, C, st, ksplit,   \
                                             cplane);                                             \
        if (bank_fmt == 3)                                                                        \
            return launch_wave<p, nz, nw, 3>(M, bank, kscale, scales, nk, Nx, B, C, st, ksplit,   \
                                             cplane);                                             \
        return launch_wave<p, nz, nw, 0>(M, bank, kscale, scales, nk, Nx, B, C, st, ksplit,      \
                                         cplane);                                                 \
    }
    if (shape == 2) {
        LITHO_WAVE(6, 6, 4)
        LITHO_WAVE(9, 4, 8)
        LITHO_WAVE(9, 9, 8)
        LITHO_WAVE(12, 6, 8)
        return LITHO_EARG;
    }
    if (shape == 1) {
        LITHO_BAND(12, 6, 24)
    }
    LITHO_BAND(6, 6, 36)
    LITHO_BAND(9, 4, 27)
    LITHO_BAND(9, 9, 27)
    LITHO_BAND(12, 6, 48)
    LITHO_BAND(12, 12, 48)
#undef LITHO_BAND
#undef LITHO_WAVE
    return LITHO_EARG;
}

// 16 x Q band row kernels: every plan uses them (the same bits in every plan)
template <int Q, int NTH>
static int launch_rows_band16(const void* in, void* out, long nrows, int count, bool h16, bool u8,
                              hipStream_t st) {
    constexpr int FPB = NTH / 16;
    const unsigned grid = cdiv((size_t)(nrows / 2), FPB);
    if (h16 && u8)
        hipLaunchKernelGGL((rows_band16_kernel<Q, true, true, NTH>), dim3(grid), dim3(NTH), 0, st,
                           in, out, (int)nrows, count);
    else if (h16)
        hipLaunchKernelGGL((rows_band16_kernel<Q, true, false, NTH>), dim3(grid), dim3(NTH), 0, st,
                           in, out, (int)nrows, count);
    else if (u8)
        hipLaunchKernelGGL((rows_band16_kernel<Q, false, true, NTH>), dim3(grid), dim3(NTH), 0, st,
                           in, out, (int)nrows, count);
    else
        hipLaunchKernelGGL((rows_band16_kernel<Q, false, false, NTH>), dim3(grid), dim3(NTH), 0,
                           st, in, out, (int)nrows, count);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

template <int Q, int NTH>
static int launch_image_band16(const void* U, void* img, long nrows, int band, bool h16,
                               bool out16, hipStream_t st) {
    constexpr int FPB = NTH / 16;
    const unsigned grid = cdiv((size_t)(nrows / 2), FPB);
    if (h16 && out16)
        hipLaunchKernelGGL((image_band16_kernel<Q, 3, true, true, NTH>), dim3(grid), dim3(NTH), 0,
                           st, U, img, (int)nrows, band);
    else if (h16)
        hipLaunchKernelGGL((image_band16_kernel<Q, 3, true, false, NTH>), dim3(grid), dim3(NTH), 0,
                           st, U, img, (int)nrows, band);
    else if (out16)
        hipLaunchKernelGGL((image_band16_kernel<Q, 3, false, true, NTH>), dim3(grid), dim3(NTH), 0,
                           st, U, img, (int)nrows, band);
    else
        hipLaunchKernelGGL((image_band16_kernel<Q, 3, false, false, NTH>), dim3(grid), dim3(NTH),
                           0, st, U, img, (int)nrows, band);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_rows_dft_t(int L, const void* in, void* out, long nrows, int kmin, int count, bool h16,
                      bool lean, bool u8, hipStream_t st) {
    if (nrows % L || nrows % 2 || count <= 0) return LITHO_EARG;
#define LITHO_RT1(p, q, nth, hh, uu)                                                            \
    hipLaunchKernelGGL((rows_fwd2_kernel<p, q, nth, true, true, hh, uu>), dim3(grid), dim3(nth), \
                       0, st, in, out, (int)nrows, kmin, count, nullptr)
#define LITHO_RT(p, q, nth)                                                                   \
    if (L == p * q) {                                                                         \
        constexpr int FPB = nth / Fft2Geom<p, q>::T;                                          \
        unsigned grid = cdiv((size_t)(nrows / 2), FPB);                                       \
        if (h16 && u8) LITHO_RT1(p, q, nth, true, true);                                      \
        else if (h16) LITHO_RT1(p, q, nth, true, false);                                      \
        else if (u8) LITHO_RT1(p, q, nth, false, true);                                       \
        else LITHO_RT1(p, q, nth, false, false);                                              \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                       \
    }
    // the band window of 1024-point rows: one kernel for every plan (the same
    // bits); 512-thread blocks share one twiddle table, so two blocks (76 KB of
    // LDS each) keep 16 waves per CU where 256-thread blocks fit 12
    if (L == 1024 && kmin == 0 && count <= 64) {
        unsigned grid = cdiv((size_t)(nrows / 2), 16);
        if (h16 && u8)
            hipLaunchKernelGGL((rows_band1024_kernel<true, true, 512>), dim3(grid), dim3(512), 0,
                               st, in, out, (int)nrows, count);
        else if (h16)
            hipLaunchKernelGGL((rows_band1024_kernel<true, false, 512>), dim3(grid), dim3(512), 0,
                               st, in, out, (int)nrows, count);
        else if (u8)
            hipLaunchKernelGGL((rows_band1024_kernel<false, true, 512>), dim3(grid), dim3(512), 0,
                               st, in, out, (int)nrows, count);
        else
            hipLaunchKernelGGL((rows_band1024_kernel<false, false, 512>), dim3(grid), dim3(512), 0,
                               st, in, out, (int)nrows, count);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
    if (kmin == 0 && count <= 32) {
        if (L == 256) return launch_rows_band16<16, 256>(in, out, nrows, count, h16, u8, st);
        if (L == 512) return launch_rows_band16<32, 512>(in, out, nrows, count, h16, u8, st);
        if (L == 768) return launch_rows_band16<48, 128>(in, out, nrows, count, h16, u8, st);
    }
    LITHO_RT(16, 16, 256)
    LITHO_RT(16, 32, 256)
    LITHO_RT(16, 48, 128)
    if (lean) {
        LITHO_RT(32, 32, 128)
    }
    LITHO_RT(32, 32, 256)
#undef LITHO_RT
#undef LITHO_RT1
    return LITHO_EARG;
}

int launch_mask_cols_t(int N, const void* R, float2* M, int B, int cx, int kymin, int cy, bool h16,
                       bool lean, hipStream_t st, int mw) {
    if (B <= 0 || B > 65535) return LITHO_EARG;
    if (mw && (mw % 2 == 0 || cx != mw / 2 + 1 || kymin != -(mw / 2) || cy != mw))
        return LITHO_EARG;
    const bool win16 = kymin > -32 && kymin + cy - 1 < 32;  // |ky| < 32
#define LITHO_MC16(q)                                                                        \
    if (N == 16 * q && win16) {                                                              \
        if (h16)                                                                             \
            hipLaunchKernelGGL((mask_cols16_kernel<q, true, 128>), dim3(cdiv(cx, 8), B),     \
                               dim3(128), 0, st, R, M, cx, kymin, cy, mw);                   \
        else                                                                                 \
            hipLaunchKernelGGL((mask_cols16_kernel<q, false, 128>), dim3(cdiv(cx, 8), B),    \
                               dim3(128), 0, st, R, M, cx, kymin, cy, mw);                   \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                      \
    }
    LITHO_MC16(16)  // every plan: one kernel, the same bits
    LITHO_MC16(32)
    LITHO_MC16(48)
#undef LITHO_MC16
    if (N == 1024 && kymin > -64 && kymin + cy - 1 < 64) {  // every plan: one kernel, the same bits
        if (h16)
            hipLaunchKernelGGL((mask_cols_band1024_kernel<true>), dim3(cdiv(cx, 8), B), dim3(256), 0,
                               st, R, M, cx, kymin, cy, mw);
        else
            hipLaunchKernelGGL((mask_cols_band1024_kernel<false>), dim3(cdiv(cx, 8), B), dim3(256),
                               0, st, R, M, cx, kymin, cy, mw);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
#define LITHO_MC(p, q, nth)                                                                    \
    if (N == p * q) {                                                                          \
        constexpr int FPB = nth / Fft2Geom<p, q>::T;                                           \
        if (h16)                                                                               \
            hipLaunchKernelGGL((mask_cols2_kernel<p, q, nth, true>), dim3(cdiv(cx, FPB), B),   \
                               dim3(nth), 0, st, R, M, cx, kymin, cy, mw);                     \
        else                                                                                   \
            hipLaunchKernelGGL((mask_cols2_kernel<p, q, nth, false>), dim3(cdiv(cx, FPB), B),  \
                               dim3(nth), 0, st, R, M, cx, kymin, cy, mw);                     \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                        \
    }
    LITHO_MC(16, 16, 256)
    LITHO_MC(16, 32, 256)
    LITHO_MC(16, 48, 128)
    if (lean) {
        LITHO_MC(32, 32, 128)
    }
    LITHO_MC(32, 32, 256)
#undef LITHO_MC
    return LITHO_EARG;
}

int launch_fi_cols_split(int N, const float2* C, void* U, int B, int band, int ns,
                         hipStream_t st, int nsum, long cplane) {
    if (B <= 0 || B > 65535 || N != 1024 || 2 * band + 1 > 160 || ns < 1 || ns > band + 1 ||
        nsum < 1 || nsum > 2)
        return LITHO_EARG;
    const dim3 grid(cdiv(band + 1, 8), B);
    hipLaunchKernelGGL((fi_cols_band1024_kernel<false, 256, true>), grid, dim3(256), 0, st, C, U,
                       band, ns, nsum, cplane);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_fi_cols_t(int N, const float2* C, void* U, int B, int band, bool h16, bool lean,
                     hipStream_t st, int nsum, long cplane) {
    if (B <= 0 || B > 65535 || 2 * band + 1 > N || nsum < 1 || nsum > 2) return LITHO_EARG;
#define LITHO_FC16(q)                                                                          \
    if (N == 16 * q && 2 * band + 1 <= 3 * q) {                                                \
        const dim3 grid(cdiv(band + 1, 8), B);                                                 \
        if (h16)                                                                               \
            hipLaunchKernelGGL((fi_cols16_kernel<q, 3, true, 128>), grid, dim3(128), 0, st, C, \
                               U, band, nsum, cplane);                                         \
        else                                                                                   \
            hipLaunchKernelGGL((fi_cols16_kernel<q, 3, false, 128>), grid, dim3(128), 0, st, C, \
                               U, band, nsum, cplane);                                         \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                        \
    }
    LITHO_FC16(16)  // every plan: one kernel, the same bits
    LITHO_FC16(32)
    LITHO_FC16(48)
#undef LITHO_FC16
    if (N == 1024 && 2 * band + 1 <= 160) {  // every plan: one kernel, the same bits
        const dim3 grid(cdiv(band + 1, 8), B);
        if (h16)
            hipLaunchKernelGGL((fi_cols_band1024_kernel<true>), grid, dim3(256), 0, st, C, U, band, 0,
                               nsum, cplane);
        else
            hipLaunchKernelGGL((fi_cols_band1024_kernel<false>), grid, dim3(256), 0, st, C, U, band,
                               0, nsum, cplane);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
#define LITHO_FC(p, q, nz, nth)                                                                 \
    if (N == p * q && 2 * band + 1 <= nz * q) {                                                 \
        constexpr int FPB = nth / Fft2Geom<p, q>::T;                                            \
        if (h16)                                                                                \
            hipLaunchKernelGGL((fi_cols2_kernel<p, q, nz, nth, true>),                          \
                               dim3(cdiv(band + 1, FPB), B), dim3(nth), 0, st, C, U, band, nsum, \
                               cplane);                                                         \
        else                                                                                    \
            hipLaunchKernelGGL((fi_cols2_kernel<p, q, nz, nth, false>),                         \
                               dim3(cdiv(band + 1, FPB), B), dim3(nth), 0, st, C, U, band, nsum, \
                               cplane);                                                         \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                         \
    }
    LITHO_FC(16, 16, 3, 256)
    LITHO_FC(16, 16, 16, 256)
    LITHO_FC(16, 32, 3, 256)
    LITHO_FC(16, 32, 16, 256)
    LITHO_FC(16, 48, 3, 128)
    LITHO_FC(16, 48, 16, 128)
    if (lean) {
        LITHO_FC(32, 32, 5, 128)
    }
    LITHO_FC(32, 32, 5, 256)
    LITHO_FC(32, 32, 32, 256)
#undef LITHO_FC
    return LITHO_EARG;
}

int launch_image_rows_t(int N, const void* U, void* img, long nrows, int band, bool h16,
                        bool lean, bool out16, hipStream_t st) {
    if (nrows % N || nrows % 2 || 2 * band + 1 > N) return LITHO_EARG;
#define LITHO_IT1(p, q, nz, nth, hh, oo)                                                         \
    hipLaunchKernelGGL((image_rows2_kernel<p, q, nz, nth, true, hh, oo>), dim3(grid), dim3(nth), \
                       0, st, U, img, (int)nrows, band, nullptr)
#define LITHO_IT(p, q, nz, nth)                                                                \
    if (N == p * q && 2 * band + 1 <= nz * q) {                                                \
        constexpr int FPB = nth / Fft2Geom<p, q>::T;                                           \
        unsigned grid = cdiv((size_t)(nrows / 2), FPB);                                        \
        if (h16 && out16) LITHO_IT1(p, q, nz, nth, true, true);                                \
        else if (h16) LITHO_IT1(p, q, nz, nth, true, false);                                   \
        else if (out16) LITHO_IT1(p, q, nz, nth, false, true);                                 \
        else LITHO_IT1(p, q, nz, nth, false, false);                                           \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                        \
    }
    if (N == 1024 && 2 * band + 1 <= 160) {  // every plan: one kernel, the same bits
        unsigned grid = cdiv((size_t)(nrows / 2), 8);
        if (h16 && out16)
            hipLaunchKernelGGL((image_band1024_kernel<true, true>), dim3(grid), dim3(256), 0, st, U,
                               img, (int)nrows, band);
        else if (h16)
            hipLaunchKernelGGL((image_band1024_kernel<true, false>), dim3(grid), dim3(256), 0, st,
                               U, img, (int)nrows, band);
        else if (out16)
            hipLaunchKernelGGL((image_band1024_kernel<false, true>), dim3(grid), dim3(256), 0, st,
                               U, img, (int)nrows, band);
        else
            hipLaunchKernelGGL((image_band1024_kernel<false, false>), dim3(grid), dim3(256), 0, st,
                               U, img, (int)nrows, band);
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
    }
    if (N == 256 && 2 * band + 1 <= 3 * 16)
        return launch_image_band16<16, 256>(U, img, nrows, band, h16, out16, st);
    if (N == 512 && 2 * band + 1 <= 3 * 32)
        return launch_image_band16<32, 256>(U, img, nrows, band, h16, out16, st);
    if (N == 768 && 2 * band + 1 <= 3 * 48)
        return launch_image_band16<48, 128>(U, img, nrows, band, h16, out16, st);
    LITHO_IT(16, 16, 3, 256)
    LITHO_IT(16, 16, 16, 256)
    LITHO_IT(16, 32, 3, 256)
    LITHO_IT(16, 32, 16, 256)
    LITHO_IT(16, 48, 3, 128)
    LITHO_IT(16, 48, 16, 128)
    if (lean) {
        LITHO_IT(32, 32, 5, 128)
    }
    LITHO_IT(32, 32, 5, 256)
    LITHO_IT(32, 32, 32, 256)
#undef LITHO_IT
#undef LITHO_IT1
    return LITHO_EARG;
}

// ---------------------------------------------------------------------------
// tmpImgp helpers of the accelerator-shaped entry (v18 top, socs_2048_stream_
// refactored_v18_hls_fft.cpp L58-131, L380-387).
//
// spectrum_window: M[(y+m)*W + x+m] = mskf[Ly/2+y][Lx/2+x] for |x|, |y| <= m
// (W = 2m+1), zero where the index leaves the Lx x Ly planes (the embed's
// bounds check, v18 .cpp L114-116).  One thread per window entry.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void spectrum_window_kernel(const float* __restrict__ re,
                                                              const float* __restrict__ im,
                                                              int Lx, int Ly, int m,
                                                              float2* __restrict__ M) {
    const int W = 2 * m + 1;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= W * W) return;
    const int y = i / W - m, x = i % W - m;
    const int my = Ly / 2 + y, mx = Lx / 2 + x;
    float2 v = make_float2(0.f, 0.f);
    if (my >= 0 && my < Ly && mx >= 0 && mx < Lx) {
        const long j = (long)my * Lx + mx;
        v = make_float2(re[j], im[j]);
    }
    M[i] = v;
}

// grid_sample_shift: T is the unshifted intensity on a G x G grid; the Gy x Gx
// grid (G / Gx, G / Gy integers) samples the same band-limited intensity at
// every (G / Gx)-th column and (G / Gy)-th row.  Writes it fftshifted
// (v18 .cpp L199-216: out[y][x] = in[(y + Gy/2) % Gy][(x + Gx/2) % Gx]) to
// o1 and o2 (either may be null).
__global__ __launch_bounds__(256) void grid_sample_shift_kernel(const float* __restrict__ T,
                                                                int G, int Gx, int Gy,
                                                                float* __restrict__ o1,
                                                                float* __restrict__ o2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Gx * Gy) return;
    const int y = i / Gx, x = i - y * Gx;
    const int sy = (y + Gy / 2) % Gy, sx = (x + Gx / 2) % Gx;
    const float v = T[(long)sy * (G / Gy) * G + (long)sx * (G / Gx)];
    if (o1) o1[i] = v;
    if (o2) o2[i] = v;
}

int launch_spectrum_window(const float* re, const float* im, int Lx, int Ly, int m, float2* M,
                           hipStream_t st) {
    if (!re || !im || !M || Lx <= 0 || Ly <= 0 || m < 0) return LITHO_EARG;
    const int W = 2 * m + 1;
    hipLaunchKernelGGL(spectrum_window_kernel, dim3(cdiv((size_t)W * W, 256)), dim3(256), 0, st,
                       re, im, Lx, Ly, m, M);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

int launch_grid_sample_shift(const float* T, int G, int Gx, int Gy, float* o1, float* o2,
                             hipStream_t st) {
    if (!T || Gx <= 0 || Gy <= 0 || G % Gx || G % Gy) return LITHO_EARG;
    hipLaunchKernelGGL(grid_sample_shift_kernel, dim3(cdiv((size_t)Gx * Gy, 256)), dim3(256), 0,
                       st, T, G, Gx, Gy, o1, o2);
    return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;
}

}  // namespace litho
