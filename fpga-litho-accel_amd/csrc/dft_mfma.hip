// Band DFTs of the mask rows and of the image rows as fp16 matrix products on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_f16, FP32 accumulate).
//
// Only a narrow band of the N-point row spectrum is live on both ends of the
// SOCS path (the 2Nx+1 window on the mask side, the 2·(2Nx)+1 intensity band on
// the image side), so each of these transforms is a dense N x (band) product:
//
//   mask side   R[b][j][y] = sum_x mask[b][y][x] * s_k e^{-2 pi i k x / N},  k = kmin + j
//               (s_k = (-1)^k folds the centring shift of FT_r2c_padded,
//                litho_thread.cpp L255-331, into the table)
//   image side  img[b][y][x] = sum_{kx=0..nb-1} w_kx Re(U[b][kx][y] e^{+2 pi i kx x / N})
//               (w_0 = 1, w_kx = 2: the c2r of FI, litho_thread.cpp L692-772)
//
// Numerics.  The mask is exact in fp16 (u8 levels 0..255, or 0/1 floats); the
// mask-side table is one fp16 rounding of cos/sin.  On the image side U is
// split on the fly into hi + lo fp16 parts (lo scaled by 2^12 so it stays in
// fp16's normal range) for kx up to half the band, hi alone above; the table
// rows for the lo parts carry the 2^-12.
//
// Tiling.  A wave owns RT row tiles of 16 rows, so that one 16x32 B fragment
// read from LDS feeds RT MFMAs: on the mask side a 256-row block is 4 waves of
// 4 tiles for windows of up to 16 columns and 8 waves of 2 tiles for wider
// ones; on the image side a 256-row block is 8 waves of 2 tiles.  The B operand
// (the cos/sin table) is pre-arranged on the host in fragment order: lane l
// reads 16 contiguous bytes, so LDS reads are conflict-free and the global ->
// LDS copy is a straight uint4 copy.  Within one k-step, lane l of the MFMA
// holds the 8 k-values 8*(l/16) .. +7 of its A row (l%16) and B column (l%16);
// the D fragment holds rows 4*(l/16) + i, column l%16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "dft_mfma.hpp"

namespace litho {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma16(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// two bytes of d (byte 2*hi, 2*hi+1) as exact fp16 values packed in a dword:
// 0x6400 | b is the fp16 value 1024 + b, minus 1024 is exact.  One byte
// permute builds both halves (selector bytes 0-3 pick bytes of d, 4 picks
// 0x64 from the constant word).
__device__ __forceinline__ unsigned u8pair_to_h2(unsigned d, int hi) {
    const unsigned bits = __builtin_amdgcn_perm(0x64646464u, d, hi ? 0x04030402u : 0x04010400u);
    h2v v = __builtin_bit_cast(h2v, bits);
    v = v - (h2v){(_Float16)1024.f, (_Float16)1024.f};
    return __builtin_bit_cast(unsigned, v);
}

// LDS written by some lanes of a wave and read by others: order it within the
// wavefront (no block barrier needed)
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned f2_to_h2(float a, float b) {
    h2v v = {(_Float16)a, (_Float16)b};
    return __builtin_bit_cast(unsigned, v);
}

constexpr int ROWS_PER_BLOCK = 256;  // mask side: rows per block
constexpr int BITS_MAXCH = 8;        // bit-packed masks: up to 1024 pixels per row
constexpr int KCH = 128;             // mask-side K (pixel) chunk staged in LDS
constexpr int XCH = 64;              // image-side output columns per LDS chunk
// image side: row tiles per wave (32 rows; 64 rows measured 2 % slower with
// fp16 images, 13 % with f32)
constexpr int IMAGE_RT = 2;
// image side: threads per block.  8 waves share one table chunk, so each lane
// prefetches 4 table fragments instead of 7: 126 VGPRs and 64 KB of LDS, two
// blocks and four waves per SIMD (4-wave blocks: 135 VGPRs, three waves per
// SIMD, 3 % slower per step)
constexpr int IMAGE_NTH = 512;

// ---------------------------------------------------------------------------
// Mask rows -> transposed row-DFT window R [B][count][N] (same contract as the
// FFT row kernel's transposed output).  NG = ceil(count / 16) frequency groups,
// each as a real-part and an imaginary-part column tile.
// ---------------------------------------------------------------------------
// one bit per pixel, pixel 8j + i at bit i of byte j: the two pixels at
// bits sh, sh + 1 of x as exact fp16 0 / 1 values packed in a dword
// (t * 0x8001 puts bit 0 at 0 and bit 1 at 16; times 0x3C00 = fp16 1.0)
__device__ __forceinline__ unsigned bitpair_to_h2(unsigned x, int sh) {
    const unsigned t = (x >> sh) & 3u;
    return ((t * 0x8001u) & 0x10001u) * 0x3C00u;
}

// MF: mask format, 0 f32, 1 u8 levels, 2 one bit per pixel
template <int NG, int RTW, int MF, bool H16>
__global__ __launch_bounds__(ROWS_PER_BLOCK * 4 / RTW, 2) void mask_rows_mfma_kernel(
    const void* __restrict__ mask_, const u4v* __restrict__ Fpk, void* __restrict__ R_, int N,
    int count, float oscale) {
    constexpr bool U8 = MF == 1, BITS = MF == 2;
    constexpr int NT = 2 * NG;
    constexpr int NTH = ROWS_PER_BLOCK * 4 / RTW;  // 64 threads per 16 rows x RTW
    constexpr int RT = RTW;
    constexpr int CH = 4 * NT * 64;  // uint4 per K chunk (4 k-steps)
    constexpr int PER = CH / NTH;
    static_assert(CH % NTH == 0, "chunk copy");
    // double-buffered table chunk: chunk c+1 is fetched while chunk c is used.
    // Two separate arrays (not one [2][CH]) let the compiler see that reads of
    // one buffer never alias the DMA into the other, so it need not wait.
    __shared__ u4v fsA[CH], fsB[CH];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int b = blockIdx.y;
    const int ybase = blockIdx.x * ROWS_PER_BLOCK + w * 16 * RT;  // row within the image
    const long rbase = (long)b * N + ybase;
    f4 acc[RT][NT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[rt][t] = (f4){0.f, 0.f, 0.f, 0.f};
    // A fragments: 32 consecutive pixels of row (16 rt + r16) per lane and
    // chunk; u8 pixels stay packed until their k-step and are prefetched.
    auto load_u8 = [&](int c, u4v (&raw)[RT][2]) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            const long row = rbase + 16 * rt + r16;
            const u4v* p = (const u4v*)((const unsigned char*)mask_ + row * N + c * KCH + 32 * g);
            raw[rt][0] = p[0];
            raw[rt][1] = p[1];
        }
    };
    // one bit per pixel: the 32 pixels of lane group g in one dword per row
    auto load_bits = [&](int c, unsigned (&raw)[RT]) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            const long row = rbase + 16 * rt + r16;
            raw[rt] = ((const unsigned*)mask_)[row * (N / 32) + c * (KCH / 32) + g];
        }
    };
    // two register sets for the u8 / bit fragments (ping-pong, no loop-carried
    // copy): chunk c+1 loads into one set while chunk c computes from the other
    u4v rawA[RT][2], rawB[RT][2];
    unsigned bitA[RT];
    unsigned bitAll[BITS ? BITS_MAXCH : 1][RT];
    // table chunk c -> dst straight from global memory into LDS (16 bytes per
    // lane, the wave's 64 lanes contiguous), so no registers hold it
    auto load_table = [&](int c, u4v* dst) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            __builtin_amdgcn_global_load_lds(
                (const void*)(Fpk + (long)c * CH + NTH * i + tid),
                (__attribute__((address_space(3))) void*)(dst + NTH * i + 64 * w), 16, 0, 0);
    };
    // N is a multiple of ROWS_PER_BLOCK (checked by the launcher), so there
    // are at least two chunks; stating it keeps the first chunk's loads ahead
    // of the barrier below (a guarded loop would let them sink past it and
    // leave them pending at the loop head, costing a full wait per pass)
    const int nch = N / KCH;
    __builtin_assume(nch >= ROWS_PER_BLOCK / KCH);
    load_table(0, fsA);
    if (U8) load_u8(0, rawA);
    if constexpr (BITS) load_bits(0, bitAll[0]);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    auto chunk = [&](int c, const u4v (&raw)[RT][2], const unsigned (&bits)[RT], const u4v* fb,
                     u4v* fnext) {
        if (c + 1 < nch) load_table(c + 1, fnext);
        u4v af[RT][4];
        if (MF == 0) {
#pragma unroll
            for (int rt = 0; rt < RT; rt++) {
                const long row = rbase + 16 * rt + r16;
                const float4* p = (const float4*)((const float*)mask_ + row * N + c * KCH + 32 * g);
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const float4 x0 = p[2 * s], x1 = p[2 * s + 1];
                    af[rt][s] = (u4v){f2_to_h2(x0.x, x0.y), f2_to_h2(x0.z, x0.w),
                                      f2_to_h2(x1.x, x1.y), f2_to_h2(x1.z, x1.w)};
                }
            }
        }
        // B fragments one (k-step, tile) ahead: each LDS read overlaps the
        // MFMAs of the previous fragment
        u4v bnext = fb[lane];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            u4v as[RT];
#pragma unroll
            for (int rt = 0; rt < RT; rt++)
                as[rt] = BITS ? (u4v){bitpair_to_h2(bits[rt], 8 * s),
                                      bitpair_to_h2(bits[rt], 8 * s + 2),
                                      bitpair_to_h2(bits[rt], 8 * s + 4),
                                      bitpair_to_h2(bits[rt], 8 * s + 6)}
                       : U8 ? (u4v){u8pair_to_h2(raw[rt][s >> 1][2 * (s & 1)], 0),
                                    u8pair_to_h2(raw[rt][s >> 1][2 * (s & 1)], 1),
                                    u8pair_to_h2(raw[rt][s >> 1][2 * (s & 1) + 1], 0),
                                    u8pair_to_h2(raw[rt][s >> 1][2 * (s & 1) + 1], 1)}
                            : af[rt][s];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const u4v bcur = bnext;
                const int nx = s * NT + t + 1;
                if (nx < 4 * NT) bnext = fb[nx * 64 + lane];
                const h8 bf = __builtin_bit_cast(h8, bcur);
#pragma unroll
                for (int rt = 0; rt < RT; rt++)
                    acc[rt][t] = mfma16(__builtin_bit_cast(h8, as[rt]), bf, acc[rt][t]);
            }
        }
        // the next chunk's table (and u8 fragments) have landed before any
        // wave reads them; every wave is done with fb before it is refilled
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    if constexpr (BITS) {
        // the whole row tile is N / 8 bytes: every chunk's bits are fetched
        // here (chunk 0's already waited for), so the loop waits only on the
        // table, which comes from L2
#pragma unroll
        for (int c = 1; c < BITS_MAXCH; c++)
            if (c < nch) load_bits(c, bitAll[c]);
#pragma unroll
        for (int c = 0; c < BITS_MAXCH; c++) {
            if (c >= nch) break;
            if (c % 2 == 0)
                chunk(c, rawA, bitAll[c], fsA, fsB);
            else
                chunk(c, rawA, bitAll[c], fsB, fsA);
        }
    } else {
        for (int c = 0; c < nch; c += 2) {
            if (U8 && c + 1 < nch) load_u8(c + 1, rawB);
            chunk(c, rawA, bitA, fsA, fsB);
            if (c + 1 >= nch) break;
            if (U8 && c + 2 < nch) load_u8(c + 2, rawA);
            chunk(c + 1, rawB, bitA, fsB, fsA);
        }
    }
    // D fragment: rows 4g + i of the tile, column r16 = frequency 16 fg + r16
#pragma unroll
    for (int fg = 0; fg < NG; fg++) {
        const int j = 16 * fg + r16;
        if (j >= count) continue;
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            const int y = ybase + 16 * rt + 4 * g;
            const f4 re = acc[rt][2 * fg] * oscale, im = acc[rt][2 * fg + 1] * oscale;
            const long o = ((long)b * count + j) * N + y;
            if (H16) {
                uint4 v;
                v.x = f2_to_h2(re[0], im[0]);
                v.y = f2_to_h2(re[1], im[1]);
                v.z = f2_to_h2(re[2], im[2]);
                v.w = f2_to_h2(re[3], im[3]);
                *(uint4*)((unsigned*)R_ + o) = v;
            } else {
                float4* q = (float4*)((float2*)R_ + o);
                q[0] = make_float4(re[0], im[0], re[1], im[1]);
                q[1] = make_float4(re[2], im[2], re[3], im[3]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// U [B][nb][N] (float2, FP32) -> image [B][N][N] (f32 or fp16).  The first KS
// k-steps of 32 cover 8 kx values each, (re, im) x (hi, lo) for two kx per
// lane group (kx < 8 KS, at least kx <= nb / 2); the other NKS - KS steps cover
// 16 kx each as (re, im) hi parts only.  The intensity spectrum above half the
// band is small (it comes only from products of band-edge kernel components):
// rounding it to fp16 alone adds below 1e-7 of the image at 1024^2 (measured on
// the FP64 images of Manhattan and random grey-level masks), where rounding all
// of U would add 3-5e-4.
// U arrives as FP32 [nb][N], or (USPLIT) in a compact form from the column
// pass: per image [ns][N] FP32 for the split columns kx < ns = min(nb, 8 KS),
// then [nb - ns][N] fp16 for the rest (rounded as the product would).
// The product is formed transposed (table as the A operand, U as B): a lane's
// D fragment is 4 consecutive pixels of one image row.  Fragments are staged
// per wave in LDS until every row has a full 128-byte segment, which is then
// stored 16 bytes per lane (8 rows per instruction), so HBM sees whole lines.
// ---------------------------------------------------------------------------
template <int NKS, int KS, bool OUT16, int RT, int NTH, bool USPLIT>
__global__ __launch_bounds__(NTH, 2) void image_rows_mfma_kernel(const void* __restrict__ U_,
                                                                 const u4v* __restrict__ Hpk,
                                                                 void* __restrict__ img_, int N,
                                                                 int nb) {
    constexpr int NW = NTH / 64;      // waves per block, sharing one table chunk
    constexpr int CH = 4 * NKS * 64;  // u4v per XCH-column chunk (4 column tiles)
    constexpr int PER = (CH + NTH - 1) / NTH;
    // CH is a multiple of 256, so the last copy round is whole waves
    static_assert(CH % 256 == 0 && NTH % 256 == 0, "chunk copy");
    auto copy_lane = [](int tid, int i) { return CH % NTH == 0 || tid + NTH * i < CH; };
    constexpr int OB = OUT16 ? 2 : 4;         // bytes per pixel
    constexpr int TPS = OUT16 ? 4 : 2;        // column tiles per 128-byte row segment
    constexpr int SROW = 16 * TPS * OB + 16;  // staged row pitch in bytes (padded)
    __shared__ u4v fs[CH];
    constexpr int RW = 16 * RT;  // rows per wave
    __shared__ __attribute__((aligned(16))) unsigned char stg[NW][RW * SROW];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int b = blockIdx.y;
    const int ybase = blockIdx.x * NW * RW + w * RW;
    // FP32 U [nb][N], or the compact form: [ns][N] FP32 for the columns the
    // product splits, then [nb - ns][N] fp16 (its hi parts, already rounded)
    const int ns = USPLIT ? min(nb, 8 * KS) : 0;
    const unsigned char* Us =
        (const unsigned char*)U_ + (long)b * ((long)ns * N * 8 + (long)(nb - ns) * N * 4);
    const float2* Ub = USPLIT ? (const float2*)Us : (const float2*)U_ + (long)b * nb * N;
    const unsigned* Uhi = (const unsigned*)(Us + (long)ns * N * 8);
    u4v pre[PER];
#pragma unroll
    for (int i = 0; i < PER; i++)
        if (copy_lane(tid, i)) pre[i] = Hpk[tid + NTH * i];
    // U fragments for the whole K extent stay in registers
    u4v a[RT][NKS];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        const int y = ybase + 16 * rt + r16;
#pragma unroll
        for (int s = 0; s < NKS; s++) {
            unsigned q[4];
            if (USPLIT && s >= KS) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int kx = 8 * KS + 16 * (s - KS) + 4 * g + e;  // >= ns here
                    unsigned v = Uhi[(long)min(kx - ns, nb - ns - 1) * N + y];
                    if (kx >= nb) v = 0u;
                    q[e] = v;
                }
            } else if (s < KS) {
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int kx = 8 * s + 2 * g + e;  // < ns in the compact form
                    // clamped index and a select, not a branch: every load is in
                    // flight at once
                    float2 v = Ub[(long)min(kx, nb - 1) * N + y];
                    if (kx >= nb) v = make_float2(0.f, 0.f);
                    const _Float16 hr = (_Float16)v.x, hi = (_Float16)v.y;
                    const float lr = (v.x - (float)hr) * 4096.f, li = (v.y - (float)hi) * 4096.f;
                    q[2 * e] = __builtin_bit_cast(unsigned, (h2v){hr, hi});
                    q[2 * e + 1] = f2_to_h2(lr, li);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int kx = 8 * KS + 16 * (s - KS) + 4 * g + e;
                    float2 v = Ub[(long)min(kx, nb - 1) * N + y];
                    if (kx >= nb) v = make_float2(0.f, 0.f);
                    q[e] = f2_to_h2(v.x, v.y);
                }
            }
            a[rt][s] = (u4v){q[0], q[1], q[2], q[3]};
        }
    }
#pragma unroll
    for (int i = 0; i < PER; i++)
        if (copy_lane(tid, i)) fs[tid + NTH * i] = pre[i];
    __syncthreads();
    unsigned char* sw = stg[w];
    unsigned char* out = (unsigned char*)img_ + ((long)b * N * N + (long)ybase * N) * OB;
    const int nch = N / XCH;
    for (int c = 0; c < nch; c++) {
        const bool more = c + 1 < nch;
        if (more) {
#pragma unroll
            for (int i = 0; i < PER; i++)
                if (copy_lane(tid, i)) pre[i] = Hpk[(long)(c + 1) * CH + tid + NTH * i];
        }
        u4v tnext = fs[lane];
#pragma unroll
        for (int tp = 0; tp < 4 / TPS; tp++) {
#pragma unroll
            for (int tt = 0; tt < TPS; tt++) {
                const int t = tp * TPS + tt;
                f4 acc[RT];
#pragma unroll
                for (int rt = 0; rt < RT; rt++) acc[rt] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < NKS; s++) {
                    const u4v tcur = tnext;
                    const int nx = t * NKS + s + 1;
                    if (nx < 4 * NKS) tnext = fs[nx * 64 + lane];
                    const h8 tf = __builtin_bit_cast(h8, tcur);
#pragma unroll
                    for (int rt = 0; rt < RT; rt++)
                        acc[rt] = mfma16(tf, __builtin_bit_cast(h8, a[rt][s]), acc[rt]);
                }
                // D[x][y]: pixels x = 4g..4g+3 of the column tile, row y = 16 rt + r16
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    unsigned char* d = sw + (16 * rt + r16) * SROW + (16 * tt + 4 * g) * OB;
                    if (OUT16) {
                        uint2 v;
                        v.x = f2_to_h2(acc[rt][0], acc[rt][1]);
                        v.y = f2_to_h2(acc[rt][2], acc[rt][3]);
                        *(uint2*)d = v;
                    } else {
                        *(float4*)d = make_float4(acc[rt][0], acc[rt][1], acc[rt][2], acc[rt][3]);
                    }
                }
            }
            wave_sync_lds();
            // RW rows x 128 bytes: 16 bytes per lane, 8 rows per store; the
            // image is not read again here, so the stores stream past the
            // caches (nontemporal)
            const long x0 = (long)c * XCH + 16 * TPS * tp;
#pragma unroll
            for (int q = 0; q < RW / 8; q++) {
                const int row = 8 * q + (lane >> 3), seg = (lane & 7) * 16;
                const u4v v = *(const u4v*)(sw + row * SROW + seg);
                __builtin_nontemporal_store(v, (u4v*)(out + ((long)row * N + x0) * OB + seg));
            }
            wave_sync_lds();
        }
        __syncthreads();  // every wave is done with this table chunk
        if (more) {
#pragma unroll
            for (int i = 0; i < PER; i++)
                if (copy_lane(tid, i)) fs[tid + NTH * i] = pre[i];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Host: tables in fragment order, launchers
// ---------------------------------------------------------------------------
static inline unsigned short f16_bits(double v) {
    return __builtin_bit_cast(unsigned short, (_Float16)(float)v);
}

std::vector<unsigned short> mask_rows_table(int N, int kmin, int count) {
    const int NG = (count + 15) / 16, NT = 2 * NG;
    const int nch = N / KCH;
    std::vector<unsigned short> t((size_t)nch * 4 * NT * 64 * 8);
    for (int c = 0; c < nch; c++)
        for (int s = 0; s < 4; s++)
            for (int tt = 0; tt < NT; tt++)
                for (int l = 0; l < 64; l++)
                    for (int i = 0; i < 8; i++) {
                        const int x = c * KCH + 32 * (l >> 4) + 8 * s + i;
                        const int j = 16 * (tt >> 1) + (l & 15);
                        double v = 0.0;
                        if (j < count) {
                            const long k = kmin + j;
                            const long m = (((k * x) % N) + N) % N;  // exact phase index
                            const double th = 2.0 * M_PI * (double)m / N;
                            const double sg = (k & 1) ? -1.0 : 1.0;
                            v = (tt & 1) ? -sg * std::sin(th) : sg * std::cos(th);
                        }
                        t[((((size_t)c * 4 + s) * NT + tt) * 64 + l) * 8 + i] = f16_bits(v);
                    }
    return t;
}

static void image_rows_steps(int nb, int* nks, int* ks) {
    const int split = std::min(nb, nb / 2 + 1);  // kx <= nb / 2 keep their lo parts
    *ks = (split + 7) / 8;
    const int rest = std::max(0, nb - 8 * *ks);
    *nks = *ks + (rest + 15) / 16;
}

std::vector<unsigned short> image_rows_table(int N, int nb) {
    int NKS, KS;
    image_rows_steps(nb, &NKS, &KS);
    const int nch = N / XCH;
    std::vector<unsigned short> t((size_t)nch * 4 * NKS * 64 * 8);
    for (int c = 0; c < nch; c++)
        for (int tt = 0; tt < 4; tt++)
            for (int s = 0; s < NKS; s++)
                for (int l = 0; l < 64; l++)
                    for (int i = 0; i < 8; i++) {
                        const int x = c * XCH + 16 * tt + (l & 15);
                        // split steps: (re, im) x (hi, lo) of two kx per lane
                        // group; hi steps: (re, im) hi of four kx
                        const int kx = s < KS ? 8 * s + 2 * (l >> 4) + (i >> 2)
                                              : 8 * KS + 16 * (s - KS) + 4 * (l >> 4) + (i >> 1);
                        const int part = s < KS ? (i & 3) : (i & 1);
                        double v = 0.0;
                        if (kx < nb) {
                            const double wk = kx ? 2.0 : 1.0;
                            const long m = ((long)kx * x) % N;
                            const double th = 2.0 * M_PI * (double)m / N;
                            v = (part & 1) ? -wk * std::sin(th) : wk * std::cos(th);
                            if (kx == 0 && (part & 1)) v = 0.0;
                            if (part & 2) v /= 4096.0;
                        }
                        t[((((size_t)c * 4 + tt) * NKS + s) * 64 + l) * 8 + i] = f16_bits(v);
                    }
    return t;
}

int launch_mask_rows_mfma(int N, const void* mask, int mfmt, const void* table, void* R, int B,
                          int count, bool h16, hipStream_t st) {
    if (N % ROWS_PER_BLOCK || N % KCH || B <= 0 || B > 65535 || count <= 0 || count > 80 ||
        mfmt < 0 || mfmt > 2 || (mfmt == 2 && N > BITS_MAXCH * KCH))
        return LITHO_EARG;
    const int NG = (count + 15) / 16;
    const dim3 grid(N / ROWS_PER_BLOCK, B);
    const float osc = mfmt == 1 ? 1.f / 255.f : 1.f;
    const u4v* F = (const u4v*)table;
#define LITHO_MR1(ng, rtw, mf)                                                                  \
    if (mfmt == mf) {                                                                           \
        if (h16)                                                                                \
            hipLaunchKernelGGL((mask_rows_mfma_kernel<ng, rtw, mf, true>), grid, dim3(nth), 0, st, \
                               mask, F, R, N, count, osc);                                      \
        else                                                                                    \
            hipLaunchKernelGGL((mask_rows_mfma_kernel<ng, rtw, mf, false>), grid, dim3(nth), 0,   \
                               st, mask, F, R, N, count, osc);                                  \
    }
#define LITHO_MR(ng, rtw)                                                                       \
    if (NG == ng) {                                                                             \
        constexpr int nth = ROWS_PER_BLOCK * 4 / rtw;                                           \
        LITHO_MR1(ng, rtw, 0)                                                                   \
        LITHO_MR1(ng, rtw, 1)                                                                   \
        LITHO_MR1(ng, rtw, 2)                                                                   \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                         \
    }
    LITHO_MR(1, 4)
    // 32 rows per wave from two frequency groups on (two 8-wave blocks per CU):
    // 7 % faster than 64 rows at 768 and 1024 points, the same at 512
    LITHO_MR(2, 2)
    LITHO_MR(3, 2)
    LITHO_MR(4, 2)
    LITHO_MR(5, 2)
#undef LITHO_MR
#undef LITHO_MR1
    return LITHO_EARG;
}

int image_rows_split_cols(int nb) {
    int NKS, KS;
    image_rows_steps(nb, &NKS, &KS);
    return std::min(nb, 8 * KS);
}

int launch_image_rows_mfma(int N, const void* U, int nb, const void* table, void* img, int B,
                           bool out16, bool usplit, hipStream_t st) {
    if (N % ROWS_PER_BLOCK || N % XCH || B <= 0 || B > 65535 || nb <= 0 || nb > 72)
        return LITHO_EARG;
    int NKS, KS;
    image_rows_steps(nb, &NKS, &KS);
    constexpr int IRT = IMAGE_RT;  // row tiles per wave
    constexpr int INTH = IMAGE_NTH;
    static_assert(ROWS_PER_BLOCK % (INTH / 4 * IRT) == 0, "image rows per block divide N");
    const dim3 grid(N / (INTH / 4 * IRT), B);
    const u4v* H = (const u4v*)table;
#define LITHO_IR1(nks, ks, o16, us)                                                            \
    hipLaunchKernelGGL((image_rows_mfma_kernel<nks, ks, o16, IRT, INTH, us>), grid, dim3(INTH), \
                       0, st, U, H, img, N, nb)
#define LITHO_IR(nks, ks)                                                                      \
    if (NKS == nks && KS == ks) {                                                              \
        if (out16 && usplit)                                                                   \
            LITHO_IR1(nks, ks, true, true);                                                    \
        else if (out16)                                                                        \
            LITHO_IR1(nks, ks, true, false);                                                   \
        else if (usplit)                                                                       \
            LITHO_IR1(nks, ks, false, true);                                                   \
        else                                                                                   \
            LITHO_IR1(nks, ks, false, false);                                                  \
        return hipGetLastError() == hipSuccess ? LITHO_OK : LITHO_EHIP;                        \
    }
    // every (steps, split steps) pair of nb <= 72
    LITHO_IR(1, 1)
    LITHO_IR(2, 1)
    LITHO_IR(2, 2)
    LITHO_IR(3, 2)
    LITHO_IR(4, 3)
    LITHO_IR(5, 3)
    LITHO_IR(5, 4)
    LITHO_IR(6, 4)
    LITHO_IR(7, 5)
#undef LITHO_IR
#undef LITHO_IR1
    return LITHO_EARG;
}

}  // namespace litho
