// LDS-resident batched FFT for gfx950: Stockham autosort passes (radix 8,
// then 4 or 2 for the remainder) over `nf` sequences of power-of-two length L
// held in LDS, all threads of the block cooperating.  The pass structure is
// the published Stockham formulation (Govindaraju et al., SC'08, "High
// performance discrete Fourier transforms on graphics processors"): input
// element j + r*L/R, twiddle e^{s 2 pi i r (j mod Ns)/(Ns R)}, output index
// (j div Ns)*Ns*R + (j mod Ns) + r*Ns, Ns growing by R every pass.
//
// Twiddles come from one table tw[m] = e^{-2 pi i m / TW_MAX} built in double
// on the host and rounded once to float, so every twiddle is within 0.5 ulp.
#pragma once

#include "common.hpp"

namespace litho {

constexpr int TW_MAX = 8192;

template <bool INV>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int idx) {
    float2 w = tw[idx];
    return INV ? make_float2(w.x, -w.y) : w;
}

// forward: X[q] = sum_r v[r] e^{-2 pi i r q / R}; inverse: e^{+...}
template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
    float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    float2 s13 = cadd(v[1], v[3]), d13 = rot90<INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    // o[k] *= w8^k, w8 = e^{-+ i pi/4}
    float2 o1 = o[1], o3 = o[3];
    if (!INV) {
        o[1] = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
        o[3] = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    } else {
        o[1] = make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y));
        o[3] = make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y));
    }
    o[2] = rot90<INV>(o[2]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dftR(float2* v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else dft8<INV>(v);
}

template <int L, int R, bool INV>
__device__ __forceinline__ void stockham_pass(const float2* __restrict__ src, float2* __restrict__ dst,
                                              int nf, int Ns, const float2* __restrict__ tw,
                                              int tid, int nthr) {
    constexpr int Q = L / R;
    const int total = nf * Q;
    const int step = TW_MAX / (Ns * R);
    for (int b = tid; b < total; b += nthr) {
        const int f = b / Q;
        const int j = b - f * Q;
        const int k = j & (Ns - 1);
        const float2* s = src + f * L;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = s[j + r * Q];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], twiddle<INV>(tw, r * k * step));
        }
        dftR<R, INV>(v);
        float2* d = dst + f * L + (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) d[r * Ns] = v[r];
    }
}

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v / 2); }

// In-LDS FFT of nf sequences a[f*L + i].  The caller syncs after filling `a`;
// on return the result is in the returned buffer (a or b) and the block is
// synchronised.
template <int L, bool INV>
__device__ float2* fft_lds(float2* a, float2* b, int nf, const float2* __restrict__ tw, int tid,
                           int nthr) {
    static_assert((L & (L - 1)) == 0 && L >= 2 && L <= TW_MAX, "power-of-two L");
    constexpr int LG = ilog2_c(L);
    constexpr int N8 = LG / 3;
    constexpr int REM = LG % 3;
    float2* src = a;
    float2* dst = b;
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < N8; p++) {
        stockham_pass<L, 8, INV>(src, dst, nf, Ns, tw, tid, nthr);
        __syncthreads();
        float2* t = src;
        src = dst;
        dst = t;
        Ns *= 8;
    }
    if constexpr (REM == 2) {
        stockham_pass<L, 4, INV>(src, dst, nf, Ns, tw, tid, nthr);
        __syncthreads();
        src = dst;
    } else if constexpr (REM == 1) {
        stockham_pass<L, 2, INV>(src, dst, nf, Ns, tw, tid, nthr);
        __syncthreads();
        src = dst;
    }
    return src;
}

}  // namespace litho
