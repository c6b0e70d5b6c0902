// Register-resident DFTs and a two-stage (four-step) FFT for gfx950.
//
// dft_reg<N>(v): in-register DFT of N = 2, 4, 8, 16 or 32 points by
// decimation in time (radix 4 / 2 recursion) with compile-time twiddles, so
// inputs known to be zero at compile time fold away.
//
// Two-stage FFT of G = P * Q points (n = Q n1 + n2, k = k1 + P k2):
//   X[k1 + P k2] = sum_n2 W_Q^{n2 k2} W_G^{n2 k1} sum_n1 x[Q n1 + n2] W_P^{n1 k1}
// Stage 1 (per n2): P-point DFT over n1, times W_G^{n2 k1}; transpose through
// LDS; stage 2 (per k1): Q-point DFT over n2.  T = min(P, Q) threads per
// transform, each holding G / T values.
#pragma once

#include "common.hpp"

namespace litho {

// cos / sin of 2*pi*j/32, j = 0..31 (exact decimal values of the roots)
__device__ constexpr float kCos32[32] = {
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
    0.70710678118654752f, 0.55557023301960218f, 0.38268343236508977f, 0.19509032201612827f,
    0.0f, -0.19509032201612827f, -0.38268343236508977f, -0.55557023301960218f,
    -0.70710678118654752f, -0.83146961230254524f, -0.92387953251128674f, -0.98078528040323043f,
    -1.0f, -0.98078528040323043f, -0.92387953251128674f, -0.83146961230254524f,
    -0.70710678118654752f, -0.55557023301960218f, -0.38268343236508977f, -0.19509032201612827f,
    0.0f, 0.19509032201612827f, 0.38268343236508977f, 0.55557023301960218f,
    0.70710678118654752f, 0.83146961230254524f, 0.92387953251128674f, 0.98078528040323043f};

// W_N^e = e^{-+2 pi i e / N} for N <= 32 (forward: minus sign)
template <int N, bool INV>
__device__ __forceinline__ float2 wroot(int e) {
    const int j = ((e * (32 / N)) % 32 + 32) % 32;
    const float c = kCos32[j];
    const float s = kCos32[(j + 24) % 32];  // sin(2 pi j/32) = cos(2 pi (j-8)/32)
    return INV ? make_float2(c, s) : make_float2(c, -s);
}

template <bool INV>
__device__ __forceinline__ float2 cmul_w(float2 a, float2 w) {
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

template <int N, bool INV>
__device__ __forceinline__ void dft_reg(float2* v);

template <bool INV>
__device__ __forceinline__ void dft_reg2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
}

template <bool INV>
__device__ __forceinline__ void dft_reg4(float2* v) {
    float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    float2 t = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    float2 d13 = INV ? make_float2(-t.y, t.x) : make_float2(t.y, -t.x);
    v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
    v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
    v[1] = make_float2(d02.x + d13.x, d02.y + d13.y);
    v[3] = make_float2(d02.x - d13.x, d02.y - d13.y);
}

// N = 4 M by decimation in time: four interleaved M-point DFTs, then twiddle
// and radix-4 across them.
template <int N, bool INV>
__device__ __forceinline__ void dft_reg(float2* v) {
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
        dft_reg2<INV>(v);
    } else if constexpr (N == 4) {
        dft_reg4<INV>(v);
    } else if constexpr (N == 8) {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        dft_reg4<INV>(e);
        dft_reg4<INV>(o);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float2 t = cmul_w<INV>(o[k], wroot<8, INV>(k));
            v[k] = make_float2(e[k].x + t.x, e[k].y + t.y);
            v[k + 4] = make_float2(e[k].x - t.x, e[k].y - t.y);
        }
    } else {
        constexpr int M = N / 4;
        float2 a[4][M];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int m = 0; m < M; m++) a[r][m] = v[4 * m + r];
#pragma unroll
        for (int r = 0; r < 4; r++) dft_reg<M, INV>(a[r]);
#pragma unroll
        for (int k = 0; k < M; k++) {
            float2 t[4];
            t[0] = a[0][k];
#pragma unroll
            for (int r = 1; r < 4; r++) t[r] = cmul_w<INV>(a[r][k], wroot<N, INV>(r * k));
            dft_reg4<INV>(t);
#pragma unroll
            for (int q = 0; q < 4; q++) v[k + M * q] = t[q];
        }
    }
}

// Two-stage FFT geometry.
template <int P_, int Q_>
struct Fft2Geom {
    static constexpr int P = P_;
    static constexpr int Q = Q_;
    static constexpr int G = P * Q;
    static constexpr int T = P < Q ? P : Q;  // threads per transform
    static constexpr int S1 = Q / T;         // stage-1 DFTs (size P) per thread
    static constexpr int S2 = P / T;         // stage-2 DFTs (size Q) per thread
    static constexpr int LDQ = Q + 1;        // padded row of the transpose buffer
    static constexpr int XBUF = P * LDQ;     // float2 per transform in LDS
    static constexpr int PER = G / T;        // outputs held per thread
};

// Stage 1 for one transform: thread t of the transform computes the S1
// P-point DFTs n2 = t + T s, with x(n) supplying input n (called only for
// n1 < NZ, i.e. n < NZ*Q; larger inputs are zero), multiplies by the
// inter-stage twiddles (twG[m] = W_G^m forward, from LDS) and writes the
// transpose buffer xb[k1 * LDQ + n2].
template <class Geo, int NZ, bool INV, class Load>
__device__ __forceinline__ void fft2_stage1(int t, float2* __restrict__ xb,
                                            const float2* __restrict__ twG, Load x) {
    constexpr int P = Geo::P, Q = Geo::Q, G = Geo::G;
#pragma unroll
    for (int s = 0; s < Geo::S1; s++) {
        const int n2 = t + Geo::T * s;
        float2 v[P];
#pragma unroll
        for (int n1 = 0; n1 < P; n1++) v[n1] = n1 < NZ ? x(Q * n1 + n2) : make_float2(0.f, 0.f);
        dft_reg<P, INV>(v);
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) {
            float2 w = twG[(n2 * k1) & (G - 1)];
            if (INV) w.y = -w.y;
            float2 y = k1 == 0 ? v[k1] : cmul_w<false>(v[k1], w);
            xb[k1 * Geo::LDQ + n2] = y;
        }
    }
}

// Stage 2 for one transform: thread t computes k1 = t + T s and calls
// out(k1 + P k2, value) for every k2; `sidx` is the output slot index
// (s * Q + k2) for register bookkeeping.
template <class Geo, bool INV, class Store>
__device__ __forceinline__ void fft2_stage2(int t, const float2* __restrict__ xb, Store out) {
    constexpr int P = Geo::P, Q = Geo::Q;
#pragma unroll
    for (int s = 0; s < Geo::S2; s++) {
        const int k1 = t + Geo::T * s;
        float2 u[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) u[n2] = xb[k1 * Geo::LDQ + n2];
        dft_reg<Q, INV>(u);
#pragma unroll
        for (int k2 = 0; k2 < Q; k2++) out(s * Q + k2, k1 + P * k2, u[k2]);
    }
}

}  // namespace litho
