// Register-resident DFTs and a two-stage (four-step) FFT for gfx950.
//
// dft_reg<N>(v): in-register DFT of any N that divides 288 (factors 2 and 3)
// by decimation in time (radix 4, 2 or 3 recursion) with compile-time
// twiddles, so inputs known to be zero at compile time fold away.
//
// Two-stage FFT of G = P * Q points (n = Q n1 + n2, k = k1 + P k2):
//   X[k1 + P k2] = sum_n2 W_Q^{n2 k2} W_G^{n2 k1} sum_n1 x[Q n1 + n2] W_P^{n1 k1}
// Stage 1 (per n2): P-point DFT over n1, times W_G^{n2 k1}; transpose through
// LDS; stage 2 (per k1): Q-point DFT over n2.  T = min(P, Q) threads per
// transform, each holding G / T values.
#pragma once

#include "common.hpp"

namespace litho {

// cos(2*pi*j/288), j = 0..287 (roots for every N dividing 288 = 2^5 3^2)
__device__ constexpr float kCosTab[288] = {
    1.0f, 0.99976202707990913f, 0.9990482215818578f, 0.99785892323860348f,
    0.99619469809174555f, 0.99405633822231965f, 0.99144486137381038f, 0.98836151046776066f,
    0.98480775301220802f, 0.98078528040323043f, 0.97629600711993336f, 0.97134206981326143f,
    0.96592582628906831f, 0.96004985438592871f, 0.95371695074822693f, 0.94693012949510569f,
    0.93969262078590843f, 0.93200786928279855f, 0.92387953251128674f, 0.91531147911944721f,
    0.90630778703664994f, 0.89687274153268826f, 0.88701083317822171f, 0.87672675570750769f,
    0.86602540378443871f, 0.85491187067294661f, 0.84339144581288572f, 0.83146961230254524f,
    0.8191520442889918f, 0.80644460426748255f, 0.79335334029123528f, 0.77988448309288183f,
    0.76604444311897801f, 0.75183980747897738f, 0.73727733681012397f, 0.72236396205975562f,
    0.70710678118654757f, 0.6915130557822694f, 0.67559020761566035f, 0.65934581510006895f,
    0.64278760968653936f, 0.62592347218405919f, 0.60876142900872066f, 0.59130964836358246f,
    0.57357643635104616f, 0.55557023301960229f, 0.53729960834682389f, 0.51877325816052133f,
    0.50000000000000011f, 0.48098876891938774f, 0.46174861323503386f, 0.44228869021900125f,
    0.42261826174069944f, 0.40274668985873724f, 0.38268343236508984f, 0.36243803828370169f,
    0.34202014332566882f, 0.3214394653031617f, 0.30070579950427306f, 0.2798290140309922f,
    0.25881904510252096f, 0.2376858923261731f, 0.2164396139381029f, 0.19509032201612833f,
    0.17364817766693041f, 0.1521233861899168f, 0.13052619222005149f, 0.10886687485196474f,
    0.087155742747658138f, 0.065403129230143048f, 0.043619387365336007f, 0.021814885034561159f,
    6.123233995736766e-17f, -0.021814885034561034f, -0.043619387365335889f, -0.065403129230142923f,
    -0.087155742747658013f, -0.10886687485196461f, -0.13052619222005138f, -0.15212338618991669f,
    -0.1736481776669303f, -0.19509032201612819f, -0.21643961393810257f, -0.23768589232617318f,
    -0.25881904510252085f, -0.27982901403099208f, -0.30070579950427295f, -0.32143946530316136f,
    -0.34202014332566849f, -0.36243803828370158f, -0.38268343236508973f, -0.40274668985873713f,
    -0.42261826174069933f, -0.44228869021900113f, -0.46174861323503419f, -0.48098876891938747f,
    -0.49999999999999978f, -0.51877325816052122f, -0.53729960834682355f, -0.55557023301960229f,
    -0.57357643635104616f, -0.59130964836358246f, -0.60876142900872066f, -0.62592347218405875f,
    -0.64278760968653936f, -0.65934581510006884f, -0.67559020761566024f, -0.69151305578226929f,
    -0.70710678118654746f, -0.7223639620597555f, -0.73727733681012397f, -0.75183980747897727f,
    -0.7660444431189779f, -0.77988448309288161f, -0.79335334029123505f, -0.80644460426748266f,
    -0.81915204428899191f, -0.83146961230254535f, -0.8433914458128855f, -0.85491187067294638f,
    -0.86602540378443849f, -0.87672675570750769f, -0.88701083317822171f, -0.89687274153268826f,
    -0.90630778703664994f, -0.9153114791194471f, -0.92387953251128674f, -0.93200786928279844f,
    -0.93969262078590832f, -0.94693012949510558f, -0.95371695074822682f, -0.96004985438592871f,
    -0.96592582628906831f, -0.97134206981326132f, -0.97629600711993325f, -0.98078528040323043f,
    -0.98480775301220802f, -0.98836151046776066f, -0.99144486137381038f, -0.99405633822231965f,
    -0.99619469809174555f, -0.99785892323860348f, -0.9990482215818578f, -0.99976202707990913f,
    -1.0f, -0.99976202707990913f, -0.9990482215818578f, -0.99785892323860348f,
    -0.99619469809174555f, -0.99405633822231965f, -0.99144486137381049f, -0.98836151046776077f,
    -0.98480775301220813f, -0.98078528040323043f, -0.97629600711993336f, -0.97134206981326143f,
    -0.96592582628906842f, -0.96004985438592882f, -0.95371695074822693f, -0.94693012949510569f,
    -0.93969262078590843f, -0.93200786928279855f, -0.92387953251128685f, -0.91531147911944721f,
    -0.90630778703665027f, -0.89687274153268859f, -0.8870108331782216f, -0.87672675570750758f,
    -0.8660254037844386f, -0.85491187067294649f, -0.84339144581288561f, -0.83146961230254546f,
    -0.81915204428899202f, -0.80644460426748277f, -0.79335334029123539f, -0.77988448309288205f,
    -0.76604444311897835f, -0.75183980747897716f, -0.73727733681012408f, -0.72236396205975562f,
    -0.70710678118654768f, -0.69151305578226951f, -0.67559020761566035f, -0.65934581510006907f,
    -0.64278760968653947f, -0.62592347218405964f, -0.60876142900872088f, -0.59130964836358224f,
    -0.57357643635104572f, -0.55557023301960218f, -0.53729960834682422f, -0.51877325816052144f,
    -0.50000000000000044f, -0.48098876891938769f, -0.46174861323503436f, -0.44228869021900136f,
    -0.42261826174069994f, -0.40274668985873735f, -0.3826834323650895f, -0.3624380382837018f,
    -0.34202014332566855f, -0.32143946530316181f, -0.30070579950427295f, -0.27982901403099236f,
    -0.25881904510252063f, -0.23768589232617343f, -0.21643961393810368f, -0.19509032201612866f,
    -0.17364817766693033f, -0.15212338618991628f, -0.13052619222005163f, -0.1088668748519642f,
    -0.087155742747658249f, -0.065403129230143617f, -0.043619387365336132f, -0.021814885034561724f,
    -1.8369701987210297e-16f, 0.021814885034560468f, 0.043619387365335764f, 0.065403129230143256f,
    0.087155742747657888f, 0.10886687485196471f, 0.13052619222005127f, 0.15212338618991678f,
    0.17364817766692997f, 0.1950903220161283f, 0.21643961393810246f, 0.23768589232617307f,
    0.2588190451025203f, 0.27982901403099114f, 0.30070579950427345f, 0.32143946530316148f,
    0.34202014332566899f, 0.36243803828370147f, 0.38268343236509f, 0.40274668985873702f,
    0.42261826174069883f, 0.44228869021900102f, 0.46174861323503325f, 0.48098876891938736f,
    0.49999999999999933f, 0.51877325816052189f, 0.53729960834682389f, 0.55557023301960184f,
    0.57357643635104605f, 0.59130964836358202f, 0.60876142900872054f, 0.62592347218405864f,
    0.64278760968653925f, 0.6593458151000684f, 0.67559020761566013f, 0.69151305578226951f,
    0.70710678118654735f, 0.72236396205975573f, 0.73727733681012386f, 0.7518398074789775f,
    0.76604444311897779f, 0.77988448309288183f, 0.79335334029123494f, 0.8064446042674821f,
    0.81915204428899158f, 0.83146961230254479f, 0.84339144581288594f, 0.85491187067294661f,
    0.86602540378443882f, 0.87672675570750769f, 0.88701083317822149f, 0.89687274153268826f,
    0.90630778703664971f, 0.9153114791194471f, 0.92387953251128652f, 0.93200786928279844f,
    0.93969262078590843f, 0.94693012949510558f, 0.95371695074822693f, 0.9600498543859286f,
    0.96592582628906831f, 0.97134206981326132f, 0.97629600711993336f, 0.98078528040323032f,
    0.98480775301220802f, 0.98836151046776066f, 0.99144486137381027f, 0.99405633822231965f,
    0.99619469809174555f, 0.99785892323860348f, 0.9990482215818578f, 0.99976202707990913f};

// W_N^e = e^{-+2 pi i e / N} for N dividing 288 (forward: minus sign)
template <int N, bool INV>
__device__ __forceinline__ float2 wroot(int e) {
    static_assert(288 % N == 0, "root table covers divisors of 288");
    const int j = ((e * (288 / N)) % 288 + 288) % 288;
    const float c = kCosTab[j];
    const float sn = kCosTab[(j + 216) % 288];  // sin(2 pi j/288) = cos(2 pi (j-72)/288)
    return INV ? make_float2(c, sn) : make_float2(c, -sn);
}

template <bool INV>
__device__ __forceinline__ float2 cmul_w(float2 a, float2 w) {
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// a * W_N^e (the conjugate root when INV).  The roots 1, -i, -1, +i are applied
// exactly as swaps and sign changes: the table's quarter-period entries are
// round-off-sized, not zero, so a full complex product there would cost four
// VALU operations for nothing.  e is a constant once the callers' loops are
// unrolled, so only one branch survives.
template <int N, bool INV>
__device__ __forceinline__ float2 twiddle(float2 a, int e) {
    e = ((e % N) + N) % N;
    if (e == 0) return a;
    if (2 * e == N) return make_float2(-a.x, -a.y);
    if (4 * e == N)  // forward -i, inverse +i
        return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
    if (4 * e == 3 * N)  // forward +i, inverse -i
        return INV ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
    return cmul_w<INV>(a, wroot<N, INV>(e));
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

template <bool INV>
__device__ __forceinline__ void dft_reg3(float2* v) {
    const float h = 0.86602540378443865f;  // sqrt(3)/2
    float2 a = v[0], b = v[1], c = v[2];
    float2 sm = make_float2(b.x + c.x, b.y + c.y);
    float2 d = make_float2(b.x - c.x, b.y - c.y);
    float2 t = make_float2(a.x - 0.5f * sm.x, a.y - 0.5f * sm.y);
    // forward: X1 = t - i h d, X2 = t + i h d ; inverse swaps
    float2 ihd = make_float2(-h * d.y, h * d.x);
    v[0] = make_float2(a.x + sm.x, a.y + sm.y);
    if (!INV) {
        v[1] = make_float2(t.x - ihd.x, t.y - ihd.y);
        v[2] = make_float2(t.x + ihd.x, t.y + ihd.y);
    } else {
        v[1] = make_float2(t.x + ihd.x, t.y + ihd.y);
        v[2] = make_float2(t.x - ihd.x, t.y - ihd.y);
    }
}

// Split of N into coprime A = 3^a and B = 2^b (both > 1), or A = 1.
constexpr int pfa_three_part(int n) {
    int a = 1;
    while (n % 3 == 0) {
        n /= 3;
        a *= 3;
    }
    return n > 1 && a > 1 ? a : 1;
}

constexpr int inverse_mod(int x, int m) {
    for (int y = 1; y < m; y++)
        if ((x * y) % m == 1) return y;
    return 1;
}

// N = A B with gcd(A, B) = 1 by the prime-factor (Good-Thomas) mapping:
// input n = (B n1 + A n2) mod N, output k = (B (B^-1 mod A) k1 + A (A^-1 mod B) k2)
// mod N, so W_N^{nk} = W_A^{n1 k1} W_B^{n2 k2}: A B-point DFTs, then B
// A-point DFTs, and no twiddle products between them.  The index maps are
// compile-time after unrolling.
template <int N, int A, bool INV>
__device__ __forceinline__ void dft_reg_pfa(float2* v) {
    constexpr int B = N / A;
    constexpr int CA = B * inverse_mod(B % A, A), CB = A * inverse_mod(A % B, B);
    float2 y[A][B];
#pragma unroll
    for (int n1 = 0; n1 < A; n1++) {
#pragma unroll
        for (int n2 = 0; n2 < B; n2++) y[n1][n2] = v[(B * n1 + A * n2) % N];
        dft_reg<B, INV>(y[n1]);
    }
#pragma unroll
    for (int k2 = 0; k2 < B; k2++) {
        float2 z[A];
#pragma unroll
        for (int n1 = 0; n1 < A; n1++) z[n1] = y[n1][k2];
        dft_reg<A, INV>(z);
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) v[(CA * k1 + CB * k2) % N] = z[k1];
    }
}

// N = R M by decimation in time (R = 4 when 4 | N and N > 4, else 2, else 3):
// R interleaved M-point DFTs, twiddles W_N^{r k}, then R-point DFTs across.
// Sizes with both factors 2 and 3 take the prime-factor split instead.
template <int N, bool INV>
__device__ __forceinline__ void dft_reg(float2* v) {
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
        dft_reg2<INV>(v);
    } else if constexpr (N == 3) {
        dft_reg3<INV>(v);
    } else if constexpr (N == 4) {
        dft_reg4<INV>(v);
    } else if constexpr (pfa_three_part(N) > 1) {
        dft_reg_pfa<N, pfa_three_part(N), INV>(v);
    } else {
        constexpr int R = (N % 4 == 0) ? 4 : (N % 2 == 0 ? 2 : 3);
        constexpr int M = N / R;
        float2 a[R][M];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < M; m++) a[r][m] = v[R * m + r];
#pragma unroll
        for (int r = 0; r < R; r++) dft_reg<M, INV>(a[r]);
#pragma unroll
        for (int k = 0; k < M; k++) {
            float2 t[R];
            t[0] = a[0][k];
#pragma unroll
            for (int r = 1; r < R; r++) t[r] = twiddle<N, INV>(a[r][k], r * k);
            dft_reg<R, INV>(t);
#pragma unroll
            for (int q = 0; q < R; q++) v[k + M * q] = t[q];
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
// W_G^m = e^{-2 pi i m / G} for m < G (G need not be a power of two), from
// single-precision sincospi of the reduced argument 2m/G (the pi-scaled
// argument keeps the reduction exact; results within ~1 ulp).  For G a
// multiple of 4 one sincospi gives four entries: W^(m + G/4) = -i W^m,
// W^(m + G/2) = -W^m, W^(m + 3G/4) = i W^m are exact sign and part swaps.
template <int G>
__device__ __forceinline__ void fill_twiddles(float2* __restrict__ twG, int tid, int nth) {
    if constexpr (G % 4 == 0) {
        constexpr int G4 = G / 4;
        for (int i = tid; i < G4; i += nth) {
            float sn, cs;
            sincospif(2.0f * (float)i / (float)G, &sn, &cs);
            twG[i] = make_float2(cs, -sn);
            twG[i + G4] = make_float2(-sn, -cs);
            twG[i + 2 * G4] = make_float2(-cs, sn);
            twG[i + 3 * G4] = make_float2(sn, cs);
        }
    } else {
        for (int i = tid; i < G; i += nth) {
            float sn, cs;
            sincospif(2.0f * (float)i / (float)G, &sn, &cs);
            twG[i] = make_float2(cs, -sn);
        }
    }
}

template <class Geo, int NZ, bool INV, class Load>
__device__ __forceinline__ void fft2_stage1(int t, float2* __restrict__ xb,
                                            const float2* __restrict__ twG, Load x) {
    constexpr int P = Geo::P, Q = Geo::Q;
#pragma unroll
    for (int s = 0; s < Geo::S1; s++) {
        const int n2 = t + Geo::T * s;
        float2 v[P];
#pragma unroll
        for (int n1 = 0; n1 < P; n1++) v[n1] = n1 < NZ ? x(Q * n1 + n2) : make_float2(0.f, 0.f);
        dft_reg<P, INV>(v);
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) {
            float2 w = twG[n2 * k1];
            if (INV) w.y = -w.y;
            float2 y = k1 == 0 ? v[k1] : cmul_w<false>(v[k1], w);
            xb[k1 * Geo::LDQ + n2] = y;
        }
    }
}

// Stage 1 with the loader called as x(n1, n2) (compile-time n1 after
// unrolling), for inputs held in registers.
template <class Geo, int NZ, bool INV, class Load>
__device__ __forceinline__ void fft2_stage1_ij(int t, float2* __restrict__ xb,
                                               const float2* __restrict__ twG, Load x) {
    constexpr int P = Geo::P;
#pragma unroll
    for (int s = 0; s < Geo::S1; s++) {
        const int n2 = t + Geo::T * s;
        float2 v[P];
#pragma unroll
        for (int n1 = 0; n1 < P; n1++) v[n1] = n1 < NZ ? x(n1, s) : make_float2(0.f, 0.f);
        dft_reg<P, INV>(v);
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) {
            float2 w = twG[n2 * k1];
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
