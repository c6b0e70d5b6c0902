"""The register DFTs of csrc/fft_reg.hpp (every transform of the band and
large-grid kernels is built from them) compiled for the host with g++: each
size that divides 288, forward and inverse, against a double-precision
direct DFT, and with zero inputs past the first half (the pruned windows).
Sizes with factors 2 and 3 take the prime-factor split, the others radix
4 / 2 / 3 decimation in time; both are pinned here."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "fpga-litho-accel_amd", "csrc", "fft_reg.hpp")

# stand-ins for the HIP vector type and qualifiers the header uses
SHIM = r"""
#pragma once
#include <cmath>
#define __device__
#define __forceinline__ inline
struct float2 { float x, y; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline void sincospif(float a, float* s, float* c) {
    *s = (float)std::sin(a * 3.14159265358979323846);
    *c = (float)std::cos(a * 3.14159265358979323846);
}
"""

HARNESS = r"""
#include "fft_reg.hpp"
#include <complex>
#include <cstdio>
#include <random>
using namespace litho;
template <int N, bool INV, bool HALF>
double check() {
    std::mt19937 g(7 * N + 2 * INV + HALF);
    std::normal_distribution<float> d;
    float2 v[N];
    std::complex<double> x[N];
    for (int i = 0; i < N; i++) {
        const bool z = HALF && 2 * i >= N;
        v[i] = z ? make_float2(0.f, 0.f) : make_float2(d(g), d(g));
        x[i] = {v[i].x, v[i].y};
    }
    dft_reg<N, INV>(v);
    double e = 0, m = 0;
    for (int k = 0; k < N; k++) {
        std::complex<double> s = 0;
        for (int n = 0; n < N; n++)
            s += x[n] * std::polar(1.0, (INV ? 2.0 : -2.0) * M_PI * n * k / N);
        e = std::max(e, std::abs(s - std::complex<double>(v[k].x, v[k].y)));
        m = std::max(m, std::abs(s));
    }
    printf("N=%d inv=%d half=%d %.3e\n", N, INV, HALF, e / m);
    return e / m;
}
template <int N>
double all() {
    double w = 0;
    w = std::max(w, check<N, false, false>());
    w = std::max(w, check<N, true, false>());
    w = std::max(w, check<N, false, true>());
    w = std::max(w, check<N, true, true>());
    return w;
}
int main() {
    double w = 0;
    w = std::max(w, all<2>()); w = std::max(w, all<3>()); w = std::max(w, all<4>());
    w = std::max(w, all<6>()); w = std::max(w, all<8>()); w = std::max(w, all<9>());
    w = std::max(w, all<12>()); w = std::max(w, all<16>()); w = std::max(w, all<18>());
    w = std::max(w, all<24>()); w = std::max(w, all<32>()); w = std::max(w, all<36>());
    w = std::max(w, all<48>()); w = std::max(w, all<72>()); w = std::max(w, all<96>());
    w = std::max(w, all<144>()); w = std::max(w, all<288>());
    printf("worst %.3e\n", w);
    return w < 1e-6 ? 0 : 1;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_register_dfts_match_direct_dft(tmp_path):
    shutil.copy(SRC, tmp_path / "fft_reg.hpp")
    (tmp_path / "common.hpp").write_text(SHIM)
    (tmp_path / "t.cpp").write_text(HARNESS)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", str(tmp_path), str(tmp_path / "t.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout[-400:])
    assert r.returncode == 0, r.stdout
