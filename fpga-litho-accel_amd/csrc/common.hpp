// Shared helpers of the MI355X SOCS engine: status handling, float2 complex
// arithmetic and launch bookkeeping.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/litho_socs.h"

#define LITHO_HIP(expr)                                  \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) {                          \
            litho::set_last_hip_error(e_, #expr, __FILE__, __LINE__); \
            return LITHO_EHIP;                           \
        }                                                \
    } while (0)

#define LITHO_TRY(expr)             \
    do {                            \
        int s_ = (expr);            \
        if (s_ != LITHO_OK) return s_; \
    } while (0)

namespace litho {

void set_last_hip_error(hipError_t e, const char* what, const char* file, int line);

__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

__host__ __device__ inline double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

inline int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) r++;
    return r;
}
inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
inline int next_pow2(int v) {
    int g = 1;
    while (g < v) g <<= 1;
    return g;
}
inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace litho
