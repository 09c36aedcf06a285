// Device helpers shared by the LK pass and the control-point kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wr {

// 16-bit storage of a level image: the float32 value rounded to nearest even.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);   // NaN stays NaN
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ uint16_t to_f16(float f) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(f));
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float pick4(float a0, float a1, float a2, float a3, int o) {
    return o == 0 ? a0 : (o == 1 ? a1 : (o == 2 ? a2 : a3));
}

// Bilinear CP interpolation of one axis: coordinate (v - r)/(2r+1) in float32
// (calFlow3d_Wei_v1.py:281-282), clamped to [0, n-1] (clamp to [0, n] then
// map_coordinates' nearest mode).
struct Axis1 {
    int i0, i1;
    float t;
};
__device__ __forceinline__ Axis1 cp_axis(int v, int r, int p, int n) {
    float c = __fdiv_rn(static_cast<float>(v - r), static_cast<float>(p));
    c = clampf(c, 0.0f, static_cast<float>(n - 1));
    Axis1 a;
    a.i0 = static_cast<int>(floorf(c));
    a.t = c - static_cast<float>(a.i0);
    a.i1 = min(a.i0 + 1, n - 1);
    return a;
}

__device__ __forceinline__ float cp_bilinear(const float* __restrict__ plane, int ncx, const Axis1& ax,
                                             const Axis1& ay) {
    const float v00 = plane[ay.i0 * ncx + ax.i0];
    const float v10 = plane[ay.i0 * ncx + ax.i1];
    const float v01 = plane[ay.i1 * ncx + ax.i0];
    const float v11 = plane[ay.i1 * ncx + ax.i1];
    return lerpf(lerpf(v00, v10, ax.t), lerpf(v01, v11, ax.t), ay.t);
}

}  // namespace wr
