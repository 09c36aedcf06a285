// Final full-resolution pull warp, correctMotion (src/wholistic_registration/
// utils/calFlow3d_Wei_v1.py:676-706): out(p) = trilinear(in, p + motion(p))
// with map_coordinates' nearest mode (coordinates clamped to the volume).
//
// HBM traffic per voxel: motion 12 B + output 4 B + the input's share of the
// gathers (s_in B when the field is smooth: neighbouring threads gather
// neighbouring voxels, so the 8 corner loads are served by L1/L2).  The
// motion is read as three dword loads per thread; consecutive threads read
// consecutive 12-byte records, so every wavefront load is fully coalesced.

#include "wr_internal.h"

namespace wr {

namespace {

template <int IN_T>
__device__ __forceinline__ float ld(const void* p, size_t i) {
    if constexpr (IN_T == E_F32) return static_cast<const float*>(p)[i];
    else return static_cast<float>(static_cast<const uint16_t*>(p)[i]);
}

template <int IN_T>
__global__ __launch_bounds__(256) void k_warp(const void* __restrict__ in, int nx, int ny, int nz,
                                               const float* __restrict__ motion, float* __restrict__ out) {
    const size_t plane = static_cast<size_t>(nx) * ny;
    const size_t nvox = plane * nz;
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= nvox) return;
    const int x = static_cast<int>(v % nx);
    const int y = static_cast<int>((v / nx) % ny);
    const int z = static_cast<int>(v / plane);
    const float cx = fminf(fmaxf(static_cast<float>(x) + motion[3 * v + 0], 0.0f), static_cast<float>(nx - 1));
    const float cy = fminf(fmaxf(static_cast<float>(y) + motion[3 * v + 1], 0.0f), static_cast<float>(ny - 1));
    const float cz = fminf(fmaxf(static_cast<float>(z) + motion[3 * v + 2], 0.0f), static_cast<float>(nz - 1));
    const int fx = static_cast<int>(floorf(cx)), fy = static_cast<int>(floorf(cy)), fz = static_cast<int>(floorf(cz));
    const float tx = cx - fx, ty = cy - fy, tz = cz - fz;
    const int x1 = min(fx + 1, nx - 1), y1 = min(fy + 1, ny - 1), z1 = min(fz + 1, nz - 1);
    const size_t r00 = (static_cast<size_t>(fz) * ny + fy) * nx, r01 = (static_cast<size_t>(fz) * ny + y1) * nx;
    const size_t r10 = (static_cast<size_t>(z1) * ny + fy) * nx, r11 = (static_cast<size_t>(z1) * ny + y1) * nx;
    auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
    const float a = lerp(lerp(ld<IN_T>(in, r00 + fx), ld<IN_T>(in, r00 + x1), tx),
                         lerp(ld<IN_T>(in, r01 + fx), ld<IN_T>(in, r01 + x1), tx), ty);
    const float b = lerp(lerp(ld<IN_T>(in, r10 + fx), ld<IN_T>(in, r10 + x1), tx),
                         lerp(ld<IN_T>(in, r11 + fx), ld<IN_T>(in, r11 + x1), tx), ty);
    out[v] = lerp(a, b, tz);
}

}  // namespace

void launch_warp(const void* in, Elem in_t, int nx, int ny, int nz, const float* motion, float* out,
                 hipStream_t s) {
    const size_t nvox = static_cast<size_t>(nx) * ny * nz;
    const unsigned blocks = static_cast<unsigned>((nvox + 255) / 256);
    if (in_t == E_U16) k_warp<E_U16><<<blocks, 256, 0, s>>>(in, nx, ny, nz, motion, out);
    else k_warp<E_F32><<<blocks, 256, 0, s>>>(in, nx, ny, nz, motion, out);
}

}  // namespace wr
