// Final full-resolution pull warp, correctMotion (src/wholistic_registration/
// utils/calFlow3d_Wei_v1.py:676-706): out(p) = trilinear(in, p + motion(p))
// with map_coordinates' nearest mode (coordinates clamped to the volume).
//
// HBM traffic per voxel: motion 12 B + output 4 B + the input's share of the
// gathers (s_in B when the field is smooth: neighbouring threads gather
// neighbouring voxels, so the 8 corner loads are served by L1/L2).

#include "wr_internal.h"

namespace wr {

namespace {

template <int IN_T>
__device__ __forceinline__ float ld(const void* p, size_t i) {
    if constexpr (IN_T == E_F32) return static_cast<const float*>(p)[i];
    else return static_cast<float>(static_cast<const uint16_t*>(p)[i]);
}

typedef float f3u __attribute__((ext_vector_type(3), aligned(4)));

// One workgroup per 256 voxels of kWarpRows consecutive rows (y, z): a thread
// issues the 12-byte motion loads of its kWarpRows voxels, then all their
// corner gathers, before any result is used (row positions from uniform
// scalar arithmetic; rows past the volume are clamped and not stored).
constexpr int kWarpRows = 4;

template <int IN_T>
__global__ __launch_bounds__(256) void k_warp(const void* __restrict__ in, int nx, int ny, int nz, int xblocks,
                                               const float* __restrict__ motion, float* __restrict__ out) {
    const unsigned b = blockIdx.x;
    const unsigned rg = b / static_cast<unsigned>(xblocks);      // group of kWarpRows rows
    const int x = static_cast<int>(b - rg * static_cast<unsigned>(xblocks)) * 256 + static_cast<int>(threadIdx.x);
    if (x >= nx) return;
    const unsigned nrows = static_cast<unsigned>(ny) * static_cast<unsigned>(nz);
    f3u m[kWarpRows];
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = min(rg * kWarpRows + j, nrows - 1);
        m[j] = *reinterpret_cast<const f3u*>(motion + 3 * (static_cast<size_t>(row) * nx + x));
    }
    float val[kWarpRows][8];
    float tt[kWarpRows][3];
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = min(rg * kWarpRows + j, nrows - 1);
        const int y = static_cast<int>(row % static_cast<unsigned>(ny));
        const int z = static_cast<int>(row / static_cast<unsigned>(ny));
        const float cx = fminf(fmaxf(static_cast<float>(x) + m[j].x, 0.0f), static_cast<float>(nx - 1));
        const float cy = fminf(fmaxf(static_cast<float>(y) + m[j].y, 0.0f), static_cast<float>(ny - 1));
        const float cz = fminf(fmaxf(static_cast<float>(z) + m[j].z, 0.0f), static_cast<float>(nz - 1));
        const int fx = static_cast<int>(floorf(cx)), fy = static_cast<int>(floorf(cy)), fz = static_cast<int>(floorf(cz));
        tt[j][0] = cx - fx;
        tt[j][1] = cy - fy;
        tt[j][2] = cz - fz;
        const int x1 = min(fx + 1, nx - 1), y1 = min(fy + 1, ny - 1), z1 = min(fz + 1, nz - 1);
        const size_t r00 = (static_cast<size_t>(fz) * ny + fy) * nx, r01 = (static_cast<size_t>(fz) * ny + y1) * nx;
        const size_t r10 = (static_cast<size_t>(z1) * ny + fy) * nx, r11 = (static_cast<size_t>(z1) * ny + y1) * nx;
        val[j][0] = ld<IN_T>(in, r00 + fx);
        val[j][1] = ld<IN_T>(in, r00 + x1);
        val[j][2] = ld<IN_T>(in, r01 + fx);
        val[j][3] = ld<IN_T>(in, r01 + x1);
        val[j][4] = ld<IN_T>(in, r10 + fx);
        val[j][5] = ld<IN_T>(in, r10 + x1);
        val[j][6] = ld<IN_T>(in, r11 + fx);
        val[j][7] = ld<IN_T>(in, r11 + x1);
    }
    auto lerp = [](float a, float c, float t) { return a + t * (c - a); };
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = rg * kWarpRows + j;
        if (row >= nrows) break;
        const float a = lerp(lerp(val[j][0], val[j][1], tt[j][0]), lerp(val[j][2], val[j][3], tt[j][0]), tt[j][1]);
        const float c = lerp(lerp(val[j][4], val[j][5], tt[j][0]), lerp(val[j][6], val[j][7], tt[j][0]), tt[j][1]);
        out[static_cast<size_t>(row) * nx + x] = lerp(a, c, tt[j][2]);
    }
}

}  // namespace

void launch_warp(const void* in, Elem in_t, int nx, int ny, int nz, const float* motion, float* out,
                 hipStream_t s) {
    const int xblocks = (nx + 255) / 256;
    const long long groups = (static_cast<long long>(ny) * nz + kWarpRows - 1) / kWarpRows;
    const unsigned blocks = static_cast<unsigned>(static_cast<long long>(xblocks) * groups);
    if (in_t == E_U16) k_warp<E_U16><<<blocks, 256, 0, s>>>(in, nx, ny, nz, xblocks, motion, out);
    else k_warp<E_F32><<<blocks, 256, 0, s>>>(in, nx, ny, nz, xblocks, motion, out);
}

}  // namespace wr
