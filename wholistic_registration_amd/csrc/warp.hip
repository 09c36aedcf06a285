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

typedef float f3u __attribute__((ext_vector_type(3), aligned(4)));

// One workgroup per 256 voxels of kWarpRows consecutive rows (y, z): a thread
// issues the 12-byte motion loads of its kWarpRows voxels, then all their
// corner gathers, before any result is used (row positions from uniform
// scalar arithmetic; rows past the volume are clamped and not stored).
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
    WarpTaps w[kWarpRows];
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = min(rg * kWarpRows + j, nrows - 1);
        const int y = static_cast<int>(row % static_cast<unsigned>(ny));
        const int z = static_cast<int>(row / static_cast<unsigned>(ny));
        w[j] = warp_taps<IN_T>(in, nx, ny, nz, x, y, z, m[j].x, m[j].y, m[j].z);
    }
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = rg * kWarpRows + j;
        if (row >= nrows) break;
        out[static_cast<size_t>(row) * nx + x] = warp_value(w[j]);
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
