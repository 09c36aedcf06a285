// Smoothness-penalty normaliser on gfx950: getSmPnltNormFctr
// (src/wholistic_registration/utils/preprocess.py:42-66).
//
//   Ix = (ref[x+1] - ref[x-1]) / 2,  Iy = (ref[y+1] - ref[y-1]) / 2
//   factor = (mean(Ix^2 over unmasked x in [1, nx-2]) +
//             mean(Iy^2 over unmasked y in [1, ny-2])) / 2
//
// in the reference's arithmetic for the input dtype: uint16 differences wrap
// modulo 2^16 before the true division (numpy uint16 - uint16, then / 2 as
// float64); float32 differences, halves and squares are float32; float64 is
// float64 throughout.  Squares are summed in float64 and counted, one
// workgroup per row, and the row partials are reduced in a fixed order by one
// workgroup, so the result does not depend on scheduling.

#include <cmath>
#include <cstdint>

#include "wr_internal.h"

// every square is rounded before it is summed, as in the reference
#pragma clang fp contract(off)

namespace wr {

namespace {

constexpr int kThreads = 256;

template <int IN_T>
__device__ __forceinline__ double sq_half_diff(const void* p, size_t hi, size_t lo) {
    if constexpr (IN_T == E_U16) {
        const uint16_t* v = static_cast<const uint16_t*>(p);
        const uint16_t d = static_cast<uint16_t>(v[hi] - v[lo]);
        const double h = static_cast<double>(d) / 2.0;
        return h * h;
    } else if constexpr (IN_T == E_F32) {
        const float* v = static_cast<const float*>(p);
        const float h = (v[hi] - v[lo]) / 2.0f;
        return static_cast<double>(h * h);
    } else {
        const double* v = static_cast<const double*>(p);
        const double h = (v[hi] - v[lo]) / 2.0;
        return h * h;
    }
}

// One row (y, z) per workgroup: partial sums and counts of both directions.
template <int IN_T>
__global__ __launch_bounds__(kThreads) void k_penalty_rows(const void* __restrict__ ref,
                                                          const float* __restrict__ mask, int nx, int ny,
                                                          double* __restrict__ part) {
    const unsigned row = blockIdx.x;             // z * ny + y
    const int y = static_cast<int>(row % static_cast<unsigned>(ny));
    const size_t base = static_cast<size_t>(row) * nx;
    const bool yin = y >= 1 && y <= ny - 2;
    double sx = 0.0, sy = 0.0;
    unsigned cx = 0, cy = 0;
    for (int x = threadIdx.x; x < nx; x += kThreads) {
        const size_t i = base + x;
        const bool keep = mask == nullptr || mask[i] == 0.0f;
        if (keep && x >= 1 && x <= nx - 2) {
            sx += sq_half_diff<IN_T>(ref, i + 1, i - 1);
            ++cx;
        }
        if (keep && yin) {
            sy += sq_half_diff<IN_T>(ref, i + nx, i - nx);
            ++cy;
        }
    }
    __shared__ double s_d[2][kThreads];
    __shared__ unsigned s_c[2][kThreads];
    s_d[0][threadIdx.x] = sx;
    s_d[1][threadIdx.x] = sy;
    s_c[0][threadIdx.x] = cx;
    s_c[1][threadIdx.x] = cy;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            s_d[0][threadIdx.x] += s_d[0][threadIdx.x + w];
            s_d[1][threadIdx.x] += s_d[1][threadIdx.x + w];
            s_c[0][threadIdx.x] += s_c[0][threadIdx.x + w];
            s_c[1][threadIdx.x] += s_c[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* p = part + 4 * static_cast<size_t>(row);
        p[0] = s_d[0][0];
        p[1] = s_d[1][0];
        p[2] = static_cast<double>(s_c[0][0]);
        p[3] = static_cast<double>(s_c[1][0]);
    }
}

// Fixed-order reduction of the row partials; out = {mean_x, mean_y, factor},
// and the factor also to `factor` when it is not null.
__global__ __launch_bounds__(1024) void k_penalty_finish(const double* __restrict__ part, long long nrows,
                                                         double* __restrict__ out, double* __restrict__ factor) {
    __shared__ double s[4][1024];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (long long r = threadIdx.x; r < nrows; r += 1024)
        for (int q = 0; q < 4; ++q) a[q] += part[4 * r + q];
    for (int q = 0; q < 4; ++q) s[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // an empty selection gives NaN, as numpy's mean of an empty array
        const double mx = s[2][0] > 0.0 ? s[0][0] / s[2][0] : NAN;
        const double my = s[3][0] > 0.0 ? s[1][0] / s[3][0] : NAN;
        out[0] = mx;
        out[1] = my;
        out[2] = (mx + my) / 2.0;
        if (factor) *factor = out[2];
    }
}

}  // namespace

void launch_penalty(const void* ref, Elem t, const float* mask, int nx, int ny, int nz, double* part,
                    double* out, double* factor, hipStream_t s) {
    const unsigned rows = static_cast<unsigned>(static_cast<long long>(ny) * nz);
    switch (t) {
        case E_U16:
            k_penalty_rows<E_U16><<<rows, kThreads, 0, s>>>(ref, mask, nx, ny, part);
            break;
        case E_F32:
            k_penalty_rows<E_F32><<<rows, kThreads, 0, s>>>(ref, mask, nx, ny, part);
            break;
        default:
            k_penalty_rows<E_F64><<<rows, kThreads, 0, s>>>(ref, mask, nx, ny, part);
    }
    k_penalty_finish<<<1, 1024, 0, s>>>(part, rows, out, factor);
}

}  // namespace wr
