// Elementwise and reduction kernels of the registration harness on gfx950:
//
//   transform(image, k, method)      src/wholistic_registration/utils/registration.py:26-36
//   normalize_std(mean, std, image)  src/wholistic_registration/utils/preprocess.py:288-293
//   dat_mem + dat_ca_tran            src/wholistic_registration/utils/registration.py:306
//   pick_initial_reference           src/wholistic_registration/utils/reference.py:19-99
//     (frame slices at a spatial stride, per-frame means, the demeaned Gram
//      matrix of the frames, and the mean of the chosen frames)
//
// Float32 steps are rounded one at a time (no contraction), as numpy rounds
// them.  Reductions are per-workgroup partials in a fixed order plus one
// finishing workgroup, so results do not depend on scheduling.

#include <cmath>
#include <cstdint>

#include "wr_internal.h"

#pragma clang fp contract(off)

namespace wr {

namespace {

constexpr int kT = 256;

template <int T>
__device__ __forceinline__ float as_f32(const void* p, size_t i) {
    if constexpr (T == E_U16)
        return static_cast<float>(static_cast<const uint16_t*>(p)[i]);
    else
        return static_cast<const float*>(p)[i];
}

// k * f(image): f = identity (0), sqrt (1), log2(1 + x) (2), log10(1 + x) (3).
// uint16 input: 1 + x wraps modulo 2^16 and f is evaluated in float32, as
// numpy does for uint16 arrays; an integer k in [0, 65535] with the identity
// multiplies modulo 2^16 (numpy keeps uint16), a non-integer k in float64.
template <int T>
__global__ __launch_bounds__(kT) void k_transform(const void* __restrict__ in, long long n, float kf, double kd,
                                                  int method, int k_int, float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= static_cast<size_t>(n)) return;
    float r;
    if constexpr (T == E_U16) {
        const uint16_t x = static_cast<const uint16_t*>(in)[i];
        const float xp1 = static_cast<float>(static_cast<uint16_t>(x + 1u));
        switch (method) {
            case 0:
                r = k_int ? static_cast<float>(static_cast<uint16_t>(static_cast<unsigned>(x) *
                                                                     static_cast<unsigned>(kd)))
                          : static_cast<float>(kd * static_cast<double>(x));
                break;
            case 1: r = kf * sqrtf(static_cast<float>(x)); break;
            case 2: r = kf * log2f(xp1); break;
            default: r = kf * log10f(xp1); break;
        }
    } else {
        const float x = static_cast<const float*>(in)[i];
        switch (method) {
            case 0: r = kf * x; break;
            case 1: r = kf * sqrtf(x); break;
            case 2: r = kf * log2f(1.0f + x); break;
            default: r = kf * log10f(1.0f + x); break;
        }
    }
    out[i] = r;
}

// Sum of x (pass 0) or of (x - mean)^2 in float32 steps (pass 1), per workgroup.
template <bool SQ>
__global__ __launch_bounds__(kT) void k_sum(const float* __restrict__ in, long long n, const float* __restrict__ ms,
                                            double* __restrict__ part) {
    const float mu = SQ ? ms[0] : 0.0f;
    double acc = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * kT;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x; i < static_cast<size_t>(n); i += stride) {
        const float v = in[i];
        if (SQ) {
            const float d = v - mu;
            acc += static_cast<double>(d * d);
        } else {
            acc += static_cast<double>(v);
        }
    }
    __shared__ double s[kT];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// ms[0] = mean (pass 0) or ms[1] = std (pass 1), rounded to float32.
template <bool SQ>
__global__ __launch_bounds__(kT) void k_sum_finish(const double* __restrict__ part, int nparts, long long n,
                                                   float* __restrict__ ms) {
    __shared__ double s[kT];
    double a = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kT) a += part[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = s[0] / static_cast<double>(n);
        if (SQ)
            ms[1] = static_cast<float>(sqrt(m));
        else
            ms[0] = static_cast<float>(m);
    }
}

// ((x - mean_prev) / std_prev) * std + mean, float32 steps.
__global__ __launch_bounds__(kT) void k_normalize(const float* __restrict__ in, long long n,
                                                  const float* __restrict__ prev, float mean, float std_,
                                                  float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= static_cast<size_t>(n)) return;
    const float z = (in[i] - prev[0]) / prev[1];
    out[i] = z * std_ + mean;
}

template <int T>
__global__ __launch_bounds__(kT) void k_add(const void* __restrict__ a, const float* __restrict__ b, long long n,
                                            float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= static_cast<size_t>(n)) return;
    out[i] = as_f32<T>(a, i) + b[i];
}

// Running sum of frames in the order they are added: uint16 frames into
// float64 (exact), float32 frames into float32 (numpy's sequential sum over
// the leading axis).
template <int T>
__global__ __launch_bounds__(kT) void k_accumulate(const void* __restrict__ frame, long long n, int first,
                                                   void* __restrict__ acc) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= static_cast<size_t>(n)) return;
    if constexpr (T == E_U16) {
        double* a = static_cast<double*>(acc);
        const double v = static_cast<double>(static_cast<const uint16_t*>(frame)[i]);
        a[i] = first ? v : a[i] + v;
    } else {
        float* a = static_cast<float*>(acc);
        const float v = static_cast<const float*>(frame)[i];
        a[i] = first ? v : a[i] + v;
    }
}

template <int T>
__global__ __launch_bounds__(kT) void k_divide(void* __restrict__ acc, long long n, int count) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= static_cast<size_t>(n)) return;
    if constexpr (T == E_U16) {
        double* a = static_cast<double*>(acc);
        a[i] = a[i] / static_cast<double>(count);
    } else {
        float* a = static_cast<float*>(acc);
        a[i] = a[i] / static_cast<float>(count);
    }
}

// Row t of m (T, D): frame t's slab (nzs, ny, nx) sampled at [::ds, ::ds],
// as float32, in (z, y, x) order.
template <int T>
__global__ __launch_bounds__(kT) void k_pick_gather(const void* __restrict__ slab, int nzs, int ny, int nx, int ds,
                                                    int sy, int sx, float* __restrict__ m) {
    const long long d_per = static_cast<long long>(nzs) * sy * sx;
    const long long d = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    const int t = blockIdx.y;
    if (d >= d_per) return;
    const int x = static_cast<int>(d % sx) * ds;
    const int y = static_cast<int>((d / sx) % sy) * ds;
    const int z = static_cast<int>(d / (static_cast<long long>(sx) * sy));
    const size_t src = ((static_cast<size_t>(t) * nzs + z) * ny + y) * nx + x;
    m[static_cast<size_t>(t) * d_per + d] = as_f32<T>(slab, src);
}

// Per-row mean of m (T, D) in float64, rounded to float32.
__global__ __launch_bounds__(kT) void k_row_mean(const float* __restrict__ m, long long d, float* __restrict__ mean) {
    const float* row = m + static_cast<size_t>(blockIdx.x) * d;
    double acc = 0.0;
    for (long long i = threadIdx.x; i < d; i += kT) acc += static_cast<double>(row[i]);
    __shared__ double s[kT];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[blockIdx.x] = static_cast<float>(s[0] / static_cast<double>(d));
}

// cc[i][j] = sum_d (m[i,d] - mean_i) * (m[j,d] - mean_j) for a 16 x 16 tile of
// frame pairs per workgroup: 256-column chunks of both row blocks are staged
// in LDS, each thread owns one pair and accumulates in float64 in column order.
constexpr int kGT = 16, kGC = 256;

__global__ __launch_bounds__(kT) void k_gram(const float* __restrict__ m, const float* __restrict__ mean, int nt,
                                             long long d, double* __restrict__ cc) {
    const int bi = blockIdx.x, bj = blockIdx.y;
    if (bj < bi) return;   // symmetric: the upper tiles only
    __shared__ float a[kGT][kGC + 1];
    __shared__ float b[kGT][kGC + 1];
    const int ti = threadIdx.x / kGT, tj = threadIdx.x % kGT;
    const int gi = bi * kGT + ti, gj = bj * kGT + tj;
    double acc = 0.0;
    for (long long c0 = 0; c0 < d; c0 += kGC) {
        for (int q = threadIdx.x; q < kGT * kGC; q += kT) {
            const int r = q / kGC, c = q % kGC;
            const long long col = c0 + c;
            const int ri = bi * kGT + r, rj = bj * kGT + r;
            a[r][c] = (ri < nt && col < d) ? m[static_cast<size_t>(ri) * d + col] - mean[ri] : 0.0f;
            b[r][c] = (rj < nt && col < d) ? m[static_cast<size_t>(rj) * d + col] - mean[rj] : 0.0f;
        }
        __syncthreads();
        for (int c = 0; c < kGC; ++c) acc += static_cast<double>(a[ti][c] * b[tj][c]);
        __syncthreads();
    }
    if (gi < nt && gj < nt) {
        cc[static_cast<size_t>(gi) * nt + gj] = acc;
        cc[static_cast<size_t>(gj) * nt + gi] = acc;
    }
}

unsigned blocks_of(long long n) { return static_cast<unsigned>((n + kT - 1) / kT); }

unsigned stream_grid(long long n) {
    const long long b = (n + kT - 1) / kT;
    return static_cast<unsigned>(b < kStreamBlocks ? (b > 0 ? b : 1) : kStreamBlocks);
}

}  // namespace

void launch_transform(const void* in, Elem t, long long n, double k, int method, int k_int, float* out,
                      hipStream_t s) {
    if (t == E_U16)
        k_transform<E_U16><<<blocks_of(n), kT, 0, s>>>(in, n, static_cast<float>(k), k, method, k_int, out);
    else
        k_transform<E_F32><<<blocks_of(n), kT, 0, s>>>(in, n, static_cast<float>(k), k, method, k_int, out);
}

void launch_mean_std(const float* in, long long n, double* part, float* ms, hipStream_t s) {
    const unsigned g = stream_grid(n);
    k_sum<false><<<g, kT, 0, s>>>(in, n, ms, part);
    k_sum_finish<false><<<1, kT, 0, s>>>(part, static_cast<int>(g), n, ms);
    k_sum<true><<<g, kT, 0, s>>>(in, n, ms, part);
    k_sum_finish<true><<<1, kT, 0, s>>>(part, static_cast<int>(g), n, ms);
}

void launch_normalize(const float* in, long long n, const float* prev, float mean, float std_, float* out,
                      hipStream_t s) {
    k_normalize<<<blocks_of(n), kT, 0, s>>>(in, n, prev, mean, std_, out);
}

void launch_add(const void* a, Elem ta, const float* b, long long n, float* out, hipStream_t s) {
    if (ta == E_U16)
        k_add<E_U16><<<blocks_of(n), kT, 0, s>>>(a, b, n, out);
    else
        k_add<E_F32><<<blocks_of(n), kT, 0, s>>>(a, b, n, out);
}

void launch_accumulate(const void* frame, Elem t, long long n, int first, void* acc, hipStream_t s) {
    if (t == E_U16)
        k_accumulate<E_U16><<<blocks_of(n), kT, 0, s>>>(frame, n, first, acc);
    else
        k_accumulate<E_F32><<<blocks_of(n), kT, 0, s>>>(frame, n, first, acc);
}

void launch_divide(void* acc, Elem t, long long n, int count, hipStream_t s) {
    if (t == E_U16)
        k_divide<E_U16><<<blocks_of(n), kT, 0, s>>>(acc, n, count);
    else
        k_divide<E_F32><<<blocks_of(n), kT, 0, s>>>(acc, n, count);
}

void launch_frame_gram(const void* slab, Elem t, int nt, int nzs, int ny, int nx, int ds, float* m, float* mean,
                       double* cc, hipStream_t s) {
    const int sy = (ny + ds - 1) / ds, sx = (nx + ds - 1) / ds;
    const long long d = static_cast<long long>(nzs) * sy * sx;
    const dim3 g(blocks_of(d), nt);
    if (t == E_U16)
        k_pick_gather<E_U16><<<g, kT, 0, s>>>(slab, nzs, ny, nx, ds, sy, sx, m);
    else
        k_pick_gather<E_F32><<<g, kT, 0, s>>>(slab, nzs, ny, nx, ds, sy, sx, m);
    k_row_mean<<<nt, kT, 0, s>>>(m, d, mean);
    const int tiles = (nt + kGT - 1) / kGT;
    k_gram<<<dim3(tiles, tiles), kT, 0, s>>>(m, mean, nt, d, cc);
}

}  // namespace wr
