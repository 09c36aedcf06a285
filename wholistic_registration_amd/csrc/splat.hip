// Projection of warped samples onto fixed reference z planes on gfx950: the
// two splat kernels of the cross-resolution path and the arithmetic of their
// host wrappers.
//
//   project_coords_to_fixed_planes_gpu           calFlowCrossResolution.py:480-568,878-1141
//   project_coords_to_fixed_planes_weighted_gpu  calFlowCrossResolution.py:571-669,700-875
//
// Every sample (value v, reference coordinate x, y, z) contributes to each
// target plane z0 with |z - z0| <= z_window:
//   max splat:      to the pixel nearest (x, y), or to the footprint
//                   [floor(x) - R, floor(x) + 1 + R] x [floor(y) - R, floor(y) + 1 + R]
//                   of reference pixels, each mapped to output pixel
//                   (xi / ds, yi / ds); the output keeps the maximum of the
//                   shifted values v + shift, shift = max(1 - min(v), 0), so
//                   that the float bits order like integers (atomicMax on the
//                   bits), and the shift is taken off afterwards;
//   weighted splat: a z weight (1, triangular or Gaussian in |z - z0|) times
//                   the bilinear weights of (x / ds, y / ds) on the output
//                   grid, accumulated as sum(v * w) and sum(w); the output is
//                   their ratio where sum(w) > eps.
// Unlike the reference's launch (one thread per sample and plane), a thread
// takes one sample and walks the planes, so each coordinate is read once.
// Samples are stored with z fastest, so the lanes of a wave usually hold the
// z column of one (x, y) and hit the same output pixels: when every
// contributing lane of a wave has the same footprint, the wave reduces the
// contributions first (maximum, or sums) and one lane issues one atomic per
// pixel; otherwise every lane issues its own.  Outputs are (plane, x_out,
// y_out), y fastest.

#include <climits>
#include <cmath>
#include <cstdint>

#include "wr_internal.h"

namespace wr {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ unsigned okey(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_value(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_fill_f(float* p, long long n, float v) {
    const long long i = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

// Smallest finite value (as an order-preserving key; all-ones = none).
__global__ __launch_bounds__(kT) void k_min_finite(const float* __restrict__ v, long long n,
                                                   unsigned* __restrict__ kmin) {
    unsigned m = 0xffffffffu;
    for (long long i = static_cast<long long>(blockIdx.x) * kT + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kT) {
        const float f = v[i];
        if (isfinite(f)) m = min(m, okey(f));
    }
    for (int off = 32; off > 0; off >>= 1) m = min(m, static_cast<unsigned>(__shfl_down(static_cast<int>(m), off, 64)));
    if ((threadIdx.x & 63) == 0 && m != 0xffffffffu) atomicMin(kmin, m);
}

// shift = max(1 - vmin, 0) in float32 (0 without finite values).
__global__ void k_shift(const unsigned* kmin, float* shift) {
    const unsigned k = *kmin;
    if (k == 0xffffffffu) {
        *shift = 0.0f;
    } else {
        const float s = -okey_value(k) + 1.0f;
        *shift = fmaxf(s, 0.0f);
    }
}

__global__ void k_init_key(unsigned* kmin) { *kmin = 0xffffffffu; }

struct SplatGeom {
    long long n;
    int nplanes, xref, yref, xout, yout, ds;
    float z_window;
};

__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The lane of the wave holding the first set bit of `mask`, and whether every
// set lane has the same (a, b) as that lane.
__device__ __forceinline__ bool same_key(unsigned long long mask, bool on, int a, int b, int& first) {
    first = __ffsll(static_cast<long long>(mask)) - 1;
    const int ka = __shfl(a, first, 64), kb = __shfl(b, first, 64);
    return __ballot(on && (a != ka || b != kb)) == 0ull;
}

__global__ __launch_bounds__(kT) void k_splat_max(const float* __restrict__ values, const float* __restrict__ coords,
                                                  const float* __restrict__ target_z, SplatGeom g, int nearest,
                                                  int radius, const float* __restrict__ shift,
                                                  float* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    bool ok = i < g.n;
    const float val = ok ? values[i] + *shift : 0.0f;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (ok) {
        x = coords[3 * i];
        y = coords[3 * i + 1];
        z = coords[3 * i + 2];
    }
    ok = ok && isfinite(val) && isfinite(x) && isfinite(y) && isfinite(z);
    if (!__ballot(ok)) return;           // uniform: no sample of this wave contributes
    int xs = 0, xe = -1, ys = 0, ye = -1;
    if (nearest) {
        xs = xe = static_cast<int>(floorf(x + 0.5f));
        ys = ye = static_cast<int>(floorf(y + 0.5f));
    } else {
        const int xf = static_cast<int>(floorf(x)), yf = static_cast<int>(floorf(y));
        xs = xf - radius;
        xe = xf + 1 + radius;
        ys = yf - radius;
        ye = yf + 1 + radius;
    }
    const int bits = __float_as_int(val);
    const long long plane_sz = static_cast<long long>(g.xout) * g.yout;
    const int lane = threadIdx.x & 63;
    for (int p = 0; p < g.nplanes; ++p) {
        const bool on = ok && fabsf(z - target_z[p]) <= g.z_window;
        const unsigned long long mask = __ballot(on);
        if (!mask) continue;
        int first;
        if (same_key(mask, on, xs, ys, first)) {
            // one footprint for the wave: its maximum per pixel, one atomic each
            const int xs0 = __shfl(xs, first, 64), ys0 = __shfl(ys, first, 64);
            const int xe0 = __shfl(xe, first, 64), ye0 = __shfl(ye, first, 64);
            for (int xi = xs0; xi <= xe0; ++xi) {
                if (xi < 0 || xi >= g.xref) continue;
                const int xo = xi / g.ds;
                if (xo >= g.xout) continue;
                for (int yi = ys0; yi <= ye0; ++yi) {
                    if (yi < 0 || yi >= g.yref) continue;
                    const int yo = yi / g.ds;
                    if (yo >= g.yout) continue;
                    const int m = wave_max(on ? bits : INT_MIN);
                    if (lane == first)
                        atomicMax(reinterpret_cast<int*>(out) + p * plane_sz + static_cast<long long>(xo) * g.yout + yo,
                                  m);
                }
            }
            continue;
        }
        if (!on) continue;
        for (int xi = xs; xi <= xe; ++xi) {
            if (xi < 0 || xi >= g.xref) continue;
            const int xo = xi / g.ds;
            if (xo >= g.xout) continue;
            for (int yi = ys; yi <= ye; ++yi) {
                if (yi < 0 || yi >= g.yref) continue;
                const int yo = yi / g.ds;
                if (yo >= g.yout) continue;
                atomicMax(reinterpret_cast<int*>(out) + p * plane_sz + static_cast<long long>(xo) * g.yout + yo, bits);
            }
        }
    }
}

// Shift off, empty (-inf) pixels to the fill value.
__global__ void k_max_finish(float* out, long long n, const float* shift, int fill, float fill_value) {
    const long long i = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    const float v = out[i];
    if (isfinite(v))
        out[i] = v - *shift;
    else if (fill)
        out[i] = fill_value;
}

__global__ __launch_bounds__(kT) void k_splat_weighted(const float* __restrict__ values,
                                                       const float* __restrict__ coords,
                                                       const float* __restrict__ target_z, SplatGeom g, float z_sigma,
                                                       int mode, float* __restrict__ sum_val,
                                                       float* __restrict__ sum_w) {
    const long long i = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    bool ok = i < g.n;
    const float val = ok ? values[i] : 0.0f;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (ok) {
        x = coords[3 * i];
        y = coords[3 * i + 1];
        z = coords[3 * i + 2];
    }
    ok = ok && isfinite(val) && isfinite(x) && isfinite(y) && isfinite(z) && x >= 0.0f &&
         x <= static_cast<float>(g.xref - 1) && y >= 0.0f && y <= static_cast<float>(g.yref - 1);
    if (!__ballot(ok)) return;           // uniform: no sample of this wave contributes
    const float dsf = static_cast<float>(g.ds);
    const float xo_f = x / dsf, yo_f = y / dsf;
    const int xo0 = static_cast<int>(floorf(xo_f)), yo0 = static_cast<int>(floorf(yo_f));
    const float dx = xo_f - static_cast<float>(xo0), dy = yo_f - static_cast<float>(yo0);
    const long long plane_sz = static_cast<long long>(g.xout) * g.yout;
    const int lane = threadIdx.x & 63;
    for (int p = 0; p < g.nplanes; ++p) {
        const float dz = fabsf(z - target_z[p]);
        float wz = 1.0f;
        if (mode == 1) {
            wz = 1.0f - dz / (g.z_window + 1e-6f);
            if (wz < 0.0f) wz = 0.0f;
        } else if (mode == 2) {
            const float sigma = z_sigma > 0.0f ? z_sigma : g.z_window * 0.5f;
            wz = expf(-0.5f * dz * dz / (sigma * sigma + 1e-6f));
        }
        const bool on = ok && dz <= g.z_window && wz > 0.0f && isfinite(wz);
        const unsigned long long mask = __ballot(on);
        if (!mask) continue;
        int first;
        if (same_key(mask, on, xo0, yo0, first)) {
            // one 2x2 footprint for the wave: per pixel the wave's sums, one atomic pair
            const int bx = __shfl(xo0, first, 64), by = __shfl(yo0, first, 64);
            for (int ax = 0; ax <= 1; ++ax) {
                const int xo = bx + ax;
                if (xo < 0 || xo >= g.xout) continue;
                const float wx = ax == 0 ? 1.0f - dx : dx;
                for (int ay = 0; ay <= 1; ++ay) {
                    const int yo = by + ay;
                    if (yo < 0 || yo >= g.yout) continue;
                    const float wy = ay == 0 ? 1.0f - dy : dy;
                    const float w = on ? wx * wy * wz : 0.0f;
                    const bool use = w > 0.0f && isfinite(w);
                    const float sv = wave_sum(use ? val * w : 0.0f);
                    const float sw = wave_sum(use ? w : 0.0f);
                    if (lane == first && sw > 0.0f) {
                        const long long o = p * plane_sz + static_cast<long long>(xo) * g.yout + yo;
                        atomicAdd(sum_val + o, sv);
                        atomicAdd(sum_w + o, sw);
                    }
                }
            }
            continue;
        }
        if (!on) continue;
        for (int ax = 0; ax <= 1; ++ax) {
            const int xo = xo0 + ax;
            if (xo < 0 || xo >= g.xout) continue;
            const float wx = ax == 0 ? 1.0f - dx : dx;
            for (int ay = 0; ay <= 1; ++ay) {
                const int yo = yo0 + ay;
                if (yo < 0 || yo >= g.yout) continue;
                const float wy = ay == 0 ? 1.0f - dy : dy;
                const float w = wx * wy * wz;
                if (w <= 0.0f || !isfinite(w)) continue;
                const long long o = p * plane_sz + static_cast<long long>(xo) * g.yout + yo;
                atomicAdd(sum_val + o, val * w);
                atomicAdd(sum_w + o, w);
            }
        }
    }
}

__global__ void k_weighted_finish(const float* sum_val, const float* sum_w, long long n, float eps, float fill_value,
                                  float* out) {
    const long long i = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    const float w = sum_w[i];
    out[i] = w > eps ? sum_val[i] / fmaxf(w, eps) : fill_value;
}

unsigned blocks_for(long long n) { return static_cast<unsigned>((n + kT - 1) / kT); }

}  // namespace

void launch_splat_max(const float* values, const float* coords, long long n, const float* target_z, int nplanes,
                      int xref, int yref, int xout, int yout, float z_window, int ds, int nearest, int radius,
                      int fill, float fill_value, unsigned* kmin, float* shift, float* out, hipStream_t s) {
    const SplatGeom g{n, nplanes, xref, yref, xout, yout, ds, z_window};
    const long long nout = static_cast<long long>(nplanes) * xout * yout;
    k_init_key<<<1, 1, 0, s>>>(kmin);
    const long long b = (n + kT - 1) / kT;
    k_min_finite<<<static_cast<unsigned>(b < kStreamBlocks ? (b > 0 ? b : 1) : kStreamBlocks), kT, 0, s>>>(values, n,
                                                                                                         kmin);
    k_shift<<<1, 1, 0, s>>>(kmin, shift);
    k_fill_f<<<blocks_for(nout), kT, 0, s>>>(out, nout, -INFINITY);
    if (n > 0) k_splat_max<<<blocks_for(n), kT, 0, s>>>(values, coords, target_z, g, nearest, radius, shift, out);
    k_max_finish<<<blocks_for(nout), kT, 0, s>>>(out, nout, shift, fill, fill_value);
}

void launch_splat_weighted(const float* values, const float* coords, long long n, const float* target_z,
                           int nplanes, int xref, int yref, int xout, int yout, float z_window, float z_sigma, int ds,
                           int mode, float eps, float fill_value, float* sum_val, float* sum_w, float* out,
                           hipStream_t s) {
    const SplatGeom g{n, nplanes, xref, yref, xout, yout, ds, z_window};
    const long long nout = static_cast<long long>(nplanes) * xout * yout;
    k_fill_f<<<blocks_for(nout), kT, 0, s>>>(sum_val, nout, 0.0f);
    k_fill_f<<<blocks_for(nout), kT, 0, s>>>(sum_w, nout, 0.0f);
    if (n > 0) k_splat_weighted<<<blocks_for(n), kT, 0, s>>>(values, coords, target_z, g, z_sigma, mode, sum_val, sum_w);
    k_weighted_finish<<<blocks_for(nout), kT, 0, s>>>(sum_val, sum_w, nout, eps, fill_value, out);
}

}  // namespace wr
