// Cross-resolution registration (getMotion_v2) on gfx950.
//
//   spline interpolant     calFlowCrossResolution.py:463-477, 1144-1153
//   one layer's solver     :2031-2237 (optimize_layer_cross_resolution)
//   wrong-region detection :1190-1222, 1570-1787, 2306-2351
//   reference trap mask    :1363-1567
//
// The reference reads its reference level image through scipy's
// map_coordinates(order=3, mode="nearest"): the volume is padded by 12
// replicated voxels per side, prefiltered into cubic B-spline coefficients
// along each axis (Unser's recursive filter: gain 6, pole sqrt(3) - 2,
// half-sample symmetric initial conditions), and read with 4 taps per axis
// whose indices are clamped to the padded volume (coordinates themselves are
// not clamped).  Here the coefficients are built once per level (one thread
// per line, float64) and every read is a 64-tap gather.
//
// One iteration of a layer is five launches: sample (phase, reference value,
// validity, raw residual), residual (3x3x1 mean, mask, squared-sum partials),
// neighbour differences at the control points (penalty partials), then after
// the host's stop decision the per-control-point structure tensor and solve
// (one wave per control point walks its (2r+1)^2 window, recomputing the
// trilinear gradient of the reference at each voxel's phase) and the dense
// field (bilinear control-point interpolant, |delta motion| maximum).  The
// host reads two sums and one maximum per iteration: the stop rules are the
// reference's, in its float32 arithmetic.  Sums are formed per workgroup in
// a fixed order and reduced by one workgroup, so results do not depend on
// scheduling.
//
// Wrong regions: per z plane one workgroup takes the control points' local
// error medians (bitonic sort in global memory), flags the robust outliers,
// labels 4-connected components by min-label propagation, and lists each
// kept component's widened box; a voxel kernel marks the boxes.  The trap
// mask works inside the box around the seeds: closing, Gaussian smoothing
// (scipy's symmetric correlate order, float32 per axis), seed-intensity
// median and MAD by radix selection, then a bounded region growth.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "wr_internal.h"

// float32 and float64 steps are rounded one at a time, as numpy rounds them
#pragma clang fp contract(off)

namespace wr {

namespace {

constexpr int kT = 256;
constexpr int kSplinePad = 12;
constexpr long long kMaxElems = 4294967295ll;   // the kernels' 32-bit index arithmetic

unsigned blocks_for(long long n, int t = kT) {
    const long long b = (n + t - 1) / t;
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>(b, 1 << 20)));
}

__device__ __forceinline__ long long gtid() {
    return static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ long long gstride() { return static_cast<long long>(gridDim.x) * blockDim.x; }

// grids of kernels that finish with an atomic per workgroup on one address:
// few workgroups, grid-stride loops (the atomics to one address serialise)
unsigned blocks_atomic(long long n) { return std::min(blocks_for(n), 1024u); }

// the index of this lane's entry when the lanes with `pred` append to a list
// (one atomic per wave), -1 for the others
__device__ __forceinline__ long long wave_append(unsigned long long* count, bool pred) {
    const unsigned long long m = __ballot(pred);
    if (!m) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, leader, 64);
    if (!pred) return -1;
    const unsigned long long below = m & ((1ull << lane) - 1ull);
    return static_cast<long long>(base + __popcll(below));
}

__device__ __forceinline__ int clampi_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// floor of a coordinate as an int, bounded so that non-finite or huge values
// still give an in-range (clamped) tap
__device__ __forceinline__ int floor_i(double x, double* frac) {
    double f = floor(x);
    if (!(f > -1.0e8)) f = -1.0e8;
    if (f > 1.0e8) f = 1.0e8;
    *frac = x - f;
    return static_cast<int>(f);
}

// (x, y, z) of flat index i of a (nz, ny, nx) array, in 32-bit arithmetic
// (every volume here holds fewer than 2^32 elements, checked at the entry
// points; 64-bit division is many times slower)
__device__ __forceinline__ void unflat(long long i, int nx, int ny, int& x, int& y, int& z) {
    const unsigned u = static_cast<unsigned>(i);
    const unsigned t = u / static_cast<unsigned>(nx);
    x = static_cast<int>(u - t * static_cast<unsigned>(nx));
    const unsigned zz = t / static_cast<unsigned>(ny);
    y = static_cast<int>(t - zz * static_cast<unsigned>(ny));
    z = static_cast<int>(zz);
}

// ------------------------------------------------------------ block reductions

__device__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t += sh[i];
    __syncthreads();
    return t;
}

__device__ float block_max(float v, float* sh) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    float t = 0.0f;
    if (threadIdx.x == 0)
        for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t = fmaxf(t, sh[i]);
    __syncthreads();
    return t;
}

// one workgroup adds n partials in index order
__global__ void k_finish_sum(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double sh[16];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) *out = v;
}

__global__ void k_finish_max(const float* __restrict__ part, int n, float* __restrict__ out) {
    __shared__ float sh[16];
    float v = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = fmaxf(v, part[i]);
    v = block_max(v, sh);
    if (threadIdx.x == 0) *out = v;
}

__global__ void k_count_u8(const uint8_t* __restrict__ m, long long n, unsigned long long* __restrict__ out) {
    __shared__ double sh[16];
    double c = 0.0;
    for (long long i = gtid(); i < n; i += gstride()) c += m[i] ? 1.0 : 0.0;
    c = block_sum(c, sh);
    if (threadIdx.x == 0 && c > 0.0) atomicAdd(out, static_cast<unsigned long long>(c));
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------ spline interpolant

template <typename T>
__global__ void k_spline_pad(const T* __restrict__ in, int nx, int ny, int nz, double* __restrict__ c) {
    const int px = nx + 2 * kSplinePad, py = ny + 2 * kSplinePad, pz = nz + 2 * kSplinePad;
    const long long n = static_cast<long long>(px) * py * pz;
    for (long long i = gtid(); i < n; i += gstride()) {
        const int x = static_cast<int>(i % px);
        const long long t = i / px;
        const int y = static_cast<int>(t % py), z = static_cast<int>(t / py);
        const int sx = clampi_(x - kSplinePad, 0, nx - 1), sy = clampi_(y - kSplinePad, 0, ny - 1),
                  sz = clampi_(z - kSplinePad, 0, nz - 1);
        c[i] = static_cast<double>(in[(static_cast<long long>(sz) * ny + sy) * nx + sx]);
    }
}

// The cubic B-spline prefilter of every line along one axis: line l starts at
// (l / inner) * outer_stride + (l % inner) * inner_stride, elements `stride`
// apart.  Causal pass with the half-sample symmetric initial value, then the
// anticausal pass (scipy's mode "nearest" after its edge padding).
__global__ void k_spline_line(double* __restrict__ c, long long nlines, int len, long long stride, long long inner,
                              long long inner_stride, long long outer_stride, double z, double gain, double zn) {
    const long long l = gtid();
    if (l >= nlines) return;
    double* p = c + (l / inner) * outer_stride + (l % inner) * inner_stride;
    for (int i = 0; i < len; ++i) p[i * stride] *= gain;
    if (len == 1) return;
    // the causal initial value; terms past 48 are below 1e-27 of the first
    // (|z|^48) and the mirrored terms vanish with z^len for long lines
    const double c0 = p[0];
    double s = c0 + zn * p[(len - 1) * stride];
    double zi = z;
    const int nsum = len < 48 ? len : 48;
    for (int i = 1; i < nsum; ++i) {
        s += zi * (p[i * stride] + zn * p[(len - 1 - i) * stride]);
        zi *= z;
    }
    p[0] = s * z / (1.0 - zn * zn) + c0;
    for (int i = 1; i < len; ++i) p[i * stride] += z * p[(i - 1) * stride];
    p[(len - 1) * stride] *= z / (z - 1.0);
    for (int i = len - 2; i >= 0; --i) p[i * stride] = z * (p[(i + 1) * stride] - p[i * stride]);
}

struct SplineVol {
    const double* c;
    int px, py, pz;   // padded extents
};

__device__ __forceinline__ void bspline_w(double t, double w[4]) {
    const double t2 = t * t, t3 = t2 * t, u = 1.0 - t;
    w[0] = u * u * u / 6.0;
    w[1] = (4.0 - 6.0 * t2 + 3.0 * t3) / 6.0;
    w[2] = (1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3) / 6.0;
    w[3] = t3 / 6.0;
}

typedef double d2v __attribute__((ext_vector_type(2), aligned(8)));

// map_coordinates(order=3, mode="nearest") at (x, y, z) in unpadded voxel
// units.  The four x taps of a row are two 16-byte loads when none of them is
// clamped (every voxel within 10 of the volume), four 8-byte loads otherwise;
// the sum is the same either way.
__device__ double spline_at(const SplineVol& v, float fx, float fy, float fz) {
    double tx, ty, tz;
    const int x0 = floor_i(static_cast<double>(fx) + kSplinePad, &tx) - 1;
    const int y0 = floor_i(static_cast<double>(fy) + kSplinePad, &ty) - 1;
    const int z0 = floor_i(static_cast<double>(fz) + kSplinePad, &tz) - 1;
    double wx[4], wy[4], wz[4];
    bspline_w(tx, wx);
    bspline_w(ty, wy);
    bspline_w(tz, wz);
    const bool inner = x0 >= 0 && x0 + 3 <= v.px - 1;
    int ix[4];
    for (int i = 0; i < 4; ++i) ix[i] = clampi_(x0 + i, 0, v.px - 1);
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) {
        const long long zo = static_cast<long long>(clampi_(z0 + k, 0, v.pz - 1)) * v.py;
        for (int j = 0; j < 4; ++j) {
            const double* row = v.c + (zo + clampi_(y0 + j, 0, v.py - 1)) * v.px;
            const double wyz = wz[k] * wy[j];
            double t[4];
            if (inner) {
                const d2v a = *reinterpret_cast<const d2v*>(row + x0);
                const d2v b = *reinterpret_cast<const d2v*>(row + x0 + 2);
                t[0] = a.x;
                t[1] = a.y;
                t[2] = b.x;
                t[3] = b.y;
            } else {
                for (int i = 0; i < 4; ++i) t[i] = row[ix[i]];
            }
            for (int i = 0; i < 4; ++i) acc += wyz * wx[i] * t[i];
        }
    }
    return acc;
}

// points (n, 3) interleaved; the z coordinate divided by z_div in float32
// (generate_continuous_H_gpu's index conversion)
__global__ void k_spline_sample(SplineVol v, const float* __restrict__ pts, long long n, float z_div,
                                double* __restrict__ out) {
    for (long long i = gtid(); i < n; i += gstride())
        out[i] = spline_at(v, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2] / z_div);
}

// ------------------------------------------------------------ one iteration

struct Geo {
    int nx, ny, nz;     // moving level
    int rx, ry, rz;     // reference level
    long long n;        // nx * ny * nz
    long long plane;    // nx * ny
};

// phase = phase0 + motion; the reference at the phase (spline); validity;
// raw residual mov_l - sampled
__global__ __launch_bounds__(kT) void k_xr_sample(Geo g, const float* __restrict__ motion, const float* __restrict__ phase0,
                                                  SplineVol ref, SplineVol mask, int has_mask,
                                                  const uint8_t* __restrict__ mmov, const double* __restrict__ mov_l,
                                                  float* __restrict__ phase, double* __restrict__ sampled,
                                                  uint8_t* __restrict__ valid, double* __restrict__ raw) {
    const long long n = g.n;
    for (long long i = gtid(); i < n; i += gstride()) {
        const float px = phase0[i] + motion[i];
        const float py = phase0[n + i] + motion[n + i];
        const float pz = phase0[2 * n + i] + motion[2 * n + i];
        phase[i] = px;
        phase[n + i] = py;
        phase[2 * n + i] = pz;
        const double s = spline_at(ref, px, py, pz);
        bool ok = !(mmov && mmov[i] != 0);
        if (has_mask && static_cast<float>(spline_at(mask, px, py, pz)) > 0.5f) ok = false;
        if (sampled) sampled[i] = s;
        valid[i] = ok ? 1 : 0;
        raw[i] = mov_l[i] - s;
    }
}

// validity only, at a given phase, with its count
__global__ __launch_bounds__(kT) void k_xr_valid(Geo g, const float* __restrict__ phase, SplineVol mask, int has_mask,
                                                 const uint8_t* __restrict__ mmov, uint8_t* __restrict__ valid,
                                                 unsigned long long* __restrict__ count) {
    __shared__ double sh[16];
    const long long n = g.n;
    double c = 0.0;
    for (long long i = gtid(); i < n; i += gstride()) {
        bool ok = !(mmov && mmov[i] != 0);
        if (ok && has_mask && static_cast<float>(spline_at(mask, phase[i], phase[n + i], phase[2 * n + i])) > 0.5f)
            ok = false;
        if (valid) valid[i] = ok ? 1 : 0;
        c += ok ? 1.0 : 0.0;
    }
    c = block_sum(c, sh);
    if (threadIdx.x == 0 && c > 0.0) atomicAdd(count, static_cast<unsigned long long>(c));
}

// residual = 3x3x1 mean of the raw residual (replicated edges; the mean's
// float32 weights), 0 where not valid; per-workgroup sum of residual^2
__global__ __launch_bounds__(kT) void k_xr_residual(Geo g, const double* __restrict__ raw,
                                                    const uint8_t* __restrict__ valid, double w9,
                                                    double* __restrict__ res, double* __restrict__ first,
                                                    double* __restrict__ part) {
    __shared__ double sh[16];
    double acc = 0.0;
    for (long long i = gtid(); i < g.n; i += gstride()) {
        double v = 0.0;
        if (valid[i]) {
            int x, y, z;
            unflat(i, g.nx, g.ny, x, y, z);
            const long long zo = static_cast<long long>(z) * g.plane;
            for (int dy = -1; dy <= 1; ++dy) {
                const double* row = raw + zo + static_cast<long long>(clampi_(y + dy, 0, g.ny - 1)) * g.nx;
                for (int dx = -1; dx <= 1; ++dx) v += row[clampi_(x + dx, 0, g.nx - 1)] * w9;
            }
        }
        res[i] = v;
        if (first) first[i] = v;
        acc += v * v;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

struct Cps {
    int ncx, ncy, r, p;   // control points per axis, half-width, pitch 2r + 1
    long long n;          // ncx * ncy * nz
};

// neighbour difference of the field at the control points (mean of the 8
// x-y neighbours minus the centre, replicated edges, float32 like the
// reference's correlate on float32), z component times zscale in float32;
// nei_sum = float32(S) * d; per-workgroup sums of S * |d|^2 (float32 terms)
__global__ __launch_bounds__(kT) void k_xr_nei(Geo g, Cps cp, const float* __restrict__ motion, float zscale,
                                               float s_pen, float w8, float* __restrict__ nei_sum,
                                               double* __restrict__ part) {
    __shared__ double sh[16];
    double acc = 0.0;
    for (long long q = gtid(); q < cp.n; q += gstride()) {
        int i, j, z;
        unflat(q, cp.ncx, cp.ncy, i, j, z);
        float d[3];
        for (int c = 0; c < 3; ++c) {
            const float* m = motion + c * g.n + static_cast<long long>(z) * g.plane;
            double s = 0.0;
            for (int dj = -1; dj <= 1; ++dj) {
                const int yy = cp.r + clampi_(j + dj, 0, cp.ncy - 1) * cp.p;
                for (int di = -1; di <= 1; ++di) {
                    const int xx = cp.r + clampi_(i + di, 0, cp.ncx - 1) * cp.p;
                    const double w = (di == 0 && dj == 0) ? -1.0 : static_cast<double>(w8);
                    s += static_cast<double>(m[static_cast<long long>(yy) * g.nx + xx]) * w;
                }
            }
            d[c] = static_cast<float>(s);
        }
        d[2] = d[2] * zscale;
        if (nei_sum) {
            nei_sum[q] = s_pen * d[0];
            nei_sum[cp.n + q] = s_pen * d[1];
            nei_sum[2 * cp.n + q] = s_pen * d[2];
        }
        const float sq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        acc += static_cast<double>(sq * s_pen);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// trilinear map_coordinates(order=1, mode="nearest") of a float64 volume
__device__ double tri_at(const double* __restrict__ v, int rx, int ry, int rz, float fx, float fy, float fz) {
    double tx, ty, tz;
    const int x0 = floor_i(static_cast<double>(fx), &tx);
    const int y0 = floor_i(static_cast<double>(fy), &ty);
    const int z0 = floor_i(static_cast<double>(fz), &tz);
    const int xa = clampi_(x0, 0, rx - 1), xb = clampi_(x0 + 1, 0, rx - 1);
    const int ya = clampi_(y0, 0, ry - 1), yb = clampi_(y0 + 1, 0, ry - 1);
    const int za = clampi_(z0, 0, rz - 1), zb = clampi_(z0 + 1, 0, rz - 1);
    const double wx[2] = {1.0 - tx, tx}, wy[2] = {1.0 - ty, ty}, wz[2] = {1.0 - tz, tz};
    const int xs[2] = {xa, xb}, ys[2] = {ya, yb}, zs[2] = {za, zb};
    double acc = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j) {
            const double* row = v + (static_cast<long long>(zs[k]) * ry + ys[j]) * rx;
            for (int i = 0; i < 2; ++i) acc += wx[i] * wy[j] * wz[k] * row[xs[i]];
        }
    return acc;
}

// One wave per control point: the nine window sums of the gradient products
// (the reference's (2r+1)x(2r+1)x1 correlate sampled at the control point,
// :2160-2181), the regularised solve (:181-237), the step clamp, the z
// scaling and the control-point update (:2183-2190).  cp_new: 3 planes of
// float64.  Workgroups of 4 waves.
__global__ __launch_bounds__(256) void k_xr_cp_solve(Geo g, Cps cp, const float* __restrict__ phase,
                                                     const uint8_t* __restrict__ valid,
                                                     const double* __restrict__ res, const double* __restrict__ ref_l,
                                                     const float* __restrict__ motion,
                                                     const float* __restrict__ nei_sum, double s_pen,
                                                     double mov_range, double zr_hr, double* __restrict__ cp_new) {
    const long long q = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= cp.n) return;
    int i, j, z;
    unflat(q, cp.ncx, cp.ncy, i, j, z);
    const int xc = cp.r + i * cp.p, yc = cp.r + j * cp.p;
    const int w = 2 * cp.r + 1, nw = w * w;
    const float fxmax = static_cast<float>(g.rx - 1), fymax = static_cast<float>(g.ry - 1),
                fzmax = static_cast<float>(g.rz - 1);
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = lane; k < nw; k += 64) {
        const int xx = clampi_(xc - cp.r + k % w, 0, g.nx - 1);
        const int yy = clampi_(yc - cp.r + k / w, 0, g.ny - 1);
        const long long v = static_cast<long long>(z) * g.plane + static_cast<long long>(yy) * g.nx + xx;
        if (!valid[v]) continue;
        const float px = phase[v], py = phase[g.n + v], pz = phase[2 * g.n + v];
        const double gx = (tri_at(ref_l, g.rx, g.ry, g.rz, fminf(fmaxf(px + 1.0f, 0.0f), fxmax), py, pz) -
                           tri_at(ref_l, g.rx, g.ry, g.rz, fminf(fmaxf(px - 1.0f, 0.0f), fxmax), py, pz)) / 2.0;
        const double gy = (tri_at(ref_l, g.rx, g.ry, g.rz, px, fminf(fmaxf(py + 1.0f, 0.0f), fymax), pz) -
                           tri_at(ref_l, g.rx, g.ry, g.rz, px, fminf(fmaxf(py - 1.0f, 0.0f), fymax), pz)) / 2.0;
        const double gz = (tri_at(ref_l, g.rx, g.ry, g.rz, px, py, fminf(fmaxf(pz + 1.0f, 0.0f), fzmax)) -
                           tri_at(ref_l, g.rx, g.ry, g.rz, px, py, fminf(fmaxf(pz - 1.0f, 0.0f), fzmax))) / 2.0 /
                          zr_hr;
        const double it = res[v];
        s[0] += gx * gx;
        s[1] += gx * gy;
        s[2] += gx * gz;
        s[3] += gy * gy;
        s[4] += gy * gz;
        s[5] += gz * gz;
        s[6] += gx * it;
        s[7] += gy * it;
        s[8] += gz * it;
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int m = 0; m < 9; ++m) s[m] += __shfl_down(s[m], o, 64);
    if (lane != 0) return;
    const double xx = s[0] + s_pen, xy = s[1], xz = s[2], yy = s[3] + s_pen, yz = s[4], zz = s[5] + s_pen;
    const double xt = s[6] + static_cast<double>(nei_sum[q]);
    const double yt = s[7] + static_cast<double>(nei_sum[cp.n + q]);
    const double zt = s[8] + static_cast<double>(nei_sum[2 * cp.n + q]);
    const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    const double m11 = yy * zz - yz * yz;
    const double m12 = -(xy * zz - yz * xz);
    const double m13 = xy * yz - yy * xz;
    const double m22 = xx * zz - xz * xz;
    const double m23 = -(xx * yz - xy * xz);
    const double m33 = xx * yy - xy * xy;
    double v[3] = {(m11 * xt + m12 * yt + m13 * zt) / det, (m12 * xt + m22 * yt + m23 * zt) / det,
                   (m13 * xt + m23 * yt + m33 * zt) / det};
    for (int c = 0; c < 3; ++c)
        if (v[c] != v[c]) v[c] = 0.0;
    const double nrm = fmax(sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / mov_range, 1.0);
    v[0] /= nrm;
    v[1] /= nrm;
    v[2] /= nrm;
    v[2] /= zr_hr;
    const long long vc = static_cast<long long>(z) * g.plane + static_cast<long long>(yc) * g.nx + xc;
    for (int c = 0; c < 3; ++c) cp_new[c * cp.n + q] = static_cast<double>(motion[c * g.n + vc]) + v[c];
}

// The dense field from the control points: bilinear interpolation at
// (clamp((x - r)/(2r+1), 0, ncx), clamp((y - r)/(2r+1), 0, ncy), z) with the
// coordinate in float32 (:240-267) and the interpolation of the float64
// field (map_coordinates order 1, taps clamped); per-workgroup max |delta|.
__global__ __launch_bounds__(kT) void k_xr_dense(Geo g, Cps cp, const double* __restrict__ cpf,
                                                 float* __restrict__ motion, float* __restrict__ part) {
    __shared__ float sh[16];
    float mx = 0.0f;
    const float fp = static_cast<float>(cp.p), fr = static_cast<float>(cp.r);
    for (long long i = gtid(); i < g.n; i += gstride()) {
        int x, y, z;
        unflat(i, g.nx, g.ny, x, y, z);
        float cx = __fdiv_rn(static_cast<float>(x) - fr, fp);
        float cy = __fdiv_rn(static_cast<float>(y) - fr, fp);
        cx = fminf(fmaxf(cx, 0.0f), static_cast<float>(cp.ncx));
        cy = fminf(fmaxf(cy, 0.0f), static_cast<float>(cp.ncy));
        double tx, ty;
        const int x0 = floor_i(cx, &tx), y0 = floor_i(cy, &ty);
        const int xa = min(x0, cp.ncx - 1), xb = min(x0 + 1, cp.ncx - 1);
        const int ya = min(y0, cp.ncy - 1), yb = min(y0 + 1, cp.ncy - 1);
        const long long zo = static_cast<long long>(z) * cp.ncy;
        for (int c = 0; c < 3; ++c) {
            const double* f = cpf + c * cp.n;
            const double v = (1.0 - tx) * (1.0 - ty) * f[(zo + ya) * cp.ncx + xa] +
                             tx * (1.0 - ty) * f[(zo + ya) * cp.ncx + xb] +
                             (1.0 - tx) * ty * f[(zo + yb) * cp.ncx + xa] + tx * ty * f[(zo + yb) * cp.ncx + xb];
            const float nv = static_cast<float>(v);
            mx = fmaxf(mx, fabsf(nv - motion[c * g.n + i]));
            motion[c * g.n + i] = nv;
        }
    }
    mx = block_max(mx, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

// ------------------------------------------------------------ wrong regions

// local error at the control points: the k x k x 1 mean of residual^2
// (float32 weights), :1190-1222; one wave per control point
__global__ __launch_bounds__(256) void k_xr_cp_err(Geo g, Cps cp, const double* __restrict__ res, double wk, int mae,
                                                   double* __restrict__ err) {
    const long long q = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= cp.n) return;
    int i, j, z;
    unflat(q, cp.ncx, cp.ncy, i, j, z);
    const int xc = cp.r + i * cp.p, yc = cp.r + j * cp.p;
    const int w = 2 * cp.r + 1, nw = w * w;
    double s = 0.0;
    for (int k = lane; k < nw; k += 64) {
        const int xx = clampi_(xc - cp.r + k % w, 0, g.nx - 1);
        const int yy = clampi_(yc - cp.r + k / w, 0, g.ny - 1);
        const double v = res[static_cast<long long>(z) * g.plane + static_cast<long long>(yy) * g.nx + xx];
        s += (mae ? fabs(v) : v * v) * wk;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) err[q] = s;
}

// in-workgroup ascending bitonic sort of p (a power of two) doubles
__device__ void bitonic(double* a, int p) {
    for (int k = 2; k <= p; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < p; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const double x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ double sorted_median(const double* a, int n) {
    return (n & 1) ? a[n / 2] : (a[n / 2 - 1] + a[n / 2]) / 2.0;
}

struct Box {
    int x0, x1, y0, y1;
};

// One workgroup per z plane: median / MAD outliers over the (inner) control
// points, 4-connected components (min-label propagation), kept components'
// boxes.  Scratch per plane: srt (pow2 doubles), lab / cnt / 4 bbox ints per
// control point; boxes per plane up to ncx * ncy.
__global__ __launch_bounds__(1024) void k_xr_planes(Cps cp, int nx, int ny, const double* __restrict__ err,
                                                    double thr, int min_size, int pow2, double* __restrict__ srt,
                                                    int* __restrict__ lab, int* __restrict__ cnt,
                                                    int* __restrict__ bb, Box* __restrict__ boxes,
                                                    int* __restrict__ nbox, int* __restrict__ total) {
    const int z = blockIdx.x;
    const int ncx = cp.ncx, ncy = cp.ncy, npl = ncx * ncy;
    const double* e = err + static_cast<long long>(z) * npl;
    double* a = srt + static_cast<long long>(z) * pow2;
    int* L = lab + static_cast<long long>(z) * npl;
    int* C = cnt + static_cast<long long>(z) * npl;
    int* B = bb + static_cast<long long>(z) * npl * 4;
    Box* bx = boxes + static_cast<long long>(z) * npl;
    const bool inner = npl > 144 && ncx > 2 && ncy > 2;
    const int ux = inner ? ncx - 2 : ncx, uy = inner ? ncy - 2 : ncy, nu = ux * uy, off = inner ? 1 : 0;
    __shared__ double s_med, s_mad;
    __shared__ int s_changed;
    // median of the used control points
    for (int k = threadIdx.x; k < pow2; k += blockDim.x)
        a[k] = k < nu ? e[(k / ux + off) * ncx + (k % ux + off)] : INFINITY;
    __syncthreads();
    bitonic(a, pow2);
    if (threadIdx.x == 0) s_med = sorted_median(a, nu);
    __syncthreads();
    const double med = s_med;
    for (int k = threadIdx.x; k < pow2; k += blockDim.x)
        a[k] = k < nu ? fabs(e[(k / ux + off) * ncx + (k % ux + off)] - med) : INFINITY;
    __syncthreads();
    bitonic(a, pow2);
    if (threadIdx.x == 0) s_mad = sorted_median(a, nu);
    __syncthreads();
    const double mad = s_mad;
    const double den = 1.4826 * mad + 1e-12;
    for (int c = threadIdx.x; c < npl; c += blockDim.x) {
        const int i = c % ncx, j = c / ncx;
        bool sig = false;
        if (mad >= 1e-12 && i >= off && i < off + ux && j >= off && j < off + uy)
            sig = fabs(e[c] - med) / den > thr;
        L[c] = sig ? c : -1;
        C[c] = 0;
        B[4 * c] = INT32_MAX;
        B[4 * c + 1] = -1;
        B[4 * c + 2] = INT32_MAX;
        B[4 * c + 3] = -1;
    }
    __syncthreads();
    // min-label propagation over the 4-neighbourhood, with pointer jumping
    for (int it = 0; it < npl + 2; ++it) {
        if (threadIdx.x == 0) s_changed = 0;
        __syncthreads();
        for (int c = threadIdx.x; c < npl; c += blockDim.x) {
            int l = L[c];
            if (l < 0) continue;
            const int i = c % ncx, j = c / ncx;
            int m = l;
            if (i > 0 && L[c - 1] >= 0) m = min(m, L[c - 1]);
            if (i + 1 < ncx && L[c + 1] >= 0) m = min(m, L[c + 1]);
            if (j > 0 && L[c - ncx] >= 0) m = min(m, L[c - ncx]);
            if (j + 1 < ncy && L[c + ncx] >= 0) m = min(m, L[c + ncx]);
            m = min(m, L[m]);
            if (m < l) {
                L[c] = m;
                s_changed = 1;
            }
        }
        __syncthreads();
        const int ch = s_changed;
        __syncthreads();
        if (!ch) break;
    }
    for (int c = threadIdx.x; c < npl; c += blockDim.x) {
        const int l = L[c];
        if (l < 0) continue;
        atomicAdd(&C[l], 1);
        atomicMin(&B[4 * l], c % ncx);
        atomicMax(&B[4 * l + 1], c % ncx);
        atomicMin(&B[4 * l + 2], c / ncx);
        atomicMax(&B[4 * l + 3], c / ncx);
    }
    __syncthreads();
    const int dx = static_cast<int>(ceil(1.5 * cp.r));
    for (int c = threadIdx.x; c < npl; c += blockDim.x) {
        if (L[c] != c || C[c] < min_size) continue;
        Box b;
        b.x0 = max(0, cp.r + B[4 * c] * cp.p - dx);
        b.x1 = min(nx - 1, cp.r + B[4 * c + 1] * cp.p + dx);
        b.y0 = max(0, cp.r + B[4 * c + 2] * cp.p - dx);
        b.y1 = min(ny - 1, cp.r + B[4 * c + 3] * cp.p + dx);
        bx[atomicAdd(&nbox[z], 1)] = b;
        atomicAdd(total, 1);
    }
}

// bad voxels: inside a kept box of their plane (and, with mode 0, valid)
__global__ __launch_bounds__(kT) void k_xr_boxes(Geo g, int npl, const Box* __restrict__ boxes,
                                                 const int* __restrict__ nbox, const uint8_t* __restrict__ valid,
                                                 int and_valid, uint8_t* __restrict__ bad) {
    for (long long i = gtid(); i < g.n; i += gstride()) {
        int x, y, z;
        unflat(i, g.nx, g.ny, x, y, z);
        const Box* bz = boxes + static_cast<long long>(z) * npl;
        const int nb = nbox[z];
        uint8_t b = 0;
        for (int k = 0; k < nb && !b; ++k) b = (x >= bz[k].x0 && x <= bz[k].x1 && y >= bz[k].y0 && y <= bz[k].y1);
        bad[i] = (b && (!and_valid || valid[i])) ? 1 : 0;
    }
}

// 'highresidual': dense = 3x3x1 mean of (first^2 - last^2), 0 outside bad;
// the values inside bad are listed (any order) for the quantile
__global__ __launch_bounds__(kT) void k_xr_dense_err(Geo g, const double* __restrict__ first,
                                                     const double* __restrict__ last, const uint8_t* __restrict__ bad,
                                                     double w9, double* __restrict__ dense, double* __restrict__ list,
                                                     unsigned long long* __restrict__ nlist) {
    for (long long i0 = static_cast<long long>(blockIdx.x) * blockDim.x; i0 < g.n; i0 += gstride()) {
        const long long i = i0 + threadIdx.x;   // whole waves iterate together (wave_append)
        const bool in = i < g.n;
        double v = 0.0;
        if (in && bad[i]) {
            int x, y, z;
            unflat(i, g.nx, g.ny, x, y, z);
            const long long zo = static_cast<long long>(z) * g.plane;
            for (int dy = -1; dy <= 1; ++dy) {
                const long long ro = zo + static_cast<long long>(clampi_(y + dy, 0, g.ny - 1)) * g.nx;
                for (int dx = -1; dx <= 1; ++dx) {
                    const long long k = ro + clampi_(x + dx, 0, g.nx - 1);
                    v += (first[k] * first[k] - last[k] * last[k]) * w9;
                }
            }
        }
        const long long at = wave_append(nlist, in && bad[i]);
        if (at >= 0) list[at] = v;
        if (in) dense[i] = v;
    }
}

__global__ __launch_bounds__(kT) void k_xr_keep_high(Geo g, const double* __restrict__ dense,
                                                     const double* __restrict__ th, const uint8_t* __restrict__ valid,
                                                     uint8_t* __restrict__ bad) {
    const double t = *th;
    for (long long i = gtid(); i < g.n; i += gstride()) bad[i] = (bad[i] && dense[i] >= t && valid[i]) ? 1 : 0;
}

// ------------------------------------------------------------ radix selection (float64 keys)

__device__ __forceinline__ unsigned long long key64(double d) {
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(d));
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unkey64(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double(static_cast<long long>(u));
}

struct SelState {
    unsigned long long prefix, mask, rank;
    double value;
};

__global__ void k_sel_init(SelState* st, const unsigned long long* n, double q_or_neg, long long rank) {
    // rank >= 0: that rank; otherwise floor((n - 1) * q) clipped to [0, n - 1]
    long long k = rank;
    if (rank < 0) {
        const long long nn = static_cast<long long>(*n);
        k = static_cast<long long>(floor(static_cast<double>(nn - 1) * q_or_neg));
        k = k < 0 ? 0 : (k > nn - 1 ? nn - 1 : k);
    }
    st->prefix = 0;
    st->mask = 0;
    st->rank = static_cast<unsigned long long>(k);
}

constexpr int kSelBins = 2048;   // 11-bit digits: 64 = 5 x 11 + 9

// per-workgroup LDS histogram of the digit at `shift` (`bits` wide) over the
// values whose higher digits equal the prefix, added to the global one
__global__ __launch_bounds__(kT) void k_sel_hist(const double* __restrict__ v, const unsigned long long* __restrict__ n,
                                                 const SelState* __restrict__ st, int shift, int bits,
                                                 unsigned* __restrict__ hist) {
    __shared__ unsigned h[kSelBins];
    for (int b = threadIdx.x; b < kSelBins; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const long long nn = static_cast<long long>(*n);
    const unsigned long long pre = st->prefix, msk = st->mask, dm = (1ull << bits) - 1ull;
    for (long long i = gtid(); i < nn; i += gstride()) {
        const unsigned long long k = key64(v[i]);
        if ((k & msk) == pre) atomicAdd(&h[(k >> shift) & dm], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSelBins; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// one workgroup of 1024: 2 bins per thread, prefix over the threads, the bin
// holding the rank; clears the histogram for the next pass
__global__ __launch_bounds__(1024) void k_sel_pick(unsigned* __restrict__ hist, int shift, int bits, SelState* st,
                                                   int last) {
    constexpr int kPer = kSelBins / 1024;
    __shared__ unsigned long long sc[1024];
    const int t = threadIdx.x;
    const unsigned long long rank = st->rank;   // read before any thread can update it
    unsigned long long local = 0;
    for (int b = 0; b < kPer; ++b) local += hist[t * kPer + b];
    sc[t] = local;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned long long v = t >= o ? sc[t - o] : 0ull;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    const unsigned long long before = t > 0 ? sc[t - 1] : 0ull;
    if (rank >= before && rank < sc[t]) {
        unsigned long long acc = before;
        for (int b = 0; b < kPer; ++b) {
            const unsigned h = hist[t * kPer + b];
            if (rank < acc + h) {
                const unsigned long long bin = static_cast<unsigned long long>(t * kPer + b);
                st->prefix |= bin << shift;
                st->mask |= ((1ull << bits) - 1ull) << shift;
                st->rank = rank - acc;
                if (last) st->value = unkey64(st->prefix);
                break;
            }
            acc += h;
        }
    }
    __syncthreads();
    for (int b = 0; b < kPer; ++b) hist[t * kPer + b] = 0u;
}

// the rank-th smallest of the n (device count) values into st->value; the
// histogram (kSelBins uints) must be zero on entry and is zero on exit
void select_rank(const double* v, const unsigned long long* n, long long rank, double q, SelState* st,
                 unsigned* hist, long long n_bound, hipStream_t s) {
    k_sel_init<<<1, 1, 0, s>>>(st, n, q, rank);
    const unsigned g = std::min(blocks_for(n_bound), 512u);
    const int shifts[6] = {53, 42, 31, 20, 9, 0}, bits[6] = {11, 11, 11, 11, 11, 9};
    for (int p = 0; p < 6; ++p) {
        k_sel_hist<<<g, kT, 0, s>>>(v, n, st, shifts[p], bits[p], hist);
        k_sel_pick<<<1, 1024, 0, s>>>(hist, shifts[p], bits[p], st, p == 5);
    }
}

// ------------------------------------------------------------ trap mask

struct Roi {
    int x0, y0, z0, wx, wy, wz;
    long long n;
};

// the seeds' bounding box and count: per-thread, then per-workgroup in LDS,
// then one atomic per quantity and workgroup
__global__ __launch_bounds__(kT) void k_trap_bbox(Geo g, const uint8_t* __restrict__ bad, const float* __restrict__ phase,
                                                  int* __restrict__ bb) {
    int v[7] = {INT32_MAX, -1, INT32_MAX, -1, INT32_MAX, -1, 0};
    for (long long i = gtid(); i < g.n; i += gstride()) {
        if (!bad[i]) continue;
        const int sx = static_cast<int>(rintf(phase[i]));
        const int sy = static_cast<int>(rintf(phase[g.n + i]));
        const int sz = static_cast<int>(rintf(phase[2 * g.n + i]));
        if (sx < 0 || sx >= g.rx || sy < 0 || sy >= g.ry || sz < 0 || sz >= g.rz) continue;
        v[0] = min(v[0], sx);
        v[1] = max(v[1], sx);
        v[2] = min(v[2], sy);
        v[3] = max(v[3], sy);
        v[4] = min(v[4], sz);
        v[5] = max(v[5], sz);
        v[6] += 1;
    }
    __shared__ int sh[7];
    if (threadIdx.x < 6) sh[threadIdx.x] = (threadIdx.x & 1) ? -1 : INT32_MAX;
    if (threadIdx.x == 6) sh[6] = 0;
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) {   // wave reduction first: one LDS update per wave
        v[0] = min(v[0], __shfl_down(v[0], o, 64));
        v[1] = max(v[1], __shfl_down(v[1], o, 64));
        v[2] = min(v[2], __shfl_down(v[2], o, 64));
        v[3] = max(v[3], __shfl_down(v[3], o, 64));
        v[4] = min(v[4], __shfl_down(v[4], o, 64));
        v[5] = max(v[5], __shfl_down(v[5], o, 64));
        v[6] += __shfl_down(v[6], o, 64);
    }
    if ((threadIdx.x & 63) == 0 && v[6] > 0) {
        atomicMin(&sh[0], v[0]);
        atomicMax(&sh[1], v[1]);
        atomicMin(&sh[2], v[2]);
        atomicMax(&sh[3], v[3]);
        atomicMin(&sh[4], v[4]);
        atomicMax(&sh[5], v[5]);
        atomicAdd(&sh[6], v[6]);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sh[6] > 0) {
        atomicMin(&bb[0], sh[0]);
        atomicMax(&bb[1], sh[1]);
        atomicMin(&bb[2], sh[2]);
        atomicMax(&bb[3], sh[3]);
        atomicMin(&bb[4], sh[4]);
        atomicMax(&bb[5], sh[5]);
        atomicAdd(&bb[6], sh[6]);
    }
}

// bb: min x, max x, min y, max y, min z, max z, seed count
__global__ void k_trap_bbox_init(int* bb) {
    const int t = threadIdx.x;
    if (t < 6) bb[t] = (t & 1) ? -1 : INT32_MAX;
    if (t == 6) bb[6] = 0;
}

__global__ __launch_bounds__(kT) void k_trap_scatter(Geo g, const uint8_t* __restrict__ bad, const float* __restrict__ phase,
                                                     Roi roi, uint8_t* __restrict__ seed) {
    for (long long i = gtid(); i < g.n; i += gstride()) {
        if (!bad[i]) continue;
        const int sx = static_cast<int>(rintf(phase[i]));
        const int sy = static_cast<int>(rintf(phase[g.n + i]));
        const int sz = static_cast<int>(rintf(phase[2 * g.n + i]));
        if (sx < 0 || sx >= g.rx || sy < 0 || sy >= g.ry || sz < 0 || sz >= g.rz) continue;
        seed[(static_cast<long long>(sz - roi.z0) * roi.wy + (sy - roi.y0)) * roi.wx + (sx - roi.x0)] = 1;
    }
}

// binary dilation (erode = 0) or erosion (erode = 1) of a ROI byte volume by a
// symmetric structuring element given as offsets; outside the ROI counts as
// 0 (scipy's border_value 0).  With `allowed`, the dilation is intersected
// with it and `changed` is set where the result differs from the input.
__global__ __launch_bounds__(kT) void k_morph(Roi roi, const uint8_t* __restrict__ in, const int* __restrict__ off,
                                              int noff, int erode, const uint8_t* __restrict__ allowed,
                                              uint8_t* __restrict__ out, int* __restrict__ changed) {
    for (long long i = gtid(); i < roi.n; i += gstride()) {
        int x, y, z;
        unflat(i, roi.wx, roi.wy, x, y, z);
        bool r = erode ? true : false;
        for (int k = 0; k < noff; ++k) {
            const int xx = x + off[3 * k], yy = y + off[3 * k + 1], zz = z + off[3 * k + 2];
            const bool inside = xx >= 0 && xx < roi.wx && yy >= 0 && yy < roi.wy && zz >= 0 && zz < roi.wz;
            const bool v = inside && in[(static_cast<long long>(zz) * roi.wy + yy) * roi.wx + xx];
            if (erode) {
                if (!v) {
                    r = false;
                    break;
                }
            } else if (v) {
                r = true;
                break;
            }
        }
        if (allowed) r = r && allowed[i];
        const uint8_t o = r ? 1 : 0;
        if (changed) {
            const unsigned long long m = __ballot(o != in[i]);
            if (m && (threadIdx.x & 63) == static_cast<unsigned>(__ffsll(static_cast<long long>(m)) - 1))
                atomicOr(changed, 1);
        }
        out[i] = o;
    }
}

__global__ __launch_bounds__(kT) void k_trap_crop(Roi roi, const double* __restrict__ ref_l, int rx, int ry,
                                                  float* __restrict__ out) {
    for (long long i = gtid(); i < roi.n; i += gstride()) {
        int x, y, z;
        unflat(i, roi.wx, roi.wy, x, y, z);
        out[i] = static_cast<float>(
            ref_l[(static_cast<long long>(z + roi.z0) * ry + (y + roi.y0)) * rx + (x + roi.x0)]);
    }
}

__device__ __forceinline__ int reflect_idx(int i, int n) {
    const int p = 2 * n;
    i %= p;
    if (i < 0) i += p;
    return i >= n ? p - 1 - i : i;
}

// one axis of the Gaussian filter: scipy's symmetric correlate order
// (centre term, then the pairs from the outermost in), float64 sum, float32 out,
// mode "reflect"
__global__ __launch_bounds__(kT) void k_gauss_axis(Roi roi, const float* __restrict__ in, int axis, const double* w,
                                                   int rad, float* __restrict__ out) {
    const int dims[3] = {roi.wx, roi.wy, roi.wz};
    const long long strides[3] = {1, roi.wx, static_cast<long long>(roi.wx) * roi.wy};
    const int n = dims[axis];
    const long long st = strides[axis];
    for (long long i = gtid(); i < roi.n; i += gstride()) {
        int x, y, z;
        unflat(i, roi.wx, roi.wy, x, y, z);
        const int pos = axis == 0 ? x : (axis == 1 ? y : z);
        const long long base = i - static_cast<long long>(pos) * st;
        double acc = static_cast<double>(in[i]) * w[rad];
        for (int j = rad; j >= 1; --j)
            acc += (static_cast<double>(in[base + reflect_idx(pos - j, n) * st]) +
                    static_cast<double>(in[base + reflect_idx(pos + j, n) * st])) *
                   w[rad + j];
        out[i] = static_cast<float>(acc);
    }
}

// values (as float64) of `vals` where `m` is set, listed in any order
__global__ __launch_bounds__(kT) void k_list_masked(Roi roi, const float* __restrict__ vals, const uint8_t* __restrict__ m,
                                                    double* __restrict__ list, unsigned long long* __restrict__ n) {
    for (long long i0 = static_cast<long long>(blockIdx.x) * blockDim.x; i0 < roi.n; i0 += gstride()) {
        const long long i = i0 + threadIdx.x;   // whole waves iterate together (wave_append)
        const bool set = i < roi.n && m[i];
        const long long at = wave_append(n, set);
        if (at >= 0) list[at] = static_cast<double>(vals[i]);
    }
}

// |v - med| in float32 for the listed values, in place
__global__ __launch_bounds__(kT) void k_abs_dev(double* __restrict__ list, const unsigned long long* __restrict__ n,
                                                const float* __restrict__ med) {
    const long long nn = static_cast<long long>(*n);
    const float m = *med;
    for (long long i = gtid(); i < nn; i += gstride()) list[i] = static_cast<double>(fabsf(static_cast<float>(list[i]) - m));
}

// numpy's float32 median from the two middle order statistics (or one)
__global__ void k_median_f32(const SelState* lo, const SelState* hi, const unsigned long long* n, float* out) {
    const float a = static_cast<float>(lo->value), b = static_cast<float>(hi->value);
    *out = (*n & 1ull) ? b : (a + b) / 2.0f;
}

__global__ __launch_bounds__(kT) void k_trap_allowed(Roi roi, const uint8_t* __restrict__ near,
                                                     const float* __restrict__ sm, const float* __restrict__ med,
                                                     float tol, const uint8_t* __restrict__ seed,
                                                     uint8_t* __restrict__ allowed) {
    const float m = *med;
    for (long long i = gtid(); i < roi.n; i += gstride())
        allowed[i] = ((near[i] && fabsf(sm[i] - m) <= tol) || seed[i]) ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_copy_mask(long long n, const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
    for (long long i = gtid(); i < n; i += gstride()) out[i] = (in && in[i]) ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_trap_write(Roi roi, const uint8_t* __restrict__ region,
                                                   const uint8_t* __restrict__ seed, int rx, int ry,
                                                   uint8_t* __restrict__ out, unsigned long long* __restrict__ count) {
    __shared__ double sh[16];
    double c = 0.0;
    for (long long i = gtid(); i < roi.n; i += gstride()) {
        if (!(region[i] || seed[i])) continue;
        int x, y, z;
        unflat(i, roi.wx, roi.wy, x, y, z);
        out[(static_cast<long long>(z + roi.z0) * ry + (y + roi.y0)) * rx + (x + roi.x0)] = 1;
        c += 1.0;
    }
    c = block_sum(c, sh);
    if (threadIdx.x == 0 && c > 0.0) atomicAdd(count, static_cast<unsigned long long>(c));
}

// ------------------------------------------------------------ element-wise set-up

__global__ __launch_bounds__(kT) void k_xr_scale(const double* __restrict__ in, long long n, double mul, double div,
                                                 float* __restrict__ out) {
    for (long long i = gtid(); i < n; i += gstride()) out[i] = static_cast<float>(in[i] * mul / div);
}

__global__ __launch_bounds__(kT) void k_xr_threshold(const double* __restrict__ in, long long n, double thr,
                                                     uint8_t* __restrict__ out) {
    for (long long i = gtid(); i < n; i += gstride()) out[i] = in[i] > thr ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_xr_phase(int nx, int ny, int nz, double zr, double zr_hr, float* __restrict__ ph) {
    const long long n = static_cast<long long>(nx) * ny * nz;
    for (long long i = gtid(); i < n; i += gstride()) {
        int x, y, z;
        unflat(i, nx, ny, x, y, z);
        ph[i] = static_cast<float>(x);
        ph[n + i] = static_cast<float>(y);
        ph[2 * n + i] = static_cast<float>(static_cast<double>(z) * zr / zr_hr);
    }
}

// ------------------------------------------------------------ z-offset search

// Sobel gradient magnitude of every slice (calculate.py:109-112): slices of h
// (fastest) x w, each step rounded to float32 as scipy writes it: the
// difference along one axis, then the [1, 2, 1] smoothing along the other
// (replicated edges), then sqrt(gx^2 + gy^2).
__global__ __launch_bounds__(kT) void k_fz_gradmag(const float* __restrict__ in, int h, int w, long long n,
                                                   float* __restrict__ out) {
    for (long long i = gtid(); i < n; i += gstride()) {
        int a, b, sl;
        unflat(i, h, w, a, b, sl);   // a along h (axis 0), b along w (axis 1)
        const float* p = in + static_cast<long long>(sl) * h * w;
        auto at = [&](int aa, int bb) { return p[static_cast<long long>(clampi_(bb, 0, w - 1)) * h + clampi_(aa, 0, h - 1)]; };
        // derivative along w, smoothed along h
        float d1[3];
        for (int k = -1; k <= 1; ++k) d1[k + 1] = at(a + k, b + 1) - at(a + k, b - 1);
        const float gx = static_cast<float>(static_cast<double>(d1[1]) * 2.0 +
                                            (static_cast<double>(d1[0]) + static_cast<double>(d1[2])));
        // derivative along h, smoothed along w
        float d0[3];
        for (int k = -1; k <= 1; ++k) d0[k + 1] = at(a + 1, b + k) - at(a - 1, b + k);
        const float gy = static_cast<float>(static_cast<double>(d0[1]) * 2.0 +
                                            (static_cast<double>(d0[0]) + static_cast<double>(d0[2])));
        out[i] = sqrtf(gx * gx + gy * gy);
    }
}

// per slice: weighted mean and weighted variance about it (float64 sums)
__global__ __launch_bounds__(1024) void k_fz_stats(const float* __restrict__ f, const float* __restrict__ wt, long long hw,
                                                   double* __restrict__ mean, double* __restrict__ var) {
    __shared__ double sh[16];
    const float* p = f + static_cast<long long>(blockIdx.x) * hw;
    double s = 0.0;
    for (long long i = threadIdx.x; i < hw; i += blockDim.x) s += static_cast<double>(wt[i]) * p[i];
    s = block_sum(s, sh);
    __shared__ double m;
    if (threadIdx.x == 0) m = s;
    __syncthreads();
    const float mf = static_cast<float>(m);
    double v = 0.0;
    for (long long i = threadIdx.x; i < hw; i += blockDim.x) {
        const float c = p[i] - mf;
        v += static_cast<double>(wt[i]) * c * c;
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        mean[blockIdx.x] = m;
        var[blockIdx.x] = v;
    }
}

// one workgroup per (moving slice k, reference slice z): the weighted ZNCC
__global__ __launch_bounds__(1024) void k_fz_zncc(const float* __restrict__ mf, const float* __restrict__ rf,
                                                  const float* __restrict__ wt, long long hw, int nz,
                                                  const double* __restrict__ mmean, const double* __restrict__ mvar,
                                                  const double* __restrict__ rmean, const double* __restrict__ rvar,
                                                  double eps, float* __restrict__ scores) {
    __shared__ double sh[16];
    const int k = blockIdx.x / nz, z = blockIdx.x % nz;
    const float* a = mf + static_cast<long long>(k) * hw;
    const float* b = rf + static_cast<long long>(z) * hw;
    const float ma = static_cast<float>(mmean[k]), mb = static_cast<float>(rmean[z]);
    double s = 0.0;
    for (long long i = threadIdx.x; i < hw; i += blockDim.x)
        s += static_cast<double>(wt[i]) * static_cast<double>(a[i] - ma) * static_cast<double>(b[i] - mb);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const float mv = static_cast<float>(mvar[k]), rv = static_cast<float>(rvar[z]);
        const float num = static_cast<float>(s);
        float sc = num / (sqrtf(mv * rv) + static_cast<float>(eps));
        sc = fminf(fmaxf(sc, -1.0f), 1.0f);
        scores[blockIdx.x] = sc;
    }
}

// ------------------------------------------------------------ host helpers

int n_cp_axis(int n, int r) {
    const int p = 2 * r + 1;
    const int span = (n - 1) - r;
    return span > 0 ? (span + p - 1) / p : 0;
}

SplineVol spline_vol(const double* c, int nx, int ny, int nz) {
    return SplineVol{c, nx + 2 * kSplinePad, ny + 2 * kSplinePad, nz + 2 * kSplinePad};
}

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// carve consecutive 256-byte aligned pieces of a DevBuf
struct Carve {
    char* base;
    size_t off = 0;
    template <class T> T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base + off);
        off += align256(count * sizeof(T));
        return p;
    }
};

// offsets of scipy-style ball structuring elements (see make_ball in the
// Python module): |dx|, |dy| <= R, |dz| <= Rz, dx^2 + dy^2 + (dz zr)^2 <=
// R^2 + 1e-6 evaluated in float32 as the reference does
std::vector<int> ball_offsets(int radius_xy, double z_ratio) {
    const int R = std::max(1, radius_xy);
    const int rz = static_cast<int>(std::max(1.0, std::min(std::ceil(R / z_ratio), 64.0)));
    std::vector<int> off;
    const float zr = static_cast<float>(z_ratio);
    for (int dz = -rz; dz <= rz; ++dz)
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                const float fx = static_cast<float>(dx), fy = static_cast<float>(dy), fz = static_cast<float>(dz);
                const float zz = fz * zr;
                const float d2 = fx * fx + fy * fy + zz * zz;
                if (d2 <= static_cast<float>(static_cast<double>(R) * R + 1e-6)) {
                    off.push_back(dx);
                    off.push_back(dy);
                    off.push_back(dz);
                }
            }
    return off;
}

// numpy's sum of the 9 or fewer Gaussian weights: pairwise in blocks of 8,
// then the rest in order
double np_sum(const std::vector<double>& a) {
    const size_t n = a.size();
    if (n < 8) {
        double s = 0.0;
        for (double v : a) s += v;
        return s;
    }
    double r[8];
    for (int k = 0; k < 8; ++k) r[k] = a[k];
    size_t i = 8;
    for (; i + 8 <= n - (n % 8) && i < n - (n % 8); i += 8)
        for (int k = 0; k < 8; ++k) r[k] += a[i + k];
    double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) s += a[i];
    return s;
}

}  // namespace

}  // namespace wr

using namespace wr;

extern "C" {

int wr_spline_coef(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz, double* coef, void* stream) {
    WR_ARG(ctx && in && coef, "wr_spline_coef: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_spline_coef: empty shape");
    WR_ARG(dtype == WR_F64 || dtype == WR_F32 || dtype == WR_U8, "wr_spline_coef: dtype must be WR_F64, WR_F32 or WR_U8");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int px = nx + 2 * kSplinePad, py = ny + 2 * kSplinePad, pz = nz + 2 * kSplinePad;
    const long long n = static_cast<long long>(px) * py * pz;
    WR_ARG(n <= kMaxElems, "wr_spline_coef: padded volume above 2^32 elements");
    const unsigned g = blocks_for(n);
    if (dtype == WR_F64)
        k_spline_pad<double><<<g, kT, 0, s>>>(static_cast<const double*>(in), nx, ny, nz, coef);
    else if (dtype == WR_F32)
        k_spline_pad<float><<<g, kT, 0, s>>>(static_cast<const float*>(in), nx, ny, nz, coef);
    else
        k_spline_pad<uint8_t><<<g, kT, 0, s>>>(static_cast<const uint8_t*>(in), nx, ny, nz, coef);
    const double z = std::sqrt(3.0) - 2.0;
    const double gain = (1.0 - z) * (1.0 - 1.0 / z);
    // lines along x, y, z
    const long long plane = static_cast<long long>(px) * py;
    struct Ax {
        long long nlines;
        int len;
        long long stride, inner, inner_stride, outer_stride;
    } axes[3] = {{static_cast<long long>(py) * pz, px, 1, 1, 0, px},
                 {static_cast<long long>(pz) * px, py, px, px, 1, plane},
                 {plane, pz, plane, plane, 1, 0}};
    for (const Ax& a : axes) {
        const double zn = std::pow(z, a.len);
        k_spline_line<<<blocks_for(a.nlines), kT, 0, s>>>(coef, a.nlines, a.len, a.stride, a.inner, a.inner_stride,
                                                           a.outer_stride, z, gain, zn);
    }
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_spline_sample(wr_ctx* ctx, const double* coef, int nx, int ny, int nz, const float* pts, long long n,
                     double z_ratio, double* out, void* stream) {
    WR_ARG(ctx && coef && pts && out, "wr_spline_sample: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0 && n >= 0, "wr_spline_sample: empty shape");
    WR_ARG(z_ratio != 0.0, "wr_spline_sample: zRatio must be nonzero");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    if (n == 0) return WR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    k_spline_sample<<<blocks_for(n), kT, 0, s>>>(spline_vol(coef, nx, ny, nz), pts, n, static_cast<float>(z_ratio),
                                                  out);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_xres_optimize(wr_ctx* ctx, const wr_xres_layer* L, const double* mov_l, const double* ref_l,
                     const double* ref_coef, const double* mask_coef, const uint8_t* mask_mov,
                     const float* phase0, const float* motion_init, const int* forced, float* motion,
                     float* phase, double* sampled, double* residual, uint8_t* valid, double* first_residual,
                     wr_xres_trace* trace, void* stream) {
    WR_ARG(ctx && L && mov_l && ref_l && ref_coef && phase0 && motion_init && motion && phase && residual && valid &&
               trace,
           "wr_xres_optimize: null argument");
    WR_ARG(L->nx > 0 && L->ny > 0 && L->nz > 0 && L->rx > 0 && L->ry > 0 && L->rz > 0,
           "wr_xres_optimize: empty level");
    WR_ARG(L->r >= 1 && L->iters >= 1, "wr_xres_optimize: r and iter must be positive");
    WR_ARG(L->z_ratio_hr > 0.0 && L->mov_range > 0.0, "wr_xres_optimize: zRatio_HR and movRange must be positive");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    Geo g{L->nx, L->ny, L->nz, L->rx, L->ry, L->rz, static_cast<long long>(L->nx) * L->ny * L->nz,
          static_cast<long long>(L->nx) * L->ny};
    Cps cp{n_cp_axis(L->nx, L->r), n_cp_axis(L->ny, L->r), L->r, 2 * L->r + 1, 0};
    WR_ARG(cp.ncx > 0 && cp.ncy > 0, "wr_xres_optimize: the level holds no control point (x or y below r + 2)");
    WR_ARG(g.n <= kMaxElems, "wr_xres_optimize: level above 2^32 voxels");
    cp.n = static_cast<long long>(cp.ncx) * cp.ncy * L->nz;
    const unsigned gv = blocks_for(g.n), gc = blocks_for(cp.n);
    const unsigned gw = static_cast<unsigned>((cp.n + 3) / 4);
    // scratch: raw residual, nei_sum, cp field, partials, scalars
    size_t bytes = align256(g.n * sizeof(double)) + align256(3 * cp.n * sizeof(float)) +
                   align256(3 * cp.n * sizeof(double)) + align256(gv * sizeof(double)) +
                   align256(gc * sizeof(double)) + align256(gv * sizeof(float)) + align256(64);
    WR_TRY(grow(c.xws, bytes));
    Carve cv{c.xws.as<char>()};
    double* raw = cv.take<double>(g.n);
    float* nei = cv.take<float>(3 * cp.n);
    double* cpf = cv.take<double>(3 * cp.n);
    double* pres = cv.take<double>(gv);
    double* ppen = cv.take<double>(gc);
    float* pmax = cv.take<float>(gv);
    double* scal = cv.take<double>(4);
    const SplineVol rv = spline_vol(ref_coef, L->rx, L->ry, L->rz);
    const SplineVol mv = spline_vol(mask_coef ? mask_coef : ref_coef, L->rx, L->ry, L->rz);
    const double w9 = static_cast<double>(1.0f / 9.0f);
    const float w8 = 1.0f / 8.0f;
    const float s_pen_f = static_cast<float>(L->smooth_penalty_sum);
    const float zr_f = static_cast<float>(L->z_ratio_hr);
    WR_HIP_CHECK(hipMemcpyAsync(motion, motion_init, 3 * g.n * sizeof(float), hipMemcpyDeviceToDevice, s));

    // errors of the current field: residual + penalty (zscale: the z
    // component's factor of the neighbour difference)
    auto errors = [&](bool first, float zscale, bool want_nei, double* d_err, double* p_err) -> int {
        k_xr_sample<<<gv, kT, 0, s>>>(g, motion, phase0, rv, mv, mask_coef ? 1 : 0, mask_mov, mov_l, phase, sampled,
                                      valid, raw);
        k_xr_residual<<<gv, kT, 0, s>>>(g, raw, valid, w9, residual, first ? first_residual : nullptr, pres);
        k_xr_nei<<<gc, kT, 0, s>>>(g, cp, motion, zscale, s_pen_f, w8, want_nei ? nei : nullptr, ppen);
        k_finish_sum<<<1, 1024, 0, s>>>(pres, static_cast<int>(gv), scal);
        k_finish_sum<<<1, 1024, 0, s>>>(ppen, static_cast<int>(gc), scal + 1);
        WR_HIP_CHECK(hipGetLastError());
        double h[2];
        WR_HIP_CHECK(hipMemcpyAsync(h, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        WR_HIP_CHECK(hipStreamSynchronize(s));
        *d_err = h[0] / static_cast<double>(g.n);
        // numpy: float32 sum of float32 terms, divided by the voxel count in float32
        *p_err = static_cast<double>(static_cast<float>(h[1]) / static_cast<float>(g.n));
        return WR_OK;
    };

    float old_err[3] = {INFINITY, INFINITY, INFINITY};
    int records = 0, updates = 0, reason = 0;
    for (int k = 0; k < L->iters; ++k) {
        double de, pe;
        WR_TRY(errors(k == 0, zr_f, true, &de, &pe));
        const double err = de + pe;
        if (trace->diff_error) trace->diff_error[k] = de;
        if (trace->penalty_error) trace->penalty_error[k] = pe;
        if (trace->current_error) trace->current_error[k] = err;
        ++records;
        // the stop rules before the update, in the reference's float32 arithmetic
        int stop = -1;
        if (forced) {
            if (k == forced[0] - 1 && forced[1] == forced[0] - 1) stop = 4;
        } else {
            const float ef = static_cast<float>(err);
            int n_le = 0;
            for (float o : old_err) n_le += (o <= ef) ? 1 : 0;
            if (k == L->iters - 1)
                stop = 0;
            else if (n_le > 1)
                stop = 1;
            else if (std::fabs(old_err[2] - ef) < static_cast<float>(L->tol))
                stop = 2;
            else {
                old_err[0] = old_err[1];
                old_err[1] = old_err[2];
                old_err[2] = ef;
            }
        }
        if (stop >= 0) {
            reason = stop;
            break;
        }
        k_xr_cp_solve<<<gw, 256, 0, s>>>(g, cp, phase, valid, residual, ref_l, motion, nei, L->smooth_penalty_sum,
                                          L->mov_range, L->z_ratio_hr, cpf);
        k_xr_dense<<<gv, kT, 0, s>>>(g, cp, cpf, motion, pmax);
        k_finish_max<<<1, 1024, 0, s>>>(pmax, static_cast<int>(gv), reinterpret_cast<float*>(scal + 2));
        WR_HIP_CHECK(hipGetLastError());
        float mx = 0.0f;
        WR_HIP_CHECK(hipMemcpyAsync(&mx, scal + 2, sizeof(float), hipMemcpyDeviceToHost, s));
        WR_HIP_CHECK(hipStreamSynchronize(s));
        ++updates;
        if (forced) {
            if (k == forced[0] - 1) {
                reason = 4;
                break;
            }
        } else if (mx < static_cast<float>(L->tol)) {
            reason = 3;
            break;
        }
        if (k == L->iters - 1) reason = 0;
    }
    // final recompute: no z scaling of the neighbour difference (:2214)
    double de, pe;
    WR_TRY(errors(false, 1.0f, false, &de, &pe));
    trace->records = records;
    trace->updates = updates;
    trace->reason = reason;
    trace->final_diff = de;
    trace->final_penalty = pe;
    trace->final_error = de + pe;
    return WR_OK;
}

int wr_xres_valid(wr_ctx* ctx, const float* phase, int nx, int ny, int nz, const double* mask_coef, int rx, int ry,
                  int rz, const uint8_t* mask_mov, uint8_t* valid, long long* count, void* stream) {
    WR_ARG(ctx && phase && count, "wr_xres_valid: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0 && rx > 0 && ry > 0 && rz > 0, "wr_xres_valid: empty shape");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    Geo g{nx, ny, nz, rx, ry, rz, static_cast<long long>(nx) * ny * nz, static_cast<long long>(nx) * ny};
    WR_ARG(g.n <= kMaxElems, "wr_xres_valid: volume above 2^32 voxels");
    WR_TRY(grow(c.xsel, 256));
    auto* cnt = c.xsel.as<unsigned long long>();
    k_zero_u64<<<1, 64, 0, s>>>(cnt, 1);
    const SplineVol mv = spline_vol(mask_coef ? mask_coef : nullptr, rx, ry, rz);
    k_xr_valid<<<blocks_atomic(g.n), kT, 0, s>>>(g, phase, mv, mask_coef ? 1 : 0, mask_mov, valid, cnt);
    WR_HIP_CHECK(hipGetLastError());
    unsigned long long h = 0;
    WR_HIP_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    *count = static_cast<long long>(h);
    return WR_OK;
}

int wr_xres_bad_regions(wr_ctx* ctx, const double* residual, const uint8_t* valid, const double* first_residual,
                        int nx, int ny, int nz, int r, int metric, double mad_threshold, int min_size, int mode,
                        double q, uint8_t* bad, int* n_regions, void* stream) {
    WR_ARG(ctx && residual && valid && bad && n_regions, "wr_xres_bad_regions: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0 && r >= 1, "wr_xres_bad_regions: empty shape");
    WR_ARG(mode == 0 || mode == 1, "wr_xres_bad_regions: mode must be 0 (direct) or 1 (highresidual)");
    WR_ARG(metric == 0 || metric == 1, "metric must be 'mse' or 'mae'");
    WR_ARG(mode == 0 || first_residual, "wr_xres_bad_regions: the highresidual mode needs first_residual");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    Geo g{nx, ny, nz, 0, 0, 0, static_cast<long long>(nx) * ny * nz, static_cast<long long>(nx) * ny};
    WR_ARG(g.n <= kMaxElems, "wr_xres_bad_regions: volume above 2^32 voxels");
    Cps cp{n_cp_axis(nx, r), n_cp_axis(ny, r), r, 2 * r + 1, 0};
    *n_regions = 0;
    if (cp.ncx == 0 || cp.ncy == 0) {
        k_copy_mask<<<blocks_for(g.n), kT, 0, s>>>(g.n, nullptr, bad);
        WR_HIP_CHECK(hipGetLastError());
        return WR_OK;
    }
    cp.n = static_cast<long long>(cp.ncx) * cp.ncy * nz;
    const int npl = cp.ncx * cp.ncy;
    int pow2 = 1;
    while (pow2 < npl) pow2 <<= 1;
    const size_t bytes = align256(cp.n * sizeof(double)) + align256(static_cast<size_t>(nz) * pow2 * sizeof(double)) +
                         3 * align256(static_cast<size_t>(cp.n) * 4 * sizeof(int)) +
                         align256(static_cast<size_t>(cp.n) * sizeof(Box)) + align256((nz + 2) * sizeof(int)) +
                         (mode == 1 ? 2 * align256(g.n * sizeof(double)) + align256(kSelBins * sizeof(unsigned)) +
                                          align256(2 * sizeof(SelState)) + align256(64)
                                    : 0);
    WR_TRY(grow(c.xsel, bytes));
    Carve cv{c.xsel.as<char>()};
    double* err = cv.take<double>(cp.n);
    double* srt = cv.take<double>(static_cast<size_t>(nz) * pow2);
    int* lab = cv.take<int>(cp.n * 4);
    int* cnt = cv.take<int>(cp.n * 4);
    int* bb = cv.take<int>(cp.n * 4);
    Box* boxes = cv.take<Box>(cp.n);
    int* nbox = cv.take<int>(nz + 2);
    WR_HIP_CHECK(hipMemsetAsync(nbox, 0, (nz + 2) * sizeof(int), s));
    const int k = 2 * r + 1;
    const double wk = static_cast<double>(1.0f / static_cast<float>(k * k));
    k_xr_cp_err<<<static_cast<unsigned>((cp.n + 3) / 4), 256, 0, s>>>(g, cp, residual, wk, metric, err);
    k_xr_planes<<<nz, 1024, 0, s>>>(cp, nx, ny, err, mad_threshold, min_size, pow2, srt, lab, cnt, bb, boxes, nbox,
                                    nbox + nz);
    k_xr_boxes<<<blocks_for(g.n), kT, 0, s>>>(g, npl, boxes, nbox, valid, mode == 0 ? 1 : 0, bad);
    WR_HIP_CHECK(hipGetLastError());
    int total = 0;
    WR_HIP_CHECK(hipMemcpyAsync(&total, nbox + nz, sizeof(int), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    *n_regions = total;
    if (mode == 1 && total > 0) {
        double* dense = cv.take<double>(g.n);
        double* list = cv.take<double>(g.n);
        unsigned* hist = cv.take<unsigned>(kSelBins);
        SelState* st = cv.take<SelState>(2);
        auto* nl = cv.take<unsigned long long>(4);
        WR_HIP_CHECK(hipMemsetAsync(hist, 0, kSelBins * sizeof(unsigned), s));
        k_zero_u64<<<1, 64, 0, s>>>(nl, 1);
        k_xr_dense_err<<<blocks_for(g.n), kT, 0, s>>>(g, first_residual, residual, bad,
                                                       static_cast<double>(1.0f / 9.0f), dense, list, nl);
        unsigned long long hn = 0;
        WR_HIP_CHECK(hipMemcpyAsync(&hn, nl, sizeof(hn), hipMemcpyDeviceToHost, s));
        WR_HIP_CHECK(hipStreamSynchronize(s));
        if (hn > 0) {
            select_rank(list, nl, -1, q, st, hist, static_cast<long long>(hn), s);
            k_xr_keep_high<<<blocks_for(g.n), kT, 0, s>>>(g, dense, &st->value, valid, bad);
        } else {
            k_copy_mask<<<blocks_for(g.n), kT, 0, s>>>(g.n, nullptr, bad);
        }
        WR_HIP_CHECK(hipGetLastError());
    } else if (mode == 1) {
        k_copy_mask<<<blocks_for(g.n), kT, 0, s>>>(g.n, nullptr, bad);
        WR_HIP_CHECK(hipGetLastError());
    }
    return WR_OK;
}

int wr_xres_trap_mask(wr_ctx* ctx, const uint8_t* bad, const float* phase, int nx, int ny, int nz,
                      const double* ref_l, int rx, int ry, int rz, const uint8_t* mask_ref_l, double z_ratio_ref,
                      double expand_radius_xy, double sigma, double intensity_k, uint8_t* corrected,
                      long long* trap_count, void* stream) {
    WR_ARG(ctx && bad && phase && ref_l && corrected && trap_count, "wr_xres_trap_mask: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0 && rx > 0 && ry > 0 && rz > 0, "wr_xres_trap_mask: empty shape");
    WR_ARG(std::isfinite(z_ratio_ref) && z_ratio_ref > 0.0,
           "_make_ball: z_ratio must be a positive, finite physical-units-per-z-index ratio");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    Geo g{nx, ny, nz, rx, ry, rz, static_cast<long long>(nx) * ny * nz, static_cast<long long>(nx) * ny};
    const long long nref = static_cast<long long>(rx) * ry * rz;
    WR_ARG(g.n <= kMaxElems && nref <= kMaxElems, "wr_xres_trap_mask: volume above 2^32 voxels");
    *trap_count = 0;
    k_copy_mask<<<blocks_for(nref), kT, 0, s>>>(nref, mask_ref_l, corrected);
    WR_TRY(grow(c.xsel, 4096));
    int* bb = c.xsel.as<int>();
    k_trap_bbox_init<<<1, 64, 0, s>>>(bb);
    k_trap_bbox<<<blocks_atomic(g.n), kT, 0, s>>>(g, bad, phase, bb);
    WR_HIP_CHECK(hipGetLastError());
    int h[7];
    WR_HIP_CHECK(hipMemcpyAsync(h, bb, sizeof(h), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    if (h[6] == 0) return WR_OK;
    // the box around the seeds (:1483-1502)
    const int mxy = static_cast<int>(std::max(1.0, expand_radius_xy + 2));
    const int mz = static_cast<int>(std::max(1.0, std::ceil((expand_radius_xy + 2) / std::max(z_ratio_ref, 1e-6))));
    Roi roi;
    roi.x0 = std::max(0, h[0] - mxy);
    roi.y0 = std::max(0, h[2] - mxy);
    roi.z0 = std::max(0, h[4] - mz);
    roi.wx = std::min(rx, h[1] + mxy + 1) - roi.x0;
    roi.wy = std::min(ry, h[3] + mxy + 1) - roi.y0;
    roi.wz = std::min(rz, h[5] + mz + 1) - roi.z0;
    roi.n = static_cast<long long>(roi.wx) * roi.wy * roi.wz;
    const std::vector<int> ball1 = ball_offsets(1, z_ratio_ref);
    // Python's round(): halves to even
    const std::vector<int> ballE =
        ball_offsets(std::max(1, static_cast<int>(std::nearbyint(expand_radius_xy))), z_ratio_ref);
    // Gaussian weights (scipy: radius int(4 sigma + 0.5), exp(-0.5 / sigma^2 x^2) / sum)
    const int rad = sigma > 0 ? static_cast<int>(4.0 * sigma + 0.5) : 0;
    std::vector<double> gw(2 * rad + 1);
    if (sigma > 0) {
        for (int k = -rad; k <= rad; ++k) gw[k + rad] = std::exp(-0.5 / (sigma * sigma) * static_cast<double>(k * k));
        const double sum = np_sum(gw);
        for (double& v : gw) v /= sum;
    }
    const size_t nb = static_cast<size_t>(roi.n);
    const size_t bytes = 5 * align256(nb) + 3 * align256(nb * sizeof(float)) + align256(nb * sizeof(double)) +
                         align256((ball1.size() + ballE.size()) * sizeof(int)) + align256(gw.size() * sizeof(double)) +
                         align256(kSelBins * sizeof(unsigned)) + align256(4 * sizeof(SelState)) + align256(256);
    WR_TRY(grow(c.xroi, bytes));
    Carve cv{c.xroi.as<char>()};
    uint8_t* seed = cv.take<uint8_t>(nb);
    uint8_t* tmp = cv.take<uint8_t>(nb);
    uint8_t* near = cv.take<uint8_t>(nb);
    uint8_t* allowed = cv.take<uint8_t>(nb);
    uint8_t* region = cv.take<uint8_t>(nb);
    float* crop = cv.take<float>(nb);
    float* sm = cv.take<float>(nb);
    float* sm2 = cv.take<float>(nb);
    double* list = cv.take<double>(nb);
    int* offs = cv.take<int>(ball1.size() + ballE.size());
    double* dgw = cv.take<double>(gw.size());
    unsigned* hist = cv.take<unsigned>(kSelBins);
    SelState* st = cv.take<SelState>(4);
    auto* misc = cv.take<unsigned long long>(8);   // [0] list count, [1] region count, [2] changed, [3..4] medians
    int* changed = reinterpret_cast<int*>(misc + 2);
    float* meds = reinterpret_cast<float*>(misc + 3);
    WR_HIP_CHECK(hipMemcpyAsync(offs, ball1.data(), ball1.size() * sizeof(int), hipMemcpyHostToDevice, s));
    WR_HIP_CHECK(hipMemcpyAsync(offs + ball1.size(), ballE.data(), ballE.size() * sizeof(int), hipMemcpyHostToDevice, s));
    if (!gw.empty()) WR_HIP_CHECK(hipMemcpyAsync(dgw, gw.data(), gw.size() * sizeof(double), hipMemcpyHostToDevice, s));
    WR_HIP_CHECK(hipMemsetAsync(seed, 0, nb, s));
    WR_HIP_CHECK(hipMemsetAsync(hist, 0, kSelBins * sizeof(unsigned), s));
    k_zero_u64<<<1, 64, 0, s>>>(misc, 8);
    const unsigned gr = blocks_for(roi.n);
    const int n1 = static_cast<int>(ball1.size() / 3), nE = static_cast<int>(ballE.size() / 3);
    k_trap_scatter<<<blocks_for(g.n), kT, 0, s>>>(g, bad, phase, roi, seed);
    // closing with the radius-1 ball
    k_morph<<<gr, kT, 0, s>>>(roi, seed, offs, n1, 0, nullptr, tmp, nullptr);
    k_morph<<<gr, kT, 0, s>>>(roi, tmp, offs, n1, 1, nullptr, seed, nullptr);
    // smoothed reference inside the box; the seeds' values listed
    k_trap_crop<<<gr, kT, 0, s>>>(roi, ref_l, rx, ry, crop);
    const float* smoothed = crop;
    if (sigma > 0) {
        k_gauss_axis<<<gr, kT, 0, s>>>(roi, crop, 0, dgw, rad, sm);
        k_gauss_axis<<<gr, kT, 0, s>>>(roi, sm, 1, dgw, rad, sm2);
        k_gauss_axis<<<gr, kT, 0, s>>>(roi, sm2, 2, dgw, rad, sm);
        smoothed = sm;
    }
    k_list_masked<<<gr, kT, 0, s>>>(roi, smoothed, seed, list, misc);
    WR_HIP_CHECK(hipGetLastError());
    unsigned long long nseed = 0;
    WR_HIP_CHECK(hipMemcpyAsync(&nseed, misc, sizeof(nseed), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    if (nseed == 0) return WR_OK;   // the closing removed every seed
    const long long ns = static_cast<long long>(nseed);
    // median, then the median of |v - median|, both numpy float32 medians
    const long long lo = (ns & 1) ? ns / 2 : ns / 2 - 1, hi = ns / 2;
    select_rank(list, misc, lo, 0.0, st, hist, ns, s);
    select_rank(list, misc, hi, 0.0, st + 1, hist, ns, s);
    k_median_f32<<<1, 1, 0, s>>>(st, st + 1, misc, meds);
    k_abs_dev<<<blocks_for(ns), kT, 0, s>>>(list, misc, meds);
    select_rank(list, misc, lo, 0.0, st + 2, hist, ns, s);
    select_rank(list, misc, hi, 0.0, st + 3, hist, ns, s);
    k_median_f32<<<1, 1, 0, s>>>(st + 2, st + 3, misc, meds + 1);
    WR_HIP_CHECK(hipGetLastError());
    float hm[2];
    WR_HIP_CHECK(hipMemcpyAsync(hm, meds, sizeof(hm), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    const double tol = std::max(intensity_k * 1.4826 * static_cast<double>(hm[1]), 1e-3);
    // admissible voxels: near the seeds and of similar intensity, or seeds
    k_morph<<<gr, kT, 0, s>>>(roi, seed, offs + ball1.size(), nE, 0, nullptr, near, nullptr);
    k_trap_allowed<<<gr, kT, 0, s>>>(roi, near, smoothed, meds, static_cast<float>(tol), seed, allowed);
    // bounded growth from the seeds (:1399-1410)
    WR_HIP_CHECK(hipMemcpyAsync(region, seed, nb, hipMemcpyDeviceToDevice, s));
    const int steps = static_cast<int>(expand_radius_xy);
    for (int it = 0; it < steps; ++it) {
        k_zero_u64<<<1, 64, 0, s>>>(misc + 2, 1);
        k_morph<<<gr, kT, 0, s>>>(roi, region, offs, n1, 0, allowed, tmp, changed);
        WR_HIP_CHECK(hipGetLastError());
        int ch = 0;
        WR_HIP_CHECK(hipMemcpyAsync(&ch, changed, sizeof(int), hipMemcpyDeviceToHost, s));
        WR_HIP_CHECK(hipStreamSynchronize(s));
        if (!ch) break;
        std::swap(region, tmp);
    }
    k_trap_write<<<blocks_atomic(roi.n), kT, 0, s>>>(roi, region, seed, rx, ry, corrected, misc + 1);
    WR_HIP_CHECK(hipGetLastError());
    unsigned long long tc = 0;
    WR_HIP_CHECK(hipMemcpyAsync(&tc, misc + 1, sizeof(tc), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    *trap_count = static_cast<long long>(tc);
    return WR_OK;
}

int wr_count_u8(wr_ctx* ctx, const uint8_t* m, long long n, long long* count, void* stream) {
    WR_ARG(ctx && m && count && n >= 0, "wr_count_u8: null argument");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    WR_TRY(grow(c.xsel, 256));
    auto* cnt = c.xsel.as<unsigned long long>();
    k_zero_u64<<<1, 64, 0, s>>>(cnt, 1);
    if (n > 0) k_count_u8<<<blocks_atomic(n), kT, 0, s>>>(m, n, cnt);
    WR_HIP_CHECK(hipGetLastError());
    unsigned long long h = 0;
    WR_HIP_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    *count = static_cast<long long>(h);
    return WR_OK;
}

int wr_find_init_z(wr_ctx* ctx, const float* mov, int nk, const float* ref, int nz, int h, int w, double delta,
                    int use_gradient, int use_hann, int direction, double eps, uint16_t* z_init, float* scores_out,
                    double* best_score, int* best_z0, void* stream) {
    WR_ARG(ctx && mov && ref && z_init, "wr_find_init_z: null argument");
    WR_ARG(nk >= 1, "Moving input must have at least one slice.");
    WR_ARG(nz >= 1, "Reference input must have at least one slice.");
    WR_ARG(h >= 1 && w >= 1, "wr_find_init_z: empty slice");
    const long long hw = static_cast<long long>(h) * w;
    WR_ARG(hw * (nk + nz) <= kMaxElems, "wr_find_init_z: stacks above 2^32 voxels");
    // offsets (float32 as the reference forms them) and the admissible starts
    std::vector<float> off(nk);
    for (int k = 0; k < nk; ++k) off[k] = static_cast<float>(k) * static_cast<float>(delta) * static_cast<float>(direction);
    const float omin = *std::min_element(off.begin(), off.end()), omax = *std::max_element(off.begin(), off.end());
    const int z0_min = static_cast<int>(std::ceil(-omin));
    const int z0_max = static_cast<int>(std::floor(static_cast<float>(nz - 1) - omax));
    if (z0_max < z0_min) {
        char buf[256];
        std::snprintf(buf, sizeof(buf),
                      "No valid z0 exists. Z=%d, K=%d, delta_ref_idx=%g. Need reference z range >= %g.", nz, nk, delta,
                      (nk - 1) * std::fabs(delta) + 1);
        set_error(buf);
        return WR_E_ARG;
    }
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the normalised weights: Hann window (numpy hanning, outer product rounded to
    // float32, floored at 1e-3) or ones, divided by their sum + eps in float32
    std::vector<float> wt(hw);
    std::vector<double> wa(h, 1.0), wb(w, 1.0);
    if (use_hann) {
        const double pi = 3.14159265358979323846;
        if (h > 1)
            for (int i = 0; i < h; ++i) wa[i] = 0.5 + 0.5 * std::cos(pi * (1 - h + 2 * i) / (h - 1));
        if (w > 1)
            for (int j = 0; j < w; ++j) wb[j] = 0.5 + 0.5 * std::cos(pi * (1 - w + 2 * j) / (w - 1));
    }
    double sum = 0.0;
    for (int j = 0; j < w; ++j)
        for (int i = 0; i < h; ++i) {
            float v = use_hann ? std::max(static_cast<float>(wa[i] * wb[j]), 1e-3f) : 1.0f;
            wt[static_cast<size_t>(j) * h + i] = v;
            sum += v;
        }
    const float den = static_cast<float>(sum) + static_cast<float>(eps);
    for (float& v : wt) v = v / den;
    const size_t bytes = align256(hw * sizeof(float)) + (use_gradient ? align256(hw * (nk + nz) * sizeof(float)) : 0) +
                         4 * align256((nk + nz) * sizeof(double)) + align256(static_cast<size_t>(nk) * nz * sizeof(float));
    WR_TRY(grow(c.xsel, bytes));
    Carve cv{c.xsel.as<char>()};
    float* dwt = cv.take<float>(hw);
    const float* mf = mov;
    const float* rf = ref;
    if (use_gradient) {
        float* g = cv.take<float>(hw * (nk + nz));
        k_fz_gradmag<<<blocks_for(hw * nk), kT, 0, s>>>(mov, h, w, hw * nk, g);
        k_fz_gradmag<<<blocks_for(hw * nz), kT, 0, s>>>(ref, h, w, hw * nz, g + hw * nk);
        mf = g;
        rf = g + hw * nk;
    }
    double* mm = cv.take<double>(nk);
    double* mv = cv.take<double>(nk);
    double* rm = cv.take<double>(nz);
    double* rv = cv.take<double>(nz);
    float* sc = cv.take<float>(static_cast<size_t>(nk) * nz);
    WR_HIP_CHECK(hipMemcpyAsync(dwt, wt.data(), hw * sizeof(float), hipMemcpyHostToDevice, s));
    k_fz_stats<<<nk, 1024, 0, s>>>(mf, dwt, hw, mm, mv);
    k_fz_stats<<<nz, 1024, 0, s>>>(rf, dwt, hw, rm, rv);
    k_fz_zncc<<<nk * nz, 1024, 0, s>>>(mf, rf, dwt, hw, nz, mm, mv, rm, rv, eps, sc);
    WR_HIP_CHECK(hipGetLastError());
    std::vector<float> hs(static_cast<size_t>(nk) * nz);
    WR_HIP_CHECK(hipMemcpyAsync(hs.data(), sc, hs.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    // the search (:432-446): first start with the largest float32 score sum
    float best = -INFINITY;
    int bz = z0_min;
    for (int z0 = z0_min; z0 <= z0_max; ++z0) {
        float acc = 0.0f;
        for (int k = 0; k < nk; ++k) {
            const int zi = static_cast<int>(std::nearbyint(static_cast<float>(z0) + off[k]));
            acc += hs[static_cast<size_t>(k) * nz + zi];
        }
        if (acc > best) {
            best = acc;
            bz = z0;
        }
    }
    for (int k = 0; k < nk; ++k) z_init[k] = static_cast<uint16_t>(static_cast<float>(bz) + off[k]);
    if (scores_out) std::copy(hs.begin(), hs.end(), scores_out);
    if (best_score) *best_score = best;
    if (best_z0) *best_z0 = bz;
    return WR_OK;
}

int wr_xres_scale(wr_ctx* ctx, const double* in, long long n, double mul, double div, float* out, void* stream) {
    WR_ARG(ctx && in && out && n >= 0, "wr_xres_scale: null argument");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    if (n == 0) return WR_OK;
    k_xr_scale<<<blocks_for(n), kT, 0, static_cast<hipStream_t>(stream)>>>(in, n, mul, div, out);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_xres_threshold(wr_ctx* ctx, const double* in, long long n, double thr, uint8_t* out, void* stream) {
    WR_ARG(ctx && in && out && n >= 0, "wr_xres_threshold: null argument");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    if (n == 0) return WR_OK;
    k_xr_threshold<<<blocks_for(n), kT, 0, static_cast<hipStream_t>(stream)>>>(in, n, thr, out);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_xres_phase_identity(wr_ctx* ctx, int nx, int ny, int nz, double z_ratio, double z_ratio_hr, float* phase,
                           void* stream) {
    WR_ARG(ctx && phase, "wr_xres_phase_identity: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0 && z_ratio_hr != 0.0, "wr_xres_phase_identity: empty shape");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    const long long n = static_cast<long long>(nx) * ny * nz;
    k_xr_phase<<<blocks_for(n), kT, 0, static_cast<hipStream_t>(stream)>>>(nx, ny, nz, z_ratio, z_ratio_hr, phase);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

}  // extern "C"
