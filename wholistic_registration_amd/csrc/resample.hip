// Separable MATLAB-style resampling on gfx950: the pyramid levels, the mask
// levels, the warm-start field and the level-to-level field upsample.
//
// Tables follow contributions() of src/wholistic_registration/utils/
// imresize.py:60-82 (Keys cubic a=-0.5 :49-57, triangle :41-46, support 4,
// antialiasing by kernel stretching when the scale is below one,
// half-sample-symmetric mirroring, zero tap columns dropped).  They are built
// on the host in float64 with the same operation order as the reference,
// including numpy's pairwise row sum used for the normalisation, so that the
// weights are the reference's weights bit for bit.  The gather-and-weight
// passes accumulate in float64 with unfused multiply and add, in tap order,
// as numpy's reduction over a non-innermost axis does (imresize.py:121-137).
//
// Memory: both passes stream along x, the fastest axis.  The x pass gathers
// along a row that the L1/L2 caches hold (taps of neighbouring outputs
// overlap); the y pass reads whole rows per tap, fully coalesced.

#include <cmath>

#include "wr_internal.h"

namespace wr {

// ------------------------------------------------------------------ tables

#pragma clang fp contract(off)

static double kern_cubic(double t) {
    const double a = std::fabs(t);
    const double a2 = a * a;
    const double a3 = a2 * a;
    const double inner = (1.5 * a3 - 2.5 * a2 + 1) * (a <= 1 ? 1.0 : 0.0);
    const double outer = (-0.5 * a3 + 2.5 * a2 - 4 * a + 2) * ((a > 1 && a <= 2) ? 1.0 : 0.0);
    return inner + outer;
}

static double kern_tent(double t) {
    const double neg = (t >= -1 && t < 0) ? 1.0 : 0.0;
    const double pos = (t <= 1 && t >= 0) ? 1.0 : 0.0;
    return (t + 1) * neg + (1 - t) * pos;
}

// numpy's pairwise summation of a contiguous float64 run (the order used by
// np.sum along the innermost axis).
static double np_pairwise(const double* a, long n) {
    if (n < 8) {
        double res = 0.0;
        for (long i = 0; i < n; ++i) res += a[i];
        return res;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        long i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    long n2 = n / 2;
    n2 -= n2 % 8;
    return np_pairwise(a, n2) + np_pairwise(a + n2, n - n2);
}

Table build_table(int n_in, int n_out, Kernel kern, std::vector<double>& w_arena,
                  std::vector<int>& i_arena) {
    Table t;
    t.n_in = n_in;
    t.n_out = n_out;
    if (n_in == n_out) {
        // scale 1: every row has the single weight kernel(0) = 1 at its own
        // index (kernel(+-1) = kernel(+-2) = 0 for both kernels), i.e. exactly
        // the identity.
        t.identity = true;
        t.k = 1;
        return t;
    }
    const double scale = 1.0 * n_out / n_in;
    const bool shrink = scale < 1;
    const double width = shrink ? 1.0 * 4.0 / scale : 4.0;
    const int ntap = static_cast<int>(std::ceil(width)) + 2;
    const double shift = 0.5 * (1 - 1 / scale);
    double (*h)(double) = kern == KERNEL_CUBIC ? kern_cubic : kern_tent;

    std::vector<double> w(static_cast<size_t>(n_out) * ntap);
    std::vector<int> ind(static_cast<size_t>(n_out) * ntap);
    for (int j = 0; j < n_out; ++j) {
        const double u = static_cast<double>(j + 1) / scale + shift;
        const double left = std::floor(u - width / 2);
        double* row = &w[static_cast<size_t>(j) * ntap];
        for (int k = 0; k < ntap; ++k) {
            const int ii = static_cast<int>((left + k) - 1);
            const double arg = (u - ii) - 1;
            row[k] = shrink ? scale * h(scale * arg) : h(arg);
            ind[static_cast<size_t>(j) * ntap + k] = ii;
        }
        const double s = np_pairwise(row, ntap);
        for (int k = 0; k < ntap; ++k) row[k] = row[k] / s;
    }
    const int m = 2 * n_in;
    for (auto& ii : ind) {
        int q = ii % m;
        if (q < 0) q += m;
        ii = q < n_in ? q : (m - 1 - q);
    }
    std::vector<int> live;
    for (int k = 0; k < ntap; ++k) {
        bool any = false;
        for (int j = 0; j < n_out && !any; ++j) any = w[static_cast<size_t>(j) * ntap + k] != 0.0;
        if (any) live.push_back(k);
    }
    t.k = static_cast<int>(live.size());
    t.w_off = w_arena.size();
    t.i_off = i_arena.size();
    for (int j = 0; j < n_out; ++j)
        for (int k : live) {
            w_arena.push_back(w[static_cast<size_t>(j) * ntap + k]);
            i_arena.push_back(ind[static_cast<size_t>(j) * ntap + k]);
        }
    return t;
}

#pragma clang fp contract(on)

// ------------------------------------------------------------------ kernels

template <int IN_T>
__device__ __forceinline__ double load_d(const void* p, size_t i) {
    if constexpr (IN_T == E_F32) return static_cast<double>(static_cast<const float*>(p)[i]);
    else if constexpr (IN_T == E_U16) return static_cast<double>(static_cast<const uint16_t*>(p)[i]);
    else return static_cast<const double*>(p)[i];
}

__device__ __forceinline__ void store_out(void* out, size_t i, double v, int om, double divisor) {
    if (om == OUT_F64) static_cast<double*>(out)[i] = v;
    else if (om == OUT_F32) static_cast<float*>(out)[i] = static_cast<float>(v / divisor);
    else static_cast<uint8_t*>(out)[i] = v > 0.0 ? 1 : 0;
}

template <int IN_T>
__global__ __launch_bounds__(256) void k_resample_x(const void* __restrict__ in, int in_stride,
                                                     int in_comp, void* __restrict__ out, int om,
                                                     double divisor, const double* __restrict__ w,
                                                     const int* __restrict__ idx, int n_in,
                                                     int n_out, int k, int xblocks) {
    const long long b = blockIdx.x;
    const long long row = b / xblocks;
    const int xo = static_cast<int>(b % xblocks) * 256 + threadIdx.x;
    if (xo >= n_out) return;
    const size_t in_row = static_cast<size_t>(row) * n_in;
    const double* wr = w + static_cast<size_t>(xo) * k;
    const int* ir = idx + static_cast<size_t>(xo) * k;
    double acc = __dmul_rn(wr[0], load_d<IN_T>(in, (in_row + ir[0]) * in_stride + in_comp));
    for (int t = 1; t < k; ++t)
        acc = __dadd_rn(acc, __dmul_rn(wr[t], load_d<IN_T>(in, (in_row + ir[t]) * in_stride + in_comp)));
    store_out(out, static_cast<size_t>(row) * n_out + xo, acc, om, divisor);
}

template <int IN_T>
__global__ __launch_bounds__(256) void k_resample_y(const void* __restrict__ in, int in_stride,
                                                     int in_comp, void* __restrict__ out, int om,
                                                     double divisor, const double* __restrict__ w,
                                                     const int* __restrict__ idx, int n_in,
                                                     int n_out, int k, int nx, int xblocks) {
    const long long b = blockIdx.x;
    const int x = static_cast<int>(b % xblocks) * 256 + threadIdx.x;
    const long long zy = b / xblocks;          // z * n_out + yo
    if (x >= nx) return;
    const int yo = static_cast<int>(zy % n_out);
    const long long z = zy / n_out;
    const double* wr = w + static_cast<size_t>(yo) * k;
    const int* ir = idx + static_cast<size_t>(yo) * k;
    const size_t base = static_cast<size_t>(z) * n_in;
    double acc = __dmul_rn(wr[0], load_d<IN_T>(in, ((base + ir[0]) * nx + x) * in_stride + in_comp));
    for (int t = 1; t < k; ++t)
        acc = __dadd_rn(acc, __dmul_rn(wr[t], load_d<IN_T>(in, ((base + ir[t]) * nx + x) * in_stride + in_comp)));
    store_out(out, static_cast<size_t>(zy) * nx + x, acc, om, divisor);
}

template <int IN_T>
__global__ __launch_bounds__(256) void k_convert(const void* __restrict__ in, int in_stride,
                                                  int in_comp, void* __restrict__ out, int om,
                                                  double divisor, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    store_out(out, i, load_d<IN_T>(in, i * in_stride + in_comp), om, divisor);
}

// ------------------------------------------------------------------ launchers

void launch_resample_x(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int ny, int nz, hipStream_t s) {
    const int xblocks = (t.n_out + 255) / 256;
    const long long blocks = static_cast<long long>(xblocks) * ny * nz;
    const double* wp = w + t.w_off;
    const int* ip = idx + t.i_off;
    switch (in_t) {
        case E_F32:
            k_resample_x<E_F32><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, xblocks);
            break;
        case E_U16:
            k_resample_x<E_U16><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, xblocks);
            break;
        default:
            k_resample_x<E_F64><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, xblocks);
    }
}

void launch_resample_y(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int nx, int nz, hipStream_t s) {
    const int xblocks = (nx + 255) / 256;
    const long long blocks = static_cast<long long>(xblocks) * t.n_out * nz;
    const double* wp = w + t.w_off;
    const int* ip = idx + t.i_off;
    switch (in_t) {
        case E_F32:
            k_resample_y<E_F32><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, nx, xblocks);
            break;
        case E_U16:
            k_resample_y<E_U16><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, nx, xblocks);
            break;
        default:
            k_resample_y<E_F64><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, nx, xblocks);
    }
}

void launch_convert(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                    OutMode om, double divisor, size_t n, hipStream_t s) {
    const long long blocks = static_cast<long long>((n + 255) / 256);
    switch (in_t) {
        case E_F32:
            k_convert<E_F32><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, n);
            break;
        case E_U16:
            k_convert<E_U16><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, n);
            break;
        default:
            k_convert<E_F64><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, n);
    }
}

}  // namespace wr
