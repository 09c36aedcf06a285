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
// x pass (the fast axis): a workgroup owns a group of output columns and up
// to 64 rows; it stages the group's table slice and each row's span of
// source columns in LDS, one row per lane.
// y pass (a slow axis): one thread per output voxel along x; every tap is a
// coalesced row read and the tap table is uniform per workgroup.

#include <algorithm>
#include <cmath>
#include <type_traits>

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
    // x-pass column groups: the largest power of two <= 64 whose source span
    // fits kXSpan inputs (and at most 2048 table entries); the span of each
    // group
    auto span_of = [&](int xo_try) {
        int worst = 0;
        for (int g = 0; g * xo_try < n_out; ++g) {
            int lo = n_in, hi = 0;
            for (int j = g * xo_try; j < std::min(n_out, (g + 1) * xo_try); ++j)
                for (int k = 0; k < t.k; ++k) {
                    const int ii = i_arena[t.i_off + static_cast<size_t>(j) * t.k + k];
                    lo = std::min(lo, ii);
                    hi = std::max(hi, ii);
                }
            worst = std::max(worst, hi - lo + 1);
        }
        return worst;
    };
    int xo = 1;
    while (xo * 2 <= 64 && xo * 2 * t.k <= 2048 && span_of(xo * 2) <= kXSpan) xo *= 2;
    t.xo_blk = xo;
    t.n_cb = (n_out + xo - 1) / xo;
    t.s_off = i_arena.size();
    t.span_max = 0;
    for (int g = 0; g < t.n_cb; ++g) {
        int lo = n_in, hi = 0;
        for (int j = g * xo; j < std::min(n_out, (g + 1) * xo); ++j)
            for (int k = 0; k < t.k; ++k) {
                const int ii = i_arena[t.i_off + static_cast<size_t>(j) * t.k + k];
                lo = std::min(lo, ii);
                hi = std::max(hi, ii);
            }
        i_arena.push_back(lo);
        i_arena.push_back(hi);
        t.span_max = std::max(t.span_max, hi - lo + 1);
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

struct XPass {
    const double* w;
    const int* idx;
    const int* span;
    int n_in, n_out, k, xo_blk, n_cb;
    int sp;             // odd LDS row pitch (>= every group's span)
    int rt;             // rows per workgroup (<= 64, one per lane)
    long long nrows;
    OutLayout ol;
};

// x pass: a workgroup stages the table slice of one column group and the
// source span of `rt` rows in LDS; lane = row, wave = column (round robin),
// so every table read is a wave-uniform broadcast and the per-row source
// reads hit distinct banks (odd row pitch).  Each output is stored directly;
// a row's columns are written by consecutive iterations and merge in L2.
template <int IN_T>
__global__ __launch_bounds__(256) void k_resample_x(const void* __restrict__ in, int in_stride, int in_comp,
                                                     void* __restrict__ out, int om, double divisor, XPass p) {
    using T = typename std::conditional<IN_T == E_F64, double, float>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tabn = p.xo_blk * p.k;
    double* sw = reinterpret_cast<double*>(smem_raw);                 // [c][t]
    int* si = reinterpret_cast<int*>(sw + tabn);                        // [c][t]
    T* sin = reinterpret_cast<T*>(smem_raw + ((static_cast<size_t>(tabn) * 12 + 15) & ~size_t(15)));
    const int cb = static_cast<int>(blockIdx.x % p.n_cb);
    const long long r0 = static_cast<long long>(blockIdx.x / p.n_cb) * p.rt;
    const int nr = static_cast<int>(min(static_cast<long long>(p.rt), p.nrows - r0));
    const int xo0 = cb * p.xo_blk;
    const int nxo = min(p.xo_blk, p.n_out - xo0);
    const int lo = p.span[2 * cb];
    const int span = p.span[2 * cb + 1] - lo + 1;
    const size_t tab0 = static_cast<size_t>(xo0) * p.k;
    for (int i = threadIdx.x; i < nxo * p.k; i += 256) {
        sw[i] = p.w[tab0 + i];
        si[i] = p.idx[tab0 + i] - lo;
    }
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    // wave w stages rows w, w+4, ..., lanes along x; the (up to) 16 rows of a
    // lane are loaded before any is stored, so the loads overlap
    for (int xx = lane; xx < span; xx += 64) {
        T vals[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int rr = wave + 4 * j;
            vals[j] = rr < nr ? static_cast<T>(load_d<IN_T>(
                                    in, ((static_cast<size_t>(r0 + rr) * p.n_in + lo + xx) * in_stride) + in_comp))
                              : T(0);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int rr = wave + 4 * j;
            if (rr < nr) sin[rr * p.sp + xx] = vals[j];
        }
    }
    __syncthreads();
    if (lane < nr) {
        const T* row = sin + lane * p.sp;
        const size_t obase = static_cast<size_t>(r0 + lane) * p.ol.pitch + p.ol.pad + xo0;
        for (int c = wave; c < nxo; c += 4) {
            const double* wc = sw + c * p.k;
            const int* ic = si + c * p.k;
            double acc = __dmul_rn(wc[0], static_cast<double>(row[ic[0]]));
            for (int t = 1; t < p.k; ++t) acc = __dadd_rn(acc, __dmul_rn(wc[t], static_cast<double>(row[ic[t]])));
            store_out(out, obase + c, acc, om, divisor);
        }
    }
}

template <int IN_T>
__global__ __launch_bounds__(256) void k_resample_y(const void* __restrict__ in, int in_stride,
                                                     int in_comp, void* __restrict__ out, int om,
                                                     double divisor, const double* __restrict__ w,
                                                     const int* __restrict__ idx, int n_in,
                                                     int n_out, int k, int nx, int xblocks, OutLayout ol) {
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
#pragma unroll 8
    for (int t = 1; t < k; ++t)
        acc = __dadd_rn(acc, __dmul_rn(wr[t], load_d<IN_T>(in, ((base + ir[t]) * nx + x) * in_stride + in_comp)));
    store_out(out, static_cast<size_t>(zy) * ol.pitch + ol.pad + x, acc, om, divisor);
}

template <int IN_T>
__global__ __launch_bounds__(256) void k_convert(const void* __restrict__ in, int in_stride,
                                                  int in_comp, void* __restrict__ out, int om,
                                                  double divisor, int nx, size_t n, OutLayout ol) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t row = i / nx, x = i - row * nx;
    store_out(out, row * ol.pitch + ol.pad + x, load_d<IN_T>(in, i * in_stride + in_comp), om, divisor);
}

__global__ __launch_bounds__(256) void k_pad_rows(float* __restrict__ buf, long long nrows, int nx, int pitch) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= nrows * 2 * kPad) return;
    const long long row = i / (2 * kPad);
    const int j = static_cast<int>(i - row * 2 * kPad);
    float* r = buf + row * pitch;
    if (j < kPad) r[j] = r[kPad];
    else r[kPad + nx + (j - kPad)] = r[kPad + nx - 1];
}

// ------------------------------------------------------------------ launchers

void launch_resample_x(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int ny, int nz, OutLayout ol, hipStream_t s) {
    XPass p;
    p.w = w + t.w_off;
    p.idx = idx + t.i_off;
    p.span = idx + t.s_off;
    p.n_in = t.n_in;
    p.n_out = t.n_out;
    p.k = t.k;
    p.xo_blk = t.xo_blk;
    p.n_cb = t.n_cb;
    p.sp = t.span_max | 1;
    p.nrows = static_cast<long long>(ny) * nz;
    p.ol = ol;
    const size_t esz = in_t == E_F64 ? sizeof(double) : sizeof(float);
    const size_t tab = (static_cast<size_t>(t.xo_blk) * t.k * 12 + 15) & ~size_t(15);
    // at most 64 KiB of LDS: the table slice plus one source span per row
    p.rt = static_cast<int>(std::max<size_t>(1, std::min<size_t>(64, (65536 - tab) / (esz * p.sp))));
    const size_t lds = tab + esz * static_cast<size_t>(p.sp) * p.rt;
    const long long groups = (p.nrows + p.rt - 1) / p.rt;
    const unsigned blocks = static_cast<unsigned>(groups * t.n_cb);
    switch (in_t) {
        case E_F32:
            k_resample_x<E_F32><<<blocks, 256, lds, s>>>(in, in_stride, in_comp, out, om, divisor, p);
            break;
        case E_U16:
            k_resample_x<E_U16><<<blocks, 256, lds, s>>>(in, in_stride, in_comp, out, om, divisor, p);
            break;
        default:
            k_resample_x<E_F64><<<blocks, 256, lds, s>>>(in, in_stride, in_comp, out, om, divisor, p);
    }
}

void launch_resample_y(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int nx, int nz, OutLayout ol, hipStream_t s) {
    const int xblocks = (nx + 255) / 256;
    const long long blocks = static_cast<long long>(xblocks) * t.n_out * nz;
    const double* wp = w + t.w_off;
    const int* ip = idx + t.i_off;
    switch (in_t) {
        case E_F32:
            k_resample_y<E_F32><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, nx, xblocks, ol);
            break;
        case E_U16:
            k_resample_y<E_U16><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, nx, xblocks, ol);
            break;
        default:
            k_resample_y<E_F64><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor,
                                                            wp, ip, t.n_in, t.n_out, t.k, nx, xblocks, ol);
    }
}

void launch_convert(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                    OutMode om, double divisor, int nx, long long nrows, OutLayout ol, hipStream_t s) {
    const size_t n = static_cast<size_t>(nx) * nrows;
    const long long blocks = static_cast<long long>((n + 255) / 256);
    switch (in_t) {
        case E_F32:
            k_convert<E_F32><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, nx, n, ol);
            break;
        case E_U16:
            k_convert<E_U16><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, nx, n, ol);
            break;
        default:
            k_convert<E_F64><<<dim3(blocks), 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, nx, n, ol);
    }
}

void launch_pad_rows(float* buf, long long nrows, int nx, int pitch, hipStream_t s) {
    const long long n = nrows * 2 * kPad;
    k_pad_rows<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(buf, nrows, nx, pitch);
}

}  // namespace wr
