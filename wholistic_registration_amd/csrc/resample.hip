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
// y pass (a slow axis): a thread owns one or two voxels along x of eight
// output rows; every tap is a coalesced row read and the tap table is
// uniform per workgroup.

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "lk_common.h"
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
    // x-pass column groups: the largest power of two <= kXMaxCols whose
    // source span fits kXSpan inputs (and at most 2048 table entries); the
    // span of each group, then whether every output of the group takes its
    // taps from consecutive inputs (no mirrored index)
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
    while (xo * 2 <= kXMaxCols && xo * 2 * t.k <= 2048 && span_of(xo * 2) <= kXSpan) xo *= 2;
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
    t.a_off = i_arena.size();
    for (int g = 0; g < t.n_cb; ++g) {
        bool aff = true;
        for (int j = g * xo; j < std::min(n_out, (g + 1) * xo) && aff; ++j) {
            const int* row = &i_arena[t.i_off + static_cast<size_t>(j) * t.k];
            for (int k = 1; k < t.k; ++k) aff = aff && row[k] == row[0] + k;
        }
        i_arena.push_back(aff ? 1 : 0);
    }
    return t;
}

// ------------------------------------------------------------------ kernels
//
// Contraction stays off for the kernels as well: every tap is a rounded
// float64 product followed by a rounded add (no FMA), as in the reference's
// weights * values then sum.

template <int IN_T>
__device__ __forceinline__ double load_d(const void* p, size_t i) {
    if constexpr (IN_T == E_F32) return static_cast<double>(static_cast<const float*>(p)[i]);
    else if constexpr (IN_T == E_U16) return static_cast<double>(static_cast<const uint16_t*>(p)[i]);
    else return static_cast<const double*>(p)[i];
}

// Stored row of logical row r (see OutLayout).
__device__ __forceinline__ size_t out_row(const OutLayout& ol, long long r) {
    if (ol.plane_rows == 0) return static_cast<size_t>(r);
    const long long z = r / ol.ny;
    return static_cast<size_t>(z * ol.plane_rows + (r - z * ol.ny) + ol.lead);
}

// v / divisor: a multiply when the divisor is a power of two (its reciprocal
// is exact, so the product equals the quotient), else the division.
__device__ __forceinline__ double scaled(double v, double divisor) {
    if (divisor == 1.0) return v;
    if (__builtin_amdgcn_frexp_mant(divisor) == 0.5) return v * (1.0 / divisor);
    return v / divisor;
}

__device__ __forceinline__ void store_out(void* out, size_t i, double v, int om, double divisor) {
    if (om == OUT_F64) static_cast<double*>(out)[i] = v;
    else if (om == OUT_F32) static_cast<float*>(out)[i] = static_cast<float>(scaled(v, divisor));
    else if (om == OUT_BF16) static_cast<uint16_t*>(out)[i] = to_bf16(static_cast<float>(scaled(v, divisor)));
    else if (om == OUT_F16) static_cast<uint16_t*>(out)[i] = to_f16(static_cast<float>(scaled(v, divisor)));
    else static_cast<uint8_t*>(out)[i] = v > 0.0 ? 1 : 0;
}

struct XPass {
    const double* w;
    const int* idx;
    const int* span;
    const int* aff;
    int n_in, n_out, k, xo_blk, n_cb;
    int sp;             // LDS row pitch in staged elements (see launch_resample_x)
    int rt;             // rows per workgroup (<= 64, one per lane)
    int paired;         // dense 2- or 4-byte input aligned to two elements
    long long nrows;
    size_t n_total;     // elements of the input (even for paired staging)
    OutLayout ol;
};

// Output columns of a column group handled by one of the 4 waves.
constexpr int kXColsPerWave = kXMaxCols / 4;

// Staging and x compute of one column group, shared by the x pass and the
// fused x-then-y pass.  The group's source span of each row is staged in LDS
// in the input's own type (uint16, float32 or float64); the compute runs with
// lane = row and wave = column (round robin): the table row of a column is
// wave-uniform (scalar loads), the per-row reads hit distinct banks (odd row
// pitch in dwords), and a group whose outputs read consecutive inputs uses
// immediate LDS offsets.
//
// Paired staging (dense 2- or 4-byte input): each lane loads two elements
// from an even element index, so a row's span starts at LDS element o (0 or
// 1, the parity of its first element) and the pair d lands at elements 2d,
// 2d + 1.  The loads carry no branch (row index clamped to the block's last
// row, pair index to the input's last pair), so all of them are in flight
// together.
//
// Stage the span of column group cb of the rows r0 .. r0 + nr - 1 (nr <= 64)
// and compute the group's outputs of row r0 + lane into res (lane < nr):
// columns wave, wave + 4, ...  Ends after every wave has read the staged rows.
template <int IN_T>
__device__ __forceinline__ void x_group(const void* __restrict__ in, int in_stride, int in_comp, const XPass& p,
                                        long long r0, int nr, int cb, unsigned char* smem_raw, int wave, int lane,
                                        double (&res)[kXColsPerWave]) {
    using S = typename std::conditional<IN_T == E_U16, uint16_t,
                                        typename std::conditional<IN_T == E_F32, float, double>::type>::type;
    S* sin = reinterpret_cast<S*>(smem_raw);
    const int xo0 = cb * p.xo_blk;
    const int nxo = min(p.xo_blk, p.n_out - xo0);
    const int lo = p.span[2 * cb];
    const int span = p.span[2 * cb + 1] - lo + 1;
    const size_t tab0 = static_cast<size_t>(xo0) * p.k;
    const bool paired = IN_T != E_F64 && p.paired;

    if constexpr (IN_T != E_F64) {
        if (paired) {
            typedef S S2 __attribute__((ext_vector_type(2)));
            const S* src = static_cast<const S*>(in);
            const size_t last_pair = p.n_total - 2;
            const int npair = (span + 2) >> 1;
            for (int q = 0; 64 * q < npair; ++q) {
                const int d = lane + 64 * q;
                S2 v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int rr = min(wave + 4 * j, nr - 1);
                    const size_t e0 = static_cast<size_t>(r0 + rr) * p.n_in + lo;
                    const size_t e = min((e0 & ~static_cast<size_t>(1)) + 2 * static_cast<size_t>(d), last_pair);
                    v[j] = *reinterpret_cast<const S2*>(src + e);
                }
                if (d < npair) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int rr = wave + 4 * j;
                        if (rr < nr) {
                            S* dst = sin + rr * p.sp + 2 * d;
                            if constexpr (IN_T == E_U16) {
                                *reinterpret_cast<S2*>(dst) = v[j];   // 4-byte aligned: sp is even
                            } else {
                                dst[0] = v[j].x;
                                dst[1] = v[j].y;
                            }
                        }
                    }
                }
            }
        }
    }
    if (!paired) {
        // strided or float64 input: element loads, 16 rows of a lane in flight
        for (int xx = lane; xx < span; xx += 64) {
            S vals[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int rr = min(wave + 4 * j, nr - 1);
                vals[j] = static_cast<S>(
                    load_d<IN_T>(in, ((static_cast<size_t>(r0 + rr) * p.n_in + lo + xx) * in_stride) + in_comp));
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int rr = wave + 4 * j;
                if (rr < nr) sin[rr * p.sp + xx] = vals[j];
            }
        }
    }
    __syncthreads();
    // the outputs of a lane (its row, columns wave, wave + 4, ...) stay in
    // registers until every wave has finished reading the staged rows
    if (lane < nr) {
        const int o = paired ? static_cast<int>((static_cast<size_t>(r0 + lane) * p.n_in + lo) & 1) : 0;
        const S* rl = sin + lane * p.sp + o - lo;     // rl[i] = input column i of this row
        const bool aff = p.aff[cb] != 0;
#pragma unroll
        for (int q = 0; q < kXColsPerWave; ++q) {
            const int c = wave + 4 * q;
            if (c < nxo) {
                const double* wc = p.w + tab0 + static_cast<size_t>(c) * p.k;
                const int* ic = p.idx + tab0 + static_cast<size_t>(c) * p.k;
                double acc;
                if (aff) {
                    // taps ic[0] + t: one base address, immediate offsets
                    const S* b = rl + ic[0];
                    acc = wc[0] * static_cast<double>(b[0]);
                    int t = 1;
                    for (; t + 8 <= p.k; t += 8) {
#pragma unroll
                        for (int u = 0; u < 8; ++u) acc = acc + wc[t + u] * static_cast<double>(b[t + u]);
                    }
                    for (; t < p.k; ++t) acc = acc + wc[t] * static_cast<double>(b[t]);
                } else {
                    acc = wc[0] * static_cast<double>(rl[ic[0]]);
#pragma unroll 8
                    for (int t = 1; t < p.k; ++t) acc = acc + wc[t] * static_cast<double>(rl[ic[t]]);
                }
                res[q] = acc;
            }
        }
    }
    __syncthreads();
}

// x pass: a workgroup owns `rt` rows (one per lane) and one column group;
// the results are transposed through LDS and stored along the rows.
template <int IN_T>
__global__ __launch_bounds__(256) void k_resample_x(const void* __restrict__ in, int in_stride, int in_comp,
                                                     void* __restrict__ out, int om, double divisor, XPass p) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int cb = static_cast<int>(blockIdx.x % p.n_cb);
    const long long r0 = static_cast<long long>(blockIdx.x / p.n_cb) * p.rt;
    const int nr = static_cast<int>(min(static_cast<long long>(p.rt), p.nrows - r0));
    const int xo0 = cb * p.xo_blk;
    const int nxo = min(p.xo_blk, p.n_out - xo0);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    double res[kXColsPerWave];
    x_group<IN_T>(in, in_stride, in_comp, p, r0, nr, cb, smem_raw, wave, lane, res);
    // transpose through LDS (odd pitch po, in doubles, over the staged rows)
    // so that the stores run along the rows
    double* so = reinterpret_cast<double*>(smem_raw);
    const int po = p.xo_blk | 1;
    if (lane < nr) {
#pragma unroll
        for (int q = 0; q < kXColsPerWave; ++q) {
            const int c = wave + 4 * q;
            if (c < nxo) so[lane * po + c] = res[q];
        }
    }
    __syncthreads();
    const int lg = __builtin_ctz(p.xo_blk);
    for (int i = threadIdx.x; i < (nr << lg); i += 256) {
        const int rr = i >> lg, c = i & (p.xo_blk - 1);
        if (c < nxo)
            store_out(out, out_row(p.ol, r0 + rr) * p.ol.pitch + p.ol.pad + xo0 + c, so[rr * po + c], om,
                      divisor);
    }
}

// Fused x-then-y pass (both axes resampled, x first): a workgroup owns one
// column group of the x pass and `ty_rows` output rows of one z slice.  It
// runs the x pass on the source rows those outputs read (at most 64, one per
// lane), keeps the float64 results in LDS, and applies the y taps from there,
// so the float64 intermediate never goes to HBM.  Same arithmetic, in the
// same order, as the two separate passes.
struct YFused {
    const double* w;
    const int* idx;
    const int* span;    // per block of output rows: first and last source row
    int n_in, n_out, k, rows, nblk;
};

template <int IN_T>
__global__ __launch_bounds__(256) void k_resample_xy(const void* __restrict__ in, void* __restrict__ out, int om,
                                                      double divisor, XPass p, YFused q, OutLayout ol,
                                                      unsigned ytab_off) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const unsigned b = blockIdx.x;
    const int cb = static_cast<int>(b % static_cast<unsigned>(p.n_cb));
    const unsigned rest = b / static_cast<unsigned>(p.n_cb);
    const int yb = static_cast<int>(rest % static_cast<unsigned>(q.nblk));
    const int z = static_cast<int>(rest / static_cast<unsigned>(q.nblk));
    const int ylo = q.span[2 * yb];
    const int nr = q.span[2 * yb + 1] - ylo + 1;
    const long long r0 = static_cast<long long>(z) * q.n_in + ylo;
    const int xo0 = cb * p.xo_blk;
    const int nxo = min(p.xo_blk, p.n_out - xo0);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    // the y taps of the block's output rows, staged in LDS (the two output
    // rows of a wave read them as broadcasts); visible after x_group's barrier
    const int yo0 = yb * q.rows;
    const int yo1 = min(yo0 + q.rows, q.n_out);
    double* swy = reinterpret_cast<double*>(smem_raw + ytab_off);
    int* siy = reinterpret_cast<int*>(swy + q.rows * q.k);
    for (int i = threadIdx.x; i < (yo1 - yo0) * q.k; i += 256) {
        swy[i] = q.w[static_cast<size_t>(yo0) * q.k + i];
        siy[i] = q.idx[static_cast<size_t>(yo0) * q.k + i] - ylo;
    }
    double res[kXColsPerWave];
    x_group<IN_T>(in, 1, 0, p, r0, nr, cb, smem_raw, wave, lane, res);
    // x results of the block's source rows: xr[row][column], odd pitch
    double* xr = reinterpret_cast<double*>(smem_raw);
    const int po = p.xo_blk | 1;
    if (lane < nr) {
#pragma unroll
        for (int u = 0; u < kXColsPerWave; ++u) {
            const int c = wave + 4 * u;
            if (c < nxo) xr[lane * po + c] = res[u];
        }
    }
    __syncthreads();
    // y taps: lanes along the columns, 8 output rows at a time
    const int c = threadIdx.x & (kXMaxCols - 1);
    if (c >= nxo) return;
    for (int yo = yo0 + static_cast<int>(threadIdx.x) / kXMaxCols; yo < yo1; yo += 256 / kXMaxCols) {
        const double* wy = swy + (yo - yo0) * q.k;
        const int* iy = siy + (yo - yo0) * q.k;
        double acc = wy[0] * xr[iy[0] * po + c];
        for (int t = 1; t < q.k; ++t) acc = acc + wy[t] * xr[iy[t] * po + c];
        store_out(out, out_row(ol, static_cast<long long>(z) * q.n_out + yo) * ol.pitch + ol.pad + xo0 + c, acc, om,
                  divisor);
    }
}

// Upsample (both axes, x first, few taps per output): a thread owns one
// output column x of a block of output rows of one z slice.  It forms the x
// pass of each source row the block reads (at most kUpRows) into its own LDS
// column, then applies the y taps of every output row from there.  Same
// float64 arithmetic and order as the two separate passes; no intermediate in
// HBM and no barrier (a thread reads back only what it wrote).
struct UpPass {
    const double* wx;
    const int* ix;
    const double* wy;
    const int* iy;
    const int* span;    // per block of output rows: first and last source row
    int nx_in, nx_out, kx;
    int ny_in, ny_out, ky;
    int rows, nblk, xblocks;
};

template <int IN_T, int KX>
__global__ __launch_bounds__(256) void k_resample_up(const void* __restrict__ in, void* __restrict__ out, int om,
                                                      double divisor, UpPass u, OutLayout ol) {
    __shared__ double xr[kUpRows][256];
    const unsigned b = blockIdx.x;
    const int xb = static_cast<int>(b % static_cast<unsigned>(u.xblocks));
    const unsigned rest = b / static_cast<unsigned>(u.xblocks);
    const int yb = static_cast<int>(rest % static_cast<unsigned>(u.nblk));
    const int z = static_cast<int>(rest / static_cast<unsigned>(u.nblk));
    const int x = xb * 256 + static_cast<int>(threadIdx.x);
    if (x >= u.nx_out) return;
    const int ylo = u.span[2 * yb];
    const int nr = u.span[2 * yb + 1] - ylo + 1;
    double w[KX];
    int ix[KX];
#pragma unroll
    for (int t = 0; t < KX; ++t) {
        w[t] = u.wx[static_cast<size_t>(x) * KX + t];
        ix[t] = u.ix[static_cast<size_t>(x) * KX + t];
    }
    // four source rows at a time, every load issued before any is used (row
    // index clamped to the block's last source row)
    for (int j0 = 0; j0 < nr; j0 += 4) {
        double v[4][KX];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const size_t row = (static_cast<size_t>(z) * u.ny_in + ylo + min(j0 + jj, nr - 1)) * u.nx_in;
#pragma unroll
            for (int t = 0; t < KX; ++t) v[jj][t] = load_d<IN_T>(in, row + ix[t]);
        }
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            double acc = w[0] * v[jj][0];
#pragma unroll
            for (int t = 1; t < KX; ++t) acc = acc + w[t] * v[jj][t];
            if (j0 + jj < nr) xr[j0 + jj][threadIdx.x] = acc;
        }
    }
    const int y0 = yb * u.rows;
    const int y1 = min(y0 + u.rows, u.ny_out);
    for (int yo = y0; yo < y1; ++yo) {
        const double* wy = u.wy + static_cast<size_t>(yo) * u.ky;
        const int* iy = u.iy + static_cast<size_t>(yo) * u.ky;
        double acc = wy[0] * xr[iy[0] - ylo][threadIdx.x];
        for (int t = 1; t < u.ky; ++t) acc = acc + wy[t] * xr[iy[t] - ylo][threadIdx.x];
        store_out(out, out_row(ol, static_cast<long long>(z) * u.ny_out + yo) * ol.pitch + ol.pad + x, acc, om,
                  divisor);
    }
}

// V consecutive elements of a dense row (V > 1: aligned to V elements).
template <int IN_T, int V>
__device__ __forceinline__ void load_v(const void* p, size_t i, int stride, int comp, double (&v)[V]) {
    if constexpr (V == 1) {
        v[0] = load_d<IN_T>(p, i * stride + comp);
    } else {
        using E = typename std::conditional<IN_T == E_U16, uint16_t,
                                            typename std::conditional<IN_T == E_F32, float, double>::type>::type;
        typedef E EV __attribute__((ext_vector_type(V)));
        const EV e = *reinterpret_cast<const EV*>(static_cast<const E*>(p) + i);
#pragma unroll
        for (int q = 0; q < V; ++q) v[q] = static_cast<double>(e[q]);
    }
}

template <int V>
__device__ __forceinline__ void store_v(void* out, size_t i, const double (&v)[V], int om, double divisor) {
    if constexpr (V == 1) {
        store_out(out, i, v[0], om, divisor);
    } else {
        if (om == OUT_F64) {
            typedef double DV __attribute__((ext_vector_type(V)));
            DV o;
#pragma unroll
            for (int q = 0; q < V; ++q) o[q] = v[q];
            *reinterpret_cast<DV*>(static_cast<double*>(out) + i) = o;
        } else if (om == OUT_F32) {
            typedef float FV __attribute__((ext_vector_type(V)));
            FV o;
#pragma unroll
            for (int q = 0; q < V; ++q) o[q] = static_cast<float>(scaled(v[q], divisor));
            *reinterpret_cast<FV*>(static_cast<float*>(out) + i) = o;
        } else if (om == OUT_BF16 || om == OUT_F16) {
            typedef uint16_t HV __attribute__((ext_vector_type(V)));
            HV o;
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const float f = static_cast<float>(scaled(v[q], divisor));
                o[q] = om == OUT_BF16 ? to_bf16(f) : to_f16(f);
            }
            *reinterpret_cast<HV*>(static_cast<uint16_t*>(out) + i) = o;
        } else {
            typedef uint8_t BV __attribute__((ext_vector_type(V)));
            BV o;
#pragma unroll
            for (int q = 0; q < V; ++q) o[q] = v[q] > 0.0 ? 1 : 0;
            *reinterpret_cast<BV*>(static_cast<uint8_t*>(out) + i) = o;
        }
    }
}

// Output rows of one y-pass workgroup.
constexpr int kYRows = 8;

// y pass: a thread owns V consecutive x of kYRows consecutive output rows of
// one z slice; the tap table of a row is wave-uniform.
template <int IN_T, int V>
__global__ __launch_bounds__(256) void k_resample_y(const void* __restrict__ in, int in_stride,
                                                     int in_comp, void* __restrict__ out, int om,
                                                     double divisor, const double* __restrict__ w,
                                                     const int* __restrict__ idx, int n_in,
                                                     int n_out, int k, int nx, int xblocks, int yblocks,
                                                     OutLayout ol) {
    const unsigned b = blockIdx.x;
    const int xb = static_cast<int>(b % static_cast<unsigned>(xblocks));
    const unsigned rest = b / static_cast<unsigned>(xblocks);
    const int yb = static_cast<int>(rest % static_cast<unsigned>(yblocks));
    const int z = static_cast<int>(rest / static_cast<unsigned>(yblocks));
    const int x = (xb * 256 + static_cast<int>(threadIdx.x)) * V;
    if (x >= nx) return;
    const size_t base = static_cast<size_t>(z) * n_in;
    const int y0 = yb * kYRows;
    const int y1 = min(y0 + kYRows, n_out);
    for (int yo = y0; yo < y1; ++yo) {
        const double* wr = w + static_cast<size_t>(yo) * k;
        const int* ir = idx + static_cast<size_t>(yo) * k;
        double acc[V], v[V];
        load_v<IN_T, V>(in, (base + ir[0]) * nx + x, in_stride, in_comp, v);
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = wr[0] * v[q];
#pragma unroll 4
        for (int t = 1; t < k; ++t) {
            load_v<IN_T, V>(in, (base + ir[t]) * nx + x, in_stride, in_comp, v);
#pragma unroll
            for (int q = 0; q < V; ++q) acc[q] = acc[q] + wr[t] * v[q];
        }
        store_v<V>(out, out_row(ol, static_cast<long long>(z) * n_out + yo) * ol.pitch + ol.pad + x, acc, om, divisor);
    }
}

// Elementwise conversion: a thread owns V consecutive x of one row.
template <int IN_T, int V>
__global__ __launch_bounds__(256) void k_convert(const void* __restrict__ in, int in_stride,
                                                  int in_comp, void* __restrict__ out, int om,
                                                  double divisor, int nx, int xblocks, OutLayout ol) {
    const unsigned b = blockIdx.x;
    const unsigned row = b / static_cast<unsigned>(xblocks);
    const int x = (static_cast<int>(b - row * static_cast<unsigned>(xblocks)) * 256 + static_cast<int>(threadIdx.x)) * V;
    if (x >= nx) return;
    double v[V];
    load_v<IN_T, V>(in, static_cast<size_t>(row) * nx + x, in_stride, in_comp, v);
    store_v<V>(out, out_row(ol, row) * ol.pitch + ol.pad + x, v, om, divisor);
}

// Pads of a padded volume (padded_layout): x ends of the logical rows, then
// (next launch) the y pad rows of every plane, then (next launch) the z pad
// planes, each a copy of the nearest logical row or plane.
template <class T>
__global__ __launch_bounds__(256) void k_pad_x(T* __restrict__ buf, OutLayout ol, long long nrows, int nx) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= nrows * 2 * kPad) return;
    const long long r = i / (2 * kPad);
    const int j = static_cast<int>(i - r * 2 * kPad);
    T* row = buf + out_row(ol, r) * ol.pitch;
    if (j < kPad) row[j] = row[kPad];
    else row[kPad + nx + (j - kPad)] = row[kPad + nx - 1];
}

template <class T>
__global__ __launch_bounds__(256) void k_pad_y(T* __restrict__ buf, OutLayout ol, int nz) {
    // (plane, pad row, element): kPadLo + kPadHi pad rows of `pitch` floats
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const int per_plane = (kPadLo + kPadHi) * ol.pitch;
    if (i >= static_cast<long long>(nz) * per_plane) return;
    const int z = static_cast<int>(i / per_plane);
    const int k = static_cast<int>(i - static_cast<long long>(z) * per_plane);
    const int pr = k / ol.pitch, e = k - pr * ol.pitch;
    const size_t plane0 = static_cast<size_t>(z + kPadLo) * ol.plane_rows;    // stored row of pad row 0
    const int dst = pr < kPadLo ? pr : kPadLo + ol.ny + (pr - kPadLo);
    const int src = pr < kPadLo ? kPadLo : kPadLo + ol.ny - 1;
    buf[(plane0 + dst) * ol.pitch + e] = buf[(plane0 + src) * ol.pitch + e];
}

template <class T>
__global__ __launch_bounds__(256) void k_pad_z(T* __restrict__ buf, OutLayout ol, int nz) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const long long plane = static_cast<long long>(ol.plane_rows) * ol.pitch;
    if (i >= (kPadLo + kPadHi) * plane) return;
    const int pp = static_cast<int>(i / plane);
    const long long e = i - pp * plane;
    const int dst = pp < kPadLo ? pp : kPadLo + nz + (pp - kPadLo);
    const int src = pp < kPadLo ? kPadLo : kPadLo + nz - 1;
    buf[dst * plane + e] = buf[src * plane + e];
}

// ------------------------------------------------------------------ launchers

namespace {

// x-pass LDS row pitch (staged elements): room for the span shifted by one
// element plus the second element of the last pair, and an odd number of
// dwords (conflict-free reads with one row per lane); uint16 rows hold an even
// element count.
int x_pitch(Elem in_t, int span_max) {
    if (in_t == E_U16) return 2 * (((span_max + 3) / 2) | 1);
    if (in_t == E_F32) return (span_max + 2) | 1;
    return span_max | 1;
}

// LDS bytes per staged row: the larger of its span and its float64 outputs.
size_t x_lds_per_row(Elem in_t, const Table& t) {
    const size_t esz = in_t == E_U16 ? 2 : (in_t == E_F32 ? 4 : 8);
    return std::max<size_t>(esz * x_pitch(in_t, t.span_max), sizeof(double) * (t.xo_blk | 1));
}

XPass make_xpass(const void* in, Elem in_t, int in_stride, int in_comp, const Table& t, const double* w,
                 const int* idx, long long nrows, OutLayout ol, size_t* lds_per_row) {
    XPass p;
    p.w = w + t.w_off;
    p.idx = idx + t.i_off;
    p.span = idx + t.s_off;
    p.aff = idx + t.a_off;
    p.n_in = t.n_in;
    p.n_out = t.n_out;
    p.k = t.k;
    p.xo_blk = t.xo_blk;
    p.n_cb = t.n_cb;
    p.nrows = nrows;
    p.ol = ol;
    p.n_total = static_cast<size_t>(p.nrows) * t.n_in;
    const size_t esz = in_t == E_U16 ? 2 : (in_t == E_F32 ? 4 : 8);
    p.paired = in_t != E_F64 && in_stride == 1 && in_comp == 0 && p.n_total % 2 == 0 &&
               reinterpret_cast<uintptr_t>(in) % (2 * esz) == 0;
    p.sp = x_pitch(in_t, t.span_max);
    *lds_per_row = x_lds_per_row(in_t, t);
    p.rt = static_cast<int>(std::max<size_t>(1, std::min<size_t>(64, 65536 / *lds_per_row)));
    return p;
}

}  // namespace

void launch_resample_x(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int ny, int nz, OutLayout ol, hipStream_t s) {
    size_t per_row = 0;
    const XPass p = make_xpass(in, in_t, in_stride, in_comp, t, w, idx, static_cast<long long>(ny) * nz, ol,
                               &per_row);
    const size_t lds = per_row * p.rt;
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

FusedY plan_fused_y(const Table& ty, std::vector<int>& i_arena) {
    FusedY f;
    // downsampling only: an upsample's y pass writes more than it reads, and
    // the separate y pass (two x per thread, eight rows per workgroup) is the
    // faster way to stream it
    if (ty.identity || ty.n_out >= ty.n_in) return f;
    const int* idx = &i_arena[ty.i_off];
    auto worst_span = [&](int rows) {
        int worst = 0;
        for (int b = 0; b * rows < ty.n_out; ++b) {
            int lo = ty.n_in, hi = 0;
            for (int j = b * rows; j < std::min(ty.n_out, (b + 1) * rows); ++j)
                for (int k = 0; k < ty.k; ++k) {
                    lo = std::min(lo, idx[static_cast<size_t>(j) * ty.k + k]);
                    hi = std::max(hi, idx[static_cast<size_t>(j) * ty.k + k]);
                }
            worst = std::max(worst, hi - lo + 1);
        }
        return worst;
    };
    int rows = std::min(ty.n_out, 64);
    while (rows >= kFusedMinRows && worst_span(rows) > 64) --rows;
    if (rows < kFusedMinRows) return f;
    f.rows = rows;
    f.nblk = (ty.n_out + rows - 1) / rows;
    f.s_off = i_arena.size();
    for (int b = 0; b < f.nblk; ++b) {
        int lo = ty.n_in, hi = 0;
        for (int j = b * rows; j < std::min(ty.n_out, (b + 1) * rows); ++j)
            for (int k = 0; k < ty.k; ++k) {
                lo = std::min(lo, i_arena[ty.i_off + static_cast<size_t>(j) * ty.k + k]);
                hi = std::max(hi, i_arena[ty.i_off + static_cast<size_t>(j) * ty.k + k]);
            }
        i_arena.push_back(lo);
        i_arena.push_back(hi);
    }
    return f;
}


bool fused_ok(const Table& tx, const Table& ty, const FusedY& fy, Elem in_t) {
    // 64 staged source rows and the y taps of a block must fit 64 KiB of LDS
    const size_t rows = (x_lds_per_row(in_t, tx) * 64 + 15) & ~static_cast<size_t>(15);
    return fy.rows > 0 && !tx.identity &&
           rows + static_cast<size_t>(fy.rows) * ty.k * (sizeof(double) + sizeof(int)) <= 65536;
}

FusedY plan_up_y(const Table& tx, const Table& ty, std::vector<int>& i_arena) {
    FusedY f;
    if (tx.identity || ty.identity || tx.k < 2 || tx.k > kUpTaps) return f;
    const int* idx = &i_arena[ty.i_off];
    auto block_span = [&](int b, int rows, int& lo, int& hi) {
        lo = ty.n_in;
        hi = 0;
        for (int j = b * rows; j < std::min(ty.n_out, (b + 1) * rows); ++j)
            for (int k = 0; k < ty.k; ++k) {
                lo = std::min(lo, idx[static_cast<size_t>(j) * ty.k + k]);
                hi = std::max(hi, idx[static_cast<size_t>(j) * ty.k + k]);
            }
    };
    // the most output rows per block whose source rows fit kUpRows, and at
    // least as many output rows as source rows (an upsample in y)
    int rows = std::min(ty.n_out, 32);
    for (; rows >= 1; --rows) {
        int worst = 0;
        for (int b = 0; b * rows < ty.n_out; ++b) {
            int lo, hi;
            block_span(b, rows, lo, hi);
            worst = std::max(worst, hi - lo + 1);
        }
        if (worst <= kUpRows && worst <= rows) break;
    }
    if (rows < kFusedMinRows) return f;
    f.rows = rows;
    f.nblk = (ty.n_out + rows - 1) / rows;
    f.s_off = i_arena.size();
    for (int b = 0; b < f.nblk; ++b) {
        int lo, hi;
        block_span(b, rows, lo, hi);
        i_arena.push_back(lo);
        i_arena.push_back(hi);
    }
    return f;
}

namespace {

template <int IN_T>
void launch_up_k(unsigned blocks, const void* in, void* out, int om, double divisor, const UpPass& u, OutLayout ol,
                 hipStream_t s) {
    switch (u.kx) {
        case 2:
            k_resample_up<IN_T, 2><<<blocks, 256, 0, s>>>(in, out, om, divisor, u, ol);
            break;
        case 3:
            k_resample_up<IN_T, 3><<<blocks, 256, 0, s>>>(in, out, om, divisor, u, ol);
            break;
        case 4:
            k_resample_up<IN_T, 4><<<blocks, 256, 0, s>>>(in, out, om, divisor, u, ol);
            break;
        case 5:
            k_resample_up<IN_T, 5><<<blocks, 256, 0, s>>>(in, out, om, divisor, u, ol);
            break;
        default:
            k_resample_up<IN_T, 6><<<blocks, 256, 0, s>>>(in, out, om, divisor, u, ol);
    }
}

}  // namespace

void launch_resample_up(const void* in, Elem in_t, void* out, OutMode om, double divisor, const Table& tx,
                        const Table& ty, const FusedY& fu, const double* w, const int* idx, int nz, OutLayout ol,
                        hipStream_t s) {
    UpPass u;
    u.wx = w + tx.w_off;
    u.ix = idx + tx.i_off;
    u.wy = w + ty.w_off;
    u.iy = idx + ty.i_off;
    u.span = idx + fu.s_off;
    u.nx_in = tx.n_in;
    u.nx_out = tx.n_out;
    u.kx = tx.k;
    u.ny_in = ty.n_in;
    u.ny_out = ty.n_out;
    u.ky = ty.k;
    u.rows = fu.rows;
    u.nblk = fu.nblk;
    u.xblocks = (tx.n_out + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(static_cast<long long>(u.xblocks) * fu.nblk * nz);
    switch (in_t) {
        case E_F32:
            launch_up_k<E_F32>(blocks, in, out, om, divisor, u, ol, s);
            break;
        case E_U16:
            launch_up_k<E_U16>(blocks, in, out, om, divisor, u, ol, s);
            break;
        default:
            launch_up_k<E_F64>(blocks, in, out, om, divisor, u, ol, s);
    }
}

void launch_resample_xy(const void* in, Elem in_t, void* out, OutMode om, double divisor, const Table& tx,
                        const Table& ty, const FusedY& fy, const double* w, const int* idx, int nz, OutLayout ol,
                        hipStream_t s) {
    size_t per_row = 0;
    const XPass p = make_xpass(in, in_t, 1, 0, tx, w, idx, static_cast<long long>(ty.n_in) * nz, ol, &per_row);
    YFused q;
    q.w = w + ty.w_off;
    q.idx = idx + ty.i_off;
    q.span = idx + fy.s_off;
    q.n_in = ty.n_in;
    q.n_out = ty.n_out;
    q.k = ty.k;
    q.rows = fy.rows;
    q.nblk = fy.nblk;
    // 64 staged rows, then the y taps of one block of output rows
    const unsigned ytab_off = static_cast<unsigned>((per_row * 64 + 15) & ~static_cast<size_t>(15));
    const size_t lds = ytab_off + static_cast<size_t>(fy.rows) * ty.k * (sizeof(double) + sizeof(int));
    const unsigned blocks = static_cast<unsigned>(static_cast<long long>(tx.n_cb) * fy.nblk * nz);
    switch (in_t) {
        case E_F32:
            k_resample_xy<E_F32><<<blocks, 256, lds, s>>>(in, out, om, divisor, p, q, ol, ytab_off);
            break;
        case E_U16:
            k_resample_xy<E_U16><<<blocks, 256, lds, s>>>(in, out, om, divisor, p, q, ol, ytab_off);
            break;
        default:
            k_resample_xy<E_F64><<<blocks, 256, lds, s>>>(in, out, om, divisor, p, q, ol, ytab_off);
    }
}

namespace {

size_t elem_bytes(Elem e) { return e == E_U16 ? 2 : (e == E_F32 ? 4 : 8); }
size_t out_bytes(OutMode om) {
    return om == OUT_F64 ? 8 : (om == OUT_F32 ? 4 : (om == OUT_BF16 || om == OUT_F16 ? 2 : 1));
}

// V-element vector access is possible: dense input, rows and output rows that
// start on multiples of V elements, and V-aligned base pointers.
bool vec_ok(int v, const void* in, Elem in_t, int in_stride, int in_comp, const void* out, OutMode om, int nx,
            OutLayout ol) {
    return in_stride == 1 && in_comp == 0 && nx % v == 0 && ol.pitch % v == 0 && ol.pad % v == 0 &&
           reinterpret_cast<uintptr_t>(in) % (v * elem_bytes(in_t)) == 0 &&
           reinterpret_cast<uintptr_t>(out) % (v * out_bytes(om)) == 0;
}

template <int V>
void launch_y(const void* in, Elem in_t, int in_stride, int in_comp, void* out, OutMode om, double divisor,
              const Table& t, const double* wp, const int* ip, int nx, int nz, OutLayout ol, hipStream_t s) {
    const int xblocks = (nx / V + 255) / 256;
    const int yblocks = (t.n_out + kYRows - 1) / kYRows;
    const unsigned blocks = static_cast<unsigned>(static_cast<long long>(xblocks) * yblocks * nz);
    switch (in_t) {
        case E_F32:
            k_resample_y<E_F32, V><<<blocks, 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, wp, ip, t.n_in,
                                                          t.n_out, t.k, nx, xblocks, yblocks, ol);
            break;
        case E_U16:
            k_resample_y<E_U16, V><<<blocks, 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, wp, ip, t.n_in,
                                                          t.n_out, t.k, nx, xblocks, yblocks, ol);
            break;
        default:
            k_resample_y<E_F64, V><<<blocks, 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, wp, ip, t.n_in,
                                                          t.n_out, t.k, nx, xblocks, yblocks, ol);
    }
}

template <int V>
void launch_conv(const void* in, Elem in_t, int in_stride, int in_comp, void* out, OutMode om, double divisor,
                 int nx, long long nrows, OutLayout ol, hipStream_t s) {
    const int xblocks = (nx / V + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(xblocks * nrows);
    switch (in_t) {
        case E_F32:
            k_convert<E_F32, V><<<blocks, 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, nx, xblocks, ol);
            break;
        case E_U16:
            k_convert<E_U16, V><<<blocks, 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, nx, xblocks, ol);
            break;
        default:
            k_convert<E_F64, V><<<blocks, 256, 0, s>>>(in, in_stride, in_comp, out, om, divisor, nx, xblocks, ol);
    }
}

}  // namespace

void launch_resample_y(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int nx, int nz, OutLayout ol, hipStream_t s) {
    const double* wp = w + t.w_off;
    const int* ip = idx + t.i_off;
    if (vec_ok(2, in, in_t, in_stride, in_comp, out, om, nx, ol))
        launch_y<2>(in, in_t, in_stride, in_comp, out, om, divisor, t, wp, ip, nx, nz, ol, s);
    else
        launch_y<1>(in, in_t, in_stride, in_comp, out, om, divisor, t, wp, ip, nx, nz, ol, s);
}

void launch_convert(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                    OutMode om, double divisor, int nx, long long nrows, OutLayout ol, hipStream_t s) {
    if (vec_ok(4, in, in_t, in_stride, in_comp, out, om, nx, ol))
        launch_conv<4>(in, in_t, in_stride, in_comp, out, om, divisor, nx, nrows, ol, s);
    else
        launch_conv<1>(in, in_t, in_stride, in_comp, out, om, divisor, nx, nrows, ol, s);
}

namespace {
template <class T>
void pad_volume(T* buf, int nx, int ny, int nz, hipStream_t s) {
    const OutLayout ol = padded_layout(nx, ny);
    const long long nrows = static_cast<long long>(ny) * nz;
    const long long nx_ = nrows * 2 * kPad;
    k_pad_x<T><<<static_cast<unsigned>((nx_ + 255) / 256), 256, 0, s>>>(buf, ol, nrows, nx);
    const long long ny_ = static_cast<long long>(nz) * (kPadLo + kPadHi) * ol.pitch;
    k_pad_y<T><<<static_cast<unsigned>((ny_ + 255) / 256), 256, 0, s>>>(buf, ol, nz);
    const long long nz_ = static_cast<long long>(kPadLo + kPadHi) * ol.plane_rows * ol.pitch;
    k_pad_z<T><<<static_cast<unsigned>((nz_ + 255) / 256), 256, 0, s>>>(buf, ol, nz);
}
}  // namespace

void launch_pad_volume(void* buf, int esize, int nx, int ny, int nz, hipStream_t s) {
    if (esize == 2)
        pad_volume(static_cast<uint16_t*>(buf), nx, ny, nz, s);
    else
        pad_volume(static_cast<float*>(buf), nx, ny, nz, s);
}

}  // namespace wr
