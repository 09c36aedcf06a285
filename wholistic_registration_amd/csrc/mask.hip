// Automatic masks and the error metric of the registration harness on gfx950.
//
//   robust_mean_std   src/wholistic_registration/utils/preprocess.py:69-84
//   getMask           src/wholistic_registration/utils/mask.py:59-111
//   bwareafilt3_wei   src/wholistic_registration/utils/mask.py:8-56
//   mean((a - b)**2)  src/wholistic_registration/utils/registration.py:328-329
//
// Robust statistics.  The 95th percentile is numpy's linear method on the
// float32 data: the two order statistics it interpolates are found exactly by
// a three-pass radix select over order-preserving 32-bit keys (11 + 11 + 10
// bits, per-workgroup LDS histograms), plus one min pass when the upper one
// differs from the lower; the interpolation is done in float32 without
// contraction, as numpy does it.  Mean and std of the values at or below the
// threshold are accumulated in float64 (numpy sums float32 pairwise), per
// workgroup in a fixed order, then reduced by one workgroup: the result does
// not depend on scheduling.
//
// uint16 data skip the radix passes: one pass builds the exact histogram of
// the 65536 values (16-bit LDS counters, 128 KB per workgroup), from which
// one workgroup takes the order statistics, the threshold and both moments.
//
// getMask.  The z-score test |(v - mu) / sigma| > thres is monotone on each
// side of mu, so it is replaced by two exact float cut-offs found by
// bisection.  The 3x3 in-plane erosion (outside the volume counts as set),
// dilation and dilation (outside counts as clear) -- scikit-image's
// binary_opening followed by binary_dilation with a 3x3x1 footprint -- are
// separable and run on 32-voxel bit words: one thread per word of a row, a
// rolling window of rows, neighbour words through LDS; the volume is read
// once (plus a 6-row halo per 32 rows) and the byte mask written once.
//
// bwareafilt3_wei.  Union-find connected components over the 26-neighbourhood
// (each voxel unites with its 13 already-scanned neighbours; a root is linked
// under the smaller label with atomicMin, the method of Playne and Hawick,
// "A new algorithm for parallel connected-component labelling on GPUs", IEEE
// TPDS 2018), path flattening with wave-aggregated area counts per root, and
// the closed area range test; labels and areas are touched only for
// foreground voxels.  The output does not depend on which voxel becomes a
// root.

#include <cmath>
#include <cstdint>

#include "wr_internal.h"

// the float32 steps below are rounded one at a time, as numpy rounds them
#pragma clang fp contract(off)

namespace wr {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ unsigned mono_key(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_value(unsigned k) {
    const unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

template <int T>
__device__ __forceinline__ float load_f32(const void* p, size_t i) {
    if constexpr (T == E_U16)
        return static_cast<float>(static_cast<const uint16_t*>(p)[i]);
    else
        return static_cast<const float*>(p)[i];
}

// --------------------------------------------------------------- selection

__global__ void k_sel_reset(SelCtl* c, unsigned* hist, int nhist, long long prev, long long next, float gamma) {
    for (int i = threadIdx.x; i < nhist; i += blockDim.x) hist[i] = 0u;
    if (threadIdx.x == 0) {
        c->prefix = 0u;
        c->pmask = 0u;
        c->rank = prev;
        c->less = 0;
        c->want_next = next;
        c->key_lo = 0u;
        c->key_hi = 0xffffffffu;
        c->need_min = 0;
        c->gamma = gamma;
    }
}

// Histogram of key bits [shift, shift + nbits) over the keys whose already
// fixed bits match the prefix.
template <int T>
__global__ __launch_bounds__(kT) void k_hist(const void* __restrict__ in, long long n, const SelCtl* __restrict__ c,
                                             int shift, int nbits, unsigned* __restrict__ hist) {
    __shared__ unsigned h[2048];
    const int nb = 1 << nbits;
    for (int b = threadIdx.x; b < nb; b += kT) h[b] = 0u;
    __syncthreads();
    const unsigned pm = c->pmask, pf = c->prefix;
    const unsigned bm = static_cast<unsigned>(nb - 1);
    const size_t stride = static_cast<size_t>(gridDim.x) * kT;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x; i < static_cast<size_t>(n); i += stride) {
        const unsigned k = mono_key(load_f32<T>(in, i));
        if ((k & pm) == pf) atomicAdd(&h[(k >> shift) & bm], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kT)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// One workgroup: the bin holding the wanted rank; fixes its bits.
__global__ __launch_bounds__(kT) void k_select(const unsigned* __restrict__ hist, int shift, int nbits, int last,
                                               SelCtl* __restrict__ c) {
    __shared__ long long s[kT];
    const int nb = 1 << nbits;
    const int per = nb / kT;
    const int t = threadIdx.x;
    const long long rank = c->rank;
    long long local = 0;
    for (int j = 0; j < per; ++j) local += hist[t * per + j];
    s[t] = local;
    __syncthreads();
    for (int off = 1; off < kT; off <<= 1) {
        const long long v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    const long long before = t ? s[t - 1] : 0;
    if (before <= rank && rank < s[t]) {
        long long cum = before;
        int bin = t * per;
        for (int j = 0; j < per; ++j) {
            const long long hb = hist[t * per + j];
            if (rank < cum + hb) {
                bin = t * per + j;
                break;
            }
            cum += hb;
        }
        c->prefix |= static_cast<unsigned>(bin) << shift;
        c->pmask |= static_cast<unsigned>(nb - 1) << shift;
        c->rank = rank - cum;
        c->less += cum;
        if (last) {
            const long long count_le = c->less + hist[bin];
            c->key_lo = c->prefix;
            if (c->want_next < count_le) {
                c->key_hi = c->prefix;
                c->need_min = 0;
            } else {
                c->key_hi = 0xffffffffu;
                c->need_min = 1;
            }
        }
    }
}

// Smallest key above key_lo (the next order statistic) when it differs.
template <int T>
__global__ __launch_bounds__(kT) void k_min_above(const void* __restrict__ in, long long n, SelCtl* __restrict__ c) {
    if (!c->need_min) return;
    const unsigned lo = c->key_lo;
    unsigned m = 0xffffffffu;
    const size_t stride = static_cast<size_t>(gridDim.x) * kT;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x; i < static_cast<size_t>(n); i += stride) {
        const unsigned k = mono_key(load_f32<T>(in, i));
        if (k > lo && k < m) m = k;
    }
    for (int off = 32; off > 0; off >>= 1) m = min(m, static_cast<unsigned>(__shfl_down(static_cast<int>(m), off, 64)));
    if ((threadIdx.x & 63) == 0 && m != 0xffffffffu) atomicMin(&c->key_hi, m);
}

// numpy's _lerp in float32: a + (b - a) * t, or b - (b - a) * (1 - t) for t >= 0.5.
__global__ void k_lerp(SelCtl* c) {
    const float a = key_value(c->key_lo);
    const float b = key_value(c->key_hi);
    const float t = c->gamma;
    const float d = b - a;
    c->thr = t >= 0.5f ? b - d * (1.0f - t) : a + d * t;
}

// ------------------------------------------------------------ mean and std

template <int T, bool SQ>
__global__ __launch_bounds__(kT) void k_moments(const void* __restrict__ in, long long n, const SelCtl* __restrict__ c,
                                                double* __restrict__ part, long long* __restrict__ cpart) {
    const float thr = c->thr;
    const float mu = SQ ? c->mu : 0.0f;
    double acc = 0.0;
    long long cnt = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * kT;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x; i < static_cast<size_t>(n); i += stride) {
        const float v = load_f32<T>(in, i);
        if (v <= thr) {
            if (SQ) {
                const float d = v - mu;
                acc += static_cast<double>(d * d);
            } else {
                acc += static_cast<double>(v);
            }
            ++cnt;
        }
    }
    __shared__ double sd[kT];
    __shared__ long long sc[kT];
    sd[threadIdx.x] = acc;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sd[0];
        cpart[blockIdx.x] = sc[0];
    }
}

template <bool SQ>
__global__ __launch_bounds__(kT) void k_moments_finish(const double* __restrict__ part,
                                                       const long long* __restrict__ cpart, int nparts,
                                                       SelCtl* __restrict__ c) {
    __shared__ double sd[kT];
    __shared__ long long sc[kT];
    double a = 0.0;
    long long k = 0;
    for (int i = threadIdx.x; i < nparts; i += kT) {
        a += part[i];
        k += cpart[i];
    }
    sd[threadIdx.x] = a;
    sc[threadIdx.x] = k;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = sd[0] / static_cast<double>(sc[0]);
        c->count = sc[0];
        if (SQ) {
            c->sigma = static_cast<float>(sqrt(m));
        } else {
            c->mu = static_cast<float>(m);
        }
    }
}

__global__ void k_set_musig(SelCtl* c, float mu, float sigma) {
    c->mu = mu;
    c->sigma = sigma;
}

// ------------------------------------------------- uint16 robust statistics

// One pass: the exact histogram of the uint16 values.  A workgroup counts a
// chunk of at most 65535 values in 16-bit LDS counters packed two per word
// (128 KB), then adds its non-zero counts to the global histogram.
constexpr int kHistChunk = 65528;   // a multiple of 8, below 2^16

constexpr int kHistThreads = 1024;   // one workgroup per CU (its LDS): 16 waves of loads in flight

template <bool VEC>
__global__ __launch_bounds__(kHistThreads) void k_hist16(const uint16_t* __restrict__ in, long long n,
                                                         unsigned* __restrict__ ghist) {
    __shared__ unsigned h[kHist16Bins / 2];
    for (int i = threadIdx.x; i < kHist16Bins / 2; i += kHistThreads) h[i] = 0u;
    __syncthreads();
    const long long start = static_cast<long long>(blockIdx.x) * kHistChunk;
    const long long end = min(n, start + kHistChunk);
    auto add = [&](unsigned v) { atomicAdd(&h[v >> 1], 1u << ((v & 1u) * 16u)); };
    long long tail = start;
    if (VEC) {
        const long long nv = (end - start) / 8;
        const uint4* p = reinterpret_cast<const uint4*>(in + start);
        for (long long k = threadIdx.x; k < nv; k += kHistThreads) {
            const uint4 q = p[k];
            add(q.x & 0xffffu); add(q.x >> 16); add(q.y & 0xffffu); add(q.y >> 16);
            add(q.z & 0xffffu); add(q.z >> 16); add(q.w & 0xffffu); add(q.w >> 16);
        }
        tail = start + nv * 8;
    }
    for (long long i = tail + threadIdx.x; i < end; i += kHistThreads) add(in[i]);
    __syncthreads();
    for (int i = threadIdx.x; i < kHist16Bins / 2; i += kHistThreads) {
        const unsigned c = h[i];
        if (c & 0xffffu) atomicAdd(&ghist[2 * i], c & 0xffffu);
        if (c >> 16) atomicAdd(&ghist[2 * i + 1], c >> 16);
    }
}

__global__ void k_zero(unsigned* p, int n) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// From the histogram: the two order statistics, numpy's float32 lerp, and the
// mean and std of the values at or below the threshold, each value's float32
// (v - mu)^2 weighted by its count.  One workgroup of 1024 threads, 64 bins
// each, fixed-order reductions.
__global__ __launch_bounds__(1024) void k_stats16(const unsigned* __restrict__ h, long long prev, long long next,
                                                  float gamma, SelCtl* __restrict__ c) {
    __shared__ long long sc[1024];
    __shared__ double sd[1024];
    __shared__ float s_ab[2], s_thr, s_mu;
    const int t = threadIdx.x;
    const int b0 = t * 64;
    long long local = 0;
    for (int j = 0; j < 64; ++j) local += h[b0 + j];
    sc[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long v = t >= off ? sc[t - off] : 0;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    const long long before = t ? sc[t - 1] : 0;
    const long long want[2] = {prev, next};
    for (int q = 0; q < 2; ++q) {
        if (before <= want[q] && want[q] < sc[t]) {
            long long cum = before;
            for (int j = 0; j < 64; ++j) {
                cum += h[b0 + j];
                if (want[q] < cum) {
                    s_ab[q] = static_cast<float>(b0 + j);
                    break;
                }
            }
        }
    }
    __syncthreads();
    if (t == 0) {
        const float a = s_ab[0], b = s_ab[1], d = b - a;
        s_thr = gamma >= 0.5f ? b - d * (1.0f - gamma) : a + d * gamma;
        c->thr = s_thr;
    }
    __syncthreads();
    const float thr = s_thr;
    double sum = 0.0;
    long long cnt = 0;
    for (int j = 0; j < 64; ++j) {
        const float v = static_cast<float>(b0 + j);
        if (v <= thr) {
            cnt += h[b0 + j];
            sum += static_cast<double>(h[b0 + j]) * static_cast<double>(v);
        }
    }
    __syncthreads();
    sd[t] = sum;
    sc[t] = cnt;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            sd[t] += sd[t + w];
            sc[t] += sc[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        s_mu = static_cast<float>(sd[0] / static_cast<double>(sc[0]));
        c->mu = s_mu;
        c->count = sc[0];
    }
    __syncthreads();
    const float mu = s_mu;
    const long long total = sc[0];
    double sq = 0.0;
    for (int j = 0; j < 64; ++j) {
        const float v = static_cast<float>(b0 + j);
        if (v <= thr && h[b0 + j]) {
            const float d = v - mu;
            sq += static_cast<double>(h[b0 + j]) * static_cast<double>(d * d);
        }
    }
    __syncthreads();
    sd[t] = sq;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) sd[t] += sd[t + w];
        __syncthreads();
    }
    if (t == 0) c->sigma = static_cast<float>(sqrt(sd[0] / static_cast<double>(total)));
}

// ------------------------------------------------------------------ getMask

__device__ __forceinline__ bool zscore_above(float f, float mu, float sigma, float thres) {
    return fabsf((f - mu) / sigma) > thres;
}

// g(f) = (f - mu) / sigma is monotone in f (each step is rounded monotonely)
// for finite mu and finite sigma > 0, so {g > thres} and {g < -thres} are an
// upper and a lower set of the floats: their bounds, found by bisection over
// order-preserving keys, turn the test into two comparisons per voxel.
__global__ void k_cutoffs(SelCtl* c, float thres) {
    const float mu = c->mu, sigma = c->sigma;
    const bool ok = isfinite(mu) && isfinite(sigma) && sigma > 0.0f && !isnan(thres);
    c->direct = ok ? 0 : 1;
    if (!ok) return;
    const unsigned k0 = mono_key(-INFINITY), k1 = mono_key(INFINITY);
    auto g = [&](unsigned k) { return (key_value(k) - mu) / sigma; };
    // smallest key with g > thres
    float hi_cut = NAN;
    if (g(k1) > thres) {
        unsigned lo = k0, hi = k1;
        if (g(k0) > thres) {
            hi = k0;
        } else {
            while (hi - lo > 1u) {
                const unsigned mid = lo + (hi - lo) / 2u;
                if (g(mid) > thres) hi = mid; else lo = mid;
            }
        }
        hi_cut = key_value(hi);
    }
    // largest key with g < -thres
    float lo_cut = NAN;
    if (g(k0) < -thres) {
        unsigned lo = k0, hi = k1;
        if (g(k1) < -thres) {
            lo = k1;
        } else {
            while (hi - lo > 1u) {
                const unsigned mid = lo + (hi - lo) / 2u;
                if (g(mid) < -thres) lo = mid; else hi = mid;
            }
        }
        lo_cut = key_value(lo);
    }
    c->cut_lo = lo_cut;
    c->cut_hi = hi_cut;
}

// One workgroup holds `strips` strips of kMTY output rows of one z plane;
// in a strip, thread w owns the 32 columns [32w, 32w + 32) of every row, as
// one bit word.  Per input row r: the z-score bits T(r), the horizontal
// 3-erosion He(r) (neighbour words through LDS; outside the volume counts as
// set), E(r-1) = He(r-2) & He(r-1) & He(r), its horizontal 3-dilation, the
// vertical one D1(r-2), the same again for D2(r-3), which is written out.
constexpr int kMTY = 32;

template <int T>
__device__ __forceinline__ unsigned zscore_word(const void* in, size_t off, int nvalid, bool vec, bool direct,
                                                float mu, float sigma, float thres, float lo_cut,
                                                float hi_cut) {
    auto test = [&](float f) {
        return direct ? zscore_above(f, mu, sigma, thres) : (f <= lo_cut || f >= hi_cut);
    };
    unsigned w = 0u;
    if (vec && nvalid == 32) {
        if constexpr (T == E_U16) {
            const uint4* p = reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(in) + off);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint4 v = p[q];
                const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    w |= static_cast<unsigned>(test(static_cast<float>(u[e] & 0xffffu))) << (8 * q + 2 * e);
                    w |= static_cast<unsigned>(test(static_cast<float>(u[e] >> 16))) << (8 * q + 2 * e + 1);
                }
            }
        } else {
            const float4* p = reinterpret_cast<const float4*>(static_cast<const float*>(in) + off);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float4 v = p[q];
                w |= static_cast<unsigned>(test(v.x)) << (4 * q);
                w |= static_cast<unsigned>(test(v.y)) << (4 * q + 1);
                w |= static_cast<unsigned>(test(v.z)) << (4 * q + 2);
                w |= static_cast<unsigned>(test(v.w)) << (4 * q + 3);
            }
        }
    } else {
        for (int b = 0; b < nvalid; ++b) w |= static_cast<unsigned>(test(load_f32<T>(in, off + b))) << b;
    }
    return w;
}

__device__ __forceinline__ void write_bits(uint8_t* out, unsigned bits, int nvalid, bool vec) {
    if (vec && nvalid == 32) {
        unsigned q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const unsigned nb = (bits >> (4 * k)) & 15u;
            q[k] = (nb & 1u) | ((nb & 2u) << 7) | ((nb & 4u) << 14) | ((nb & 8u) << 21);
        }
        uint4* o = reinterpret_cast<uint4*>(out);
        o[0] = make_uint4(q[0], q[1], q[2], q[3]);
        o[1] = make_uint4(q[4], q[5], q[6], q[7]);
    } else {
        for (int b = 0; b < nvalid; ++b) out[b] = static_cast<uint8_t>((bits >> b) & 1u);
    }
}

template <int T>
__global__ __launch_bounds__(kT) void k_getmask_bits(const void* __restrict__ in, int nx, int ny, int W,
                                                     int strips, bool vec_in, bool vec_out,
                                                     const SelCtl* __restrict__ c, float thres,
                                                     uint8_t* __restrict__ out) {
    __shared__ unsigned ex[3][3 * kT];
    const int t = threadIdx.x;
    const int s = t / W, w = t - s * W;
    const bool act = s < strips;
    const size_t plane = static_cast<size_t>(blockIdx.y) * ny * nx;
    const int ys = (blockIdx.x * strips + s) * kMTY;
    const int slot = s * (W + 2) + w + 1;
    const int x0 = 32 * w;
    const int nvalid = act ? min(32, nx - x0) : 32;
    const unsigned valid = nvalid >= 32 ? ~0u : ((1u << nvalid) - 1u);
    const bool direct = c->direct != 0;
    const float mu = c->mu, sigma = c->sigma, lo_cut = c->cut_lo, hi_cut = c->cut_hi;
    unsigned he0 = 0u, he1 = 0u, hd0 = 0u, hd1 = 0u, hq0 = 0u, hq1 = 0u;
    for (int step = 0; step < kMTY + 6; ++step) {
        const int r = ys - 3 + step;
        unsigned tw = ~0u;
        if (act && r >= 0 && r < ny)
            tw = zscore_word<T>(in, plane + static_cast<size_t>(r) * nx + x0, nvalid, vec_in, direct, mu, sigma,
                                thres, lo_cut, hi_cut) | ~valid;
        if (act) ex[0][slot] = tw;
        __syncthreads();
        unsigned lw = ~0u, rw = ~0u;
        if (act) {
            if (w > 0) lw = ex[0][slot - 1];
            if (w < W - 1) rw = ex[0][slot + 1];
        }
        const unsigned he = tw & ((tw << 1) | (lw >> 31)) & ((tw >> 1) | (rw << 31));
        const int re = r - 1;
        const unsigned e = (re >= 0 && re < ny) ? (he0 & he1 & he) & valid : 0u;
        he0 = he1;
        he1 = he;
        if (act) ex[1][slot] = e;
        __syncthreads();
        lw = rw = 0u;
        if (act) {
            if (w > 0) lw = ex[1][slot - 1];
            if (w < W - 1) rw = ex[1][slot + 1];
        }
        const unsigned hd = e | (e << 1) | (lw >> 31) | (e >> 1) | (rw << 31);
        const int rd = r - 2;
        const unsigned d1 = (rd >= 0 && rd < ny) ? (hd0 | hd1 | hd) & valid : 0u;
        hd0 = hd1;
        hd1 = hd;
        if (act) ex[2][slot] = d1;
        __syncthreads();
        lw = rw = 0u;
        if (act) {
            if (w > 0) lw = ex[2][slot - 1];
            if (w < W - 1) rw = ex[2][slot + 1];
        }
        const unsigned hq = d1 | (d1 << 1) | (lw >> 31) | (d1 >> 1) | (rw << 31);
        const unsigned d2 = hq0 | hq1 | hq;
        hq0 = hq1;
        hq1 = hq;
        const int ro = r - 3;
        if (act && step >= 6 && ro < ny)
            write_bits(out + plane + static_cast<size_t>(ro) * nx + x0, d2 & valid, nvalid, vec_out);
    }
}

// ------------------------------------------------------ connected components
//
// The streaming kernels (init, filter) take four voxels (one 32-bit word of
// mask bytes) per thread and skip empty words; the union and flattening
// kernels take one voxel per thread, to hide the latency of their tree walks.
// Labels and areas are only written and read for foreground voxels.

// Plain loads: a stale parent read from a cache is still an ancestor of the
// voxel (parents only ever move to smaller labels of the same set), and the
// atomicMin results drive unite's progress, so the walks stay correct.
__device__ __forceinline__ int find_root(const int* L, int x) {
    int p = L[x];
    while (p != x) {       // labels only point to smaller indices: terminates
        x = p;
        p = L[x];
    }
    return x;
}

// find_root with path halving: every visited voxel is re-linked to its
// grandparent, an ancestor in any interleaving (a link made by a successful
// atomicMin is never raised again: a voxel's parent is only rewritten to an
// ancestor of the parent it was read with), so the trees stay shallow while
// the unions run.
__device__ __forceinline__ int find_halve(int* L, int x) {
    for (;;) {
        const int p = L[x];
        if (p == x) return x;
        const int g = L[p];
        if (g == p) return p;
        L[x] = g;
        x = g;
    }
}

// Link the trees of a and b: the larger root goes under the smaller.  Each
// retry lowers a + b, so the loop ends.
__device__ void unite(int* L, int a, int b) {
    for (;;) {
        a = find_halve(L, a);
        b = find_halve(L, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(L + a, b);
        if (old == a) return;   // a was still a root: linked
        a = old;                // a was linked meanwhile: retry from its parent
    }
}

__device__ __forceinline__ unsigned mask_word(const uint8_t* m, int i0, int n) {
    if (i0 + 4 <= n) return *reinterpret_cast<const unsigned*>(m + i0);
    unsigned w = 0u;
    for (int b = 0; i0 + b < n; ++b) w |= static_cast<unsigned>(m[i0 + b]) << (8 * b);
    return w;
}

__device__ __forceinline__ bool fg_of(unsigned word, int b) { return ((word >> (8 * b)) & 0xffu) != 0u; }

__global__ __launch_bounds__(kT) void k_ccl_init(const uint8_t* __restrict__ m, int n, int* __restrict__ L,
                                                 int* __restrict__ cnt, unsigned long long* __restrict__ ncomp) {
    const int q = blockIdx.x * kT + threadIdx.x;
    if (q == 0) *ncomp = 0ull;
    const int i0 = 4 * q;
    if (i0 >= n) return;
    const unsigned word = mask_word(m, i0, n);
    if (!word) return;
    for (int b = 0; b < 4; ++b)
        if (fg_of(word, b)) {
            L[i0 + b] = i0 + b;
            cnt[i0 + b] = 0;
        }
}

// Each foreground voxel unites with enough of its 13 earlier neighbours in
// (z, y, x) order that every 26-adjacent pair ends up connected.  A skipped
// neighbour q is always 26-adjacent to a neighbour r that is united, and the
// pair (q, r) lies earlier in the scan, so it is connected by induction:
//   plane z - 1: the voxel below if set (its 8 in-plane neighbours hang on
//     it); otherwise the set edge voxels of the 3x3, and a set corner only
//     when both edges beside it are clear;
//   plane z: the left voxel if set, else the one above if set; the upper
//     corners only when neither covers them (upper-left needs left and above
//     clear, upper-right needs above clear).
__global__ __launch_bounds__(kT) void k_ccl_merge(const uint8_t* __restrict__ m, int nx, int ny, int nz,
                                                  int* __restrict__ L) {
    const int n = nx * ny * nz;
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n || !m[i]) return;
    const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const bool xl = x > 0, xr = x < nx - 1, yu = y > 0, yd = y < ny - 1;
    if (z > 0) {
        const int b = i - nx * ny;
        if (m[b]) {
            unite(L, i, b);
        } else {
            const bool eu = yu && m[b - nx], ed = yd && m[b + nx], el = xl && m[b - 1], er = xr && m[b + 1];
            if (eu) unite(L, i, b - nx);
            if (ed) unite(L, i, b + nx);
            if (el) unite(L, i, b - 1);
            if (er) unite(L, i, b + 1);
            if (yu && xl && !eu && !el && m[b - nx - 1]) unite(L, i, b - nx - 1);
            if (yu && xr && !eu && !er && m[b - nx + 1]) unite(L, i, b - nx + 1);
            if (yd && xl && !ed && !el && m[b + nx - 1]) unite(L, i, b + nx - 1);
            if (yd && xr && !ed && !er && m[b + nx + 1]) unite(L, i, b + nx + 1);
        }
    }
    const bool left = xl && m[i - 1];
    const bool up = yu && m[i - nx];
    if (left) unite(L, i, i - 1);
    else if (up) unite(L, i, i - nx);
    if (!up) {
        if (!left && xl && yu && m[i - nx - 1]) unite(L, i, i - nx - 1);
        if (xr && yu && m[i - nx + 1]) unite(L, i, i - nx + 1);
    }
}

// Flatten every foreground voxel onto its root and count the areas: one
// atomic per distinct root per wave.
__global__ __launch_bounds__(kT) void k_ccl_resolve(const uint8_t* __restrict__ m, int n, int* __restrict__ L,
                                                    int* __restrict__ cnt, unsigned long long* __restrict__ ncomp) {
    const int i = blockIdx.x * kT + threadIdx.x;
    const bool fg = i < n && m[i];
    if (!__ballot(fg)) return;   // uniform: no foreground in this wave
    const int lane = threadIdx.x & 63;
    int lab = -1;
    if (fg) {
        lab = find_root(L, i);
        L[i] = lab;
        if (lab == i) atomicAdd(ncomp, 1ull);
    }
    unsigned long long active = __ballot(lab >= 0);
    while (active) {
        const int leader = __ffsll(static_cast<long long>(active)) - 1;
        const int l = __shfl(lab, leader, 64);
        const unsigned long long same = __ballot(lab == l);
        if (lane == leader) atomicAdd(cnt + l, static_cast<int>(__popcll(same)));
        if (lab == l) lab = -1;
        active = __ballot(lab >= 0);
    }
}

__global__ __launch_bounds__(kT) void k_ccl_filter(const uint8_t* __restrict__ m, const int* __restrict__ L,
                                                   const int* __restrict__ cnt, int n, double lo, double hi,
                                                   uint8_t* __restrict__ o8, float* __restrict__ of) {
    const int i0 = 4 * (blockIdx.x * kT + threadIdx.x);
    if (i0 >= n) return;
    const unsigned word = mask_word(m, i0, n);
    bool keep[4] = {false, false, false, false};
    if (word)
        for (int b = 0; b < 4; ++b)
            if (fg_of(word, b)) {
                const double a = static_cast<double>(cnt[L[i0 + b]]);
                keep[b] = lo <= a && a <= hi;
            }
    if (i0 + 4 <= n) {
        if (o8)
            *reinterpret_cast<unsigned*>(o8 + i0) = static_cast<unsigned>(keep[0]) | (static_cast<unsigned>(keep[1]) << 8) |
                                                    (static_cast<unsigned>(keep[2]) << 16) |
                                                    (static_cast<unsigned>(keep[3]) << 24);
        if (of)
            *reinterpret_cast<float4*>(of + i0) = make_float4(keep[0] ? 1.0f : 0.0f, keep[1] ? 1.0f : 0.0f,
                                                              keep[2] ? 1.0f : 0.0f, keep[3] ? 1.0f : 0.0f);
    } else {
        for (int b = 0; i0 + b < n; ++b) {
            if (o8) o8[i0 + b] = keep[b];
            if (of) of[i0 + b] = keep[b] ? 1.0f : 0.0f;
        }
    }
}

// ----------------------------------------------------- mean squared difference

template <int T>
__device__ __forceinline__ double load_d(const void* p, size_t i) {
    if constexpr (T == E_U16)
        return static_cast<double>(static_cast<const uint16_t*>(p)[i]);
    else if constexpr (T == E_F32)
        return static_cast<double>(static_cast<const float*>(p)[i]);
    else
        return static_cast<const double*>(p)[i];
}

// numpy's (a - b) ** 2 in the promoted dtype: uint16 wraps, float32 is
// rounded per step, float64 is float64; summed in float64 (exact for uint16).
template <int TA, int TB>
__device__ __forceinline__ double sq_diff(const void* a, const void* b, size_t i) {
    if constexpr (TA == E_U16 && TB == E_U16) {
        const uint16_t d = static_cast<uint16_t>(static_cast<const uint16_t*>(a)[i] - static_cast<const uint16_t*>(b)[i]);
        return static_cast<double>(static_cast<uint16_t>(static_cast<unsigned>(d) * d));
    } else if constexpr (TA == E_F64 || TB == E_F64) {
        const double d = load_d<TA>(a, i) - load_d<TB>(b, i);
        return d * d;
    } else {
        const float d = load_f32<TA>(a, i) - load_f32<TB>(b, i);
        return static_cast<double>(d * d);
    }
}

template <int TA, int TB>
__global__ __launch_bounds__(kT) void k_msd(const void* __restrict__ a, const void* __restrict__ b, long long n,
                                            double* __restrict__ part) {
    double acc = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * kT;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kT + threadIdx.x; i < static_cast<size_t>(n); i += stride)
        acc += sq_diff<TA, TB>(a, b, i);
    __shared__ double sd[kT];
    sd[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sd[threadIdx.x] += sd[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sd[0];
}

__global__ __launch_bounds__(kT) void k_msd_finish(const double* __restrict__ part, int nparts, long long n,
                                                   double* __restrict__ out) {
    __shared__ double sd[kT];
    double a = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kT) a += part[i];
    sd[threadIdx.x] = a;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sd[threadIdx.x] += sd[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sd[0] / static_cast<double>(n);
}

unsigned stream_blocks(long long n) {
    const long long b = (n + kT - 1) / kT;
    return static_cast<unsigned>(b < kStreamBlocks ? (b > 0 ? b : 1) : kStreamBlocks);
}

template <int T>
void robust_t(const void* in, long long n, long long prev, long long next, float gamma, SelCtl* c,
              unsigned* hist, double* part, long long* cpart, hipStream_t s) {
    const unsigned g = stream_blocks(n);
    k_sel_reset<<<1, kT, 0, s>>>(c, hist, 3 * 2048, prev, next, gamma);
    const int shifts[3] = {21, 10, 0}, bits[3] = {11, 11, 10};
    for (int p = 0; p < 3; ++p) {
        k_hist<T><<<g, kT, 0, s>>>(in, n, c, shifts[p], bits[p], hist + 2048 * p);
        k_select<<<1, kT, 0, s>>>(hist + 2048 * p, shifts[p], bits[p], p == 2, c);
    }
    k_min_above<T><<<g, kT, 0, s>>>(in, n, c);
    k_lerp<<<1, 1, 0, s>>>(c);
    k_moments<T, false><<<g, kT, 0, s>>>(in, n, c, part, cpart);
    k_moments_finish<false><<<1, kT, 0, s>>>(part, cpart, static_cast<int>(g), c);
    k_moments<T, true><<<g, kT, 0, s>>>(in, n, c, part, cpart);
    k_moments_finish<true><<<1, kT, 0, s>>>(part, cpart, static_cast<int>(g), c);
}

template <int TA>
void msd_a(const void* a, Elem tb, const void* b, long long n, double* part, double* out, hipStream_t s) {
    const unsigned g = stream_blocks(n);
    switch (tb) {
        case E_U16: k_msd<TA, E_U16><<<g, kT, 0, s>>>(a, b, n, part); break;
        case E_F32: k_msd<TA, E_F32><<<g, kT, 0, s>>>(a, b, n, part); break;
        default: k_msd<TA, E_F64><<<g, kT, 0, s>>>(a, b, n, part); break;
    }
    k_msd_finish<<<1, kT, 0, s>>>(part, static_cast<int>(g), n, out);
}

}  // namespace

void percentile_index(long long n, double percentile, long long* prev, long long* next, float* gamma) {
    // numpy.percentile on float32 data: q = float32(p) / float32(100),
    // virtual index = float32(n - 1) * q, linear interpolation between the
    // floor and the next index (the last element when the index reaches it).
    const float q = static_cast<float>(percentile) / 100.0f;
    const float vi = static_cast<float>(n - 1) * q;
    const float fl = std::floor(vi);
    if (vi >= static_cast<float>(n - 1)) {
        *prev = *next = n - 1;
        *gamma = 0.0f;
        return;
    }
    *prev = static_cast<long long>(fl);
    *next = *prev + 1;
    *gamma = vi - fl;
}

void launch_robust_stats(const void* in, Elem t, long long n, double percentile, SelCtl* c, unsigned* hist,
                         double* part, long long* cpart, hipStream_t s) {
    long long prev, next;
    float gamma;
    percentile_index(n, percentile, &prev, &next, &gamma);
    if (t == E_U16) {
        // one pass: the exact histogram of the 65536 values
        k_zero<<<kHist16Bins / kT, kT, 0, s>>>(hist, kHist16Bins);
        const unsigned g = static_cast<unsigned>((n + kHistChunk - 1) / kHistChunk);
        const uint16_t* p = static_cast<const uint16_t*>(in);
        if (reinterpret_cast<uintptr_t>(p) % 16 == 0)
            k_hist16<true><<<g, kHistThreads, 0, s>>>(p, n, hist);
        else
            k_hist16<false><<<g, kHistThreads, 0, s>>>(p, n, hist);
        k_stats16<<<1, 1024, 0, s>>>(hist, prev, next, gamma, c);
    } else {
        robust_t<E_F32>(in, n, prev, next, gamma, c, hist, part, cpart, s);
    }
}

void launch_set_musig(SelCtl* c, float mu, float sigma, hipStream_t s) {
    k_set_musig<<<1, 1, 0, s>>>(c, mu, sigma);
}

void launch_getmask(const void* in, Elem t, int nx, int ny, int nz, SelCtl* c, float thres, uint8_t* out,
                    hipStream_t s) {
    k_cutoffs<<<1, 1, 0, s>>>(c, thres);
    const int W = (nx + 31) / 32;
    const int strips = kT / W;
    const dim3 grid((ny + strips * kMTY - 1) / (strips * kMTY), nz);
    const bool in16 = reinterpret_cast<uintptr_t>(in) % 16 == 0;
    const bool vec_out = nx % 16 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0;
    if (t == E_U16)
        k_getmask_bits<E_U16><<<grid, kT, 0, s>>>(in, nx, ny, W, strips, in16 && nx % 8 == 0, vec_out, c, thres,
                                                  out);
    else
        k_getmask_bits<E_F32><<<grid, kT, 0, s>>>(in, nx, ny, W, strips, in16 && nx % 4 == 0, vec_out, c, thres,
                                                  out);
}

void launch_components(const uint8_t* m, int nx, int ny, int nz, double lo, double hi, int* L, int* cnt,
                       unsigned long long* ncomp, uint8_t* o8, float* of, hipStream_t s) {
    const int n = nx * ny * nz;
    const long long words = (static_cast<long long>(n) + 3) / 4;
    const unsigned g = static_cast<unsigned>((words + kT - 1) / kT);
    const unsigned g1 = static_cast<unsigned>((static_cast<long long>(n) + kT - 1) / kT);
    k_ccl_init<<<g, kT, 0, s>>>(m, n, L, cnt, ncomp);
    k_ccl_merge<<<g1, kT, 0, s>>>(m, nx, ny, nz, L);
    k_ccl_resolve<<<g1, kT, 0, s>>>(m, n, L, cnt, ncomp);
    k_ccl_filter<<<g, kT, 0, s>>>(m, L, cnt, n, lo, hi, o8, of);
}

void launch_mean_sq_diff(const void* a, Elem ta, const void* b, Elem tb, long long n, double* part, double* out,
                         hipStream_t s) {
    switch (ta) {
        case E_U16: msd_a<E_U16>(a, tb, b, n, part, out, s); break;
        case E_F32: msd_a<E_F32>(a, tb, b, n, part, out, s); break;
        default: msd_a<E_F64>(a, tb, b, n, part, out, s); break;
    }
}

}  // namespace wr
