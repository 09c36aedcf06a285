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
// getMask.  One workgroup per 64 x 16 tile of one z plane stages the z-score
// test |(v - mu) / sigma| > thres (float32, correctly rounded division) on the
// tile plus a 3-voxel halo in LDS, then applies the 3x3 in-plane erosion
// (outside the volume counts as set), dilation and dilation (outside counts
// as clear): scikit-image's binary_opening followed by binary_dilation with a
// 3x3x1 footprint.
//
// bwareafilt3_wei.  Union-find connected components over the 26-neighbourhood
// (each voxel unites with its 13 already-scanned neighbours; a root is linked
// under the smaller label with atomicMin, the method of Playne and Hawick,
// "A new algorithm for parallel connected-component labelling on GPUs", IEEE
// TPDS 2018), path flattening, wave-aggregated area counts per root, and the
// closed area range test.  The output does not depend on which voxel becomes
// a root.

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

// ------------------------------------------------------------------ getMask

constexpr int kMX = 64, kMY = 16;

template <int T>
__global__ __launch_bounds__(kT) void k_getmask(const void* __restrict__ in, int nx, int ny,
                                                const SelCtl* __restrict__ c, float thres,
                                                uint8_t* __restrict__ out) {
    __shared__ uint8_t st[kMY + 6][kMX + 6];
    __shared__ uint8_t se[kMY + 4][kMX + 4];
    __shared__ uint8_t sd[kMY + 2][kMX + 2];
    const int x0 = blockIdx.x * kMX, y0 = blockIdx.y * kMY;
    const size_t plane = static_cast<size_t>(blockIdx.z) * ny * nx;
    const float mu = c->mu, sigma = c->sigma;
    // z-score test on the tile plus a 3-voxel halo
    for (int q = threadIdx.x; q < (kMY + 6) * (kMX + 6); q += kT) {
        const int j = q / (kMX + 6), i = q % (kMX + 6);
        const int x = x0 - 3 + i, y = y0 - 3 + j;
        uint8_t v = 0;
        if (x >= 0 && x < nx && y >= 0 && y < ny) {
            const float f = load_f32<T>(in, plane + static_cast<size_t>(y) * nx + x);
            v = fabsf((f - mu) / sigma) > thres;
        }
        st[j][i] = v;
    }
    __syncthreads();
    // erosion: neighbours outside the volume count as set
    for (int q = threadIdx.x; q < (kMY + 4) * (kMX + 4); q += kT) {
        const int j = q / (kMX + 4), i = q % (kMX + 4);
        const int x = x0 - 2 + i, y = y0 - 2 + j;
        uint8_t v = 0;
        if (x >= 0 && x < nx && y >= 0 && y < ny) {
            v = 1;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = x + dx, yy = y + dy;
                    if (xx >= 0 && xx < nx && yy >= 0 && yy < ny) v &= st[j + 1 + dy][i + 1 + dx];
                }
        }
        se[j][i] = v;
    }
    __syncthreads();
    // first dilation: neighbours outside the volume count as clear
    for (int q = threadIdx.x; q < (kMY + 2) * (kMX + 2); q += kT) {
        const int j = q / (kMX + 2), i = q % (kMX + 2);
        const int x = x0 - 1 + i, y = y0 - 1 + j;
        uint8_t v = 0;
        if (x >= 0 && x < nx && y >= 0 && y < ny)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = x + dx, yy = y + dy;
                    if (xx >= 0 && xx < nx && yy >= 0 && yy < ny) v |= se[j + 1 + dy][i + 1 + dx];
                }
        sd[j][i] = v;
    }
    __syncthreads();
    // second dilation, written out
    for (int q = threadIdx.x; q < kMY * kMX; q += kT) {
        const int j = q / kMX, i = q % kMX;
        const int x = x0 + i, y = y0 + j;
        if (x >= nx || y >= ny) continue;
        uint8_t v = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx, yy = y + dy;
                if (xx >= 0 && xx < nx && yy >= 0 && yy < ny) v |= sd[j + 1 + dy][i + 1 + dx];
            }
        out[plane + static_cast<size_t>(y) * nx + x] = v;
    }
}

// ------------------------------------------------------ connected components

__device__ __forceinline__ int ld_label(const int* L, int i) {
    return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_root(const int* L, int x) {
    int p = ld_label(L, x);
    while (p != x) {       // labels only point to smaller indices: terminates
        x = p;
        p = ld_label(L, x);
    }
    return x;
}

// Link the trees of a and b: the larger root goes under the smaller.
__device__ void unite(int* L, int a, int b) {
    for (;;) {
        a = find_root(L, a);
        b = find_root(L, b);
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

__global__ __launch_bounds__(kT) void k_ccl_init(const uint8_t* __restrict__ m, int n, int* __restrict__ L,
                                                 int* __restrict__ cnt, unsigned long long* __restrict__ ncomp) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i == 0) *ncomp = 0ull;
    if (i >= n) return;
    L[i] = m[i] ? i : -1;
    cnt[i] = 0;
}

// Each foreground voxel unites with its 13 foreground neighbours that come
// earlier in (z, y, x) order: together they cover all 26 neighbour pairs.
__global__ __launch_bounds__(kT) void k_ccl_merge(const uint8_t* __restrict__ m, int nx, int ny, int nz,
                                                  int* __restrict__ L) {
    const int x = blockIdx.x * kT + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    const int i = (z * ny + y) * nx + x;
    if (!m[i]) return;
    for (int dz = -1; dz <= 0; ++dz) {
        const int zz = z + dz;
        if (zz < 0) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            if (dz == 0 && dy > 0) break;
            const int yy = y + dy;
            if (yy < 0 || yy >= ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                if (dz == 0 && dy == 0 && dx >= 0) break;
                const int xx = x + dx;
                if (xx < 0 || xx >= nx) continue;
                const int j = (zz * ny + yy) * nx + xx;
                if (m[j]) unite(L, i, j);
            }
        }
    }
}

__global__ __launch_bounds__(kT) void k_ccl_flatten(int n, int* __restrict__ L) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    if (L[i] >= 0) L[i] = find_root(L, i);
}

// Area per root: one atomic per distinct root per wave.
__global__ __launch_bounds__(kT) void k_ccl_count(const int* __restrict__ L, int n, int* __restrict__ cnt,
                                                  unsigned long long* __restrict__ ncomp) {
    const int i = blockIdx.x * kT + threadIdx.x;
    int lab = i < n ? L[i] : -1;
    if (lab >= 0 && lab == i) atomicAdd(ncomp, 1ull);
    const int lane = threadIdx.x & 63;
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

__global__ __launch_bounds__(kT) void k_ccl_filter(const int* __restrict__ L, const int* __restrict__ cnt, int n,
                                                   double lo, double hi, uint8_t* __restrict__ o8,
                                                   float* __restrict__ of) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int lab = L[i];
    bool keep = false;
    if (lab >= 0) {
        const double a = static_cast<double>(cnt[lab]);
        keep = lo <= a && a <= hi;
    }
    if (o8) o8[i] = keep;
    if (of) of[i] = keep ? 1.0f : 0.0f;
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
    if (t == E_U16)
        robust_t<E_U16>(in, n, prev, next, gamma, c, hist, part, cpart, s);
    else
        robust_t<E_F32>(in, n, prev, next, gamma, c, hist, part, cpart, s);
}

void launch_set_musig(SelCtl* c, float mu, float sigma, hipStream_t s) {
    k_set_musig<<<1, 1, 0, s>>>(c, mu, sigma);
}

void launch_getmask(const void* in, Elem t, int nx, int ny, int nz, const SelCtl* c, float thres, uint8_t* out,
                    hipStream_t s) {
    const dim3 grid((nx + kMX - 1) / kMX, (ny + kMY - 1) / kMY, nz);
    if (t == E_U16)
        k_getmask<E_U16><<<grid, kT, 0, s>>>(in, nx, ny, c, thres, out);
    else
        k_getmask<E_F32><<<grid, kT, 0, s>>>(in, nx, ny, c, thres, out);
}

void launch_components(const uint8_t* m, int nx, int ny, int nz, double lo, double hi, int* L, int* cnt,
                       unsigned long long* ncomp, uint8_t* o8, float* of, hipStream_t s) {
    const int n = nx * ny * nz;
    const unsigned g = static_cast<unsigned>((n + kT - 1) / kT);
    k_ccl_init<<<g, kT, 0, s>>>(m, n, L, cnt, ncomp);
    k_ccl_merge<<<dim3((nx + kT - 1) / kT, ny, nz), kT, 0, s>>>(m, nx, ny, nz, L);
    k_ccl_flatten<<<g, kT, 0, s>>>(n, L);
    k_ccl_count<<<g, kT, 0, s>>>(L, n, cnt, ncomp);
    k_ccl_filter<<<g, kT, 0, s>>>(L, cnt, n, lo, hi, o8, of);
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
