// Kernels of the learned initialisation network on gfx950: a 3-D U-Net with
// 3x3x3 convolutions, bfloat16 weights, activations and gradients, float32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Layout: activations channels-last (B, D, H, W, C), C a multiple of 8, so
// the 8 consecutive reduction elements a lane feeds to one MFMA are one
// 16-byte load; weights (Cout, 27, Cin) with tap = (kd * 3 + kh) * 3 + kw and
// offset (kd - 1, kh - 1, kw - 1) in (d, h, w); padding 1 (zeros).
//
// MFMA 16x16x32 bf16 operand layout (lane l of a wave64):
//   A[m = l % 16][k = 8 (l / 16) + i], B[k = 8 (l / 16) + i][n = l % 16], i < 8
//   C[m = 4 (l / 16) + i][n = l % 16], i < 4
//
// Forward: implicit GEMM with M = Cout, N = voxels, K = 27 Cin.  A wave
// computes 16 output channels x 64 voxels (four tiles sharing the weight
// operand); per k-step a lane loads 8 weights and, per tile, the 8 channels
// of one neighbour voxel.  The epilogue adds the bias, applies LeakyReLU and
// stores bfloat16.
// Data gradient: the forward kernel on flipped weights,
//   W'[ci][tap][co] = W[co][26 - tap][ci]   (offset(26 - tap) = -offset(tap)).
// Weight gradient: dW[co][tap][ci] = sum_v dY[v][co] X[v + off(tap)][ci] as
// MFMA with M = Cout, N = Cin, K = voxels, all 27 taps of a (co, ci) tile in
// one wave (the 10-voxel x window of one (kd, kh) row serves its 3 kw taps);
// float32 atomics into dW, the bias gradient from the same dY loads.

#include <cmath>
#include <cstdint>

#include "wr_internal.h"

namespace wr {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

union Pack8 {
    bf16x8 v;
    uint4 q;
    uint16_t h[8];
};

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<unsigned>(h) << 16); }

// round to nearest even; NaN stays NaN
__device__ __forceinline__ uint16_t f2bf(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    Pack8 r;
    r.q = *reinterpret_cast<const uint4*>(p);
    return r.v;
}

__device__ __forceinline__ bf16x8 zero8() {
    Pack8 r;
    r.q = make_uint4(0u, 0u, 0u, 0u);
    return r.v;
}

// Bit t of the result is set when neighbour tap t of voxel (d, h, w) lies
// inside the volume.
__device__ __forceinline__ unsigned tap_mask(int d, int h, int w, const NetGeom& g) {
    const unsigned mw = (w > 0 ? 1u : 0u) | 2u | (w + 1 < g.w ? 4u : 0u);
    const unsigned mhw = (h > 0 ? mw : 0u) | (mw << 3) | (h + 1 < g.h ? mw << 6 : 0u);
    return (d > 0 ? mhw : 0u) | (mhw << 9) | (d + 1 < g.d ? mhw << 18 : 0u);
}

// NT MFMA tiles of 16 voxels per wave (4 on large volumes, 1 on small ones
// for 4x the waves) times CT tiles of 16 output channels: each gathered
// input fragment feeds CT matrix operations (the gathers bound the kernel).
// Every output tile receives the same operations in the same order whatever
// NT and CT, so the results do not depend on them.
template <int NT, int CT>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                                  const float* __restrict__ bias, uint16_t* __restrict__ Y,
                                                  NetGeom g, int cin, int cout, int act, float slope) {
    __shared__ int toff[27];          // element offset of tap t: voxel offset * cin
    if (threadIdx.x < 27) {
        const int t = threadIdx.x;
        toff[t] = (((t / 9 - 1) * g.h + (t / 3) % 3 - 1) * g.w + t % 3 - 1) * cin;
    }
    __syncthreads();
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long nv = static_cast<long long>(g.d) * g.h * g.w;
    const int co0 = blockIdx.y * 16 * CT;
    // workgroups go round-robin to the 8 XCDs; give each XCD a contiguous run
    // of voxel chunks so the neighbour planes its taps share stay in its L2
    // (gridDim.x is a multiple of 8 on the network's volumes, so the XCD of a
    // workgroup is blockIdx.x % 8 whatever blockIdx.y, z)
    const unsigned per = gridDim.x / 8;
    const unsigned bx = blockIdx.x < per * 8 ? (blockIdx.x % 8) * per + blockIdx.x / 8 : blockIdx.x;
    const long long vb = static_cast<long long>(bx) * (64 * NT) + wave * (16 * NT);
    const uint16_t* xb = X + static_cast<size_t>(blockIdx.z) * nv * cin;
    const size_t yb = static_cast<size_t>(blockIdx.z) * nv * cout;
    const int ksz = 27 * cin;
    int vo[NT];           // element offset of the voxel's channel 0
    unsigned vm[NT];      // its in-volume taps (0: voxel past the end)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const long long v = vb + 16 * j + (l & 15);
        const bool ok = v < nv;
        const long long vv = ok ? v : 0;
        const int w = static_cast<int>(vv % g.w);
        const int h = static_cast<int>((vv / g.w) % g.h);
        const int d = static_cast<int>(vv / (static_cast<long long>(g.w) * g.h));
        vo[j] = static_cast<int>(vv) * cin;
        vm[j] = ok ? tap_mask(d, h, w, g) : 0u;
    }
    f4 acc[CT][NT];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[c][j] = f4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* wrow = Wt + static_cast<size_t>(co0 + (l & 15)) * ksz + 8 * (l >> 4);
    const size_t wct = static_cast<size_t>(16) * ksz;   // next tile of 16 output channels
    // this lane's reduction index k0 + 8 (l / 16) = tap * cin + ci, advanced by 32 per step;
    // two steps per trip so that the loads of both are in flight together
    int tap = (8 * (l >> 4)) / cin;
    int ci = 8 * (l >> 4) - tap * cin;
    for (int k0 = 0; k0 < ksz; k0 += 64) {
        int tap1 = tap, ci1 = ci + 32;
        while (ci1 >= cin) {
            ci1 -= cin;
            ++tap1;
        }
        const bool kin0 = tap < 27, kin1 = tap1 < 27;
        bf16x8 a0[CT], a1[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            a0[c] = kin0 ? load8(wrow + c * wct + k0) : zero8();
            a1[c] = kin1 ? load8(wrow + c * wct + k0 + 32) : zero8();
        }
        const int off0 = (kin0 ? toff[tap] : 0) + ci;
        const int off1 = (kin1 ? toff[tap1] : 0) + ci1;
        bf16x8 b0[NT], b1[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const bool in0 = kin0 && ((vm[j] >> (tap & 31)) & 1u);
            const bool in1 = kin1 && ((vm[j] >> (tap1 & 31)) & 1u);
            b0[j] = in0 ? load8(xb + vo[j] + off0) : zero8();
            b1[j] = in1 ? load8(xb + vo[j] + off1) : zero8();
        }
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int j = 0; j < NT; ++j)
                acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[c], b0[j], acc[c][j], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int j = 0; j < NT; ++j)
                acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[c], b1[j], acc[c][j], 0, 0, 0);
        tap = tap1;
        ci = ci1 + 32;
        while (ci >= cin) {
            ci -= cin;
            ++tap;
        }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int m0 = co0 + 16 * c + 4 * (l >> 4);
        float bs[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) bs[i] = bias ? bias[m0 + i] : 0.0f;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const long long v = vb + 16 * j + (l & 15);
            if (v >= nv) continue;
            uint16_t o[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float y = acc[c][j][i] + bs[i];
                if (act && y < 0.0f) y *= slope;
                o[i] = f2bf(y);
            }
            uint2 pk;
            pk.x = static_cast<unsigned>(o[0]) | (static_cast<unsigned>(o[1]) << 16);
            pk.y = static_cast<unsigned>(o[2]) | (static_cast<unsigned>(o[3]) << 16);
            *reinterpret_cast<uint2*>(Y + yb + static_cast<size_t>(v) * cout + m0) = pk;
        }
    }
}

__global__ void k_flip(const uint16_t* __restrict__ W, int cin, int cout, uint16_t* __restrict__ Wf) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const long long n = 27LL * cin * cout;
    if (i >= n) return;
    const int co = static_cast<int>(i % cout);
    const int tap = static_cast<int>((i / cout) % 27);
    const int ci = static_cast<int>(i / (27LL * cout));
    Wf[i] = W[(static_cast<size_t>(co) * 27 + (26 - tap)) * cin + ci];
}

// One wave per (voxel chunk, co tile, ci tile); grid (chunks, cout / 16,
// ceil(cin / 16)); the chunk index runs over batch x voxels.  Requires
// g.w % 8 == 0 (a lane's 8 voxels lie in one x row) and chunk % 32 == 0.
__global__ __launch_bounds__(64) void k_conv_wgrad(const uint16_t* __restrict__ dY, const uint16_t* __restrict__ X,
                                                   NetGeom g, int batch, int cin, int cout, int chunk,
                                                   float* __restrict__ dW, float* __restrict__ db) {
    const int l = threadIdx.x;
    const long long nv = static_cast<long long>(g.d) * g.h * g.w;
    const long long total = nv * batch;
    const int co0 = blockIdx.y * 16, ci0 = blockIdx.z * 16;
    const int co = co0 + (l & 15), ci = ci0 + (l & 15);
    const bool cok = ci < cin;
    const long long v0 = static_cast<long long>(blockIdx.x) * chunk;
    const long long v1 = v0 + chunk < total ? v0 + chunk : total;
    f4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f;
    for (long long vs = v0; vs < v1; vs += 32) {
        const long long v = vs + 8 * (l >> 4);        // first of this lane's 8 voxels
        const bool vin = v < v1;
        const long long b = vin ? v / nv : 0;
        const long long r = vin ? v - b * nv : 0;
        const int w0 = static_cast<int>(r % g.w);
        const int h = static_cast<int>((r / g.w) % g.h);
        const int d = static_cast<int>(r / (static_cast<long long>(g.w) * g.h));
        Pack8 a;
        a.q = make_uint4(0u, 0u, 0u, 0u);
        if (vin) {
            const uint16_t* p = dY + static_cast<size_t>(v) * cout + co;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                a.h[i] = p[static_cast<size_t>(i) * cout];
                bsum += bf2f(a.h[i]);
            }
        }
        const uint16_t* xb = X + static_cast<size_t>(b) * nv * cin;
#pragma unroll
        for (int kd = 0; kd < 3; ++kd) {
#pragma unroll
            for (int kh = 0; kh < 3; ++kh) {
                const int zd = d + kd - 1, zh = h + kh - 1;
                const bool rin = vin && cok && zd >= 0 && zd < g.d && zh >= 0 && zh < g.h;
                uint16_t xs[10];
                const uint16_t* row = rin ? xb + (static_cast<size_t>(zd) * g.h + zh) * g.w * cin + ci : xb;
#pragma unroll
                for (int j = 0; j < 10; ++j) {
                    const int zw = w0 - 1 + j;
                    xs[j] = (rin && zw >= 0 && zw < g.w) ? row[static_cast<size_t>(zw) * cin] : uint16_t(0);
                }
#pragma unroll
                for (int kw = 0; kw < 3; ++kw) {
                    Pack8 bb;
#pragma unroll
                    for (int i = 0; i < 8; ++i) bb.h[i] = xs[kw + i];
                    const int t = (kd * 3 + kh) * 3 + kw;
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bb.v, acc[t], 0, 0, 0);
                }
            }
        }
    }
    // C[m = co0 + 4 (l / 16) + i][n = ci0 + l % 16]
    const int n = ci0 + (l & 15);
    if (n < cin) {
#pragma unroll
        for (int t = 0; t < 27; ++t) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = co0 + 4 * (l >> 4) + i;
                atomicAdd(dW + (static_cast<size_t>(m) * 27 + t) * cin + n, acc[t][i]);
            }
        }
    }
    if (db && blockIdx.z == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (l < 16) atomicAdd(db + co, bsum);
    }
}

// g *= (y > 0 ? 1 : slope), bfloat16, 8 per thread
__global__ void k_lrelu_bwd(uint16_t* __restrict__ gr, const uint16_t* __restrict__ y, long long n8, float slope) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n8) return;
    Pack8 gv, yv;
    gv.q = reinterpret_cast<const uint4*>(gr)[i];
    yv.q = reinterpret_cast<const uint4*>(y)[i];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (!(bf2f(yv.h[k]) > 0.0f)) gv.h[k] = f2bf(bf2f(gv.h[k]) * slope);
    reinterpret_cast<uint4*>(gr)[i] = gv.q;
}

// Y (B, D/2, H/2, W/2, C) = scale * sum of the 8 children in X (B, D, H, W, C);
// a thread per (output voxel, 8 channels).
// I: the index type (32-bit unsigned where the element count allows; its
// divisions are a few instructions, 64-bit ones dozens).
template <typename I>
__global__ void k_down2(const uint16_t* __restrict__ X, NetGeom g, int c, long long n_out8, float scale,
                        uint16_t* __restrict__ Y) {
    const I i = static_cast<I>(blockIdx.x) * 256 + threadIdx.x;
    if (static_cast<long long>(i) >= n_out8) return;
    const I c8 = static_cast<I>(c / 8);
    const int q = static_cast<int>(i % c8);
    I v = i / c8;
    const I d2 = static_cast<I>(g.d / 2), h2 = static_cast<I>(g.h / 2), w2 = static_cast<I>(g.w / 2);
    const int x = static_cast<int>(v % w2);
    v /= w2;
    const int y = static_cast<int>(v % h2);
    v /= h2;
    const int z = static_cast<int>(v % d2);
    const long long b = static_cast<long long>(v / d2);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const size_t off = (((static_cast<size_t>(b) * g.d + 2 * z + (k >> 2)) * g.h + 2 * y + ((k >> 1) & 1)) * g.w +
                            2 * x + (k & 1)) * c + 8 * q;
        Pack8 p;
        p.q = *reinterpret_cast<const uint4*>(X + off);
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] += bf2f(p.h[e]);
    }
    Pack8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o.h[e] = f2bf(s[e] * scale);
    reinterpret_cast<uint4*>(Y)[i] = o.q;
}

// Y (B, D, H, W, C) (+)= scale * X at the parent voxel (B, D/2, H/2, W/2, C);
// a thread per (output voxel, 8 channels).
template <typename I>
__global__ void k_up2(const uint16_t* __restrict__ X, NetGeom g, int c, long long n_out8, float scale, int accumulate,
                      uint16_t* __restrict__ Y) {
    const I i = static_cast<I>(blockIdx.x) * 256 + threadIdx.x;
    if (static_cast<long long>(i) >= n_out8) return;
    const I c8 = static_cast<I>(c / 8);
    const int q = static_cast<int>(i % c8);
    I v = i / c8;
    const int x = static_cast<int>(v % static_cast<I>(g.w));
    v /= static_cast<I>(g.w);
    const int y = static_cast<int>(v % static_cast<I>(g.h));
    v /= static_cast<I>(g.h);
    const int z = static_cast<int>(v % static_cast<I>(g.d));
    const long long b = static_cast<long long>(v / static_cast<I>(g.d));
    const size_t src = (((static_cast<size_t>(b) * (g.d / 2) + z / 2) * (g.h / 2) + y / 2) * (g.w / 2) + x / 2) * c +
                       8 * q;
    Pack8 p, o;
    p.q = *reinterpret_cast<const uint4*>(X + src);
    if (accumulate) o.q = reinterpret_cast<const uint4*>(Y)[i];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float val = bf2f(p.h[e]) * scale;
        o.h[e] = f2bf(accumulate ? bf2f(o.h[e]) + val : val);
    }
    reinterpret_cast<uint4*>(Y)[i] = o.q;
}

// Y[v] = (A[v] (ca channels), B[v] (cb channels)); thread per (voxel, 8 channels of Y)
template <typename I>
__global__ void k_concat(const uint16_t* __restrict__ A, int ca, const uint16_t* __restrict__ B, int cb,
                         long long n8, uint16_t* __restrict__ Y) {
    const I i = static_cast<I>(blockIdx.x) * 256 + threadIdx.x;
    if (static_cast<long long>(i) >= n8) return;
    const I c8 = static_cast<I>((ca + cb) / 8);
    const int q = static_cast<int>(i % c8);
    const long long v = static_cast<long long>(i / c8);
    const uint4 val = 8 * q < ca ? *reinterpret_cast<const uint4*>(A + static_cast<size_t>(v) * ca + 8 * q)
                                 : *reinterpret_cast<const uint4*>(B + static_cast<size_t>(v) * cb + 8 * q - ca);
    reinterpret_cast<uint4*>(Y)[i] = val;
}

// Y[v] = (X[parent(v)] (ca channels), B[v] (cb channels)): the concatenation
// of the nearest 2x upsampling of X (B, D/2, H/2, W/2, ca) with B (B, D, H, W,
// cb) without writing the upsampled tensor; thread per (voxel, 8 channels of
// Y).  Equal bit for bit to k_up2 (scale 1) followed by k_concat.
template <typename I>
__global__ void k_upcat(const uint16_t* __restrict__ X, int ca, const uint16_t* __restrict__ B, int cb, NetGeom g,
                        long long n8, uint16_t* __restrict__ Y) {
    const I i = static_cast<I>(blockIdx.x) * 256 + threadIdx.x;
    if (static_cast<long long>(i) >= n8) return;
    const I c8 = static_cast<I>((ca + cb) / 8);
    const int q = static_cast<int>(i % c8);
    const I v = i / c8;
    uint4 val;
    if (8 * q < ca) {
        I t = v;
        const int x = static_cast<int>(t % static_cast<I>(g.w));
        t /= static_cast<I>(g.w);
        const int y = static_cast<int>(t % static_cast<I>(g.h));
        t /= static_cast<I>(g.h);
        const int z = static_cast<int>(t % static_cast<I>(g.d));
        const long long b = static_cast<long long>(t / static_cast<I>(g.d));
        const size_t src =
            (((static_cast<size_t>(b) * (g.d / 2) + z / 2) * (g.h / 2) + y / 2) * (g.w / 2) + x / 2) * ca + 8 * q;
        val = *reinterpret_cast<const uint4*>(X + src);
    } else {
        val = *reinterpret_cast<const uint4*>(B + static_cast<size_t>(v) * cb + 8 * q - ca);
    }
    reinterpret_cast<uint4*>(Y)[i] = val;
}

// dA[v] = G[v][0:ca], dB[v] = G[v][ca:ca+cb]
template <typename I>
__global__ void k_split(const uint16_t* __restrict__ G, int ca, int cb, long long n8, uint16_t* __restrict__ dA,
                        uint16_t* __restrict__ dB) {
    const I i = static_cast<I>(blockIdx.x) * 256 + threadIdx.x;
    if (static_cast<long long>(i) >= n8) return;
    const I c8 = static_cast<I>((ca + cb) / 8);
    const int q = static_cast<int>(i % c8);
    const long long v = static_cast<long long>(i / c8);
    const uint4 val = reinterpret_cast<const uint4*>(G)[i];
    if (8 * q < ca)
        *reinterpret_cast<uint4*>(dA + static_cast<size_t>(v) * ca + 8 * q) = val;
    else
        *reinterpret_cast<uint4*>(dB + static_cast<size_t>(v) * cb + 8 * q - ca) = val;
}

// Mean squared error over the first ct channels of P (n voxels x cp channels)
// against T (n x ct, float32): loss += sum (p - t)^2 * scale, dP = 2 (p - t) scale
// (zero in channels ct..cp-1).  A thread per voxel.
__global__ __launch_bounds__(256) void k_mse(const uint16_t* __restrict__ P, int cp, const float* __restrict__ T,
                                             int ct, long long n, float pscale, float scale,
                                             uint16_t* __restrict__ dP, float* __restrict__ loss) {
    const long long v = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    float s = 0.0f;
    if (v < n) {
        for (int c = 0; c < cp; ++c) {
            float gval = 0.0f;
            if (c < ct) {
                const float e = bf2f(P[static_cast<size_t>(v) * cp + c]) * pscale - T[static_cast<size_t>(v) * ct + c];
                s += e * e;
                gval = 2.0f * e * scale * pscale;
            }
            dP[static_cast<size_t>(v) * cp + c] = f2bf(gval);
        }
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1] + part[2] + part[3]) * scale);
}

// Adam (Kingma & Ba 2015, Algorithm 1) on float32 parameters; w16 (nullable)
// receives the bfloat16 copy the kernels read.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ w16, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

// float32 -> bfloat16 (round to nearest even)
__global__ void k_to_bf16(const float* __restrict__ in, long long n, uint16_t* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// Network input: (mov, ref) normalised with one (mean, 1/std), channels 2..7 zero.
__global__ void k_net_input(const float* __restrict__ mov, const float* __restrict__ ref, long long n,
                            const float* __restrict__ ms, uint16_t* __restrict__ out) {
    const long long v = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= n) return;
    const float mu = ms[0], inv = ms[1] > 0.0f ? 1.0f / ms[1] : 1.0f;
    Pack8 o;
    o.q = make_uint4(0u, 0u, 0u, 0u);
    o.h[0] = f2bf((mov[v] - mu) * inv);
    o.h[1] = f2bf((ref[v] - mu) * inv);
    reinterpret_cast<uint4*>(out)[v] = o.q;
}

// The same on a padded grid (dp, hp, wp) >= (d, h, w): voxels past the
// volume repeat its nearest edge voxel (a thread per padded voxel).
// Frame blockIdx.y of a batch: mov + y * d*h*w, out + y * dp*hp*wp records.
__global__ void k_net_input_padded(const float* __restrict__ mov, const float* __restrict__ ref, int d, int h,
                                   int w, int dp, int hp, int wp, const float* __restrict__ ms,
                                   uint16_t* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const long long n = static_cast<long long>(dp) * hp * wp;
    if (i >= n) return;
    mov += static_cast<size_t>(blockIdx.y) * d * h * w;
    out += static_cast<size_t>(blockIdx.y) * n * 8;
    const int x = min(static_cast<int>(i % wp), w - 1);
    const int y = min(static_cast<int>((i / wp) % hp), h - 1);
    const int z = min(static_cast<int>(i / (static_cast<long long>(wp) * hp)), d - 1);
    const size_t v = (static_cast<size_t>(z) * h + y) * w + x;
    const float mu = ms[0], inv = ms[1] > 0.0f ? 1.0f / ms[1] : 1.0f;
    Pack8 o;
    o.q = make_uint4(0u, 0u, 0u, 0u);
    o.h[0] = f2bf((mov[v] - mu) * inv);
    o.h[1] = f2bf((ref[v] - mu) * inv);
    reinterpret_cast<uint4*>(out)[i] = o.q;
}

// field[v][c] = scale * P[v][c], c < 3 (float32, the warm-start field)
__global__ void k_net_output(const uint16_t* __restrict__ P, int cp, long long n, float scale,
                             float* __restrict__ field) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= 3 * n) return;
    const long long v = i / 3;
    const int c = static_cast<int>(i - 3 * v);
    field[i] = bf2f(P[static_cast<size_t>(v) * cp + c]) * scale;
}

// The network's field at the control points of pyramid level L (P stored on
// an hs x ws grid per plane, of which h3 x w3 is the volume's level q, the
// rest edge padding): CP (i, j) of plane z sits at level voxel
// (r + p i, r + p j), i.e. full-resolution
// X = (x_L + 0.5) 2^L - 0.5, and at (X + 0.5) / 2^q - 0.5 in the network's
// level-q grid (q = log2 of full / network size); bilinear there with edge
// clamping, times scale, in level-L units (x, y divided by 2^L).  out: 3
// planes of nz * ncy * ncx.  A thread per control point.
__global__ void k_net_output_cp(const uint16_t* __restrict__ P, int cp, int h3, int w3, int hs, int ws, float inv_q,
                                int nz, int ncx, int ncy, int r, int pstep, float lscale, float scale,
                                float* __restrict__ out, long long pstride) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const long long ncp = static_cast<long long>(ncx) * ncy * nz;
    if (i >= ncp) return;
    // frame blockIdx.y of a batch (pstride elements of P per frame)
    P += static_cast<size_t>(blockIdx.y) * pstride;
    out += static_cast<size_t>(blockIdx.y) * 3 * ncp;
    const int ix = static_cast<int>(i % ncx);
    const int iy = static_cast<int>((i / ncx) % ncy);
    const int z = static_cast<int>(i / (static_cast<long long>(ncx) * ncy));
    // full-resolution position of the CP, then its coordinate in the network's grid
    const float u = fminf(fmaxf((static_cast<float>(r + pstep * ix) + 0.5f) * lscale * inv_q - 0.5f, 0.0f),
                          static_cast<float>(w3 - 1));
    const float v = fminf(fmaxf((static_cast<float>(r + pstep * iy) + 0.5f) * lscale * inv_q - 0.5f, 0.0f),
                          static_cast<float>(h3 - 1));
    const int u0 = static_cast<int>(floorf(u)), v0 = static_cast<int>(floorf(v));
    const int u1 = min(u0 + 1, w3 - 1), v1 = min(v0 + 1, h3 - 1);
    const float tu = u - static_cast<float>(u0), tv = v - static_cast<float>(v0);
    const size_t pz = static_cast<size_t>(z) * hs;      // stored grid hs x ws >= h3 x w3
    const uint16_t* p00 = P + ((pz + v0) * ws + u0) * cp;
    const uint16_t* p01 = P + ((pz + v0) * ws + u1) * cp;
    const uint16_t* p10 = P + ((pz + v1) * ws + u0) * cp;
    const uint16_t* p11 = P + ((pz + v1) * ws + u1) * cp;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float a = bf2f(p00[c]) + tu * (bf2f(p01[c]) - bf2f(p00[c]));
        const float b = bf2f(p10[c]) + tu * (bf2f(p11[c]) - bf2f(p10[c]));
        const float val = (a + tv * (b - a)) * scale;
        out[c * ncp + i] = c < 2 ? val / lscale : val;
    }
}

inline unsigned blocks_of(long long n, int per = 256) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

void launch_conv_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, NetGeom g, int batch,
                     int cin, int cout, int act, float slope, hipStream_t s) {
    const long long nv = static_cast<long long>(g.d) * g.h * g.w;
    // four voxel tiles per wave while that still gives 512 workgroups
    // (measured: 65k-voxel layers are faster with four tiles, 8k-voxel ones
    // with one), then as many output-channel tiles per wave (up to 4) as keep
    // 256 workgroups (measured at batch 1 and 16: better than 512; 128 and 64
    // gain nothing more)
    const int ncot = cout / 16;
    const long long kCtMinBlocks = 256;
    const int nt = static_cast<long long>(blocks_of(nv, 256)) * ncot * batch >= 512 ? 4 : 1;
    const long long vblocks = blocks_of(nv, 64 * nt);
    int ct = 1;
    for (int c : {4, 2}) {
        if (ncot % c == 0 && vblocks * (ncot / c) * batch >= kCtMinBlocks) {
            ct = c;
            break;
        }
    }
    dim3 grid(static_cast<unsigned>(vblocks), ncot / ct, batch);
    using Fn = void (*)(const uint16_t*, const uint16_t*, const float*, uint16_t*, NetGeom, int, int, int, float);
    static const Fn table[2][3] = {{k_conv_fwd<1, 1>, k_conv_fwd<1, 2>, k_conv_fwd<1, 4>},
                                   {k_conv_fwd<4, 1>, k_conv_fwd<4, 2>, k_conv_fwd<4, 4>}};
    hipLaunchKernelGGL(table[nt == 4 ? 1 : 0][ct == 4 ? 2 : ct - 1], grid, dim3(256), 0, s, x, w, bias, y, g, cin,
                       cout, act, slope);
}

void launch_flip_weights(const uint16_t* w, int cin, int cout, uint16_t* wf, hipStream_t s) {
    hipLaunchKernelGGL(k_flip, dim3(blocks_of(27LL * cin * cout)), dim3(256), 0, s, w, cin, cout, wf);
}

void launch_conv_wgrad(const uint16_t* dy, const uint16_t* x, NetGeom g, int batch, int cin, int cout, float* dw,
                       float* db, hipStream_t s) {
    const long long total = static_cast<long long>(g.d) * g.h * g.w * batch;
    const long long tiles = static_cast<long long>(cout / 16) * ((cin + 15) / 16);
    // about 4096 waves in all, at least 256 voxels per wave
    long long chunk = (total * tiles + 4095) / 4096;
    chunk = std::max<long long>(256, (chunk + 31) / 32 * 32);
    const long long chunks = (total + chunk - 1) / chunk;
    dim3 grid(static_cast<unsigned>(chunks), cout / 16, (cin + 15) / 16);
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(64), 0, s, dy, x, g, batch, cin, cout, static_cast<int>(chunk), dw,
                       db);
}

void launch_lrelu_bwd(uint16_t* gr, const uint16_t* y, long long n, float slope, hipStream_t s) {
    hipLaunchKernelGGL(k_lrelu_bwd, dim3(blocks_of(n / 8)), dim3(256), 0, s, gr, y, n / 8, slope);
}

void launch_down2(const uint16_t* x, NetGeom g, int batch, int c, float scale, uint16_t* y, hipStream_t s) {
    const long long n8 = static_cast<long long>(batch) * (g.d / 2) * (g.h / 2) * (g.w / 2) * (c / 8);
    hipLaunchKernelGGL(n8 + 256 < (1LL << 32) ? k_down2<unsigned> : k_down2<long long>, dim3(blocks_of(n8)), dim3(256), 0,
                       s, x, g, c, n8, scale, y);
}

void launch_up2(const uint16_t* x, NetGeom g, int batch, int c, float scale, int accumulate, uint16_t* y,
                hipStream_t s) {
    const long long n8 = static_cast<long long>(batch) * g.d * g.h * g.w * (c / 8);
    hipLaunchKernelGGL(n8 + 256 < (1LL << 32) ? k_up2<unsigned> : k_up2<long long>, dim3(blocks_of(n8)), dim3(256), 0, s,
                       x, g, c, n8, scale, accumulate, y);
}

void launch_concat(const uint16_t* a, int ca, const uint16_t* b, int cb, long long nvox, uint16_t* y, hipStream_t s) {
    const long long n8 = nvox * ((ca + cb) / 8);
    hipLaunchKernelGGL(n8 + 256 < (1LL << 32) ? k_concat<unsigned> : k_concat<long long>, dim3(blocks_of(n8)), dim3(256),
                       0, s, a, ca, b, cb, n8, y);
}

void launch_upcat(const uint16_t* x, int ca, const uint16_t* b, int cb, NetGeom g, int batch, uint16_t* y,
                  hipStream_t s) {
    const long long n8 = static_cast<long long>(batch) * g.d * g.h * g.w * ((ca + cb) / 8);
    hipLaunchKernelGGL(n8 + 256 < (1LL << 32) ? k_upcat<unsigned> : k_upcat<long long>, dim3(blocks_of(n8)), dim3(256),
                       0, s, x, ca, b, cb, g, n8, y);
}

void launch_split(const uint16_t* gr, int ca, int cb, long long nvox, uint16_t* da, uint16_t* db, hipStream_t s) {
    const long long n8 = nvox * ((ca + cb) / 8);
    hipLaunchKernelGGL(n8 + 256 < (1LL << 32) ? k_split<unsigned> : k_split<long long>, dim3(blocks_of(n8)), dim3(256),
                       0, s, gr, ca, cb, n8, da, db);
}

void launch_mse(const uint16_t* p, int cp, const float* t, int ct, long long nvox, float pscale, float scale,
                uint16_t* dp, float* loss, hipStream_t s) {
    hipLaunchKernelGGL(k_mse, dim3(blocks_of(nvox)), dim3(256), 0, s, p, cp, t, ct, nvox, pscale, scale, dp, loss);
}

void launch_adam(float* w, const float* gr, float* m, float* v, uint16_t* w16, long long n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t s) {
    const float bc1 = 1.0f - std::pow(b1, static_cast<float>(step));
    const float bc2 = 1.0f - std::pow(b2, static_cast<float>(step));
    hipLaunchKernelGGL(k_adam, dim3(blocks_of(n)), dim3(256), 0, s, w, gr, m, v, w16, n, lr, b1, b2, eps, bc1, bc2);
}

void launch_to_bf16(const float* in, long long n, uint16_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_of(n)), dim3(256), 0, s, in, n, out);
}

void launch_net_input(const float* mov, const float* ref, long long n, const float* ms, uint16_t* out,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_net_input, dim3(blocks_of(n)), dim3(256), 0, s, mov, ref, n, ms, out);
}

void launch_net_output_cp(const uint16_t* p, int cp, int h3, int w3, int hs, int ws, int q, int nz, int ncx, int ncy,
                          int r, int level, float scale, float* out, hipStream_t s, int nb, long long pstride) {
    const long long ncp = static_cast<long long>(ncx) * ncy * nz;
    hipLaunchKernelGGL(k_net_output_cp, dim3(blocks_of(ncp), nb), dim3(256), 0, s, p, cp, h3, w3, hs, ws,
                       1.0f / static_cast<float>(1 << q), nz, ncx, ncy, r, 2 * r + 1,
                       static_cast<float>(1 << level), scale, out, pstride);
}

void launch_net_input_padded(const float* mov, const float* ref, int d, int h, int w, int dp, int hp, int wp,
                             const float* ms, uint16_t* out, hipStream_t s, int nb) {
    const long long n = static_cast<long long>(dp) * hp * wp;
    hipLaunchKernelGGL(k_net_input_padded, dim3(blocks_of(n), nb), dim3(256), 0, s, mov, ref, d, h, w, dp, hp, wp, ms,
                       out);
}

void launch_net_output(const uint16_t* p, int cp, long long n, float scale, float* field, hipStream_t s) {
    hipLaunchKernelGGL(k_net_output, dim3(blocks_of(3 * n)), dim3(256), 0, s, p, cp, n, scale, field);
}

}  // namespace wr
