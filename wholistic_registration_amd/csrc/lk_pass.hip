// The fused per-voxel pass of one LK iteration (k_lk_pass) on gfx950.
//
// Per voxel of a pyramid level (calFlow3d_Wei_v1.py:442-588): dense motion
// (the CP bilinear interpolant, :265-292,627-640, or the level's initial
// field on its first iteration), trilinear warp of the moving level image at
// the deformed coordinate (:37-66), the mask warp (:473-474), the residual
// with its 3x3x1 mean and mask (:477-485), the six boundary-clipped
// central-difference samples (:134-206), and the nine per-patch
// structure-tensor sums (:546-588) plus the residual energy (:101-131).
//
// Tiles: a workgroup owns `cols` = np_x*(2r+1) columns and `rows` = 2r+1
// rows of one z slice, i.e. whole patches, so each per-patch sum is finished
// inside one workgroup and written once (no atomics, a deterministic order).
// A wave owns 62 columns plus one halo lane on each side and walks the rows
// on its own; the 3x3 residual mean uses a register ring for the vertical sum
// and cross-lane shuffles for the horizontal one.  Two extra rows recompute
// the halo.
//
// The launch also carries, after the tiles, the workgroups of the iteration's
// neighbour differences on the CP grid (they read only the CP field).
// Workgroups are assigned XCD-major (each XCD gets a contiguous range of
// tiles) with z fastest, so z-neighbour tiles, which share moving-image
// slices, run on the same XCD and hit its L2.
//
// The moving image is read by per-lane gathers around the deformed
// coordinate, and their cost is set by the vector-memory pipeline (address
// and data path of the texture units, ~80 % busy), not by HBM bytes: removing
// a quarter of the gather instructions saves more than halving their bytes.
// So the moving image is stored the way each storage needs the fewest gather
// words per voxel:
//   * float32: the level image itself, padded by kPad replicated voxels at
//     both x ends (and rows / planes in y and z), so each of the 12 tap rows
//     of the six clipped samples is one plain 16- or 8-byte load (32 taps);
//     where no used sample is clipped the gradients come from the central
//     rows' x differences (eval_inner), else from the general clipped forms.
//   * bfloat16 / float16: records (V, Dx, Dy, Dz) per voxel, D_a the central
//     difference with replicated edges, so the warp and all three gradients
//     are one trilinear interpolation of the 8 corner records: four 16-byte
//     loads (x pairs of the four (y, z) corner rows) instead of 12 gathers.
// Row addresses are 32-bit byte offsets (float32) or 64-bit (records); the CP
// interpolation weights of the tile's rows are formed once in LDS; both
// gradient scalings by 1/2 are applied to the per-patch sums (exact).

#include <cfloat>
#include <cmath>

#include "lk_common.h"
#include "lk_ctl.h"
#include "wr_internal.h"

namespace wr {

namespace {

constexpr int kMaxThreads = 256;

// Coordinate clamped to [0, hi] (map_coordinates mode='nearest') split into
// floor and fraction.
struct Samp {
    int f;
    float t;
};
__device__ __forceinline__ Samp split(float c, float hi) {
    const float u = __builtin_amdgcn_fmed3f(c, 0.0f, hi);
    const float fl = floorf(u);
    Samp s;
    s.f = static_cast<int>(fl);
    s.t = u - fl;
    return s;
}

// A level image stored padded (padded_layout): voxel (x, y, z), for x in
// [-kPad, nx - 1 + kPad], y in [-kPadLo, ny - 1 + kPadHi] and z likewise,
// is the float at byte offset origin + z*planeB + y*pitchB + x*4 of the
// allocation; the pads hold the nearest voxel of the volume, so every tap of
// a clamped gather is an unconditional load: one per-lane 32-bit offset from
// the allocation start plus the tap row's fixed offset.
struct PVol {
    const unsigned char* base;   // allocation start
    unsigned pitchB, planeB; // row and plane strides in bytes
    unsigned plane_rows;     // stored rows per plane
    unsigned origin;         // byte offset of voxel (0, 0, 0)
    int nx, ny, nz;
    float hx, hy, hz;        // n - 1
};

typedef float f4v __attribute__((ext_vector_type(4), aligned(4)));
typedef float f4a __attribute__((ext_vector_type(4), aligned(16)));
typedef float f2v __attribute__((ext_vector_type(2), aligned(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2), aligned(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4), aligned(8)));

// The moving level image as the records the pass interpolates
// (moving_records): record (x, y, z) = (V, Dx, Dy, Dz) with
// D_a = V(a + 1) - V(a - 1), edges replicated, for x in [0, nx], y in
// [0, ny], z in [0, nz] (the last of each axis repeats its neighbour, so the
// upper corner of a coordinate clamped to the last voxel is a plain load).
// Rows of nx + 1 records of kRec bytes; planes of ny + 1 rows.
struct QVol {
    const unsigned char* base;
    unsigned rowB;           // bytes per stored row
    unsigned plane_rows;     // stored rows per plane (ny + 1)
    float hx, hy, hz;        // n - 1
};

// Bytes per stored voxel of a level image of storage S (Store), and per
// record of the moving image.
template <int S>
constexpr unsigned kEs = S == ST_F32 ? 4u : 2u;
template <int S>
constexpr unsigned kRec = 4u * kEs<S>;

// The two 16-bit values of a word as float32 (bfloat16: the high half of a
// float32; float16: converted).
template <int S>
__device__ __forceinline__ float lo16(unsigned w) {
    if constexpr (S == ST_BF16) return __uint_as_float(w << 16);
    else return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu)));
}
template <int S>
__device__ __forceinline__ float hi16(unsigned w) {
    if constexpr (S == ST_BF16) return __uint_as_float(w & 0xffff0000u);
    else return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}

// Four or two consecutive voxels at byte offset off of the allocation, as
// float32: one 16- or 8-byte load for float32 storage (4-byte aligned), one
// 8- or 4-byte load for 16-bit storage (2-byte aligned: the hardware serves
// unaligned global loads).
template <int S>
__device__ __forceinline__ f4v ld4(const PVol& v, unsigned off) {
    if constexpr (S == ST_F32) {
        return *reinterpret_cast<const f4v*>(v.base + off);
    } else {
        const u2v q = *reinterpret_cast<const u2v*>(v.base + off);
        f4v r;
        r.x = lo16<S>(q.x);
        r.y = hi16<S>(q.x);
        r.z = lo16<S>(q.y);
        r.w = hi16<S>(q.y);
        return r;
    }
}
template <int S>
__device__ __forceinline__ f2v ld2(const PVol& v, unsigned off) {
    if constexpr (S == ST_F32) {
        return *reinterpret_cast<const f2v*>(v.base + off);
    } else {
        const unsigned q = *reinterpret_cast<const unsigned*>(v.base + off);
        f2v r;
        r.x = lo16<S>(q);
        r.y = hi16<S>(q);
        return r;
    }
}

// One template voxel (dense level image of storage S).
template <int S>
__device__ __forceinline__ float ld_ref(const void* ref, int i) {
    if constexpr (S == ST_F32) {
        return static_cast<const float*>(ref)[i];
    } else {
        const unsigned w = static_cast<const uint16_t*>(ref)[i];
        return lo16<S>(w);
    }
}

// Byte offset of voxel (fx, fy, fz) from the origin: (fz * plane_rows + fy)
// rows (fewer than 2^24) of pitchB (< 2^24) bytes, 24-bit products.
template <int S>
__device__ __forceinline__ unsigned voxel_off(const PVol& v, int fx, int fy, int fz) {
    const unsigned row = __umul24(static_cast<unsigned>(fz), v.plane_rows) + static_cast<unsigned>(fy);
    return __umul24(row, v.pitchB) + static_cast<unsigned>(fx) * kEs<S>;
}

struct Grad {
    float w, gx, gy, gz;
};

struct Coord {
    float cx, cy, cz;
    Samp sx, sy, sz;
};

__device__ __forceinline__ Coord make_coord(float hx, float hy, float hz, float cx, float cy, float cz) {
    Coord c;
    c.cx = cx;
    c.cy = cy;
    c.cz = cz;
    c.sx = split(cx, hx);
    c.sy = split(cy, hy);
    c.sz = split(cz, hz);
    return c;
}

// ---- float32 storage: taps of the padded level image

// Sample of the 4-point line l[0..3] = f(clamp(c0 - 1 + q)) at the clipped
// coordinate `c` (inside [0, n-1]), c0 = floor of the clamped centre.
// Whenever the fraction is non-zero the two taps lie inside the line; with a
// zero fraction the second tap has weight 0.
__device__ __forceinline__ float line_sample(float l0, float l1, float l2, float l3, int c0, float c) {
    const int f = static_cast<int>(floorf(c));
    const float t = c - static_cast<float>(f);
    const int o = clampi(f - c0 + 1, 0, 3);
    const int o1 = min(o + 1, 3);
    return lerpf(pick4(l0, l1, l2, l3, o), pick4(l0, l1, l2, l3, o1), t);
}

// T(clip(c + 1)) - T(clip(c - 1)) along one axis of the line.  Away from the
// faces (1 <= c <= n - 2) neither sample is clipped: c - 1 is exact in
// float32 and floors to c0 - 1 with the centre's fraction t0, and c + 1
// floors to c0 + 1 or, rounded up to an integer, to c0 + 2 with a zero
// fraction -- the same taps and weights the general form picks.
__device__ __forceinline__ float line_diff(float l0, float l1, float l2, float l3, int c0, float t0, float c,
                                           float hi) {
    if (c >= 1.0f && c <= hi - 1.0f) {
        const float cp = c + 1.0f;
        const float fp = floorf(cp);
        const float up = fp == static_cast<float>(c0 + 1) ? lerpf(l2, l3, cp - fp) : l3;
        return up - lerpf(l0, l1, t0);
    }
    return line_sample(l0, l1, l2, l3, c0, __builtin_amdgcn_fmed3f(c + 1.0f, 0.0f, hi)) -
           line_sample(l0, l1, l2, l3, c0, __builtin_amdgcn_fmed3f(c - 1.0f, 0.0f, hi));
}

// The 32 voxels around one deformed coordinate, as 4 x 16-byte rows holding
// X0..X3 at (Y1|Y2, Z1|Z2) and 8 x 8-byte rows holding X1, X2 for the y and
// z lines (see eval).
struct Taps {
    f4v q11, q21, q12, q22;
    f2v e01, e02, e31, e32, f10, f20, f13, f23;
};

// The 12 tap rows of the gather at c, as offsets from that of voxel
// (X0, Y0, Z0) = (fx - 1, fy - 1, fz - 1), which the pads keep inside the
// allocation.  Where the clamped reference would repeat an edge row or
// plane, the pads hold it.
template <int S>
__device__ __forceinline__ Taps gather_taps(const PVol& v, const Coord& c) {
    const unsigned py = v.pitchB, pz = v.planeB, e = kEs<S>;
    const unsigned a = v.origin + voxel_off<S>(v, c.sx.f - 1, c.sy.f, c.sz.f) - pz - py;
    Taps t;
    t.q11 = ld4<S>(v, a + pz + py);
    t.q21 = ld4<S>(v, a + pz + 2u * py);
    t.q12 = ld4<S>(v, a + 2u * pz + py);
    t.q22 = ld4<S>(v, a + 2u * pz + 2u * py);
    t.e01 = ld2<S>(v, a + pz + e);
    t.e02 = ld2<S>(v, a + 2u * pz + e);
    t.e31 = ld2<S>(v, a + pz + 3u * py + e);
    t.e32 = ld2<S>(v, a + 2u * pz + 3u * py + e);
    t.f10 = ld2<S>(v, a + py + e);
    t.f20 = ld2<S>(v, a + 2u * py + e);
    t.f13 = ld2<S>(v, a + 3u * pz + py + e);
    t.f23 = ld2<S>(v, a + 3u * pz + 2u * py + e);
    return t;
}

// Warp value and central-difference gradients at the deformed coordinate
// (calFlow3d:134-206): I_a = (T(c + e_a) - T(c - e_a)) / 2 where only the
// perturbed coordinate is clipped to [0, n_a - 1] and the other two follow
// map_coordinates' nearest mode.  With c0 = floor(clamp(c)) per axis, every
// sample's taps lie on the 4-point lines c0-1..c0+2 (clamped), so the 32
// distinct voxels are the 2x2x2 cube around the centre plus, per axis, the
// two planes one step further out.
__device__ __forceinline__ Grad eval_taps(const PVol& v, const Coord& c, const Taps& q) {
    const Samp& sx = c.sx;
    const Samp& sy = c.sy;
    const Samp& sz = c.sz;
    // central cube c[z][y][x]
    const float c111 = q.q11.y, c112 = q.q11.z, c121 = q.q21.y, c122 = q.q21.z;
    const float c211 = q.q12.y, c212 = q.q12.z, c221 = q.q22.y, c222 = q.q22.z;
    // x line: bilinear in (y, z) of columns X0..X3
    const float g1 = lerpf(lerpf(c111, c121, sy.t), lerpf(c211, c221, sy.t), sz.t);
    const float g2 = lerpf(lerpf(c112, c122, sy.t), lerpf(c212, c222, sy.t), sz.t);
    const float g0 = lerpf(lerpf(q.q11.x, q.q21.x, sy.t), lerpf(q.q12.x, q.q22.x, sy.t), sz.t);
    const float g3 = lerpf(lerpf(q.q11.w, q.q21.w, sy.t), lerpf(q.q12.w, q.q22.w, sy.t), sz.t);
    // x-interpolated central rows, shared by the y and z lines
    const float x11 = lerpf(c111, c112, sx.t), x21 = lerpf(c121, c122, sx.t);
    const float x12 = lerpf(c211, c212, sx.t), x22 = lerpf(c221, c222, sx.t);
    // y line: bilinear in (x, z) of rows Y0..Y3
    const float h1 = lerpf(x11, x12, sz.t);
    const float h2 = lerpf(x21, x22, sz.t);
    const float h0 = lerpf(lerpf(q.e01.x, q.e01.y, sx.t), lerpf(q.e02.x, q.e02.y, sx.t), sz.t);
    const float h3 = lerpf(lerpf(q.e31.x, q.e31.y, sx.t), lerpf(q.e32.x, q.e32.y, sx.t), sz.t);
    // z line: bilinear in (x, y) of slices Z0..Z3
    const float k1 = lerpf(x11, x21, sy.t);
    const float k2 = lerpf(x12, x22, sy.t);
    const float k0 = lerpf(lerpf(q.f10.x, q.f10.y, sx.t), lerpf(q.f20.x, q.f20.y, sx.t), sy.t);
    const float k3 = lerpf(lerpf(q.f13.x, q.f13.y, sx.t), lerpf(q.f23.x, q.f23.y, sx.t), sy.t);

    Grad o;
    o.w = lerpf(k1, k2, sz.t);
    o.gx = line_diff(g0, g1, g2, g3, sx.f, sx.t, c.cx, v.hx);
    o.gy = line_diff(h0, h1, h2, h3, sy.f, sy.t, c.cy, v.hy);
    o.gz = line_diff(k0, k1, k2, k3, sz.f, sz.t, c.cz, v.hz);
    return o;
}

// The same quantities where no perturbed coordinate is clipped (1 <= c <=
// n - 2 on every axis): T(c + e_a) and T(c - e_a) then share c's fractions,
// and T(c + e_a) - T(c - e_a) is the interpolant of the central differences
// V(a + 1) - V(a - 1) at c.  Each central row (y, z) gives its x sample and
// its x difference; the y and z differences reuse the central rows' x
// samples.  The sample w is formed in the same order as in eval, so it is
// equal bit for bit.
__device__ __forceinline__ Grad eval_inner(const Coord& c, const Taps& q) {
    const float tx = c.sx.t, ty = c.sy.t, tz = c.sz.t;
    // x samples and x differences of the central rows (y1|y2, z1|z2)
    const float a11 = lerpf(q.q11.y, q.q11.z, tx), a21 = lerpf(q.q21.y, q.q21.z, tx);
    const float a12 = lerpf(q.q12.y, q.q12.z, tx), a22 = lerpf(q.q22.y, q.q22.z, tx);
    const float d11 = lerpf(q.q11.z - q.q11.x, q.q11.w - q.q11.y, tx);
    const float d21 = lerpf(q.q21.z - q.q21.x, q.q21.w - q.q21.y, tx);
    const float d12 = lerpf(q.q12.z - q.q12.x, q.q12.w - q.q12.y, tx);
    const float d22 = lerpf(q.q22.z - q.q22.x, q.q22.w - q.q22.y, tx);
    // x samples of the outer rows of the y and z lines
    const float b01 = lerpf(q.e01.x, q.e01.y, tx), b02 = lerpf(q.e02.x, q.e02.y, tx);
    const float b31 = lerpf(q.e31.x, q.e31.y, tx), b32 = lerpf(q.e32.x, q.e32.y, tx);
    const float e10 = lerpf(q.f10.x, q.f10.y, tx), e20 = lerpf(q.f20.x, q.f20.y, tx);
    const float e13 = lerpf(q.f13.x, q.f13.y, tx), e23 = lerpf(q.f23.x, q.f23.y, tx);
    Grad o;
    o.w = lerpf(lerpf(a11, a21, ty), lerpf(a12, a22, ty), tz);
    o.gx = lerpf(lerpf(d11, d21, ty), lerpf(d12, d22, ty), tz);
    o.gy = lerpf(lerpf(a21 - b01, b31 - a11, ty), lerpf(a22 - b02, b32 - a12, ty), tz);
    o.gz = lerpf(lerpf(a12 - e10, e13 - a11, tz), lerpf(a22 - e20, e23 - a21, tz), ty);
    return o;
}

// 1 <= c <= hi - 1: neither perturbed sample of the axis is clipped.
__device__ __forceinline__ bool inner1(float c, float hi) { return c >= 1.0f && c <= hi - 1.0f; }

// ---- 16-bit storage: records of the level image

// The eight records around a deformed coordinate, [z][y][x] for the
// floors f and f + 1 of its clamped position, each as (V, Dx, Dy, Dz).
struct Corners {
    f4v r[2][2][2];
};

// Two records of 16-bit storage (one 16-byte word) as float32.
template <int S>
__device__ __forceinline__ void unpack2(const u4v w, f4v& a, f4v& b) {
    a.x = lo16<S>(w.x);
    a.y = hi16<S>(w.x);
    a.z = lo16<S>(w.y);
    a.w = hi16<S>(w.y);
    b.x = lo16<S>(w.z);
    b.y = hi16<S>(w.z);
    b.z = lo16<S>(w.w);
    b.w = hi16<S>(w.w);
}

// Four loads of 32 bytes (float32 records: two 16-byte loads each) or of
// 16 bytes (16-bit records): the x pairs of the four (y, z) corner rows.
template <int S>
__device__ __forceinline__ Corners gather_rec(const QVol& q, const Coord& c) {
    const unsigned row = __umul24(static_cast<unsigned>(c.sz.f), q.plane_rows) + static_cast<unsigned>(c.sy.f);
    const unsigned char* p0 = q.base + static_cast<size_t>(row) * q.rowB + static_cast<unsigned>(c.sx.f) * kRec<S>;
    const size_t pz = static_cast<size_t>(q.rowB) * q.plane_rows;
    Corners o;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const unsigned char* p = p0 + k * pz + j * static_cast<size_t>(q.rowB);
            if constexpr (S == ST_F32) {
                o.r[k][j][0] = *reinterpret_cast<const f4a*>(p);
                o.r[k][j][1] = *reinterpret_cast<const f4a*>(p + 16);
            } else {
                unpack2<S>(*reinterpret_cast<const u4v*>(p), o.r[k][j][0], o.r[k][j][1]);
            }
        }
    }
    return o;
}

__device__ __forceinline__ f4v lerp4(const f4v a, const f4v b, float t) {
    f4v o;
    o.x = lerpf(a.x, b.x, t);
    o.y = lerpf(a.y, b.y, t);
    o.z = lerpf(a.z, b.z, t);
    o.w = lerpf(a.w, b.w, t);
    return o;
}

// Warp value and central-difference gradients at the deformed coordinate c
// (calFlow3d:37-66,134-206): W = T(c) and 2 I_a = T(c + e_a) - T(c - e_a),
// where T is the clamped trilinear interpolant and only the perturbed
// coordinate is clipped to [0, n_a - 1].
//
// With c_a inside [0, n_a - 1] both perturbed samples share c's fraction
// (or sit on the clipped face, where the replicated edge of D reproduces the
// clip), so 2 I_a is the trilinear interpolant of D_a = V(a + 1) - V(a - 1)
// (edges replicated) at the clamped c.  Outside, at distance d from the
// volume along a, the clipped difference is (1 - d) times that interpolant
// (taken at the face), and 0 from d >= 1 on.  W is formed x, then y, then z,
// as the reference's order-1 interpolation at the clamped coordinate.
__device__ __forceinline__ Grad eval_rec(const Coord& c, const Corners& q) {
    const f4v y0 = lerp4(lerp4(q.r[0][0][0], q.r[0][0][1], c.sx.t), lerp4(q.r[0][1][0], q.r[0][1][1], c.sx.t),
                         c.sy.t);
    const f4v y1 = lerp4(lerp4(q.r[1][0][0], q.r[1][0][1], c.sx.t), lerp4(q.r[1][1][0], q.r[1][1][1], c.sx.t),
                         c.sy.t);
    const f4v v = lerp4(y0, y1, c.sz.t);
    Grad o;
    o.w = v.x;
    o.gx = v.y;
    o.gy = v.z;
    o.gz = v.w;
    const float ux = static_cast<float>(c.sx.f) + c.sx.t, uy = static_cast<float>(c.sy.f) + c.sy.t;
    const float uz = static_cast<float>(c.sz.f) + c.sz.t;
    if (__any(c.cx != ux || c.cy != uy || c.cz != uz)) {   // uniform: a lane samples outside the volume
        o.gx *= fmaxf(1.0f - fabsf(c.cx - ux), 0.0f);
        o.gy *= fmaxf(1.0f - fabsf(c.cy - uy), 0.0f);
        o.gz *= fmaxf(1.0f - fabsf(c.cz - uz), 0.0f);
    }
    return o;
}

// Trilinear sample only (4 x 8-byte rows of a float32 volume).
__device__ __forceinline__ float tri(const PVol& v, float cx, float cy, float cz) {
    const Samp sx = split(cx, v.hx), sy = split(cy, v.hy), sz = split(cz, v.hz);
    const unsigned a = v.origin + voxel_off<ST_F32>(v, sx.f, sy.f, sz.f);
    const unsigned py = v.pitchB, pz = v.planeB;
    const f2v r00 = ld2<ST_F32>(v, a), r01 = ld2<ST_F32>(v, a + py);
    const f2v r10 = ld2<ST_F32>(v, a + pz), r11 = ld2<ST_F32>(v, a + pz + py);
    const float p = lerpf(lerpf(r00.x, r00.y, sx.t), lerpf(r01.x, r01.y, sx.t), sy.t);
    const float q = lerpf(lerpf(r10.x, r10.y, sx.t), lerpf(r11.x, r11.y, sx.t), sy.t);
    return lerpf(p, q, sz.t);
}

struct PassArgs {
    PVol mov;              // float32 storage: the padded level image
    QVol q;                // 16-bit storage: its records
    PVol mmov;
    const void* ref;
    const uint8_t* mref;
    const float* m0;
    const float* cp;
    float* sums;
    double* errpart;
    Ctl* ctl;
    int k;                 // iteration of the level
    int check_parts;       // > 0: apply the |delta m| test of update k - 1 to mdpart[0 .. check_parts)
    const float* mdpart;
    double* log;
    int* counts;
    float* dnei;           // neighbour differences (written by the appended workgroups)
    double* penpart;       // their per-workgroup sums of |D|^2
    long long npen;        // appended neighbour-difference workgroups
    int nx, ny, nz, r, p, ncx, ncy, extra_x, extra_y;
    int cols, rows, nbx, nby, np_x;
    long long blocks;
    long long ncp;
    float z_ratio;
    float hx, hy, hz;      // level extents - 1
    long long fs;          // frame stride of the per-frame buffers (Frames)
};

// Decode the XCD-major, z-fastest tile order.
__device__ __forceinline__ void tile_of(long long nb, int nz, int nbx, int& bx, int& by, int& z,
                                        long long& logical) {
    const long long bid = blockIdx.x;
    const int xcd = static_cast<int>(bid & 7);
    const long long slot = bid >> 3;
    const long long per = nb >> 3;
    const int rem = static_cast<int>(nb & 7);
    logical = xcd * per + min(xcd, rem) + slot;
    z = static_cast<int>(logical % nz);
    const long long t2 = logical / nz;
    bx = static_cast<int>(t2 % nbx);
    by = static_cast<int>(t2 / nbx);
}

// Each wave owns kWaveCols consecutive columns of the tile plus one halo
// lane on each side; waves walk the rows independently (the 3x3 mean's
// horizontal neighbours come from cross-lane shuffles), and the next row's
// gathers are issued before the current row is evaluated.
template <bool DENSE, bool MREF, bool MMOV, int S>
__global__ __launch_bounds__(kMaxThreads) void k_lk_pass(PassArgs a) {
    a.ctl = at_frame(a.ctl, a.fs);
    if (a.ctl->done) return;
    // frame blockIdx.y's buffers (the template and mask_ref are shared)
    a.mov.base = at_frame(a.mov.base, a.fs);
    a.q.base = at_frame(a.q.base, a.fs);
    a.mmov.base = at_frame(a.mmov.base, a.fs);
    a.m0 = at_frame(a.m0, a.fs);
    a.cp = at_frame(a.cp, a.fs);
    a.sums = at_frame(a.sums, a.fs);
    a.errpart = at_frame(a.errpart, a.fs);
    a.mdpart = at_frame(a.mdpart, a.fs);
    a.log = at_frame(a.log, a.fs);
    a.counts = at_frame(a.counts, a.fs);
    a.dnei = at_frame(a.dnei, a.fs);
    a.penpart = at_frame(a.penpart, a.fs);
    if (a.check_parts > 0) {
        // the |delta motion| stop test of the previous update, decided alike
        // by every workgroup from the same maxima; workgroup 0 records it
        const float md = parts_max(a.mdpart, a.check_parts);
        const bool stop = check_stop(a.ctl, a.k - 1, md);
        if (blockIdx.x == 0 && threadIdx.x == 0) check_record(a.ctl, a.k - 1, md, stop, a.log, a.counts);
        if (stop) return;
    }
    if (static_cast<long long>(blockIdx.x) >= a.blocks) {
        // appended workgroups: neighbour differences of the CP field
        // (calFlow3d_Wei_v1.py:69-98,488-498) and their |D|^2 partials
        const long long pb = static_cast<long long>(blockIdx.x) - a.blocks;
        const long long i = pb * blockDim.x + threadIdx.x;
        double q = 0.0, unused = 0.0;
        if (i < a.ncp) q = neidiff_one(a.cp, a.dnei, i, a.ncp, a.ncx, a.ncy, a.z_ratio);
        block_sum2(q, unused);
        if (threadIdx.x == 0) a.penpart[pb] = q;
        return;
    }
    __shared__ float s_red[9][kMaxThreads];
    __shared__ double s_err[kMaxThreads / 64];
    __shared__ float s_rt[kMaxThreads];     // CP y-weight of each walked row
    __shared__ int s_rj[kMaxThreads];       // its CP rows, relative to ja

    int bx, by, z;
    long long logical;
    tile_of(a.blocks, a.nz, a.nbx, bx, by, z, logical);
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int x0 = bx * a.cols, y0 = by * a.rows;
    const int lcol = wave * kWaveCols + lane - 1;        // column within the tile
    const int col = x0 + lcol;
    // every lane samples a real (clamped) voxel; lanes outside the tile carry
    // no weight
    const int xc = clampi(col, 0, a.nx - 1);
    const float xf = static_cast<float>(xc);
    const float zf = static_cast<float>(z);
    const bool owned = lane >= 1 && lane <= kWaveCols && lcol < a.cols && col < a.nx;
    const int cover_x = a.ncx * a.p, cover_y = a.ncy * a.p;
    const int wx = (owned && col < cover_x) ? 1 + (col == a.nx - 1 ? a.extra_x : 0) : 0;
    const float wxf = static_cast<float>(wx);
    const int plane = a.nx * a.ny;
    const int nvox = plane * a.nz;
    const int yend = min(y0 + a.rows, a.ny);
    const int nsteps = (yend - y0) + 2;

    // CP field: the walked rows interpolate between at most three CP rows
    // (ja .. ja+2); the x-interpolated values of those rows are formed once
    // per lane, the y weights of every row once per workgroup.
    float vcx[3] = {0.0f, 0.0f, 0.0f}, vcy[3] = {0.0f, 0.0f, 0.0f}, vcz[3] = {0.0f, 0.0f, 0.0f};
    if (!DENSE) {
        const float* cpx = a.cp + static_cast<size_t>(z) * a.ncy * a.ncx;
        const float* cpy = cpx + a.ncp;
        const float* cpz = cpy + a.ncp;
        const Axis1 axx = cp_axis(xc, a.r, a.p, a.ncx);
        const int ja = cp_axis(clampi(y0 - 1, 0, a.ny - 1), a.r, a.p, a.ncy).i0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int jr = min(ja + q, a.ncy - 1) * a.ncx;
            vcx[q] = lerpf(cpx[jr + axx.i0], cpx[jr + axx.i1], axx.t);
            vcy[q] = lerpf(cpy[jr + axx.i0], cpy[jr + axx.i1], axx.t);
            vcz[q] = lerpf(cpz[jr + axx.i0], cpz[jr + axx.i1], axx.t);
        }
        for (int i = tid; i < nsteps; i += blockDim.x) {
            const Axis1 ay = cp_axis(clampi(y0 - 1 + i, 0, a.ny - 1), a.r, a.p, a.ncy);
            s_rt[i] = ay.t;
            s_rj[i] = (ay.i0 - ja) | ((ay.i1 - ja) << 8);
        }
        __syncthreads();
    }

    float ax = 0.0f, ay = 0.0f, az = 0.0f, bxv = 0.0f, byv = 0.0f, bzv = 0.0f;
    int cur = -1;
    // deformed coordinate of row s of this lane
    auto coord_of = [&](int s) {
        const int yc = clampi(y0 - 1 + s, 0, a.ny - 1);
        float mx, my, mz;
        if (DENSE) {
            const int v = z * plane + yc * a.nx + xc;
            mx = a.m0[v];
            my = a.m0[v + nvox];
            mz = a.m0[v + 2 * nvox];
        } else {
            const int pr = __builtin_amdgcn_readfirstlane(s_rj[s]);
            if (pr != cur) {   // uniform: the row crossed into the next CP interval
                cur = pr;
                const int j0 = pr & 255, j1 = pr >> 8;
                ax = pick4(vcx[0], vcx[1], vcx[2], vcx[2], j0);
                ay = pick4(vcy[0], vcy[1], vcy[2], vcy[2], j0);
                az = pick4(vcz[0], vcz[1], vcz[2], vcz[2], j0);
                bxv = pick4(vcx[0], vcx[1], vcx[2], vcx[2], j1);
                byv = pick4(vcy[0], vcy[1], vcy[2], vcy[2], j1);
                bzv = pick4(vcz[0], vcz[1], vcz[2], vcz[2], j1);
            }
            const float t = s_rt[s];
            mx = lerpf(ax, bxv, t);
            my = lerpf(ay, byv, t);
            mz = lerpf(az, bzv, t);
        }
        return make_coord(a.hx, a.hy, a.hz, xf + mx, static_cast<float>(yc) + my, zf + mz);
    };

    float d0 = 0.0f, d1 = 0.0f;        // residual of the two previous rows
    float pgx = 0.0f, pgy = 0.0f, pgz = 0.0f;
    bool pmask = false;
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f, acc4 = 0.0f, acc5 = 0.0f, acc6 = 0.0f,
          acc7 = 0.0f, acc8 = 0.0f;
    double err = 0.0;

    if (wave * kWaveCols < a.cols) {   // uniform: the wave has tile columns
        for (int s = 0; s < nsteps; ++s) {
            const Coord cc = coord_of(s);
            Taps tt;
            Corners tr;
            if constexpr (S == ST_F32) tt = gather_taps<ST_F32>(a.mov, cc);
            else tr = gather_rec<S>(a.q, cc);
            // template / mask row of this step: uniform base, per-lane column
            const int vrow = z * plane + clampi(y0 - 1 + s, 0, a.ny - 1) * a.nx;
            const float rc = ld_ref<S>(a.ref, vrow + xc);
            const int yy = y0 - 1 + s;
            bool mk = false;
            if (MREF) mk = (a.mref + vrow)[xc] != 0;
            if (MMOV) mk = mk || tri(a.mmov, cc.cx, cc.cy, cc.cz) > 0.0f;
            Grad g;
            if constexpr (S == ST_F32) {
                // rows outside [y0, yend) and the uncovered tail carry no gradient
                const bool grad_row = yy >= y0 && yy < yend && yy < cover_y;
                // the clipped forms are needed only where a gradient is used
                const bool inner = !(grad_row && wx > 0 && !mk) ||
                                   (inner1(cc.cx, a.hx) && inner1(cc.cy, a.hy) && inner1(cc.cz, a.hz));
                g = __all(inner) ? eval_inner(cc, tt) : eval_taps(a.mov, cc, tt);
            } else {
                g = eval_rec(cc, tr);
            }
            // a uniform branch, not a select (the empty volatile asm keeps the
            // compiler from speculating the division): skipped at zRatio 1
            float gz = g.gz;
            if (a.z_ratio != 1.0f) {
                __asm__ volatile("");
                gz = __fdiv_rn(gz, a.z_ratio);
            }
            const float d = rc - g.w;
            if (s >= 2) {
                const float vsum = d0 + d1 + d;
                const float left = __shfl_up(vsum, 1, 64);
                const float right = __shfl_down(vsum, 1, 64);
                const int yo = yy - 1;
                float it = (left + vsum + right) * (1.0f / 9.0f);
                if (pmask) it = 0.0f;
                const float ie = owned ? it : 0.0f;
                err += static_cast<double>(ie) * static_cast<double>(ie);
                // row yo's weight: 0 outside [y0, yend) and the uncovered tail
                const float wyf = (yo >= y0 && yo < yend && yo < cover_y)
                                      ? static_cast<float>(1 + (yo == a.ny - 1 ? a.extra_y : 0))
                                      : 0.0f;
                const float m = pmask ? 0.0f : wxf * wyf;
                const float mgx = m * pgx, mgy = m * pgy, mgz = m * pgz;
                acc0 += mgx * pgx;
                acc1 += mgx * pgy;
                acc2 += mgx * pgz;
                acc3 += mgy * pgy;
                acc4 += mgy * pgz;
                acc5 += mgz * pgz;
                acc6 += mgx * it;
                acc7 += mgy * it;
                acc8 += mgz * it;
            }
            d0 = d1;
            d1 = d;
            pgx = g.gx;
            pgy = g.gy;
            pgz = gz;
            pmask = mk;
        }
    }

    // residual energy of the tile
    for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
    if (lane == 0) s_err[wave] = err;
    // per-column partial sums -> per-patch sums
    if (owned) {
        s_red[0][lcol] = acc0;
        s_red[1][lcol] = acc1;
        s_red[2][lcol] = acc2;
        s_red[3][lcol] = acc3;
        s_red[4][lcol] = acc4;
        s_red[5][lcol] = acc5;
        s_red[6][lcol] = acc6;
        s_red[7][lcol] = acc7;
        s_red[8][lcol] = acc8;
    }
    __syncthreads();
    if (tid == 0) {
        double e = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) e += s_err[w];
        a.errpart[logical] = e;
    }
    if (by < a.ncy) {
        const int ntask = a.np_x * 9;
        const int lastc = min(a.cols, a.nx - x0);   // tile columns inside the volume
        for (int task = tid; task < ntask; task += blockDim.x) {
            const int pi = task / 9, k = task - 9 * (task / 9);
            const int gi = bx * a.np_x + pi;
            if (gi >= a.ncx) continue;
            double sacc = 0.0;
            const int u0 = pi * a.p;
            const int u1 = min(u0 + a.p, lastc);
            for (int u = u0; u < u1; ++u) sacc += static_cast<double>(s_red[k][u]);
            // the gradients were accumulated as T(c + e) - T(c - e), twice
            // the reference's (exact power-of-two scalings)
            sacc *= k < 6 ? 0.25 : 0.5;
            a.sums[static_cast<size_t>(k) * a.ncp + (static_cast<size_t>(z) * a.ncy + by) * a.ncx + gi] =
                static_cast<float>(sacc);
        }
    }
}

PVol make_pvol(const void* base, unsigned es, int nx, int ny, int nz) {
    const OutLayout ol = padded_layout(nx, ny);
    PVol v;
    v.base = static_cast<const unsigned char*>(base);
    v.pitchB = static_cast<unsigned>(ol.pitch) * es;
    v.plane_rows = static_cast<unsigned>(ol.plane_rows);
    v.planeB = v.pitchB * v.plane_rows;
    v.origin = static_cast<unsigned>((ol.lead * ol.pitch + kPad) * es);
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    v.hx = static_cast<float>(nx - 1);
    v.hy = static_cast<float>(ny - 1);
    v.hz = static_cast<float>(nz - 1);
    return v;
}

// Records of the moving level image (QVol) from its padded float32 image
// (padded_layout): four consecutive records per thread, rows (y, z) of
// nx + 1 records.  The pads hold the replicated edges (V(-1) = V(0),
// V(n) = V(n + 1) = V(n - 1)), so every record, the repeated last one
// included, is a plain stencil of the padded image.
template <int S>
__global__ __launch_bounds__(256) void k_make_records(PVol v, unsigned char* __restrict__ out, unsigned rowB,
                                                      int xblocks) {
    const unsigned b = blockIdx.x;
    const unsigned row = b / static_cast<unsigned>(xblocks);
    const int x = (static_cast<int>(b - row * static_cast<unsigned>(xblocks)) * 256 + static_cast<int>(threadIdx.x)) * 4;
    if (x > v.nx) return;
    const int y = static_cast<int>(row % static_cast<unsigned>(v.ny + 1));
    const int z = static_cast<int>(row / static_cast<unsigned>(v.ny + 1));
    const unsigned char* c = v.base + v.origin + voxel_off<ST_F32>(v, x, y, z);
    // x - 1 .. x + 4 of the row (the padded row ends at nx + 3; x + 4 past it
    // reads the next stored row, only for records that are not stored)
    const f2v l = *reinterpret_cast<const f2v*>(c - 4);
    const f4v m = *reinterpret_cast<const f4v*>(c);
    const float r4 = *reinterpret_cast<const float*>(c + 16);
    const f4v ym = *reinterpret_cast<const f4v*>(c - v.pitchB), yp = *reinterpret_cast<const f4v*>(c + v.pitchB);
    const f4v zm = *reinterpret_cast<const f4v*>(c - v.planeB), zp = *reinterpret_cast<const f4v*>(c + v.planeB);
    const float xv[6] = {l.x, m.x, m.y, m.z, m.w, r4};
    const float ymv[4] = {ym.x, ym.y, ym.z, ym.w}, ypv[4] = {yp.x, yp.y, yp.z, yp.w};
    const float zmv[4] = {zm.x, zm.y, zm.z, zm.w}, zpv[4] = {zp.x, zp.y, zp.z, zp.w};
    unsigned char* o = out + static_cast<size_t>(row) * rowB + static_cast<size_t>(x) * kRec<S>;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (x + i > v.nx) break;
        const float q0 = xv[i + 1], q1 = xv[i + 2] - xv[i], q2 = ypv[i] - ymv[i], q3 = zpv[i] - zmv[i];
        if constexpr (S == ST_F32) {
            f4a r;
            r.x = q0;
            r.y = q1;
            r.z = q2;
            r.w = q3;
            reinterpret_cast<f4a*>(o)[i] = r;
        } else {
            auto h = [](float f) -> unsigned { return S == ST_BF16 ? to_bf16(f) : to_f16(f); };
            u2v r;
            r.x = h(q0) | (h(q1) << 16);
            r.y = h(q2) | (h(q3) << 16);
            reinterpret_cast<u2v*>(o)[i] = r;
        }
    }
}

// Records of a level that equals the input (level 0) straight from the
// dense uint16 / float32 input, indices clamped instead of padded.  A thread
// makes the 8 consecutive records x .. x + 7 of one row (y, z) of nx + 1
// records; the last thread of a row also the repeated record nx (or, where
// nx is not a multiple of 8, the records x .. nx).  The threads of a row are
// consecutive lanes, so the x neighbours x - 1 and x + 8 come from the
// adjacent lanes' loads (loaded only at a wave's edges); each row of eight
// inputs is one 16-byte load (two for float32) and each pair of 16-bit
// records one 16-byte store.  bfloat16 pairs are rounded by the packed
// conversion instruction (round to nearest even, as to_bf16 for finite
// values).
template <int IN_T>
__device__ __forceinline__ float in_at(const void* in, size_t i) {
    if constexpr (IN_T == E_U16) return static_cast<float>(static_cast<const uint16_t*>(in)[i]);
    else return static_cast<const float*>(in)[i];
}
template <int IN_T>
__device__ __forceinline__ void in_row8(const void* in, size_t i, float o[8]) {
    if constexpr (IN_T == E_U16) {
        const u4v w = *reinterpret_cast<const u4v*>(static_cast<const uint16_t*>(in) + i);
        const unsigned q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            o[2 * k] = static_cast<float>(q[k] & 0xffffu);
            o[2 * k + 1] = static_cast<float>(q[k] >> 16);
        }
    } else {
        const f4v a = *reinterpret_cast<const f4v*>(static_cast<const float*>(in) + i);
        const f4v b = *reinterpret_cast<const f4v*>(static_cast<const float*>(in) + i + 4);
        o[0] = a.x;
        o[1] = a.y;
        o[2] = a.z;
        o[3] = a.w;
        o[4] = b.x;
        o[5] = b.y;
        o[6] = b.z;
        o[7] = b.w;
    }
}
// Inputs x .. x + 7 of a row, clamped to the row's last voxel.
template <int IN_T>
__device__ __forceinline__ void in_row8_clamped(const void* in, size_t r, int x, int nx, bool full, float o[8]) {
    if (full) {
        in_row8<IN_T>(in, r + x, o);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = in_at<IN_T>(in, r + min(x + k, nx - 1));
    }
}

// Two 16-bit values of storage S in one word (lo in the low half).
template <int S>
__device__ __forceinline__ unsigned pack2(float lo, float hi) {
    if constexpr (S == ST_BF16) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(unsigned, __builtin_convertvector(f2{lo, hi}, b2));
    } else {
        return to_f16(lo) | (static_cast<unsigned>(to_f16(hi)) << 16);
    }
}

template <int S>
__device__ __forceinline__ void put_record(unsigned char* o, float q0, float q1, float q2, float q3) {
    if constexpr (S == ST_F32) {
        f4a r;
        r.x = q0;
        r.y = q1;
        r.z = q2;
        r.w = q3;
        *reinterpret_cast<f4a*>(o) = r;
    } else {
        u2v r;
        r.x = pack2<S>(q0, q1);
        r.y = pack2<S>(q2, q3);
        *reinterpret_cast<u2v*>(o) = r;
    }
}

template <int IN_T, int S>
__global__ __launch_bounds__(256) void k_make_records_in(const void* __restrict__ in, int nx, int ny, int nz,
                                                         unsigned char* __restrict__ out, unsigned rowB, unsigned tpr,
                                                         unsigned items) {
    const unsigned item = blockIdx.x * 256u + threadIdx.x;
    const bool live = item < items;
    // lanes past the end work on the last item (no stores) so that every lane
    // takes part in the shuffles
    const unsigned it = live ? item : items - 1u;
    const unsigned row = it / tpr;
    const unsigned t = it - row * tpr;
    const int x = static_cast<int>(t) * 8;
    const int y = static_cast<int>(row % static_cast<unsigned>(ny + 1));
    const int z = static_cast<int>(row / static_cast<unsigned>(ny + 1));
    const int yc = min(y, ny - 1), zc = min(z, nz - 1);
    const size_t plane = static_cast<size_t>(nx) * ny;
    const size_t r0 = (static_cast<size_t>(zc) * ny + yc) * nx;
    const size_t rym = r0 - (yc > 0 ? nx : 0), ryp = r0 + (yc < ny - 1 ? nx : 0);
    const size_t rzm = r0 - (zc > 0 ? plane : 0), rzp = r0 + (zc < nz - 1 ? plane : 0);
    const bool full = x + 8 <= nx;   // all eight inputs inside the row
    float xv[10], ym[8], yp[8], zm[8], zp[8];
    in_row8_clamped<IN_T>(in, r0, x, nx, full, xv + 1);
    in_row8_clamped<IN_T>(in, rym, x, nx, full, ym);
    in_row8_clamped<IN_T>(in, ryp, x, nx, full, yp);
    in_row8_clamped<IN_T>(in, rzm, x, nx, full, zm);
    in_row8_clamped<IN_T>(in, rzp, x, nx, full, zp);
    // x - 1 and x + 8 (clamped): the neighbouring lanes' end values where
    // they hold the same row, else loaded
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const float from_left = __shfl_up(xv[8], 1, 64);
    const float from_right = __shfl_down(xv[1], 1, 64);
    if (x == 0) xv[0] = xv[1];
    else if (lane > 0) xv[0] = from_left;
    else xv[0] = in_at<IN_T>(in, r0 + x - 1);
    if (x + 8 > nx - 1) xv[9] = xv[8];   // inputs past the row end repeat its last voxel
    else if (lane < 63) xv[9] = from_right;
    else xv[9] = in_at<IN_T>(in, r0 + x + 8);
    if (!live) return;
    unsigned char* o = out + static_cast<size_t>(row) * rowB + static_cast<size_t>(x) * kRec<S>;
    if (S != ST_F32 && full) {
        // eight records of voxels inside the row: four 16-byte stores
        u4v w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int a = 2 * q, b = 2 * q + 1;
            w[q].x = pack2<S>(xv[a + 1], xv[a + 2] - xv[a]);
            w[q].y = pack2<S>(yp[a] - ym[a], zp[a] - zm[a]);
            w[q].z = pack2<S>(xv[b + 1], xv[b + 2] - xv[b]);
            w[q].w = pack2<S>(yp[b] - ym[b], zp[b] - zm[b]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) reinterpret_cast<u4v*>(o)[q] = w[q];
    } else {
        // records x .. min(x + 7, nx); the one past the last voxel repeats it
        const int last = min(x + 8, nx + 1) - x;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i >= last) break;
            const int j = x + i < nx ? i : i - 1;
            put_record<S>(o + static_cast<size_t>(i) * kRec<S>, xv[j + 1], xv[j + 2] - xv[j], yp[j] - ym[j],
                          zp[j] - zm[j]);
        }
    }
    if (x + 8 == nx)   // the row's last thread of a multiple-of-8 row: record nx repeats nx - 1
        put_record<S>(o + static_cast<size_t>(8) * kRec<S>, xv[8], xv[9] - xv[7], yp[7] - ym[7], zp[7] - zm[7]);
}

QVol make_qvol(const void* base, int store, int nx, int ny, int nz) {
    QVol q;
    q.base = static_cast<const unsigned char*>(base);
    q.rowB = static_cast<unsigned>(nx + 1) * static_cast<unsigned>(4 * store_bytes(store));
    q.plane_rows = static_cast<unsigned>(ny + 1);
    q.hx = static_cast<float>(nx - 1);
    q.hy = static_cast<float>(ny - 1);
    q.hz = static_cast<float>(nz - 1);
    return q;
}

}  // namespace

size_t moving_records_bytes(int nx, int ny, int nz, int store) {
    return static_cast<size_t>(nx + 1) * (ny + 1) * (nz + 1) * 4 * store_bytes(store);
}

void launch_moving_records(const float* mov_padded, int nx, int ny, int nz, int store, void* out, hipStream_t s) {
    const PVol v = make_pvol(mov_padded, 4u, nx, ny, nz);
    const QVol q = make_qvol(out, store, nx, ny, nz);
    const int xblocks = (nx + 1 + 1023) / 1024;   // 4 records per thread
    const dim3 grid(static_cast<unsigned>(static_cast<long long>(ny + 1) * (nz + 1) * xblocks));
    unsigned char* o = static_cast<unsigned char*>(out);
    if (store == ST_F32) hipLaunchKernelGGL(k_make_records<ST_F32>, grid, dim3(256), 0, s, v, o, q.rowB, xblocks);
    else if (store == ST_BF16) hipLaunchKernelGGL(k_make_records<ST_BF16>, grid, dim3(256), 0, s, v, o, q.rowB, xblocks);
    else hipLaunchKernelGGL(k_make_records<ST_F16>, grid, dim3(256), 0, s, v, o, q.rowB, xblocks);
}

void launch_moving_records_input(const void* in, Elem t, int nx, int ny, int nz, int store, void* out,
                                 hipStream_t s) {
    const QVol q = make_qvol(out, store, nx, ny, nz);
    // 8 records per thread; rows (y, z) of nx + 1 records, y <= ny, z <= nz
    const unsigned tpr = static_cast<unsigned>((nx + 7) / 8);
    const unsigned items = static_cast<unsigned>(static_cast<long long>(ny + 1) * (nz + 1) * tpr);
    const dim3 grid((items + 255u) / 256u);
    unsigned char* o = static_cast<unsigned char*>(out);
    using Fn = void (*)(const void*, int, int, int, unsigned char*, unsigned, unsigned, unsigned);
    static const Fn table[2][3] = {
        {k_make_records_in<E_F32, ST_F32>, k_make_records_in<E_F32, ST_BF16>, k_make_records_in<E_F32, ST_F16>},
        {k_make_records_in<E_U16, ST_F32>, k_make_records_in<E_U16, ST_BF16>, k_make_records_in<E_U16, ST_F16>}};
    hipLaunchKernelGGL(table[t == E_U16 ? 1 : 0][store], grid, dim3(256), 0, s, in, nx, ny, nz, o, q.rowB, tpr,
                       items);
}

long long launch_lk_pass(const LevelGeom& g, const LkTile& t, const LkBuffers& b, int k, bool dense,
                         int check_parts, hipStream_t s) {
    PassArgs a;
    a.mov = make_pvol(b.mov, 4u, g.nx, g.ny, g.nz);
    a.q = make_qvol(b.movq, b.store, g.nx, g.ny, g.nz);
    a.hx = static_cast<float>(g.nx - 1);
    a.hy = static_cast<float>(g.ny - 1);
    a.hz = static_cast<float>(g.nz - 1);
    a.mmov = make_pvol(b.mmov ? static_cast<const void*>(b.mmov) : b.mov, 4u, g.nx, g.ny, g.nz);
    a.ref = b.ref;
    a.mref = b.mref;
    a.m0 = b.m0;
    a.cp = b.cp;
    a.sums = b.sums;
    a.errpart = b.errpart;
    a.ctl = b.ctl;
    a.dnei = b.dnei;
    a.penpart = b.penpart;
    a.k = k;
    a.check_parts = check_parts;
    a.mdpart = b.mdpart;
    a.log = b.log;
    a.counts = b.counts;
    a.nx = g.nx;
    a.ny = g.ny;
    a.nz = g.nz;
    a.r = g.r;
    a.p = g.p;
    a.ncx = g.ncx;
    a.ncy = g.ncy;
    a.extra_x = g.extra_x;
    a.extra_y = g.extra_y;
    a.cols = t.cols;
    a.rows = t.rows;
    a.nbx = t.nbx;
    a.nby = t.nby;
    a.np_x = t.np_x;
    a.blocks = t.blocks;
    a.ncp = static_cast<long long>(g.ncx) * g.ncy * g.nz;
    a.z_ratio = g.z_ratio;
    a.fs = b.fr.fs;
    const int threads = 64 * t.nwx;
    a.npen = (a.ncp + threads - 1) / threads;
    const dim3 grid(static_cast<unsigned>(t.blocks + a.npen), static_cast<unsigned>(b.fr.n));
    const bool mr = b.mref != nullptr, mm = b.mmov != nullptr;
    using Fn = void (*)(PassArgs);
#define WR_LK_TABLE(S)                                                                                   \
    {k_lk_pass<false, false, false, S>, k_lk_pass<false, false, true, S>, k_lk_pass<false, true, false, S>, \
     k_lk_pass<false, true, true, S>,   k_lk_pass<true, false, false, S>, k_lk_pass<true, false, true, S>,  \
     k_lk_pass<true, true, false, S>,   k_lk_pass<true, true, true, S>}
    static const Fn table[3][8] = {WR_LK_TABLE(ST_F32), WR_LK_TABLE(ST_BF16), WR_LK_TABLE(ST_F16)};
#undef WR_LK_TABLE
    const int sel = (dense ? 4 : 0) | (mr ? 2 : 0) | (mm ? 1 : 0);
    hipLaunchKernelGGL(table[b.store][sel], grid, dim3(threads), 0, s, a);
    return a.npen;
}

}  // namespace wr
