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
// The kernel is VALU-bound (a wave64 FP32 instruction per 4 clocks per SIMD;
// packed FP32 is not used), so the per-voxel instruction count is what is
// minimised: the moving image rows are padded by kPad replicated voxels at
// both x ends so every clamped tap row is one plain 16- or 8-byte load; row
// addresses are 32-bit byte offsets from a workgroup-uniform base; the CP
// interpolation weights of the tile's rows are formed once in LDS; the six
// gradient samples take a branch-free fast path away from the faces.

#include <cfloat>
#include <cmath>

#include "lk_common.h"
#include "lk_ctl.h"
#include "wr_internal.h"

namespace wr {

namespace {

constexpr int kMaxThreads = 256;
constexpr int kWaveCols = 62;   // tile columns owned by one wave (64 lanes - 2 halo)

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
typedef float f2v __attribute__((ext_vector_type(2), aligned(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2), aligned(4)));

// Bytes per stored voxel of a level image of storage S (Store).
template <int S>
constexpr unsigned kEs = S == ST_F32 ? 4u : 2u;

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

struct Grad {
    float w, gx, gy, gz;
};

// The 32 voxels around one deformed coordinate, as 4 x 16-byte rows holding
// X0..X3 at (Y1|Y2, Z1|Z2) and 8 x 8-byte rows holding X1, X2 for the y and
// z lines (see eval).
struct Taps {
    f4v q11, q21, q12, q22;
    f2v e01, e02, e31, e32, f10, f20, f13, f23;
};

struct Coord {
    float cx, cy, cz;
    Samp sx, sy, sz;
};

__device__ __forceinline__ Coord make_coord(const PVol& v, float cx, float cy, float cz) {
    Coord c;
    c.cx = cx;
    c.cy = cy;
    c.cz = cz;
    c.sx = split(cx, v.hx);
    c.sy = split(cy, v.hy);
    c.sz = split(cz, v.hz);
    return c;
}

// The 12 tap rows of the gather at c, as offsets from that of voxel
// (X0, Y0, Z0) = (fx - 1, fy - 1, fz - 1), which the pads keep inside the
// allocation.  Where the clamped reference would repeat an edge row or
// plane, the pads hold it.
template <int S>
__device__ __forceinline__ Taps gather(const PVol& v, const Coord& c) {
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
__device__ __forceinline__ Grad eval(const PVol& v, const Coord& c, const Taps& q) {
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
    o.gx = line_diff(g0, g1, g2, g3, sx.f, sx.t, c.cx, v.hx) * 0.5f;
    o.gy = line_diff(h0, h1, h2, h3, sy.f, sy.t, c.cy, v.hy) * 0.5f;
    o.gz = line_diff(k0, k1, k2, k3, sz.f, sz.t, c.cz, v.hz) * 0.5f;
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
    PVol mov;
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
    if (a.ctl->done) return;
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
        return make_coord(a.mov, xf + mx, static_cast<float>(yc) + my, zf + mz);
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
            const Taps tc = gather<S>(a.mov, cc);
            // template / mask row of this step: uniform base, per-lane column
            const int vrow = z * plane + clampi(y0 - 1 + s, 0, a.ny - 1) * a.nx;
            const float rc = ld_ref<S>(a.ref, vrow + xc);
            const int yy = y0 - 1 + s;
            const Grad g = eval(a.mov, cc, tc);
            const float gz = a.z_ratio == 1.0f ? g.gz : __fdiv_rn(g.gz, a.z_ratio);
            const float d = rc - g.w;
            bool mk = false;
            if (MREF) mk = (a.mref + vrow)[xc] != 0;
            if (MMOV) mk = mk || tri(a.mmov, cc.cx, cc.cy, cc.cz) > 0.0f;
            if (s >= 2) {
                const float vsum = d0 + d1 + d;
                const float left = __shfl_up(vsum, 1, 64);
                const float right = __shfl_down(vsum, 1, 64);
                const int yo = yy - 1;
                float it = (left + vsum + right) * (1.0f / 9.0f);
                if (pmask) it = 0.0f;
                const float ie = owned ? it : 0.0f;
                err += static_cast<double>(ie) * static_cast<double>(ie);
                const float wyf = yo < cover_y ? static_cast<float>(1 + (yo == a.ny - 1 ? a.extra_y : 0)) : 0.0f;
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
            // rows outside [y0, yend) and the uncovered tail carry no gradient
            const bool grad_row = yy >= y0 && yy < yend && yy < cover_y;
            pgx = grad_row ? g.gx : 0.0f;
            pgy = grad_row ? g.gy : 0.0f;
            pgz = grad_row ? gz : 0.0f;
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

}  // namespace

long long launch_lk_pass(const LevelGeom& g, const LkTile& t, const LkBuffers& b, int k, bool dense,
                         int check_parts, hipStream_t s) {
    PassArgs a;
    a.mov = make_pvol(b.mov, static_cast<unsigned>(store_bytes(b.store)), g.nx, g.ny, g.nz);
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
    const int threads = 64 * ((t.cols + kWaveCols - 1) / kWaveCols);
    a.npen = (a.ncp + threads - 1) / threads;
    const dim3 grid(static_cast<unsigned>(t.blocks + a.npen));
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
