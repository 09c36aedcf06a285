// Patch-wise Lucas-Kanade iteration of one pyramid level on gfx950.
//
// One iteration of the reference (calFlow3d_Wei_v1.py:442-654) becomes a
// fixed sequence of launches on one stream, with every stop decision taken
// on the device (no host round trip, so a whole getMotion can be enqueued or
// captured at once):
//
//   k_neidiff     per control point (CP): 3x3 neighbour difference of the CP
//                 field (calFlow3d:69-98,488-498) and the penalty partials.
//   k_lk_pass     per voxel, one fused pass: dense motion (CP bilinear,
//                 calFlow3d:265-292,627-640, or the level's initial field),
//                 trilinear warp (a5), mask warp (a6), residual with its 3x3
//                 mean and mask (a7), the six boundary-clipped gradient
//                 samples (a10), and the nine per-patch structure-tensor sums
//                 (a11) plus the residual energy partials (a9).
//   k_err_finish  one workgroup: error terms, log, the error stop rule
//                 (calFlow3d:515-530).
//   k_solve       per CP: regularised 3x3 adjugate solve, NaN -> 0, step
//                 clamp, z scaling, CP update (calFlow3d:209-262,596-624),
//                 max |delta| (a15).
//   k_dense_maxdiff  first iteration of a level only: max |delta| against
//                 the level's initial dense field.
//   k_check       one thread: the |delta| < 1e-3 stop test (calFlow3d:641-654).
//
// Every kernel returns at once when the level's `done` flag is set.
//
// Tiles of k_lk_pass: a workgroup owns `cols` = np_x*(2r+1) columns and
// `rows` = 2r+1 rows of one z slice, i.e. whole patches, so each per-patch
// sum is finished inside one workgroup and written once (no atomics, a
// deterministic order).  Threads own one column each and walk the rows; the
// 3x3 residual mean uses a register ring for the vertical sum and one LDS
// exchange per row for the horizontal one.  Two extra threads and two extra
// rows recompute the halo.  Workgroups are assigned XCD-major (each XCD gets
// a contiguous range of tiles) with z fastest, so the z-neighbour tiles that
// share moving-image slices run on the same XCD and hit its L2.

#include <cfloat>
#include <cmath>

#include "wr_internal.h"

namespace wr {

namespace {

constexpr int kMaxThreads = 256;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

// Bilinear CP interpolation of one axis: coordinate (v - r)/(2r+1) in float32
// (calFlow3d:281-282), clamped to [0, n-1] (clamp to [0, n] then
// map_coordinates' nearest mode).
struct Axis1 {
    int i0, i1;
    float t;
};
__device__ __forceinline__ Axis1 cp_axis(int v, int r, int p, int n) {
    float c = __fdiv_rn(static_cast<float>(v - r), static_cast<float>(p));
    c = clampf(c, 0.0f, static_cast<float>(n - 1));
    Axis1 a;
    a.i0 = static_cast<int>(floorf(c));
    a.t = c - static_cast<float>(a.i0);
    a.i1 = min(a.i0 + 1, n - 1);
    return a;
}

__device__ __forceinline__ float cp_bilinear(const float* __restrict__ plane, int ncx, const Axis1& ax,
                                             const Axis1& ay) {
    const float v00 = plane[ay.i0 * ncx + ax.i0];
    const float v10 = plane[ay.i0 * ncx + ax.i1];
    const float v01 = plane[ay.i1 * ncx + ax.i0];
    const float v11 = plane[ay.i1 * ncx + ax.i1];
    return lerpf(lerpf(v00, v10, ax.t), lerpf(v01, v11, ax.t), ay.t);
}

// Coordinate clamped to [0, n-1] (map_coordinates mode='nearest') split into
// floor and fraction.
struct Samp {
    int f;
    float t;
};
__device__ __forceinline__ Samp split(float c, int n) {
    const float u = clampf(c, 0.0f, static_cast<float>(n - 1));
    Samp s;
    s.f = static_cast<int>(floorf(u));
    s.t = u - static_cast<float>(s.f);
    return s;
}

__device__ __forceinline__ float tri(const float* __restrict__ vol, int nx, int ny, int nz, float cx, float cy,
                                     float cz) {
    const Samp sx = split(cx, nx), sy = split(cy, ny), sz = split(cz, nz);
    const int x1 = min(sx.f + 1, nx - 1), y1 = min(sy.f + 1, ny - 1), z1 = min(sz.f + 1, nz - 1);
    const size_t p0 = static_cast<size_t>(sz.f) * ny, p1 = static_cast<size_t>(z1) * ny;
    const float* r00 = vol + (p0 + sy.f) * nx;
    const float* r01 = vol + (p0 + y1) * nx;
    const float* r10 = vol + (p1 + sy.f) * nx;
    const float* r11 = vol + (p1 + y1) * nx;
    const float a = lerpf(lerpf(r00[sx.f], r00[x1], sx.t), lerpf(r01[sx.f], r01[x1], sx.t), sy.t);
    const float b = lerpf(lerpf(r10[sx.f], r10[x1], sx.t), lerpf(r11[sx.f], r11[x1], sx.t), sy.t);
    return lerpf(a, b, sz.t);
}

__device__ __forceinline__ float pick4(float a0, float a1, float a2, float a3, int o) {
    return o == 0 ? a0 : (o == 1 ? a1 : (o == 2 ? a2 : a3));
}

// Sample of the 4-point line g[0..3] = f(clamp(c0 - 1 + q)) at the clipped
// coordinate `c` (already inside [0, n-1]), where c0 is the floor of the
// clamped centre.  Whenever the fraction is non-zero the two taps lie inside
// the line (see the gradient notes below); with a zero fraction the second
// tap has weight 0.
__device__ __forceinline__ float line_sample(float g0, float g1, float g2, float g3, int c0, float c) {
    const int f = static_cast<int>(floorf(c));
    const float t = c - static_cast<float>(f);
    const int o = clampi(f - c0 + 1, 0, 3);
    const int o1 = min(o + 1, 3);
    return lerpf(pick4(g0, g1, g2, g3, o), pick4(g0, g1, g2, g3, o1), t);
}

struct Grad {
    float w, gx, gy, gz;
};

// Warp value and the central-difference gradients at deformed coordinate
// (cx, cy, cz) (calFlow3d:134-206): I_a = (T(c + e_a) - T(c - e_a)) / 2 where
// only the perturbed coordinate is clipped to [0, n_a - 1] and the other two
// follow map_coordinates' nearest mode.  The 32 distinct voxels of the six
// samples and the warp are loaded once: the 2x2x2 cube around the clamped
// centre and, per axis, the two planes one step further out.
//
// Why the 4-point lines suffice: with c0 = floor(clamp(c)), an unclipped
// c - 1 is exact in float32 and floors to c0 - 1 (offset 0); an unclipped
// c + 1 floors to c0 + 1 (offset 2) or, after rounding up to an integer, to
// c0 + 2 with a zero fraction; a clipped sample is an integer (zero
// fraction) whose clamped index is one of the four line points.
__device__ __forceinline__ Grad tri_grad(const float* __restrict__ vol, int nx, int ny, int nz, float cx,
                                         float cy, float cz) {
    const Samp sx = split(cx, nx), sy = split(cy, ny), sz = split(cz, nz);
    const int X0 = max(sx.f - 1, 0), X1 = sx.f, X2 = min(sx.f + 1, nx - 1), X3 = min(sx.f + 2, nx - 1);
    const int Y0 = max(sy.f - 1, 0), Y1 = sy.f, Y2 = min(sy.f + 1, ny - 1), Y3 = min(sy.f + 2, ny - 1);
    const int Z0 = max(sz.f - 1, 0), Z1 = sz.f, Z2 = min(sz.f + 1, nz - 1), Z3 = min(sz.f + 2, nz - 1);
    auto at = [&](int z, int y, int x) { return vol[(static_cast<size_t>(z) * ny + y) * nx + x]; };
    // central cube c[z][y][x] over {1,2}^3
    const float c111 = at(Z1, Y1, X1), c112 = at(Z1, Y1, X2), c121 = at(Z1, Y2, X1), c122 = at(Z1, Y2, X2);
    const float c211 = at(Z2, Y1, X1), c212 = at(Z2, Y1, X2), c221 = at(Z2, Y2, X1), c222 = at(Z2, Y2, X2);
    // x line: G[q] = bilinear in (y, z) of column Xq
    const float g1 = lerpf(lerpf(c111, c121, sy.t), lerpf(c211, c221, sy.t), sz.t);
    const float g2 = lerpf(lerpf(c112, c122, sy.t), lerpf(c212, c222, sy.t), sz.t);
    const float g0 = lerpf(lerpf(at(Z1, Y1, X0), at(Z1, Y2, X0), sy.t), lerpf(at(Z2, Y1, X0), at(Z2, Y2, X0), sy.t), sz.t);
    const float g3 = lerpf(lerpf(at(Z1, Y1, X3), at(Z1, Y2, X3), sy.t), lerpf(at(Z2, Y1, X3), at(Z2, Y2, X3), sy.t), sz.t);
    // y line: H[q] = bilinear in (x, z) of row Yq
    const float h1 = lerpf(lerpf(c111, c112, sx.t), lerpf(c211, c212, sx.t), sz.t);
    const float h2 = lerpf(lerpf(c121, c122, sx.t), lerpf(c221, c222, sx.t), sz.t);
    const float h0 = lerpf(lerpf(at(Z1, Y0, X1), at(Z1, Y0, X2), sx.t), lerpf(at(Z2, Y0, X1), at(Z2, Y0, X2), sx.t), sz.t);
    const float h3 = lerpf(lerpf(at(Z1, Y3, X1), at(Z1, Y3, X2), sx.t), lerpf(at(Z2, Y3, X1), at(Z2, Y3, X2), sx.t), sz.t);
    // z line: K[q] = bilinear in (x, y) of slice Zq
    const float k1 = lerpf(lerpf(c111, c112, sx.t), lerpf(c121, c122, sx.t), sy.t);
    const float k2 = lerpf(lerpf(c211, c212, sx.t), lerpf(c221, c222, sx.t), sy.t);
    const float k0 = lerpf(lerpf(at(Z0, Y1, X1), at(Z0, Y1, X2), sx.t), lerpf(at(Z0, Y2, X1), at(Z0, Y2, X2), sx.t), sy.t);
    const float k3 = lerpf(lerpf(at(Z3, Y1, X1), at(Z3, Y1, X2), sx.t), lerpf(at(Z3, Y2, X1), at(Z3, Y2, X2), sx.t), sy.t);

    Grad o;
    o.w = lerpf(k1, k2, sz.t);
    const float fx = static_cast<float>(nx - 1), fy = static_cast<float>(ny - 1), fz = static_cast<float>(nz - 1);
    o.gx = (line_sample(g0, g1, g2, g3, sx.f, clampf(cx + 1.0f, 0.0f, fx)) -
            line_sample(g0, g1, g2, g3, sx.f, clampf(cx - 1.0f, 0.0f, fx))) * 0.5f;
    o.gy = (line_sample(h0, h1, h2, h3, sy.f, clampf(cy + 1.0f, 0.0f, fy)) -
            line_sample(h0, h1, h2, h3, sy.f, clampf(cy - 1.0f, 0.0f, fy))) * 0.5f;
    o.gz = (line_sample(k0, k1, k2, k3, sz.f, clampf(cz + 1.0f, 0.0f, fz)) -
            line_sample(k0, k1, k2, k3, sz.f, clampf(cz - 1.0f, 0.0f, fz))) * 0.5f;
    return o;
}

struct PassArgs {
    const float* mov;
    const float* ref;
    const uint8_t* mref;
    const float* mmov;
    const float* m0;
    const float* cp;
    float* sums;
    double* errpart;
    const Ctl* ctl;
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

template <bool DENSE, bool MREF, bool MMOV>
__global__ __launch_bounds__(kMaxThreads) void k_lk_pass(PassArgs a) {
    if (a.ctl->done) return;
    __shared__ float s_v[2][kMaxThreads];
    __shared__ float s_red[9][kMaxThreads];
    __shared__ double s_err[kMaxThreads / 64];

    int bx, by, z;
    long long logical;
    tile_of(a.blocks, a.nz, a.nbx, bx, by, z, logical);
    const int tid = threadIdx.x;
    const int x0 = bx * a.cols, y0 = by * a.rows;
    const int col = x0 - 1 + tid;
    const bool active = tid < a.cols + 2;
    const int xc = clampi(col, 0, a.nx - 1);
    const bool owned = tid >= 1 && tid <= a.cols && col < a.nx;
    const int cover_x = a.ncx * a.p, cover_y = a.ncy * a.p;
    const int wx = (owned && col < cover_x) ? 1 + (col == a.nx - 1 ? a.extra_x : 0) : 0;

    const size_t plane = static_cast<size_t>(a.nx) * a.ny;
    const size_t nvox = plane * a.nz;
    const float* cpx = a.cp + static_cast<size_t>(z) * a.ncy * a.ncx;
    const float* cpy = cpx + a.ncp;
    const float* cpz = cpy + a.ncp;
    Axis1 axx{0, 0, 0.0f};
    if (!DENSE) axx = cp_axis(xc, a.r, a.p, a.ncx);

    float d0 = 0.0f, d1 = 0.0f;        // residual of the two previous rows
    float pgx = 0.0f, pgy = 0.0f, pgz = 0.0f;
    bool pmask = false;
    float acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0f;
    double err = 0.0;

    const int yend = min(y0 + a.rows, a.ny);
    const int nsteps = (yend - y0) + 2;
    for (int s = 0; s < nsteps; ++s) {
        const int yy = y0 - 1 + s;
        const int yc = clampi(yy, 0, a.ny - 1);
        const bool core = yy >= y0 && yy < yend;
        const bool need_grad = core && wx > 0 && yy < cover_y;
        float d = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
        bool mk = false;
        if (active) {
            const size_t v = static_cast<size_t>(z) * plane + static_cast<size_t>(yc) * a.nx + xc;
            float mx, my, mz;
            if (DENSE) {
                mx = a.m0[v];
                my = a.m0[v + nvox];
                mz = a.m0[v + 2 * nvox];
            } else {
                const Axis1 ayy = cp_axis(yc, a.r, a.p, a.ncy);
                mx = cp_bilinear(cpx, a.ncx, axx, ayy);
                my = cp_bilinear(cpy, a.ncx, axx, ayy);
                mz = cp_bilinear(cpz, a.ncx, axx, ayy);
            }
            const float cx = static_cast<float>(xc) + mx;
            const float cy = static_cast<float>(yc) + my;
            const float cz = static_cast<float>(z) + mz;
            float w;
            if (need_grad) {
                const Grad g = tri_grad(a.mov, a.nx, a.ny, a.nz, cx, cy, cz);
                w = g.w;
                gx = g.gx;
                gy = g.gy;
                gz = __fdiv_rn(g.gz, a.z_ratio);
            } else {
                w = tri(a.mov, a.nx, a.ny, a.nz, cx, cy, cz);
            }
            d = a.ref[v] - w;
            if (MREF) mk = a.mref[v] != 0;
            if (MMOV) mk = mk || tri(a.mmov, a.nx, a.ny, a.nz, cx, cy, cz) > 0.0f;
        }
        if (s >= 2) {
            const float vsum = d0 + d1 + d;
            float* buf = s_v[s & 1];
            buf[tid] = vsum;
            __syncthreads();
            if (owned) {
                const int yo = yy - 1;
                float it = (buf[tid - 1] + vsum + buf[tid + 1]) / 9.0f;
                if (pmask) it = 0.0f;
                err += static_cast<double>(it) * static_cast<double>(it);
                const int wy = yo < cover_y ? 1 + (yo == a.ny - 1 ? a.extra_y : 0) : 0;
                if (wx > 0 && wy > 0 && !pmask) {
                    const float m = static_cast<float>(wx * wy);
                    acc[0] += m * (pgx * pgx);
                    acc[1] += m * (pgx * pgy);
                    acc[2] += m * (pgx * pgz);
                    acc[3] += m * (pgy * pgy);
                    acc[4] += m * (pgy * pgz);
                    acc[5] += m * (pgz * pgz);
                    acc[6] += m * (pgx * it);
                    acc[7] += m * (pgy * it);
                    acc[8] += m * (pgz * it);
                }
            }
        }
        d0 = d1;
        d1 = d;
        pgx = gx;
        pgy = gy;
        pgz = gz;
        pmask = mk;
    }

    // residual energy of the tile
    for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
    if ((tid & 63) == 0) s_err[tid >> 6] = err;
    // per-column partial sums -> per-patch sums
#pragma unroll
    for (int k = 0; k < 9; ++k) s_red[k][tid] = acc[k];
    __syncthreads();
    if (tid == 0) {
        double e = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) e += s_err[w];
        a.errpart[logical] = e;
    }
    if (by < a.ncy) {
        const int ntask = a.np_x * 9;
        for (int task = tid; task < ntask; task += blockDim.x) {
            const int pi = task / 9, k = task - 9 * (task / 9);
            const int gi = bx * a.np_x + pi;
            if (gi >= a.ncx) continue;
            double sacc = 0.0;
            const int u0 = 1 + pi * a.p;
            for (int u = 0; u < a.p; ++u) sacc += static_cast<double>(s_red[k][u0 + u]);
            a.sums[static_cast<size_t>(k) * a.ncp + (static_cast<size_t>(z) * a.ncy + by) * a.ncx + gi] =
                static_cast<float>(sacc);
        }
    }
}

// ---------------------------------------------------------------- CP kernels

__global__ __launch_bounds__(256) void k_ctl_reset(Ctl* c, int iter, int frec, int fupd, double tol,
                                                    double s_pen, double n_vox) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->done = 0;
    c->rec = 0;
    c->upd = 0;
    c->nan = 0;
    c->iter = iter;
    c->frec = frec;
    c->fupd = fupd;
    c->maxdiff = 0u;
    c->old_err[0] = c->old_err[1] = c->old_err[2] = INFINITY;
    c->tol = tol;
    c->s_pen = s_pen;
    c->inv_n = n_vox;   // holds N; the error terms divide by it
}

__global__ __launch_bounds__(256) void k_cp_init(const float* __restrict__ m0, float* __restrict__ cp, int nx,
                                                  int ny, int nz, int r, int p, int ncx, int ncy) {
    const long long ncp = static_cast<long long>(ncx) * ncy * nz;
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= ncp) return;
    const int ix = static_cast<int>(i % ncx);
    const int iy = static_cast<int>((i / ncx) % ncy);
    const int z = static_cast<int>(i / (static_cast<long long>(ncx) * ncy));
    const size_t nvox = static_cast<size_t>(nx) * ny * nz;
    const size_t v = (static_cast<size_t>(z) * ny + (r + iy * p)) * nx + (r + ix * p);
    cp[i] = m0[v];
    cp[i + ncp] = m0[v + nvox];
    cp[i + 2 * ncp] = m0[v + 2 * nvox];
}

// 3x3 neighbour difference on the CP grid, replicated edges, float64
// accumulation in the filter's C order then rounded to float32 (scipy
// correlate on a float32 array), z component times zRatio_L in float32.
__global__ __launch_bounds__(256) void k_neidiff(const Ctl* __restrict__ ctl, const float* __restrict__ cp,
                                                  float* __restrict__ dn, double* __restrict__ penpart,
                                                  int ncx, int ncy, int nz, float z_ratio) {
    if (ctl->done) return;
    __shared__ double s_w[4];
    const long long ncp = static_cast<long long>(ncx) * ncy * nz;
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    double q = 0.0;
    if (i < ncp) {
        const int ix = static_cast<int>(i % ncx);
        const int iy = static_cast<int>((i / ncx) % ncy);
        const long long zb = i - (static_cast<long long>(iy) * ncx + ix);
        float dd[3];
        for (int c = 0; c < 3; ++c) {
            const float* f = cp + c * ncp + zb;
            double t = 0.0;
            for (int a = -1; a <= 1; ++a) {
                const int xi = clampi(ix + a, 0, ncx - 1);
                for (int b = -1; b <= 1; ++b) {
                    const int yi = clampi(iy + b, 0, ncy - 1);
                    const double wgt = (a == 0 && b == 0) ? -1.0 : 0.125;
                    t = __dadd_rn(t, __dmul_rn(wgt, static_cast<double>(f[yi * ncx + xi])));
                }
            }
            dd[c] = static_cast<float>(t);
        }
        dd[2] = __fmul_rn(dd[2], z_ratio);
        dn[i] = dd[0];
        dn[i + ncp] = dd[1];
        dn[i + 2 * ncp] = dd[2];
        const float qq = __fadd_rn(__fadd_rn(__fmul_rn(dd[0], dd[0]), __fmul_rn(dd[1], dd[1])), __fmul_rn(dd[2], dd[2]));
        q = static_cast<double>(qq);
    }
    for (int off = 32; off > 0; off >>= 1) q += __shfl_down(q, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = q;
    __syncthreads();
    if (threadIdx.x == 0) penpart[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// Error terms, log and the error-based stop rule of iteration k.
__global__ __launch_bounds__(1024) void k_err_finish(Ctl* __restrict__ c, const double* __restrict__ errpart,
                                                      long long nerr, const double* __restrict__ penpart,
                                                      long long npen, double* __restrict__ log,
                                                      int* __restrict__ counts) {
    if (c->done) return;
    __shared__ double s_e[16], s_p[16];
    double e = 0.0, pp = 0.0;
    for (long long i = threadIdx.x; i < nerr; i += blockDim.x) e += errpart[i];
    for (long long i = threadIdx.x; i < npen; i += blockDim.x) pp += penpart[i];
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_down(e, off, 64);
        pp += __shfl_down(pp, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_e[threadIdx.x >> 6] = e;
        s_p[threadIdx.x >> 6] = pp;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double esum = 0.0, psum = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
        esum += s_e[w];
        psum += s_p[w];
    }
    const double nvox = c->inv_n;
    const double diff = esum / nvox;
    const double pen = (psum * c->s_pen) / nvox;
    const double cur = diff + pen;
    const int k = c->rec;
    double* row = log + static_cast<size_t>(k) * 4;
    row[0] = diff;
    row[1] = pen;
    row[2] = cur;
    c->rec = k + 1;
    counts[0] = k + 1;
    c->last_err = cur;
    bool stop;
    if (c->frec >= 0) {
        stop = (k == c->frec - 1) && (c->fupd == c->frec - 1);
    } else if (k == c->iter - 1) {
        stop = true;
    } else {
        const int worse = (c->old_err[0] <= cur) + (c->old_err[1] <= cur) + (c->old_err[2] <= cur);
        if (worse > 1) {
            stop = true;
        } else if (fabs(c->old_err[2] - cur) < c->tol) {
            stop = true;
        } else {
            stop = false;
            c->old_err[0] = c->old_err[1];
            c->old_err[1] = c->old_err[2];
            c->old_err[2] = cur;
        }
    }
    c->maxdiff = 0u;
    if (stop) c->done = 1;
}

// Regularised 3x3 solve per CP and the CP update (float64 solve).
__global__ __launch_bounds__(256) void k_solve(Ctl* __restrict__ c, const float* __restrict__ sums,
                                                const float* __restrict__ dn, float* __restrict__ cp,
                                                long long ncp, double mov_range, double z_ratio) {
    if (c->done) return;
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    float md = 0.0f;
    if (i < ncp) {
        const double s_pen = c->s_pen;
        const double xx = static_cast<double>(sums[i]) + s_pen;
        const double xy = sums[i + ncp];
        const double xz = sums[i + 2 * ncp];
        const double yy = static_cast<double>(sums[i + 3 * ncp]) + s_pen;
        const double yz = sums[i + 4 * ncp];
        const double zz = static_cast<double>(sums[i + 5 * ncp]) + s_pen;
        const double xt = static_cast<double>(sums[i + 6 * ncp]) + s_pen * static_cast<double>(dn[i]);
        const double yt = static_cast<double>(sums[i + 7 * ncp]) + s_pen * static_cast<double>(dn[i + ncp]);
        const double zt = static_cast<double>(sums[i + 8 * ncp]) + s_pen * static_cast<double>(dn[i + 2 * ncp]);
        const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
        const double m11 = yy * zz - yz * yz;
        const double m12 = -(xy * zz - yz * xz);
        const double m13 = xy * yz - yy * xz;
        const double m22 = xx * zz - xz * xz;
        const double m23 = -(xx * yz - xy * xz);
        const double m33 = xx * yy - xy * xy;
        double v[3];
        v[0] = (m11 * xt + m12 * yt + m13 * zt) / det;
        v[1] = (m12 * xt + m22 * yt + m23 * zt) / det;
        v[2] = (m13 * xt + m23 * yt + m33 * zt) / det;
        int bad = 0;
        for (int q = 0; q < 3; ++q)
            if (isnan(v[q])) {
                v[q] = 0.0;
                ++bad;
            }
        if (bad) atomicAdd(&c->nan, bad);
        double dist = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        dist = fmax(dist / mov_range, 1.0);
        v[0] /= dist;
        v[1] /= dist;
        v[2] /= dist;
        v[2] /= z_ratio;
        for (int q = 0; q < 3; ++q) {
            const float old = cp[i + q * ncp];
            const float nw = static_cast<float>(static_cast<double>(old) + v[q]);
            cp[i + q * ncp] = nw;
            md = fmaxf(md, fabsf(nw - old));
        }
    }
    for (int off = 32; off > 0; off >>= 1) md = fmaxf(md, __shfl_down(md, off, 64));
    if ((threadIdx.x & 63) == 0 && md > 0.0f) atomicMax(&c->maxdiff, __float_as_uint(md));
}

// First update of a level: max |m_1 - m_0| over every voxel (m_0 is the
// level's initial dense field, not a CP interpolant).
__global__ __launch_bounds__(256) void k_dense_maxdiff(Ctl* __restrict__ c, const float* __restrict__ m0,
                                                        const float* __restrict__ cp, int nx, int ny, int nz,
                                                        int r, int p, int ncx, int ncy) {
    if (c->done) return;
    const size_t plane = static_cast<size_t>(nx) * ny;
    const size_t nvox = plane * nz;
    const size_t ncp = static_cast<size_t>(ncx) * ncy * nz;
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    float md = 0.0f;
    if (v < nvox) {
        const int x = static_cast<int>(v % nx);
        const int y = static_cast<int>((v / nx) % ny);
        const int z = static_cast<int>(v / plane);
        const Axis1 ax = cp_axis(x, r, p, ncx), ay = cp_axis(y, r, p, ncy);
        const float* base = cp + static_cast<size_t>(z) * ncy * ncx;
        for (int q = 0; q < 3; ++q) {
            const float m1 = cp_bilinear(base + q * ncp, ncx, ax, ay);
            md = fmaxf(md, fabsf(m1 - m0[v + q * nvox]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) md = fmaxf(md, __shfl_down(md, off, 64));
    if ((threadIdx.x & 63) == 0 && md > 0.0f) atomicMax(&c->maxdiff, __float_as_uint(md));
}

__global__ void k_check(Ctl* __restrict__ c, double* __restrict__ log, int* __restrict__ counts) {
    if (threadIdx.x != 0 || c->done) return;
    const int k = c->rec - 1;
    const float md = __uint_as_float(c->maxdiff);
    log[static_cast<size_t>(k) * 4 + 3] = static_cast<double>(md);
    c->upd += 1;
    counts[1] = c->upd;
    counts[2] = c->nan;
    if (c->frec >= 0) {
        if (k == c->frec - 1) c->done = 1;
    } else if (md < 1e-3f) {
        c->done = 1;
    }
}

// Dense field at the end of a level: the CP interpolant when the level made
// at least one update, else its initial field.  Writes the 3-plane layout
// (for the next level's upsample) and/or the interleaved (Z,Y,X,3) layout
// (the returned motion), and optionally motion + grid.
__global__ __launch_bounds__(256) void k_level_end(const Ctl* __restrict__ c, const float* __restrict__ m0,
                                                    const float* __restrict__ cp, int nx, int ny, int nz, int r,
                                                    int p, int ncx, int ncy, float* __restrict__ soa,
                                                    float* __restrict__ aos, float* __restrict__ coords,
                                                    int* __restrict__ counts) {
    const size_t plane = static_cast<size_t>(nx) * ny;
    const size_t nvox = plane * nz;
    const size_t ncp = static_cast<size_t>(ncx) * ncy * nz;
    const size_t v = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v == 0) counts[2] = c->nan;
    if (v >= nvox) return;
    const int x = static_cast<int>(v % nx);
    const int y = static_cast<int>((v / nx) % ny);
    const int z = static_cast<int>(v / plane);
    float m[3];
    if (c->upd > 0) {
        const Axis1 ax = cp_axis(x, r, p, ncx), ay = cp_axis(y, r, p, ncy);
        const float* base = cp + static_cast<size_t>(z) * ncy * ncx;
        for (int q = 0; q < 3; ++q) m[q] = cp_bilinear(base + q * ncp, ncx, ax, ay);
    } else {
        for (int q = 0; q < 3; ++q) m[q] = m0[v + q * nvox];
    }
    if (soa)
        for (int q = 0; q < 3; ++q) soa[v + q * nvox] = m[q];
    if (aos)
        for (int q = 0; q < 3; ++q) aos[3 * v + q] = m[q];
    if (coords) {
        coords[3 * v + 0] = static_cast<float>(x) + m[0];
        coords[3 * v + 1] = static_cast<float>(y) + m[1];
        coords[3 * v + 2] = static_cast<float>(z) + m[2];
    }
}

inline unsigned blocks_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------- host side

LkTile plan_tile(const LevelGeom& g, int np_x, int nb_y) {
    LkTile t{};
    const int maxp = std::max(1, (kMaxThreads - 2) / g.p);
    int npx;
    if (np_x > 0) {
        npx = std::min(np_x, maxp);
    } else {
        // the widest tile that still gives >= 2048 workgroups (8 per CU),
        // else the narrowest one
        npx = 1;
        for (int c = std::min(maxp, std::max(g.ncx, 1)); c >= 1; --c) {
            const long long nbx = (g.nx + c * g.p - 1) / (c * g.p);
            const long long nby = (g.ny + g.p - 1) / g.p;
            if (nbx * nby * g.nz >= 2048) {
                npx = c;
                break;
            }
        }
    }
    (void)nb_y;
    t.np_x = npx;
    t.nb_y = 1;
    t.cols = npx * g.p;
    t.rows = g.p;
    t.nbx = (g.nx + t.cols - 1) / t.cols;
    t.nby = (g.ny + t.rows - 1) / t.rows;
    t.blocks = static_cast<long long>(t.nbx) * t.nby * g.nz;
    return t;
}

void launch_ctl_reset(Ctl* ctl, int iter, int frec, int fupd, double tol, double s_pen, double n_vox,
                      hipStream_t s) {
    k_ctl_reset<<<1, 64, 0, s>>>(ctl, iter, frec, fupd, tol, s_pen, n_vox);
}

void launch_cp_init(const LevelGeom& g, const float* m0, float* cp, hipStream_t s) {
    const long long ncp = static_cast<long long>(g.ncx) * g.ncy * g.nz;
    k_cp_init<<<blocks_for(ncp, 256), 256, 0, s>>>(m0, cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx, g.ncy);
}

int launch_iteration(const LevelGeom& g, const LkTile& t, const LkBuffers& b, int level, int k, int iter,
                     hipStream_t s) {
    (void)level;
    const long long ncp = static_cast<long long>(g.ncx) * g.ncy * g.nz;
    const unsigned cpb = blocks_for(ncp, 256);
    double* log_row = b.log;   // caller offsets to this level
    k_neidiff<<<cpb, 256, 0, s>>>(b.ctl, b.cp, b.dnei, b.penpart, g.ncx, g.ncy, g.nz, g.z_ratio);

    PassArgs a;
    a.mov = b.mov;
    a.ref = b.ref;
    a.mref = b.mref;
    a.mmov = b.mmov;
    a.m0 = b.m0;
    a.cp = b.cp;
    a.sums = b.sums;
    a.errpart = b.errpart;
    a.ctl = b.ctl;
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
    a.ncp = ncp;
    a.z_ratio = g.z_ratio;
    const int threads = std::min(kMaxThreads, ((t.cols + 2 + 63) / 64) * 64);
    const dim3 grid(static_cast<unsigned>(t.blocks));
    const bool mr = b.mref != nullptr, mm = b.mmov != nullptr;
    if (b.events) (void)hipEventRecord(b.events[2 * k], s);
    if (k == 0) {
        if (mr && mm) k_lk_pass<true, true, true><<<grid, threads, 0, s>>>(a);
        else if (mr) k_lk_pass<true, true, false><<<grid, threads, 0, s>>>(a);
        else if (mm) k_lk_pass<true, false, true><<<grid, threads, 0, s>>>(a);
        else k_lk_pass<true, false, false><<<grid, threads, 0, s>>>(a);
    } else {
        if (mr && mm) k_lk_pass<false, true, true><<<grid, threads, 0, s>>>(a);
        else if (mr) k_lk_pass<false, true, false><<<grid, threads, 0, s>>>(a);
        else if (mm) k_lk_pass<false, false, true><<<grid, threads, 0, s>>>(a);
        else k_lk_pass<false, false, false><<<grid, threads, 0, s>>>(a);
    }
    if (b.events) (void)hipEventRecord(b.events[2 * k + 1], s);
    k_err_finish<<<1, 1024, 0, s>>>(b.ctl, b.errpart, t.blocks, b.penpart, cpb, log_row + static_cast<size_t>(0),
                                    b.counts);
    k_solve<<<cpb, 256, 0, s>>>(b.ctl, b.sums, b.dnei, b.cp, ncp, g.mov_range, g.z_ratio_d);
    if (k == 0) {
        const size_t nvox = static_cast<size_t>(g.nx) * g.ny * g.nz;
        k_dense_maxdiff<<<blocks_for(static_cast<long long>(nvox), 256), 256, 0, s>>>(
            b.ctl, b.m0, b.cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx, g.ncy);
    }
    k_check<<<1, 64, 0, s>>>(b.ctl, log_row, b.counts);
    (void)iter;
    return WR_OK;
}

void launch_level_end(const LevelGeom& g, const LkBuffers& b, int level, int iter, float* dense_soa,
                      float* motion_aos, float* coords_aos, hipStream_t s) {
    (void)level;
    (void)iter;
    const long long nvox = static_cast<long long>(g.nx) * g.ny * g.nz;
    k_level_end<<<blocks_for(nvox, 256), 256, 0, s>>>(b.ctl, b.m0, b.cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx,
                                                      g.ncy, dense_soa, motion_aos, coords_aos, b.counts);
}

}  // namespace wr
