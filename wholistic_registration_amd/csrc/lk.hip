// Patch-wise Lucas-Kanade iteration of one pyramid level on gfx950.
//
// One iteration of the reference (calFlow3d_Wei_v1.py:442-654) becomes a
// fixed sequence of launches on one stream, with every stop decision taken
// on the device (no host round trip, so a whole getMotion can be enqueued or
// captured at once):
//
//   k_neidiff     per control point (CP): 3x3 neighbour difference of the CP
//                 field (calFlow3d:69-98,488-498) and the penalty partials.
//   k_lk_pass     per voxel, one fused pass (lk_pass.hip): dense motion,
//                 warp, mask, residual and its 3x3 mean, the six gradient
//                 samples, the nine per-patch structure-tensor sums and the
//                 residual energy partials.
//   k_err_finish  one workgroup: error terms, log, the error stop rule
//                 (calFlow3d:515-530).
//   k_solve       per CP: regularised 3x3 adjugate solve, NaN -> 0, step
//                 clamp, z scaling, CP update (calFlow3d:209-262,596-624),
//                 per-workgroup max |delta| (a15).
//   k_dense_maxdiff  first iteration of a level only: max |delta| against
//                 the level's initial dense field.
//   k_check       one workgroup: the |delta| < 1e-3 stop test
//                 (calFlow3d:641-654).
//
// Every kernel returns at once when the level's `done` flag is set.  All
// reductions combine per-workgroup partials in a fixed order, so results do
// not depend on scheduling.

#include <cfloat>
#include <algorithm>
#include <cmath>

#include "lk_common.h"
#include "lk_ctl.h"
#include "wr_internal.h"

namespace wr {

namespace {


// ---------------------------------------------------------------- CP kernels

__global__ __launch_bounds__(256) void k_ctl_reset(Ctl* c, int iter, int frec, int fupd, double tol,
                                                    double s_pen, double n_vox, long long fs, unsigned* hflags, unsigned gen) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c = at_frame(c, fs);
    c->done = 0;
    c->hflag = hflags ? hflags + 2 * blockIdx.y : nullptr;
    c->gen = gen;
    c->rec = 0;
    c->upd = 0;
    c->nan = 0;
    c->solve64 = 0;
    c->iter = iter;
    c->frec = frec;
    c->fupd = fupd;
    c->maxdiff = 0u;
    for (int q = 0; q < 3; ++q) c->old_err[0][q] = c->old_err[1][q] = INFINITY;
    c->tol = tol;
    c->s_pen = s_pen;
    c->inv_n = n_vox;   // holds N; the error terms divide by it
}

__global__ __launch_bounds__(256) void k_cp_init(const float* __restrict__ m0, float* __restrict__ cp, int nx,
                                                  int ny, int nz, int r, int p, int ncx, int ncy, long long fs) {
    m0 = at_frame(m0, fs);
    cp = at_frame(cp, fs);
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

__device__ __forceinline__ void block_max_to(float md, float* __restrict__ dst) {
    __shared__ float s_m[4];
    for (int off = 32; off > 0; off >>= 1) md = fmaxf(md, __shfl_down(md, off, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = md;
    __syncthreads();
    if (threadIdx.x == 0) *dst = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
}

// Regularised 3x3 solve per CP and the CP update.  The solve is float64
// (the reference's arithmetic); with solve32 (the 16-bit storage variants) it
// is float32, and float64 only where the system's condition number exceeds
// 1e6 (estimated in the Frobenius norm, kappa_F = |A|_F |adj A|_F / |det A|,
// an upper bound of the 2-norm condition number; a zero or non-finite
// determinant also takes the float64 path).
// With fuse_err, every workgroup first evaluates iteration k's error terms and
// error-based stop rule from the LK launch's partials (the same result in
// all of them; workgroup 0 records it) instead of a separate k_err_finish.
constexpr float kSolve32MaxCond = 1e6f;

__global__ __launch_bounds__(256) void k_solve(Ctl* __restrict__ c, const float* __restrict__ sums,
                                                const float* __restrict__ dn, float* __restrict__ cp,
                                                long long ncp, double mov_range, double z_ratio, int solve32,
                                                float* __restrict__ mdpart, int k, int fuse_err,
                                                const double* __restrict__ errpart, long long nerr,
                                                const double* __restrict__ penpart, long long npen,
                                                double* __restrict__ log, int* __restrict__ counts, long long fs) {
    c = at_frame(c, fs);
    if (c->done) return;
    // the LK pass of iteration k has ended: progress for a watching host
    if (c->hflag && blockIdx.x == 0 && threadIdx.x == 0) host_word(c->hflag, (c->gen << 10) | (k + 1));
    sums = at_frame(sums, fs);
    dn = at_frame(dn, fs);
    cp = at_frame(cp, fs);
    mdpart = at_frame(mdpart, fs);
    errpart = at_frame(errpart, fs);
    penpart = at_frame(penpart, fs);
    log = at_frame(log, fs);
    counts = at_frame(counts, fs);
    if (fuse_err) {
        __shared__ int s_stop;
        double diff, pen;
        err_terms(c, errpart, nerr, penpart, npen, diff, pen);
        if (threadIdx.x == 0) {
            const double cur = diff + pen;
            const bool stop = err_stop(c, k, cur);
            if (blockIdx.x == 0) err_record(c, k, diff, pen, cur, stop, log, counts);
            s_stop = stop ? 1 : 0;
        }
        __syncthreads();
        if (s_stop) return;
    }
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    float md = 0.0f;
    if (i < ncp) {
        const double s_pen = c->s_pen;
        double v[3];
        bool solved = false;
        if (solve32) {
            const float sp = static_cast<float>(s_pen);
            const float xx = sums[i] + sp, xy = sums[i + ncp], xz = sums[i + 2 * ncp];
            const float yy = sums[i + 3 * ncp] + sp, yz = sums[i + 4 * ncp], zz = sums[i + 5 * ncp] + sp;
            const float xt = sums[i + 6 * ncp] + sp * dn[i];
            const float yt = sums[i + 7 * ncp] + sp * dn[i + ncp];
            const float zt = sums[i + 8 * ncp] + sp * dn[i + 2 * ncp];
            const float m11 = yy * zz - yz * yz, m12 = -(xy * zz - yz * xz), m13 = xy * yz - yy * xz;
            const float m22 = xx * zz - xz * xz, m23 = -(xx * yz - xy * xz), m33 = xx * yy - xy * xy;
            const float det = xx * m11 + xy * m12 + xz * m13;
            const float na = sqrtf(xx * xx + yy * yy + zz * zz + 2.0f * (xy * xy + xz * xz + yz * yz));
            const float nadj = sqrtf(m11 * m11 + m22 * m22 + m33 * m33 + 2.0f * (m12 * m12 + m13 * m13 + m23 * m23));
            const float cond = na * nadj / fabsf(det);
            if (cond <= kSolve32MaxCond) {   // false for NaN, infinite or zero-determinant systems
                v[0] = (m11 * xt + m12 * yt + m13 * zt) / det;
                v[1] = (m12 * xt + m22 * yt + m23 * zt) / det;
                v[2] = (m13 * xt + m23 * yt + m33 * zt) / det;
                solved = true;
            } else {
                atomicAdd(&c->solve64, 1);
            }
        }
        if (!solved) {
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
        v[0] = (m11 * xt + m12 * yt + m13 * zt) / det;
        v[1] = (m12 * xt + m22 * yt + m23 * zt) / det;
        v[2] = (m13 * xt + m23 * yt + m33 * zt) / det;
        }
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
    block_max_to(md, mdpart + blockIdx.x);
}

// First update of a level: max |m_1 - m_0| over every voxel (m_0 is the
// level's initial dense field, not a CP interpolant).  Grid-stride over
// voxels with a fixed grid; one partial maximum per workgroup.
constexpr int kDenseBlocks = 2048;   // the most workgroups of k_dense_maxdiff

// Workgroups of k_dense_maxdiff on a level: one per 4096 voxels, 256 .. 2048.
inline int dense_blocks(const LevelGeom& g) {
    const long long nvox = static_cast<long long>(g.nx) * g.ny * g.nz;
    return static_cast<int>(std::min<long long>(kDenseBlocks, std::max<long long>(256, nvox >> 12)));
}

__global__ __launch_bounds__(256) void k_dense_maxdiff(const Ctl* __restrict__ c, const float* __restrict__ m0,
                                                        const float* __restrict__ cp, int nx, int ny, int nz,
                                                        int r, int p, int ncx, int ncy, float* __restrict__ mdpart,
                                                        long long fs) {
    c = at_frame(c, fs);
    if (c->done) return;
    m0 = at_frame(m0, fs);
    cp = at_frame(cp, fs);
    mdpart = at_frame(mdpart, fs);
    const size_t nvox = static_cast<size_t>(nx) * ny * nz;
    const size_t ncp = static_cast<size_t>(ncx) * ncy * nz;
    const unsigned nrows = static_cast<unsigned>(ny) * static_cast<unsigned>(nz);
    float md = 0.0f;
    // rows (y, z) round robin over the workgroups; the row's CP weights are
    // uniform, the column's per thread
    for (unsigned row = blockIdx.x; row < nrows; row += gridDim.x) {
        const int y = static_cast<int>(row % static_cast<unsigned>(ny));
        const int z = static_cast<int>(row / static_cast<unsigned>(ny));
        const Axis1 ay = cp_axis(y, r, p, ncy);
        const float* base = cp + static_cast<size_t>(z) * ncy * ncx;
        const float* mrow = m0 + static_cast<size_t>(row) * nx;
        for (int x = threadIdx.x; x < nx; x += 256) {
            const Axis1 ax = cp_axis(x, r, p, ncx);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float m1 = cp_bilinear(base + q * ncp, ncx, ax, ay);
                md = fmaxf(md, fabsf(m1 - mrow[x + q * nvox]));
            }
        }
    }
    block_max_to(md, mdpart + blockIdx.x);
}

// The same maximum with the level's column weights (one per x, formed once
// per workgroup) and each row's two control-point rows staged in LDS: a
// voxel then costs LDS reads instead of twelve cached gathers and a
// correctly rounded division.  Same arithmetic and operands as
// cp_bilinear; LDS holds 2*nx + 6*ncx floats (see dense_lds_bytes).
inline size_t dense_lds_bytes(const LevelGeom& g) {
    return (2 * static_cast<size_t>(g.nx) + 6 * static_cast<size_t>(g.ncx)) * sizeof(float);
}
constexpr size_t kDenseLdsMax = 48 * 1024;

__global__ __launch_bounds__(256) void k_dense_maxdiff_lds(const Ctl* __restrict__ c, const float* __restrict__ m0,
                                                            const float* __restrict__ cp, int nx, int ny, int nz,
                                                            int r, int p, int ncx, int ncy,
                                                            float* __restrict__ mdpart, long long fs) {
    extern __shared__ __align__(16) unsigned char dense_smem[];
    c = at_frame(c, fs);
    if (c->done) return;   // the same value for every thread of the frame's workgroups
    m0 = at_frame(m0, fs);
    cp = at_frame(cp, fs);
    mdpart = at_frame(mdpart, fs);
    int* sx0 = reinterpret_cast<int*>(dense_smem);
    float* stx = reinterpret_cast<float*>(sx0 + nx);
    float* srow = stx + nx;   // [component][row i0 | row i1][ncx]
    for (int x = threadIdx.x; x < nx; x += 256) {
        const Axis1 ax = cp_axis(x, r, p, ncx);
        sx0[x] = ax.i0;
        stx[x] = ax.t;
    }
    const size_t nvox = static_cast<size_t>(nx) * ny * nz;
    const size_t ncp = static_cast<size_t>(ncx) * ncy * nz;
    const unsigned nrows = static_cast<unsigned>(ny) * static_cast<unsigned>(nz);
    const int nst = 6 * ncx;
    float md = 0.0f;
    for (unsigned row = blockIdx.x; row < nrows; row += gridDim.x) {
        const int y = static_cast<int>(row % static_cast<unsigned>(ny));
        const int z = static_cast<int>(row / static_cast<unsigned>(ny));
        const Axis1 ay = cp_axis(y, r, p, ncy);
        const float* base = cp + static_cast<size_t>(z) * ncy * ncx;
        __syncthreads();   // the previous row's reads (and the column weights) are done
        for (int i = threadIdx.x; i < nst; i += 256) {
            const int q = i / (2 * ncx);
            const int rem = i - q * 2 * ncx;
            const int h = rem >= ncx ? 1 : 0;
            const int ix = rem - h * ncx;
            srow[i] = base[q * ncp + static_cast<size_t>(h ? ay.i1 : ay.i0) * ncx + ix];
        }
        __syncthreads();
        const float* mrow = m0 + static_cast<size_t>(row) * nx;
        for (int x = threadIdx.x; x < nx; x += 256) {
            const int i0 = sx0[x];
            const int i1 = min(i0 + 1, ncx - 1);
            const float t = stx[x];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float* a0 = srow + q * 2 * ncx;
                const float* a1 = a0 + ncx;
                const float m1 = lerpf(lerpf(a0[i0], a0[i1], t), lerpf(a1[i0], a1[i1], t), ay.t);
                md = fmaxf(md, fabsf(m1 - mrow[x + q * nvox]));
            }
        }
    }
    block_max_to(md, mdpart + blockIdx.x);
}

// Error terms, log and the error-based stop rule of iteration k.
__global__ __launch_bounds__(1024) void k_err_finish(Ctl* __restrict__ c, int k, const double* __restrict__ errpart,
                                                      long long nerr, const double* __restrict__ penpart,
                                                      long long npen, double* __restrict__ log,
                                                      int* __restrict__ counts, long long fs) {
    c = at_frame(c, fs);
    if (c->done) return;
    errpart = at_frame(errpart, fs);
    penpart = at_frame(penpart, fs);
    log = at_frame(log, fs);
    counts = at_frame(counts, fs);
    err_finish_block(c, k, errpart, nerr, penpart, npen, log, counts);
}

// The |delta motion| < 1e-3 stop test of the update just made; `nparts`
// per-workgroup maxima from k_solve (and k_dense_maxdiff on a level's first
// update) are reduced here.
__global__ __launch_bounds__(1024) void k_check(Ctl* __restrict__ c, int k, const float* __restrict__ mdpart,
                                                 int nparts, double* __restrict__ log, int* __restrict__ counts,
                                                 long long fs) {
    c = at_frame(c, fs);
    if (c->done) return;
    mdpart = at_frame(mdpart, fs);
    log = at_frame(log, fs);
    counts = at_frame(counts, fs);
    check_finish_block(c, k, mdpart, nparts, log, counts);
}

// Dense field at the end of a level: the CP interpolant when the level made
// at least one update, else its initial field.  Writes the 3-plane layout
// (for the next level's upsample) and/or the interleaved (Z,Y,X,3) layout
// (the returned motion), and optionally motion + grid; with WARP_T >= 0
// (the element type of `src`) also the warp of `src` by the field
// (correctMotion, warp_taps / warp_value as k_warp forms it), so the field is
// not read back.  A thread owns one column of kWarpRows consecutive rows
// (the rows of a workgroup are uniform); the x interpolation of the CP rows
// (the inner lerps of cp_bilinear) is formed once per CP row and z.
typedef float f3u __attribute__((ext_vector_type(3), aligned(4)));

template <int WARP_T>
__global__ __launch_bounds__(256) void k_level_end(const Ctl* __restrict__ c, const float* __restrict__ m0,
                                                    const float* __restrict__ cp, int nx, int ny, int nz, int r,
                                                    int p, int ncx, int ncy, int xblocks, float* __restrict__ soa,
                                                    float* __restrict__ aos, float* __restrict__ coords,
                                                    int* __restrict__ counts, int cp_start,
                                                    const void* __restrict__ src, float* __restrict__ wout,
                                                    long long fs, long long ofs, long long sfs, long long wfs) {
    c = at_frame(c, fs);
    m0 = at_frame(m0, fs);
    cp = at_frame(cp, fs);
    soa = at_frame(soa, fs);
    counts = at_frame(counts, fs);
    aos = at_frame(aos, ofs);
    coords = at_frame(coords, ofs);
    src = at_frame(src, sfs);
    wout = at_frame(wout, wfs);
    const unsigned b = blockIdx.x;
    const unsigned rg = b / static_cast<unsigned>(xblocks);
    const int x = static_cast<int>(b - rg * static_cast<unsigned>(xblocks)) * 256 + static_cast<int>(threadIdx.x);
    if (b == 0 && threadIdx.x == 0) {
        counts[2] = c->nan;
        counts[3] = c->solve64;
    }
    if (x >= nx) return;
    const unsigned nrows = static_cast<unsigned>(ny) * static_cast<unsigned>(nz);
    const size_t nvox = static_cast<size_t>(nx) * ny * nz;
    const size_t ncp = static_cast<size_t>(ncx) * ncy * nz;
    const bool interp = c->upd > 0 || cp_start;   // the CP interpolant (a level without updates keeps its start)
    const Axis1 ax = cp_axis(x, r, p, ncx);
    float m[kWarpRows][3];
    WarpTaps w[kWarpRows];
    int cz = -1, c0 = -1, c1 = -1;
    float h0[3], h1[3];
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = min(rg * kWarpRows + j, nrows - 1);
        const int y = static_cast<int>(row % static_cast<unsigned>(ny));
        const int z = static_cast<int>(row / static_cast<unsigned>(ny));
        if (interp) {
            const Axis1 ay = cp_axis(y, r, p, ncy);
            if (z != cz || ay.i0 != c0) {   // uniform: the rows of a workgroup are shared
                const float* base = cp + static_cast<size_t>(z) * ncy * ncx;
                const bool keep = z == cz && ay.i0 == c1;
                for (int q = 0; q < 3; ++q) {
                    const float* a0 = base + q * ncp + ay.i0 * ncx;
                    const float* a1 = base + q * ncp + ay.i1 * ncx;
                    h0[q] = keep ? h1[q] : lerpf(a0[ax.i0], a0[ax.i1], ax.t);
                    h1[q] = lerpf(a1[ax.i0], a1[ax.i1], ax.t);
                }
                cz = z;
                c0 = ay.i0;
                c1 = ay.i1;
            }
            for (int q = 0; q < 3; ++q) m[j][q] = lerpf(h0[q], h1[q], ay.t);
        } else {
            const size_t v = static_cast<size_t>(row) * nx + x;
            for (int q = 0; q < 3; ++q) m[j][q] = m0[v + q * nvox];
        }
        if constexpr (WARP_T >= 0) w[j] = warp_taps<WARP_T>(src, nx, ny, nz, x, y, z, m[j][0], m[j][1], m[j][2]);
    }
#pragma unroll
    for (int j = 0; j < kWarpRows; ++j) {
        const unsigned row = rg * kWarpRows + j;
        if (row >= nrows) break;
        const size_t v = static_cast<size_t>(row) * nx + x;
        if (soa)
            for (int q = 0; q < 3; ++q) soa[v + q * nvox] = m[j][q];
        if (aos) {
            f3u o;
            o.x = m[j][0];
            o.y = m[j][1];
            o.z = m[j][2];
            *reinterpret_cast<f3u*>(aos + 3 * v) = o;
        }
        if (coords) {
            const int y = static_cast<int>(row % static_cast<unsigned>(ny));
            const int z = static_cast<int>(row / static_cast<unsigned>(ny));
            f3u g;
            g.x = static_cast<float>(x) + m[j][0];
            g.y = static_cast<float>(y) + m[j][1];
            g.z = static_cast<float>(z) + m[j][2];
            *reinterpret_cast<f3u*>(coords + 3 * v) = g;
        }
        if constexpr (WARP_T >= 0) wout[v] = warp_value(w[j]);
    }
}

// End of a level above 0 whose successor starts with k_level_start: the
// NaN and float64-fallback counts, and, only where the level made no update
// and did not start from its CP field (its end field is then its start
// field), m0 copied to `prev`, since the next level's start overwrites m0.
__global__ __launch_bounds__(256) void k_level_keep(const Ctl* __restrict__ c, const float* __restrict__ m0,
                                                     float* __restrict__ prev, long long n, int* __restrict__ counts,
                                                     int cp_start, long long fs) {
    c = at_frame(c, fs);
    m0 = at_frame(m0, fs);
    prev = at_frame(prev, fs);
    counts = at_frame(counts, fs);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counts[2] = c->nan;
        counts[3] = c->solve64;
    }
    if (cp_start || c->upd > 0) return;
    const long long step = static_cast<long long>(gridDim.x) * 256;
    for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < n; i += step) prev[i] = m0[i];
}

// Float64 taps in tap order, each product and sum rounded on its own (no
// fused multiply-add), as the resampler forms them.
template <int K>
__device__ __forceinline__ double tap_sum(const double (&w)[K], const float (&v)[K]) {
#pragma clang fp contract(off)
    double acc = w[0] * static_cast<double>(v[0]);
#pragma unroll
    for (int t = 1; t < K; ++t) acc = acc + w[t] * static_cast<double>(v[t]);
    return acc;
}

__device__ __forceinline__ double tap_sum_y(const double* __restrict__ wy, const int* __restrict__ iy, int k,
                                            int ylo, const double (*xr)[256]) {
#pragma clang fp contract(off)
    double acc = wy[0] * xr[iy[0] - ylo][threadIdx.x];
    for (int t = 1; t < k; ++t) acc = acc + wy[t] * xr[iy[t] - ylo][threadIdx.x];
    return acc;
}

// Start field of level L from the end field of level L + 1
// (calFlow3d_Wei_v1.py:381-410): imresize(..., 'bilinear') of each
// component to level L's size, the x and y components times 2.  The end
// field of level L + 1 is its CP interpolant when that level made an update
// or started from its CP field; the interpolant is formed here at the source
// voxels the taps read, with the float32 arithmetic of k_level_end, so the
// dense level-(L+1) field is never written.  Otherwise the end field is that
// level's start field, which k_level_keep has copied to `prev`.  The taps
// are those of the resampler's upsample (k_resample_up): a thread owns one
// output column of a block of output rows of one z slice, puts the x taps
// of each source row into its own LDS column, then applies the y taps.
struct StartPass {
    const double* wx;
    const int* ix;
    const double* wy;
    const int* iy;
    const int* span;            // per block of output rows: first and last source row
    int nx, ny, ky, rows, nblk, xblocks;   // level L and its y table
    int nxp, nyp, rp, pp, ncxp, ncyp;      // level L + 1
};

template <int KX>
__global__ __launch_bounds__(256) void k_level_start(const Ctl* __restrict__ c, const float* __restrict__ cp,
                                                      int cp_start, const float* __restrict__ prev,
                                                      float* __restrict__ m0, StartPass u, int nz, long long fs) {
    __shared__ double xr[kUpRows][256];
    c = at_frame(c, fs);
    cp = at_frame(cp, fs);
    prev = at_frame(prev, fs);
    m0 = at_frame(m0, fs);
    const unsigned b = blockIdx.x;
    const int xb = static_cast<int>(b % static_cast<unsigned>(u.xblocks));
    const unsigned rest = b / static_cast<unsigned>(u.xblocks);
    const int yb = static_cast<int>(rest % static_cast<unsigned>(u.nblk));
    const int z = static_cast<int>(rest / static_cast<unsigned>(u.nblk));
    const int x = xb * 256 + static_cast<int>(threadIdx.x);
    if (x >= u.nx) return;
    const bool from_cp = cp_start || c->upd > 0;
    const int ylo = u.span[2 * yb];
    const int nr = u.span[2 * yb + 1] - ylo + 1;
    double w[KX];
    int ix[KX];
    Axis1 ax[KX];
#pragma unroll
    for (int t = 0; t < KX; ++t) {
        w[t] = u.wx[static_cast<size_t>(x) * KX + t];
        ix[t] = u.ix[static_cast<size_t>(x) * KX + t];
        ax[t] = cp_axis(ix[t], u.rp, u.pp, u.ncxp);
    }
    const size_t ncp = static_cast<size_t>(u.ncxp) * u.ncyp * nz;
    const size_t np = static_cast<size_t>(u.nxp) * u.nyp * nz;
    const size_t nv = static_cast<size_t>(u.nx) * u.ny * nz;
    const int y0 = yb * u.rows;
    const int y1 = min(y0 + u.rows, u.ny);
    for (int q = 0; q < 3; ++q) {
        const float* cq = cp + q * ncp + static_cast<size_t>(z) * u.ncyp * u.ncxp;
        const float* pq = prev + q * np + static_cast<size_t>(z) * u.nyp * u.nxp;
        // the x interpolation of CP rows c0 and c1 at the thread's taps (the
        // inner lerps of cp_bilinear); source rows visit CP rows in
        // ascending order, so each is formed once
        int c0 = -1, c1 = -1;
        float h0[KX], h1[KX];
        for (int j = 0; j < nr; ++j) {
            const int ys = ylo + j;
            float v[KX];
            if (from_cp) {
                const Axis1 ay = cp_axis(ys, u.rp, u.pp, u.ncyp);
                if (ay.i0 != c0 || ay.i1 != c1) {
#pragma unroll
                    for (int t = 0; t < KX; ++t) {
                        h0[t] = ay.i0 == c1 ? h1[t]
                                            : lerpf(cq[ay.i0 * u.ncxp + ax[t].i0], cq[ay.i0 * u.ncxp + ax[t].i1],
                                                    ax[t].t);
                        h1[t] = lerpf(cq[ay.i1 * u.ncxp + ax[t].i0], cq[ay.i1 * u.ncxp + ax[t].i1], ax[t].t);
                    }
                    c0 = ay.i0;
                    c1 = ay.i1;
                }
#pragma unroll
                for (int t = 0; t < KX; ++t) v[t] = lerpf(h0[t], h1[t], ay.t);
            } else {
#pragma unroll
                for (int t = 0; t < KX; ++t) v[t] = pq[static_cast<size_t>(ys) * u.nxp + ix[t]];
            }
            xr[j][threadIdx.x] = tap_sum<KX>(w, v);
        }
        float* out = m0 + q * nv + static_cast<size_t>(z) * u.ny * u.nx + x;
        for (int yo = y0; yo < y1; ++yo) {
            const double acc = tap_sum_y(u.wy + static_cast<size_t>(yo) * u.ky, u.iy + static_cast<size_t>(yo) * u.ky,
                                         u.ky, ylo, xr);
            // times 2 (exact) for the x and y components
            out[static_cast<size_t>(yo) * u.nx] = static_cast<float>(q < 2 ? acc * 2.0 : acc);
        }
    }
}

inline unsigned blocks_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------- host side

size_t mdpart_slots(long long ncp) { return static_cast<size_t>((ncp + 255) / 256) + kDenseBlocks; }

LkTile plan_tile(const LevelGeom& g, int np_x) {
    LkTile t{};
    // a workgroup has at most 4 waves of 62 owned columns each
    const int maxp = std::max(1, kMaxTileCols / g.p);
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
    t.np_x = npx;
    t.cols = npx * g.p;
    t.rows = g.p;
    t.nbx = (g.nx + t.cols - 1) / t.cols;
    t.nby = (g.ny + t.rows - 1) / t.rows;
    t.blocks = static_cast<long long>(t.nbx) * t.nby * g.nz;
    t.nwx = (t.cols + kWaveCols - 1) / kWaveCols;
    return t;
}

void launch_ctl_reset(Ctl* ctl, int iter, int frec, int fupd, double tol, double s_pen, double n_vox,
                      const Frames& fr, hipStream_t s, unsigned* hflags, unsigned gen) {
    k_ctl_reset<<<dim3(1, fr.n), 64, 0, s>>>(ctl, iter, frec, fupd, tol, s_pen, n_vox, fr.fs, hflags, gen);
}

void launch_cp_init(const LevelGeom& g, const float* m0, float* cp, const Frames& fr, hipStream_t s) {
    const long long ncp = static_cast<long long>(g.ncx) * g.ncy * g.nz;
    k_cp_init<<<dim3(blocks_for(ncp, 256), fr.n), 256, 0, s>>>(m0, cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx, g.ncy,
                                                               fr.fs);
}

int launch_iteration(const LevelGeom& g, const LkTile& t, const LkBuffers& b, int k, bool separate_stop,
                     hipStream_t s) {
    const long long ncp = static_cast<long long>(g.ncx) * g.ncy * g.nz;
    const unsigned cpb = blocks_for(ncp, 256);
    // The LK launch also computes the neighbour differences (appended
    // workgroups).  On levels where every workgroup can afford to read all
    // partials, the stop tests run inside the launches that follow them: the
    // error rule in the update launch, the |delta motion| test of update k - 1
    // in the LK launch of iteration k; elsewhere in single-workgroup kernels.
    const long long nerr = t.blocks;
    const long long npen = (ncp + 63) / 64;                  // upper bound of the LK launch's |D|^2 partials
    const int nd = dense_blocks(g);
    // the level's first iteration reads the dense start field unless it starts from the CP field
    const bool dense = k == 0 && !b.cp_start;
    // maxima behind the |delta motion| test of update j
    auto parts_of = [&](int j) { return static_cast<long long>(cpb) + (j == 0 && !b.cp_start ? nd : 0); };
    auto fuse_check = [&](int j) { return !separate_stop && (t.blocks + npen) * parts_of(j) <= (1LL << 22); };
    const bool fuse_err = !separate_stop && static_cast<long long>(cpb) * (nerr + npen) <= (1LL << 22);
    const int check_now = (k > 0 && fuse_check(k - 1)) ? static_cast<int>(parts_of(k - 1)) : 0;
    if (b.events) (void)hipEventRecord(b.events[2 * k], s);
    const long long np = launch_lk_pass(g, t, b, k, dense, check_now, s);
    if (b.events) (void)hipEventRecord(b.events[2 * k + 1], s);
    const long long fs = b.fr.fs;
    const unsigned nf = static_cast<unsigned>(b.fr.n);
    if (!fuse_err)
        k_err_finish<<<dim3(1, nf), 1024, 0, s>>>(b.ctl, k, b.errpart, nerr, b.penpart, np, b.log, b.counts, fs);
    k_solve<<<dim3(cpb, nf), 256, 0, s>>>(b.ctl, b.sums, b.dnei, b.cp, ncp, g.mov_range, g.z_ratio_d,
                                          b.store != ST_F32 ? 1 : 0, b.mdpart, k,
                                          fuse_err ? 1 : 0, b.errpart, nerr, b.penpart, np, b.log, b.counts, fs);
    if (dense) {
        const size_t lds = dense_lds_bytes(g);
        if (lds <= kDenseLdsMax && !b.maxdiff_gather)
            k_dense_maxdiff_lds<<<dim3(nd, nf), 256, lds, s>>>(b.ctl, b.m0, b.cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx,
                                                               g.ncy, b.mdpart + cpb, fs);
        else
            k_dense_maxdiff<<<dim3(nd, nf), 256, 0, s>>>(b.ctl, b.m0, b.cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx, g.ncy,
                                                         b.mdpart + cpb, fs);
    }
    if (!fuse_check(k))
        k_check<<<dim3(1, nf), 1024, 0, s>>>(b.ctl, k, b.mdpart, static_cast<int>(parts_of(k)), b.log, b.counts,
                                             fs);
    return WR_OK;
}

bool level_start_fits(const Table& tx, const Table& ty, bool x_first, const FusedY& fu) {
    return !tx.identity && !ty.identity && x_first && fu.rows > 0 && tx.k >= 2 && tx.k <= 4;
}

void launch_level_start(const LevelGeom& g, const LevelGeom& gp, const Table& tx, const Table& ty,
                        const FusedY& fu, const double* w, const int* idx, const Ctl* ctl, const float* cp,
                        int cp_start, const float* prev, float* m0, const Frames& fr, hipStream_t s) {
    StartPass u;
    u.wx = w + tx.w_off;
    u.ix = idx + tx.i_off;
    u.wy = w + ty.w_off;
    u.iy = idx + ty.i_off;
    u.span = idx + fu.s_off;
    u.nx = g.nx;
    u.ny = g.ny;
    u.ky = ty.k;
    u.rows = fu.rows;
    u.nblk = fu.nblk;
    u.xblocks = (g.nx + 255) / 256;
    u.nxp = gp.nx;
    u.nyp = gp.ny;
    u.rp = gp.r;
    u.pp = gp.p;
    u.ncxp = gp.ncx;
    u.ncyp = gp.ncy;
    const dim3 grid(static_cast<unsigned>(static_cast<long long>(u.xblocks) * fu.nblk * g.nz),
                    static_cast<unsigned>(fr.n));
    switch (tx.k) {
        case 2:
            k_level_start<2><<<grid, 256, 0, s>>>(ctl, cp, cp_start, prev, m0, u, g.nz, fr.fs);
            break;
        case 3:
            k_level_start<3><<<grid, 256, 0, s>>>(ctl, cp, cp_start, prev, m0, u, g.nz, fr.fs);
            break;
        default:
            k_level_start<4><<<grid, 256, 0, s>>>(ctl, cp, cp_start, prev, m0, u, g.nz, fr.fs);
    }
}

void launch_level_keep(const LevelGeom& g, const LkBuffers& b, float* prev, hipStream_t s) {
    const long long n = 3LL * g.nx * g.ny * g.nz;
    const unsigned blocks = static_cast<unsigned>(std::min<long long>(blocks_for(n, 256), 4096));
    k_level_keep<<<dim3(blocks, b.fr.n), 256, 0, s>>>(b.ctl, b.m0, prev, n, b.counts, b.cp_start, b.fr.fs);
}

void launch_level_end_warp(const LevelGeom& g, const LkBuffers& b, float* motion_aos, float* coords_aos,
                           const WarpOut& w, hipStream_t s) {
    const int xblocks = (g.nx + 255) / 256;
    const long long groups = (static_cast<long long>(g.ny) * g.nz + kWarpRows - 1) / kWarpRows;
    const dim3 grid(static_cast<unsigned>(static_cast<long long>(xblocks) * groups), static_cast<unsigned>(b.fr.n));
    auto kern = w.src == nullptr ? k_level_end<-1> : (w.src_t == E_U16 ? k_level_end<E_U16> : k_level_end<E_F32>);
    kern<<<grid, 256, 0, s>>>(b.ctl, b.m0, b.cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx, g.ncy, xblocks, nullptr,
                              motion_aos, coords_aos, b.counts, b.cp_start, w.src, w.out, b.fr.fs, b.fr.ofs,
                              w.src_fs, w.out_fs);
}

void launch_level_end(const LevelGeom& g, const LkBuffers& b, int level, int iter, float* dense_soa,
                      float* motion_aos, float* coords_aos, hipStream_t s) {
    (void)level;
    (void)iter;
    const int xblocks = (g.nx + 255) / 256;
    const long long groups = (static_cast<long long>(g.ny) * g.nz + kWarpRows - 1) / kWarpRows;
    const dim3 grid(static_cast<unsigned>(static_cast<long long>(xblocks) * groups), static_cast<unsigned>(b.fr.n));
    k_level_end<-1><<<grid, 256, 0, s>>>(b.ctl, b.m0, b.cp, g.nx, g.ny, g.nz, g.r, g.p, g.ncx, g.ncy, xblocks,
                                         dense_soa, motion_aos, coords_aos, b.counts, b.cp_start, nullptr, nullptr,
                                         b.fr.fs, b.fr.ofs, 0, 0);
}

}  // namespace wr
