// Device-side pieces of one LK iteration: the neighbour difference of a
// control point (computed by workgroups appended to the LK launch) and the
// level's two stop rules (single-workgroup kernels between launches).
#pragma once

#include <cmath>

#include "lk_common.h"
#include "wr_internal.h"

namespace wr {

// 3x3 neighbour difference of CP i (calFlow3d_Wei_v1.py:69-98), replicated
// edges, float64 accumulation in the filter's C order then rounded to
// float32 (scipy correlate on a float32 array), z component times zRatio_L in
// float32 (:488-498).  Writes D to dn and returns |D|^2 (float32) as double.
__device__ __forceinline__ double neidiff_one(const float* __restrict__ cp, float* __restrict__ dn, long long i,
                                              long long ncp, int ncx, int ncy, float z_ratio) {
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
    return static_cast<double>(qq);
}

// Block-wide sum of two float64 values in a fixed order (any blockDim that is
// a multiple of 64, at most 1024); the result is valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
    __shared__ double s_a[16], s_b[16];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_a[threadIdx.x >> 6] = a;
        s_b[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0.0;
        b = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            a += s_a[w];
            b += s_b[w];
        }
    }
}

// A word of mapped host memory, written through to the system.
__device__ __forceinline__ void host_word(unsigned* p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The level is done: every later kernel of the level exits; the host, where
// it watches, stops enqueueing them.
__device__ __forceinline__ void mark_done(Ctl* __restrict__ c) {
    c->done = 1;
    if (c->hflag) host_word(c->hflag + 1, c->gen);
}

// The error-based stop rule of iteration k (calFlow3d_Wei_v1.py:426,
// 502-530) from the current error and the last three accepted errors, kept
// in old_err[k & 1] (the shifted triple goes to old_err[(k + 1) & 1], so a
// launch whose workgroups all decide can never read a half-updated triple).
__device__ __forceinline__ bool err_stop(const Ctl* __restrict__ c, int k, double cur) {
    if (c->frec >= 0) return (k == c->frec - 1) && (c->fupd == c->frec - 1);
    if (k == c->iter - 1) return true;
    const double* o = c->old_err[k & 1];
    const int worse = (o[0] <= cur) + (o[1] <= cur) + (o[2] <= cur);
    if (worse > 1) return true;
    return fabs(o[2] - cur) < c->tol;
}

// Records iteration k's error terms (one thread of the launch).
__device__ __forceinline__ void err_record(Ctl* __restrict__ c, int k, double diff, double pen, double cur, bool stop,
                                          double* __restrict__ log, int* __restrict__ counts) {
    double* row = log + static_cast<size_t>(k) * 4;
    row[0] = diff;
    row[1] = pen;
    row[2] = cur;
    c->rec = k + 1;
    counts[0] = k + 1;
    c->last_err = cur;
    if (!stop) {
        const double* o = c->old_err[k & 1];
        double* n = c->old_err[(k + 1) & 1];
        n[0] = o[1];
        n[1] = o[2];
        n[2] = cur;
    }
    c->maxdiff = 0u;
    if (stop) mark_done(c);
}

// Sums of the It^2 and |D|^2 partials and the error terms of iteration k; the
// result is the same in every workgroup of a launch (same order).  Valid in
// thread 0.
__device__ __forceinline__ void err_terms(const Ctl* __restrict__ c, const double* __restrict__ errpart, long long nerr,
                                          const double* __restrict__ penpart, long long npen, double& diff,
                                          double& pen) {
    double e = 0.0, pp = 0.0;
    for (long long i = threadIdx.x; i < nerr; i += blockDim.x) e += errpart[i];
    for (long long i = threadIdx.x; i < npen; i += blockDim.x) pp += penpart[i];
    block_sum2(e, pp);
    const double nvox = c->inv_n;
    diff = e / nvox;
    pen = (pp * c->s_pen) / nvox;
}

// Error terms of iteration k, their log entries and the error-based stop
// rule, run by one workgroup.
__device__ __forceinline__ void err_finish_block(Ctl* __restrict__ c, int k, const double* __restrict__ errpart,
                                                 long long nerr, const double* __restrict__ penpart, long long npen,
                                                 double* __restrict__ log, int* __restrict__ counts) {
    double diff, pen;
    err_terms(c, errpart, nerr, penpart, npen, diff, pen);
    if (threadIdx.x != 0) return;
    const double cur = diff + pen;
    err_record(c, k, diff, pen, cur, err_stop(c, k, cur), log, counts);
}

// Block-wide maximum of `nparts` per-workgroup maxima; valid in every thread.
__device__ __forceinline__ float parts_max(const float* __restrict__ mdpart, int nparts) {
    __shared__ float s_m[16];
    __shared__ float s_r;
    float md = 0.0f;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) md = fmaxf(md, mdpart[i]);
    for (int off = 32; off > 0; off >>= 1) md = fmaxf(md, __shfl_down(md, off, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = md;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) md = fmaxf(md, s_m[w]);
        s_r = md;
    }
    __syncthreads();
    return s_r;
}

// The |delta motion| < 1e-3 stop test of update k (calFlow3d_Wei_v1.py:
// 641-654) and its record (one thread).
__device__ __forceinline__ bool check_stop(const Ctl* __restrict__ c, int k, float md) {
    if (c->frec >= 0) return k == c->frec - 1;
    return md < 1e-3f;
}
__device__ __forceinline__ void check_record(Ctl* __restrict__ c, int k, float md, bool stop, double* __restrict__ log,
                                            int* __restrict__ counts) {
    c->maxdiff = __float_as_uint(md);
    log[static_cast<size_t>(k) * 4 + 3] = static_cast<double>(md);
    c->upd = k + 1;
    counts[1] = k + 1;
    counts[2] = c->nan;
    counts[3] = c->solve64;
    if (stop) mark_done(c);
}

// Run by one workgroup: the test of update k over `nparts` maxima.
__device__ __forceinline__ void check_finish_block(Ctl* __restrict__ c, int k, const float* __restrict__ mdpart,
                                                   int nparts, double* __restrict__ log, int* __restrict__ counts) {
    const float md = parts_max(mdpart, nparts);
    if (threadIdx.x == 0) check_record(c, k, md, check_stop(c, k, md), log, counts);
}

}  // namespace wr
