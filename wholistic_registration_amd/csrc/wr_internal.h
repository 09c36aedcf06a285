// Internal declarations shared by the translation units of libwrhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wr_hip.h"

namespace wr {

// ---------------------------------------------------------------- errors

void set_error(const std::string& msg);

// The device address of the allocation-free entry point's scratch array on
// `device`, resolved by wr_create; null before that.
double* scratch_on(int device);

#define WR_HIP_CHECK(expr)                                                       \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::wr::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));  \
            return WR_E_HIP;                                                     \
        }                                                                        \
    } while (0)

// ---------------------------------------------------------------- buffers

// A device allocation that only grows; reused across calls of one context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    template <class T> T* as() const { return static_cast<T*>(p); }
};
int grow(DevBuf& b, size_t bytes);   // WR_OK or WR_E_MEM
void release(DevBuf& b);

// ---------------------------------------------------------------- resample

// One axis of a MATLAB-style resize (imresize.py:60-82): output j is
// sum_k w[j*K + k] * src[idx[j*K + k]].  Offsets into the table arena.
struct Table {
    int n_in = 0, n_out = 0, k = 0;
    size_t w_off = 0;     // doubles
    size_t i_off = 0;     // ints
    bool identity = false;
    // x-pass blocking: outputs in groups of xo_blk columns; for group g the
    // source taps lie in [span[2g], span[2g+1]] (ints at s_off in the arena),
    // and aff[g] (ints at a_off) is 1 when every output of the group reads
    // consecutive inputs
    int xo_blk = 1, n_cb = 0, span_max = 0;
    size_t s_off = 0, a_off = 0;
};

enum Kernel { KERNEL_CUBIC = 0, KERNEL_TENT = 1 };

// Builds the table exactly as the reference does (float64, numpy row-sum
// order) and appends it to the host arena.
Table build_table(int n_in, int n_out, Kernel kern, std::vector<double>& w_arena,
                  std::vector<int>& i_arena);

enum Elem { E_F32 = 0, E_U16 = 1, E_F64 = 2 };

// Longest x row accepted by the resampling and LK kernels.
constexpr int kMaxRow = 8192;
// Source span (inputs) of one x-pass column group staged per row in LDS, and
// the widest column group.
constexpr int kXSpan = 191;
constexpr int kXMaxCols = 32;
enum OutMode { OUT_F64 = 0, OUT_F32 = 1, OUT_MASK = 2, OUT_BF16 = 3, OUT_F16 = 4 };

// Storage of the pyramid level images read by the LK pass (wr_params.precision).
enum Store { ST_F32 = 0, ST_BF16 = 1, ST_F16 = 2 };
inline int store_bytes(int st) { return st == ST_F32 ? 4 : 2; }
inline OutMode store_mode(int st) { return st == ST_F32 ? OUT_F32 : (st == ST_BF16 ? OUT_BF16 : OUT_F16); }

// The moving level image (and mask_mov) is stored padded so that every tap
// of a clamped trilinear gather is a plain load: kPad replicated voxels at
// both x ends of a row, kPadLo replicated rows / planes before the first and
// kPadHi after the last in y and z.
constexpr int kPad = 4;
constexpr int kPadLo = 1;
constexpr int kPadHi = 2;

// Output placement of a resample: logical row r = z * ny + y goes to element
// row_index(r) * pitch + pad + x.  row_index(r) = r for dense rows
// (plane_rows = 0); for a volume padded in y and z it is
// (r / ny) * plane_rows + r % ny + lead.
struct OutLayout {
    int pitch;
    int pad;
    int ny = 0;            // logical rows per plane (padded layout)
    int plane_rows = 0;    // stored rows per plane; 0 = dense rows
    long long lead = 0;    // stored rows before logical row 0
};

// Layout of a (nz, ny, nx) volume padded in x, y and z.
inline OutLayout padded_layout(int nx, int ny) {
    OutLayout o{nx + 2 * kPad, kPad};
    o.ny = ny;
    o.plane_rows = ny + kPadLo + kPadHi;
    o.lead = static_cast<long long>(kPadLo) * o.plane_rows + kPadLo;
    return o;
}
inline size_t padded_floats(int nx, int ny, int nz) {
    return static_cast<size_t>(nx + 2 * kPad) * (ny + kPadLo + kPadHi) * (nz + kPadLo + kPadHi);
}

// Resample along x: in (nz, ny, t.n_in) with element stride `in_stride`
// (3 for one component of an interleaved field) and component offset
// `in_comp`; out (nz, ny, t.n_out).
void launch_resample_x(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int ny, int nz, OutLayout ol, hipStream_t s);
// Fused x-then-y resample: output rows in blocks of `rows` (0 = not
// possible), block b reading source rows [span[2b], span[2b+1]] (ints at
// s_off in the arena), at most 64 of them.
constexpr int kFusedMinRows = 8;
// Upsampling x-then-y blocks read at most kUpRows source rows; x tables of
// 2..kUpTaps taps.
constexpr int kUpRows = 16;
constexpr int kUpTaps = 6;
struct FusedY {
    int rows = 0, nblk = 0;
    size_t s_off = 0;
};
FusedY plan_fused_y(const Table& ty, std::vector<int>& i_arena);
// The fused pass applies: both tables non-identity, x first, dense input.
bool fused_ok(const Table& tx, const Table& ty, const FusedY& fy, Elem in_t);
// in (nz, ty.n_in, tx.n_in) dense -> out (nz, ty.n_out, tx.n_out) with layout ol.
void launch_resample_xy(const void* in, Elem in_t, void* out, OutMode om, double divisor, const Table& tx,
                        const Table& ty, const FusedY& fy, const double* w, const int* idx, int nz, OutLayout ol,
                        hipStream_t s);
// Direct x-then-y resample for few-tap x tables (upsampling): output rows in
// blocks of `rows` reading at most 16 source rows each (rows = 0: not
// possible); same arena layout as FusedY.
FusedY plan_up_y(const Table& tx, const Table& ty, std::vector<int>& i_arena);
void launch_resample_up(const void* in, Elem in_t, void* out, OutMode om, double divisor, const Table& tx,
                        const Table& ty, const FusedY& fu, const double* w, const int* idx, int nz, OutLayout ol,
                        hipStream_t s);
// Resample along y: in (nz, t.n_in, nx) -> out (nz, t.n_out, nx).
void launch_resample_y(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                       OutMode om, double divisor, const Table& t, const double* w,
                       const int* idx, int nx, int nz, OutLayout ol, hipStream_t s);
// Elementwise conversion (identity resample) of nrows rows of nx: out = convert(in).
void launch_convert(const void* in, Elem in_t, int in_stride, int in_comp, void* out,
                    OutMode om, double divisor, int nx, long long nrows, OutLayout ol, hipStream_t s);
// Fill the pads of a padded volume (padded_layout) of `esize`-byte elements
// whose logical rows are written: x ends, then the y pad rows, then the z pad
// planes.
void launch_pad_volume(void* buf, int esize, int nx, int ny, int nz, hipStream_t s);

// ---------------------------------------------------------------- LK solver

// Device control block of one level (one per context, reset per level).
struct Ctl {
    int done;            // level finished: every later kernel of the level exits
    int rec;             // error records so far (len(currentError))
    int upd;             // updates so far (len(max_diff_motion))
    int nan;             // NaN components zeroed by the solve
    int solve64;         // float32-solve systems that took the float64 path (condition > 1e6)
    int iter;            // option["iter"]
    int frec, fupd;      // forced schedule of this level, -1 = natural
    unsigned maxdiff;    // float bits of max |delta motion| of the current update
    double old_err[2][3];  // the last three accepted errors (calFlow3d:426,528-530), by iteration parity
    double tol;
    double s_pen;        // smoothPenalty * (2r+1)^2
    double inv_n;        // x*y*z of the level (the error terms divide by it)
    double last_err;
    unsigned* hflag;     // NULL, or this frame's two words of host memory: [0] the
                         // progress (gen << 10 | iterations whose LK pass ended),
                         // [1] gen once the level is done (the host stops
                         // enqueueing a level's iterations when it sees it)
    unsigned gen;        // the level's generation (21 bits, never 0)
};

struct LevelGeom {
    int nx, ny, nz;      // level size
    int r, p;            // patch half width, p = 2r+1
    int ncx, ncy;        // control points per row / column
    int extra_x, extra_y;  // clamped repeats of the last column / row
    float z_ratio;       // zRatio / 2^level
    double z_ratio_d;
    double mov_range;
};

// Widest LK tile: 4 waves, each owning 62 columns plus a halo lane per side.
constexpr int kWaveCols = 62;
constexpr int kMaxTileCols = 4 * kWaveCols;

struct LkTile {
    int np_x;            // patches per workgroup along x
    int cols;            // np_x * p owned columns
    int rows;            // p owned rows (one patch row)
    int nbx, nby;        // workgroups along x, y
    long long blocks;    // nbx * nby * nz
    int nwx;             // waves of the workgroup, one per 62 tile columns
};

LkTile plan_tile(const LevelGeom& g, int np_x);

// Records of the moving level image the LK pass interpolates: record
// (x, y, z) = (V, V(x+1) - V(x-1), V(y+1) - V(y-1), V(z+1) - V(z-1)) of the
// level image V (edges replicated), in the storage `store` (4 float32 or
// 4 16-bit values), for x in [0, nx], y in [0, ny], z in [0, nz] (the last
// of each axis repeats its neighbour): (nz + 1) planes of (ny + 1) rows of
// nx + 1 records.  Built from the padded float32 level image.
size_t moving_records_bytes(int nx, int ny, int nz, int store);
void launch_moving_records(const float* mov_padded, int nx, int ny, int nz, int store, void* out, hipStream_t s);
// The same records of a level equal to the input (level 0), straight from the
// dense float32 / uint16 input (t = E_F32 or E_U16).
void launch_moving_records_input(const void* in, Elem t, int nx, int ny, int nz, int store, void* out,
                                 hipStream_t s);

// Frames of one launch sequence: the launches of an iteration run
// gridDim.y = n rows of workgroups, row f registering frame f.  Every
// per-frame buffer of frame f sits f * fs bytes after frame 0's (one
// workspace per frame); the returned fields f * ofs bytes apart.
struct Frames {
    int n = 1;
    long long fs = 0;
    long long ofs = 0;
};

// Frame blockIdx.y's copy of a per-frame buffer `p` (frame stride fs bytes).
template <class T>
__device__ __forceinline__ T* at_frame(T* p, long long fs) {
    return p == nullptr ? p : (T*)((const char*)p + static_cast<long long>(blockIdx.y) * fs);
}

struct LkBuffers {
    int store;               // element type of ref and of movq's records (Store)
    const void* mov;         // level moving image, float32 padded volume (padded_layout), allocation start
    const void* movq;        // its records (launch_moving_records)
    const void* ref;         // level template
    const uint8_t* mref;     // level mask_ref (> 0 after resampling) or null
    const float* mmov;       // level mask_mov (continuous, padded like mov, allocation start) or null
    const float* m0;         // dense field at level start, 3 planes, or null
    float* cp;               // control-point field, 3 planes of ncx*ncy*nz
    float* dnei;             // neighbour differences, 3 planes
    float* sums;             // 9 planes of per-patch sums
    double* errpart;         // per-workgroup sum of It^2
    double* penpart;         // per-workgroup sum of |D|^2 (one per 64 CPs at most)
    Ctl* ctl;
    double* log;             // [level][iter][4]: diff, pen, cur, maxdiff
    int* counts;             // [level][4]: rec, upd, nan, float64 fallbacks of the float32 solve
    hipEvent_t* events;      // null, or 2*iter events of this level bracketing the LK passes
    float* mdpart;           // per-workgroup max |delta motion| partials
    int cp_start = 0;        // 1: the level starts from the CP field itself (no dense field m0):
                             // iteration 0 is a CP pass and its |delta motion| test uses CP maxima
    int maxdiff_gather = 0;  // 1: the first update's dense |delta motion| maximum uses the gather kernel
    Frames fr;               // frames of the launches (every pointer above but ref and mref is per frame)
};

// Partial-maximum slots needed by one iteration with `ncp` control points.
size_t mdpart_slots(long long ncp);

// The fused per-voxel pass of iteration k plus the neighbour differences of
// the CP field (lk_pass.hip), reading the motion from the dense field m0
// (dense) or the CP interpolant; with check_parts > 0 every workgroup first
// applies the |delta motion| test of update k - 1 to that many maxima in
// b.mdpart (workgroup 0 records it).  Returns the number of |D|^2 partials.
long long launch_lk_pass(const LevelGeom& g, const LkTile& t, const LkBuffers& b, int k, bool dense,
                         int check_parts, hipStream_t s);

void launch_ctl_reset(Ctl* ctl, int iter, int frec, int fupd, double tol, double s_pen,
                      double inv_n, const Frames& fr, hipStream_t s,
                      unsigned* hflags = nullptr, unsigned gen = 0);
void launch_cp_init(const LevelGeom& g, const float* m0, float* cp, const Frames& fr, hipStream_t s);
int launch_iteration(const LevelGeom& g, const LkTile& t, const LkBuffers& b, int k, bool separate_stop,
                     hipStream_t s);
void launch_level_end(const LevelGeom& g, const LkBuffers& b, int level, int iter,
                      float* dense_soa, float* motion_aos, float* coords_aos, hipStream_t s);
struct WarpOut;
// The level-0 end (the returned field and, optionally, motion + grid) with
// the warp of `w.src` by the same field (correctMotion) in the same kernel.
void launch_level_end_warp(const LevelGeom& g, const LkBuffers& b, float* motion_aos, float* coords_aos,
                           const WarpOut& w, hipStream_t s);
// Start field of level L (geometry g) from the end of level L + 1 (gp): the
// bilinear upsample (tables tx, ty, blocking fu of plan_up_y) of that level's
// CP interpolant, formed from `cp` on the fly, or of `prev` where that level
// made no update (see launch_level_keep).  level_start_fits: the tables this
// kernel handles (x first, 2..4 x taps).
bool level_start_fits(const Table& tx, const Table& ty, bool x_first, const FusedY& fu);
void launch_level_start(const LevelGeom& g, const LevelGeom& gp, const Table& tx, const Table& ty,
                        const FusedY& fu, const double* w, const int* idx, const Ctl* ctl, const float* cp,
                        int cp_start, const float* prev, float* m0, const Frames& fr, hipStream_t s);
// End of a level above 0 followed by launch_level_start: its counts, and its
// start field copied to `prev` where it made no update and did not start
// from its CP field.
void launch_level_keep(const LevelGeom& g, const LkBuffers& b, float* prev, hipStream_t s);

// ---------------------------------------------------------------- warp

// Pull warp of one voxel, correctMotion (calFlow3d_Wei_v1.py:676-706):
// trilinear interpolation of `in` at (x, y, z) + m with map_coordinates'
// nearest mode (the coordinate clamped to the volume).  warp_taps issues the
// eight corner loads, warp_value combines them; k_warp and the fused level-0
// end (launch_level_end with a warp) share both, so their results agree bit
// for bit.
struct WarpTaps {
    float v[8];
    float t[3];
};

template <int IN_T>
__device__ __forceinline__ float warp_ld(const void* p, size_t i) {
    if constexpr (IN_T == E_F32) return static_cast<const float*>(p)[i];
    else return static_cast<float>(static_cast<const uint16_t*>(p)[i]);
}

template <int IN_T>
__device__ __forceinline__ WarpTaps warp_taps(const void* __restrict__ in, int nx, int ny, int nz, int x, int y,
                                              int z, float mx, float my, float mz) {
    WarpTaps w;
    const float cx = fminf(fmaxf(static_cast<float>(x) + mx, 0.0f), static_cast<float>(nx - 1));
    const float cy = fminf(fmaxf(static_cast<float>(y) + my, 0.0f), static_cast<float>(ny - 1));
    const float cz = fminf(fmaxf(static_cast<float>(z) + mz, 0.0f), static_cast<float>(nz - 1));
    const int fx = static_cast<int>(floorf(cx)), fy = static_cast<int>(floorf(cy)), fz = static_cast<int>(floorf(cz));
    w.t[0] = cx - fx;
    w.t[1] = cy - fy;
    w.t[2] = cz - fz;
    const int x1 = min(fx + 1, nx - 1), y1 = min(fy + 1, ny - 1), z1 = min(fz + 1, nz - 1);
    const size_t r00 = (static_cast<size_t>(fz) * ny + fy) * nx, r01 = (static_cast<size_t>(fz) * ny + y1) * nx;
    const size_t r10 = (static_cast<size_t>(z1) * ny + fy) * nx, r11 = (static_cast<size_t>(z1) * ny + y1) * nx;
    w.v[0] = warp_ld<IN_T>(in, r00 + fx);
    w.v[1] = warp_ld<IN_T>(in, r00 + x1);
    w.v[2] = warp_ld<IN_T>(in, r01 + fx);
    w.v[3] = warp_ld<IN_T>(in, r01 + x1);
    w.v[4] = warp_ld<IN_T>(in, r10 + fx);
    w.v[5] = warp_ld<IN_T>(in, r10 + x1);
    w.v[6] = warp_ld<IN_T>(in, r11 + fx);
    w.v[7] = warp_ld<IN_T>(in, r11 + x1);
    return w;
}

__device__ __forceinline__ float warp_value(const WarpTaps& w) {
    const float a = __builtin_fmaf(w.t[1], __builtin_fmaf(w.t[0], w.v[3] - w.v[2], w.v[2]) -
                                               __builtin_fmaf(w.t[0], w.v[1] - w.v[0], w.v[0]),
                                   __builtin_fmaf(w.t[0], w.v[1] - w.v[0], w.v[0]));
    const float c = __builtin_fmaf(w.t[1], __builtin_fmaf(w.t[0], w.v[7] - w.v[6], w.v[6]) -
                                               __builtin_fmaf(w.t[0], w.v[5] - w.v[4], w.v[4]),
                                   __builtin_fmaf(w.t[0], w.v[5] - w.v[4], w.v[4]));
    return __builtin_fmaf(w.t[2], c - a, a);
}

// Rows of the warp handled by one thread (k_warp and the fused level-0 end).
constexpr int kWarpRows = 4;

void launch_warp(const void* in, Elem in_t, int nx, int ny, int nz, const float* motion,
                 float* out, hipStream_t s);

// The warp the level-0 end forms with the field it writes: frame f's source
// at src + f * src_fs bytes (element type src_t), its result at
// out + f * out_fs bytes.
struct WarpOut {
    const void* src = nullptr;
    Elem src_t = E_F32;
    long long src_fs = 0;
    float* out = nullptr;
    long long out_fs = 0;
};

// ---------------------------------------------------------------- penalty

// getSmPnltNormFctr on a (nz, ny, nx) volume: `part` holds 4 doubles per row
// (ny * nz rows), `out` receives {mean_x, mean_y, factor} on the device, and
// `factor` (device, nullable) the factor alone.
void launch_penalty(const void* ref, Elem t, const float* mask, int nx, int ny, int nz, double* part,
                    double* out, double* factor, hipStream_t s);

// ---------------------------------------------------------------- masks

// Device state of the robust statistics (percentile select, mean, std).
struct SelCtl {
    unsigned prefix, pmask;  // key bits fixed so far
    long long rank;          // rank of the wanted order statistic within the prefix class
    long long less;          // keys below the prefix class
    long long want_next;     // index of the upper order statistic
    unsigned key_lo, key_hi; // keys of the two order statistics
    int need_min;            // the upper one differs from the lower
    float gamma;             // interpolation weight
    float thr, mu, sigma;    // percentile, mean and std of the values <= thr
    long long count;         // number of values <= thr
    // the z-score test |(v - mu) / sigma| > thres as v <= cut_lo || v >= cut_hi
    // (NaN cut-off = empty side); direct = 1 evaluates the test itself
    float cut_lo, cut_hi;
    int direct;
};

// Counters of the uint16 histogram of the robust statistics (one per value).
constexpr int kHist16Bins = 65536;

// Workgroups of the grid-stride streaming reductions (fixed: their partials
// are reduced in a fixed order).
constexpr int kStreamBlocks = 2048;

// numpy.percentile's float32 index arithmetic (linear method).
void percentile_index(long long n, double percentile, long long* prev, long long* next, float* gamma);
// robust_mean_std of n float32 or uint16 values (as float32): fills c->thr,
// c->mu, c->sigma.  hist holds kHist16Bins counters (3 * 2048 of them for
// float32 input), part and cpart kStreamBlocks partials each.
void launch_robust_stats(const void* in, Elem t, long long n, double percentile, SelCtl* c, unsigned* hist,
                         double* part, long long* cpart, hipStream_t s);
void launch_set_musig(SelCtl* c, float mu, float sigma, hipStream_t s);
// Widest row of the getMask kernel (256 words of 32 voxels).
constexpr int kMaskMaxRow = 8192;
// getMask on a (nz, ny, nx) volume with c->mu, c->sigma: out 0/1 bytes.
void launch_getmask(const void* in, Elem t, int nx, int ny, int nz, SelCtl* c, float thres, uint8_t* out,
                    hipStream_t s);
// bwareafilt3_wei: 26-connected components of the nonzero bytes of m with an
// area in [lo, hi]; L and cnt hold one int per voxel; ncomp receives the
// number of components; o8 / of (either may be null)
// receive the kept voxels as bytes / float32.
void launch_components(const uint8_t* m, int nx, int ny, int nz, double lo, double hi, int* L, int* cnt,
                       unsigned long long* ncomp, uint8_t* o8, float* of, hipStream_t s);
// mean((a - b) ** 2) in numpy's dtypes, summed in float64; part holds
// kStreamBlocks partials; *out (device) receives the mean.
void launch_mean_sq_diff(const void* a, Elem ta, const void* b, Elem tb, long long n, double* part, double* out,
                         hipStream_t s);

// ---------------------------------------------------------------- harness

// k * f(image) as float32 (method 0 identity, 1 sqrt, 2 log2(1+x), 3 log10(1+x)).
void launch_transform(const void* in, Elem t, long long n, double k, int method, int k_int, float* out,
                      hipStream_t s);
// ms[0] = mean, ms[1] = std of n float32 values (float64 sums); part holds
// kStreamBlocks partials.
void launch_mean_std(const float* in, long long n, double* part, float* ms, hipStream_t s);
// ((x - prev[0]) / prev[1]) * std + mean in float32 steps.
void launch_normalize(const float* in, long long n, const float* prev, float mean, float std_, float* out,
                      hipStream_t s);
// out = float32(a) + b.
void launch_add(const void* a, Elem ta, const float* b, long long n, float* out, hipStream_t s);
// acc (float64 for uint16 frames, float32 for float32) = acc + frame (first: = frame).
void launch_accumulate(const void* frame, Elem t, long long n, int first, void* acc, hipStream_t s);
void launch_divide(void* acc, Elem t, long long n, int count, hipStream_t s);
// Demeaned Gram matrix of nt frames' slabs (nt, nzs, ny, nx) sampled at
// [::ds, ::ds]: m holds nt * D floats, mean nt floats, cc nt * nt doubles.
void launch_frame_gram(const void* slab, Elem t, int nt, int nzs, int ny, int nx, int ds, float* m, float* mean,
                       double* cc, hipStream_t s);

// ---------------------------------------------------------------- splats

// Max splat of n samples onto nplanes target planes (out: nplanes * xout *
// yout floats, (plane, x, y)); kmin / shift: device scratch.
void launch_splat_max(const float* values, const float* coords, long long n, const float* target_z, int nplanes,
                      int xref, int yref, int xout, int yout, float z_window, int ds, int nearest, int radius,
                      int fill, float fill_value, unsigned* kmin, float* shift, float* out, hipStream_t s);
// Weighted splat: sum_val and sum_w accumulate, out = sum_val / sum_w where sum_w > eps.
void launch_splat_weighted(const float* values, const float* coords, long long n, const float* target_z,
                           int nplanes, int xref, int yref, int xout, int yout, float z_window, float z_sigma, int ds,
                           int mode, float eps, float fill_value, float* sum_val, float* sum_w, float* out,
                           hipStream_t s);

// ---------------------------------------------------------------- learned initialisation

// Extents of one channels-last activation volume (B, d, h, w, C).
struct NetGeom {
    int d, h, w;
};
// 3x3x3 convolution, padding 1: y (B, d, h, w, cout) = act(conv(x, w) + bias);
// bfloat16 x, w (cout, 27, cin), y; float32 bias (nullable); cin % 8 == 0,
// cout % 16 == 0; act: LeakyReLU with slope.
void launch_conv_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, NetGeom g, int batch,
                     int cin, int cout, int act, float slope, hipStream_t s);
// wf (cin, 27, cout) = w (cout, 27, cin) with the taps reversed.
void launch_flip_weights(const uint16_t* w, int cin, int cout, uint16_t* wf, hipStream_t s);
// dw (cout, 27, cin) += sum over voxels of dy (x) x, db (cout, nullable) +=
// sum of dy; float32; g.w % 8 == 0.
void launch_conv_wgrad(const uint16_t* dy, const uint16_t* x, NetGeom g, int batch, int cin, int cout, float* dw,
                       float* db, hipStream_t s);
// gr *= (y > 0 ? 1 : slope) over n values (n % 8 == 0).
void launch_lrelu_bwd(uint16_t* gr, const uint16_t* y, long long n, float slope, hipStream_t s);
// y (B, d/2, h/2, w/2, c) = scale * sum of the 8 children of x (B, d, h, w, c).
void launch_down2(const uint16_t* x, NetGeom g, int batch, int c, float scale, uint16_t* y, hipStream_t s);
// y (B, d, h, w, c) (+)= scale * x at the parent voxel of (B, d/2, h/2, w/2, c).
void launch_up2(const uint16_t* x, NetGeom g, int batch, int c, float scale, int accumulate, uint16_t* y,
                hipStream_t s);
void launch_concat(const uint16_t* a, int ca, const uint16_t* b, int cb, long long nvox, uint16_t* y, hipStream_t s);
// The concatenation of the nearest 2x upsampling of x (batch, d/2, h/2, w/2,
// ca) with b (batch, d, h, w, cb) into y (batch, d, h, w, ca + cb); g holds
// the full-resolution d, h, w.
void launch_upcat(const uint16_t* x, int ca, const uint16_t* b, int cb, NetGeom g, int batch, uint16_t* y,
                  hipStream_t s);
void launch_split(const uint16_t* gr, int ca, int cb, long long nvox, uint16_t* da, uint16_t* db, hipStream_t s);
// loss += scale * sum over voxels and channels c < ct of (pscale p - t)^2;
// dp = d loss / d p (zero in channels ct..cp-1).
void launch_mse(const uint16_t* p, int cp, const float* t, int ct, long long nvox, float pscale, float scale,
                uint16_t* dp, float* loss, hipStream_t s);
// One Adam step (step counts from 1); w16 (nullable) receives bfloat16 copies.
void launch_adam(float* w, const float* gr, float* m, float* v, uint16_t* w16, long long n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t s);
void launch_to_bf16(const float* in, long long n, uint16_t* out, hipStream_t s);
// (n, 8) bfloat16: (mov - ms[0]) / ms[1], (ref - ms[0]) / ms[1], zeros.
void launch_net_input(const float* mov, const float* ref, long long n, const float* ms, uint16_t* out,
                      hipStream_t s);
// The same input on a padded (dp, hp, wp) grid, edge voxels repeated; nb
// frames of mov (frame stride d*h*w) into nb consecutive padded grids.
void launch_net_input_padded(const float* mov, const float* ref, int d, int h, int w, int dp, int hp, int wp,
                             const float* ms, uint16_t* out, hipStream_t s, int nb = 1);
// field (n, 3) float32 = scale * p[:, 0:3] of p (n, cp) bfloat16.
void launch_net_output(const uint16_t* p, int cp, long long n, float scale, float* field, hipStream_t s);
// The field of p (nz, h3, w3, cp) (level q of the full volume; stored on an
// hs x ws grid per plane, the rest edge padding) at the control
// points (ncx x ncy per plane, half-width r) of pyramid level `level`, in that
// level's units: out 3 planes of nz * ncy * ncx float32.
void launch_net_output_cp(const uint16_t* p, int cp, int h3, int w3, int hs, int ws, int q, int nz, int ncx, int ncy,
                          int r, int level, float scale, float* out, hipStream_t s,
                          int nb = 1, long long pstride = 0);

// ---------------------------------------------------------------- context

struct Context {
    int device = 0;
    DevBuf fws;                          // per-frame workspaces of wr_get_motion(_batch) (Frames)
    DevBuf ref_l, tmp, tmp2, mref_l;     // template and mask_ref levels; resampling intermediates
    DevBuf tab_w, tab_i, pen;
    std::vector<double> tab_wh;          // host copies of the tables on the device
    std::vector<int> tab_ih;
    bool tab_valid = false;
    DevBuf msel, mlab, mcnt, mtmp;       // mask statistics, component labels / areas, byte mask
    DevBuf hsmall, hbig;                 // harness: statistics / Gram scratch, sampled frame matrix
    DevBuf splat;                        // splats: shift scratch and the weighted sums
    DevBuf netw;                         // learned initialisation: flipped weights, input statistics
    DevBuf xws, xsel, xroi;              // cross-resolution solver: per-iteration scratch, selection, trap ROI
    std::vector<hipEvent_t> events;
    // progress and stop words of each frame's level in mapped host memory
    // (Ctl::hflag), read by the host to stop enqueueing a finished level
    unsigned* hflags = nullptr;          // host address
    unsigned* hflags_dev = nullptr;      // device address of the same memory
    int hflags_frames = 0;
    unsigned gen = 0;
    hipStream_t side = nullptr;          // pyramid builds, ahead of the iterations
    std::vector<hipEvent_t> lvl_events;  // [L]: level L's images are ready; [layer+1]: fork
    wr_stats stats{};
    // geometry of the template level images (and mask_ref levels) the last
    // wr_get_motion built, for wr_params.reuse_template; valid = 0: none
    struct TemplateKey {
        int valid = 0, nx = 0, ny = 0, nz = 0, layer = 0, dtype = 0, precision = 0, masked = 0;
        const void* ref = nullptr;
        const float* mref = nullptr;
    } tmpl;
};

}  // namespace wr

// The opaque handle of the C ABI.
struct wr_ctx {
    wr::Context c;
};

#define WR_TRY(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != WR_OK) return rc_; \
    } while (0)

#define WR_ARG(cond, msg)          \
    do {                           \
        if (!(cond)) {             \
            ::wr::set_error(msg);  \
            return WR_E_ARG;       \
        }                          \
    } while (0)
