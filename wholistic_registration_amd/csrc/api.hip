// C ABI of libwrhip.so and the host-side driver of getMotion.
//
// wr_get_motion enqueues the whole pyramid (calFlow3d_Wei_v1.py:327-673) on
// one stream: per level the two level images and masks (imresize from full
// resolution, :346-351,422-423), the level's initial field (warm start
// :362-380 or the bilinear x2 upsample of the previous level :381-410), and
// `iter` launches of the iteration sequence whose stop decisions are taken on
// the device.  The host waits once, at the end, for the logs.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include "wr_internal.h"

// Device scratch of the allocation-free entry point
// (wr_smooth_penalty_factor_noalloc): the four row partials of the smoothness
// normaliser for up to WR_NOALLOC_MAX_ROWS rows, then its three results.
// Each GPU holds its own copy, at its own address, until the process ends;
// host code never passes the array's name to a kernel, only the address
// resolved for the device (scratch_on).
static __device__ __align__(256) double scratch[4 * static_cast<size_t>(WR_NOALLOC_MAX_ROWS) + 4];

namespace wr {

namespace {
thread_local std::string g_error;

__global__ void k_fill(float* p, size_t n, float v) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill(float* p, size_t n, float v, hipStream_t s) {
    k_fill<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(p, n, v);
}

// n floats of every frame's copy of a per-frame buffer (Frames).
__global__ void k_fill_frames(float* p, size_t n, float v, long long fs) {
    p = at_frame(p, fs);
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256)
        p[i] = v;
}

void fill_frames(float* p, size_t n, float v, const Frames& fr, hipStream_t s) {
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 8192));
    k_fill_frames<<<dim3(blocks, fr.n), 256, 0, s>>>(p, n, v, fr.fs);
}

// Number of control points along an axis: len(range(r, n - 1, 2r + 1))
// (calFlow3d:434-435).
int n_cp(int n, int r) {
    const int p = 2 * r + 1;
    const int span = (n - 1) - r;
    return span > 0 ? (span + p - 1) / p : 0;
}

struct Plan2 {
    Table tx, ty;
    bool x_first;
    FusedY fy;           // fused x-then-y blocking for downsampling (rows = 0: none)
    FusedY fu;           // direct x-then-y blocking for upsampling (rows = 0: none)
};

// Axis order of imresize: ascending scale, stable (imresize.py:170-171,185-187).
bool x_before_y(int nx_in, int nx_out, int ny_in, int ny_out) {
    const double sx = 1.0 * nx_out / nx_in, sy = 1.0 * ny_out / ny_in;
    return sx <= sy;
}

// in (nz, ny_in, nx_in) -> out (nz, ny_out, nx_out) through `tmp` (float64);
// `ol` places the output rows (pitch and leading pad, in elements).
void resample2(const void* in, Elem in_t, int stride, int comp, void* out, OutMode om, double div,
               const Plan2& pl, const double* w, const int* idx, int nz, double* tmp, OutLayout ol,
               hipStream_t s) {
    const Table& tx = pl.tx;
    const Table& ty = pl.ty;
    if (tx.identity && ty.identity) {
        launch_convert(in, in_t, stride, comp, out, om, div, tx.n_in, static_cast<long long>(ty.n_in) * nz, ol, s);
    } else if (ty.identity) {
        launch_resample_x(in, in_t, stride, comp, out, om, div, tx, w, idx, ty.n_in, nz, ol, s);
    } else if (tx.identity) {
        launch_resample_y(in, in_t, stride, comp, out, om, div, ty, w, idx, tx.n_in, nz, ol, s);
    } else if (pl.x_first && stride == 1 && comp == 0 && pl.fu.rows > 0) {
        launch_resample_up(in, in_t, out, om, div, tx, ty, pl.fu, w, idx, nz, ol, s);
    } else if (pl.x_first && stride == 1 && comp == 0 && fused_ok(tx, ty, pl.fy, in_t)) {
        launch_resample_xy(in, in_t, out, om, div, tx, ty, pl.fy, w, idx, nz, ol, s);
    } else if (pl.x_first) {
        launch_resample_x(in, in_t, stride, comp, tmp, OUT_F64, 1.0, tx, w, idx, ty.n_in, nz, OutLayout{tx.n_out, 0}, s);
        launch_resample_y(tmp, E_F64, 1, 0, out, om, div, ty, w, idx, tx.n_out, nz, ol, s);
    } else {
        launch_resample_y(in, in_t, stride, comp, tmp, OUT_F64, 1.0, ty, w, idx, tx.n_in, nz, OutLayout{tx.n_in, 0}, s);
        launch_resample_x(tmp, E_F64, 1, 0, out, om, div, tx, w, idx, ty.n_out, nz, ol, s);
    }
}

OutLayout dense(int nx) { return OutLayout{nx, 0}; }

size_t tmp_bytes(const Plan2& pl, int nz) {
    if (pl.tx.identity || pl.ty.identity) return 0;
    const size_t a = pl.x_first ? static_cast<size_t>(pl.tx.n_out) * pl.ty.n_in
                                : static_cast<size_t>(pl.tx.n_in) * pl.ty.n_out;
    return a * nz * sizeof(double);
}

Plan2 make_plan(int nx_in, int nx_out, int ny_in, int ny_out, Kernel k, std::vector<double>& wa,
                std::vector<int>& ia) {
    Plan2 p;
    p.tx = build_table(nx_in, nx_out, k, wa, ia);
    p.ty = build_table(ny_in, ny_out, k, wa, ia);
    p.x_first = x_before_y(nx_in, nx_out, ny_in, ny_out);
    if (p.x_first) {
        p.fy = plan_fused_y(p.ty, ia);
        p.fu = plan_up_y(p.tx, p.ty, ia);
    }
    return p;
}

// The tables are replaced with host-synchronous copies, which are not ordered
// after work on a non-blocking stream: wait for the caller's stream first, so
// that no kernel of an earlier call on it still reads the old tables.  Tables
// equal to the ones on the device are not uploaded again (no host wait: a
// call with the same geometry can be captured in a graph).
int upload_tables(Context& c, const std::vector<double>& wa, const std::vector<int>& ia, hipStream_t s) {
    if (c.tab_valid && wa == c.tab_wh && ia == c.tab_ih) return WR_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    WR_HIP_CHECK(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        set_error("resampling tables differ from the previous call's on this context: make one call of this "
                  "geometry before capturing it");
        return WR_E_ARG;
    }
    c.tab_valid = false;
    WR_HIP_CHECK(hipStreamSynchronize(s));
    if (int rc = grow(c.tab_w, std::max<size_t>(wa.size(), 1) * sizeof(double))) return rc;
    if (int rc = grow(c.tab_i, std::max<size_t>(ia.size(), 1) * sizeof(int))) return rc;
    if (!wa.empty()) WR_HIP_CHECK(hipMemcpy(c.tab_w.p, wa.data(), wa.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!ia.empty()) WR_HIP_CHECK(hipMemcpy(c.tab_i.p, ia.data(), ia.size() * sizeof(int), hipMemcpyHostToDevice));
    c.tab_wh = wa;
    c.tab_ih = ia;
    c.tab_valid = true;
    return WR_OK;
}

Elem elem_of(int dtype) { return dtype == WR_U16 ? E_U16 : (dtype == WR_F64 ? E_F64 : E_F32); }

}  // namespace

void set_error(const std::string& msg) { g_error = msg; }

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), resolved once per
// device; null (with the error set) when the runtime cannot give it.
double* resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index above the scratch table");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[device] = static_cast<double*>(ptr);
    }
    return g_scratch[device];
}
}  // namespace

double* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

int grow(DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return WR_OK;
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    const size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        set_error("device allocation of " + std::to_string(want) + " bytes failed");
        return WR_E_MEM;
    }
    b.bytes = want;
    return WR_OK;
}

void release(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

}  // namespace wr

using namespace wr;


extern "C" {

wr_ctx* wr_create(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        set_error("wr_create: no HIP device " + std::to_string(device));
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_error("wr_create: hipSetDevice failed");
        return nullptr;
    }
    // set-up of the allocation-free entry point: this device's copy of the scratch array
    if (!resolve_scratch(device)) return nullptr;
    auto* ctx = new wr_ctx;
    ctx->c.device = device;
    return ctx;
}

void wr_destroy(wr_ctx* ctx) {
    if (!ctx) return;
    Context& c = ctx->c;
    (void)hipSetDevice(c.device);
    for (DevBuf* b : {&c.fws, &c.ref_l, &c.tmp, &c.tmp2, &c.mref_l, &c.tab_w, &c.tab_i, &c.pen,
                      &c.msel, &c.mlab, &c.mcnt, &c.mtmp, &c.hsmall, &c.hbig, &c.splat, &c.netw, &c.xws,
                      &c.xsel, &c.xroi})
        release(*b);
    for (hipEvent_t e : c.events) (void)hipEventDestroy(e);
    for (hipEvent_t e : c.lvl_events) (void)hipEventDestroy(e);
    if (c.side) (void)hipStreamDestroy(c.side);
    if (c.hflags) (void)hipHostFree(c.hflags);
    delete ctx;
}

const char* wr_last_error(void) { return g_error.c_str(); }

int wr_get_stats(wr_ctx* ctx, wr_stats* out) {
    WR_ARG(ctx && out, "wr_get_stats: null argument");
    *out = ctx->c.stats;
    return WR_OK;
}

int wr_imresize(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz, int nx_out, int ny_out,
                int method, double* out, void* stream) {
    WR_ARG(ctx && in && out, "wr_imresize: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0 && nx_out > 0 && ny_out > 0, "wr_imresize: empty shape");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16 || dtype == WR_F64, "wr_imresize: unsupported dtype");
    WR_ARG(method == 0 || method == 1, "unidentified kernel method supplied");
    WR_ARG(nx <= kMaxRow && nx_out <= kMaxRow, "wr_imresize: x extent above 8192 voxels");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<double> wa;
    std::vector<int> ia;
    const Plan2 pl = make_plan(nx, nx_out, ny, ny_out, method == 0 ? KERNEL_CUBIC : KERNEL_TENT, wa, ia);
    WR_TRY(upload_tables(c, wa, ia, s));
    WR_TRY(grow(c.tmp, tmp_bytes(pl, nz)));
    resample2(in, elem_of(dtype), 1, 0, out, OUT_F64, 1.0, pl, c.tab_w.as<double>(), c.tab_i.as<int>(), nz,
              c.tmp.as<double>(), dense(nx_out), s);
    WR_HIP_CHECK(hipGetLastError());
    WR_HIP_CHECK(hipStreamSynchronize(s));
    return WR_OK;
}

int wr_pyramid_level(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz, int level, float* out,
                     void* stream) {
    WR_ARG(ctx && in && out, "wr_pyramid_level: null argument");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_pyramid_level: dtype must be WR_F32 or WR_U16");
    WR_ARG(level >= 0 && level < WR_MAX_LEVELS, "wr_pyramid_level: level out of range");
    const int lx = static_cast<int>(nx / std::ldexp(1.0, level)), ly = static_cast<int>(ny / std::ldexp(1.0, level));
    WR_ARG(lx > 0 && ly > 0 && nz > 0, "wr_pyramid_level: level has an empty axis");
    WR_ARG(nx <= kMaxRow, "wr_pyramid_level: x extent above 8192 voxels");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<double> wa;
    std::vector<int> ia;
    const Plan2 pl = make_plan(nx, lx, ny, ly, KERNEL_CUBIC, wa, ia);
    WR_TRY(upload_tables(c, wa, ia, s));
    WR_TRY(grow(c.tmp, tmp_bytes(pl, nz)));
    resample2(in, elem_of(dtype), 1, 0, out, OUT_F32, 1.0, pl, c.tab_w.as<double>(), c.tab_i.as<int>(), nz,
              c.tmp.as<double>(), dense(lx), s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_smooth_penalty_factor(wr_ctx* ctx, const void* ref, int dtype, int nx, int ny, int nz,
                             const float* mask_ref, double* factor, void* stream) {
    WR_ARG(ctx && ref && factor, "wr_smooth_penalty_factor: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_smooth_penalty_factor: empty volume");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16 || dtype == WR_F64,
           "wr_smooth_penalty_factor: dtype must be WR_F32, WR_U16 or WR_F64");
    WR_ARG(static_cast<long long>(ny) * nz < (1LL << 24), "wr_smooth_penalty_factor: more than 2^24 rows");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t rows = static_cast<size_t>(ny) * nz;
    WR_TRY(grow(c.pen, (4 * rows + 3) * sizeof(double)));
    double* part = c.pen.as<double>();
    double* out = part + 4 * rows;
    launch_penalty(ref, elem_of(dtype), mask_ref, nx, ny, nz, part, out, nullptr, s);
    WR_HIP_CHECK(hipGetLastError());
    double h[3];
    WR_HIP_CHECK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    *factor = h[2];
    return WR_OK;
}

int wr_smooth_penalty_factor_noalloc(wr_ctx* ctx, const void* ref, int dtype, int nx, int ny, int nz,
                                     const float* mask_ref, double* factor, void* stream) {
    WR_ARG(ctx && ref && factor, "wr_smooth_penalty_factor_noalloc: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_smooth_penalty_factor_noalloc: empty volume");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16 || dtype == WR_F64,
           "wr_smooth_penalty_factor_noalloc: dtype must be WR_F32, WR_U16 or WR_F64");
    WR_ARG(static_cast<long long>(ny) * nz <= WR_NOALLOC_MAX_ROWS,
           "wr_smooth_penalty_factor_noalloc: " + std::to_string(static_cast<long long>(ny) * nz) +
               " rows (ny * nz) exceed the scratch maximum of " + std::to_string(WR_NOALLOC_MAX_ROWS));
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    double* part = scratch_on(c.device);
    WR_ARG(part != nullptr, "wr_smooth_penalty_factor_noalloc: scratch of this device not resolved");
    const size_t rows = static_cast<size_t>(ny) * nz;
    // every row partial is written before it is read: nothing to clear
    launch_penalty(ref, elem_of(dtype), mask_ref, nx, ny, nz, part, part + 4 * rows, factor,
                   static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_correct_motion(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz, const float* motion,
                      float* out, void* stream) {
    WR_ARG(ctx && in && motion && out, "wr_correct_motion: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_correct_motion: empty volume");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_correct_motion: dtype must be WR_F32 or WR_U16");
    WR_ARG(256.0 * ((nx + 255) / 256) * ny * nz < 4294967296.0, "wr_correct_motion: volume too large for one launch");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_warp(in, elem_of(dtype), nx, ny, nz, motion, out, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

namespace {

// Moving frames of one registration call: frame f of mov, mask_mov and
// motion_init starts f * stride elements after frame 0; frame f's field
// f * out_stride floats after frame 0's.
struct FrameArgs {
    int n = 1;
    long long mov_stride = 0, mmov_stride = 0, init_stride = 0, out_stride = 0;
    bool async = false;   // no logs: return once everything is enqueued
    WarpOut warp;         // warp.src set: correctMotion of it in the level-0 end
};

// ---- the host's view of a level's progress
//
// The stop rules run on the device, so the host enqueues a level's `iter`
// iterations without waiting, and the launches after the level has stopped
// exit at once (a few microseconds each).  Where the call waits for its
// results anyway, the host instead keeps at most kAhead iterations in flight
// and reads, from mapped host memory the kernels write (Ctl::hflag), how far
// every frame's level has come and whether it is done; it stops enqueueing a
// level once every frame's is.  The launches and their order are the same
// either way, only the ones that would exit at once are not made.

constexpr int kAhead = 2;
constexpr unsigned kGenMask = (1u << 21) - 1;

int ensure_hflags(Context& c, int frames) {
    if (c.hflags && c.hflags_frames >= frames) return WR_OK;
    if (c.hflags) WR_HIP_CHECK(hipHostFree(c.hflags));
    c.hflags = nullptr;
    c.hflags_dev = nullptr;
    c.hflags_frames = 0;
    void* h = nullptr;
    WR_HIP_CHECK(hipHostMalloc(&h, sizeof(unsigned) * 2 * frames, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(h, 0, sizeof(unsigned) * 2 * frames);
    void* d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        ::wr::set_error(std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
        return WR_E_HIP;
    }
    c.hflags = static_cast<unsigned*>(h);
    c.hflags_dev = static_cast<unsigned*>(d);
    c.hflags_frames = frames;
    return WR_OK;
}

// Waits until every frame's level has ended the LK pass of its first `need`
// iterations or is done.  Returns 1 when every frame's level is done, 0 to go
// on, -1 when nothing moved for a long time (the caller stops watching).
int wait_progress(const Context& c, int frames, unsigned gen, int need) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; ++spin) {
        bool all_done = true, ready = true;
        for (int f = 0; f < frames; ++f) {
            const unsigned pr = __atomic_load_n(c.hflags + 2 * f, __ATOMIC_ACQUIRE);
            const unsigned dn = __atomic_load_n(c.hflags + 2 * f + 1, __ATOMIC_ACQUIRE);
            if (dn == gen) continue;
            all_done = false;
            if ((pr >> 10) != gen || static_cast<int>(pr & 1023u) < need) ready = false;
        }
        if (all_done) return 1;
        if (ready) return 0;
        if ((spin & 1023u) == 1023u) {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) return -1;
            std::this_thread::yield();
        }
    }
}

static int get_motion_frames(wr_ctx* ctx, const void* mov, const void* ref, int dtype, int nx, int ny, int nz,
                      const float* mask_ref, const float* mask_mov, const float* motion_init, const wr_params* prm,
                      float* motion_out, float* coords_out, wr_log* logs, const FrameArgs& fa, void* stream) {
    WR_ARG(ctx && mov && ref && prm && motion_out, "wr_get_motion: null argument");
    WR_ARG(fa.n >= 1 && fa.n <= 65535, "wr_get_motion_batch: frames must be 1 .. 65535");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_get_motion: dtype must be WR_F32 or WR_U16");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_get_motion: empty volume");
    // the LK kernels index the three field planes with 32-bit offsets
    WR_ARG(3.0 * nx * ny * nz < 2147483647.0, "wr_get_motion: volume exceeds 715M voxels");
    WR_ARG(nx <= kMaxRow, "wr_get_motion: x extent above 8192 voxels");
    // row-structured launches: 256 threads per (started) 256 voxels of a row
    WR_ARG(256.0 * ((nx + 255) / 256) * ny * nz < 4294967296.0, "wr_get_motion: too many rows for one launch");
    const int layer = prm->layer, iter = prm->iter, r = prm->r;
    WR_ARG(layer >= 0 && layer < WR_MAX_LEVELS, "wr_get_motion: layer out of range");
    WR_ARG(iter >= 1, "wr_get_motion: iter must be at least 1");
    WR_ARG(r >= 0, "wr_get_motion: r must be non-negative");
    WR_ARG(2 * r + 1 <= kMaxTileCols, "wr_get_motion: r above 123 (patch wider than one workgroup tile)");
    WR_ARG(prm->zRatio != 0.0 && std::isfinite(prm->zRatio), "wr_get_motion: zRatio must be finite, non-zero");
    WR_ARG(prm->movRange > 0.0, "wr_get_motion: movRange must be positive");
    WR_ARG(prm->precision >= 0 && prm->precision <= 2, "wr_get_motion: precision must be 0, 1 or 2");
    const int p = 2 * r + 1;
    const int st = prm->precision;        // storage of the level images (Store)
    const size_t es = static_cast<size_t>(store_bytes(st));
    Context::TemplateKey key;
    key.valid = 1;
    key.nx = nx;
    key.ny = ny;
    key.nz = nz;
    key.layer = layer;
    key.dtype = dtype;
    key.precision = st;
    key.masked = mask_ref != nullptr;
    key.ref = ref;
    key.mref = mask_ref;
    const bool reuse = prm->reuse_template != 0;
    if (reuse) {
        const Context::TemplateKey& k = ctx->c.tmpl;
        WR_ARG(k.valid && k.nx == nx && k.ny == ny && k.nz == nz && k.layer == layer && k.dtype == dtype &&
                   k.precision == st && k.masked == key.masked && k.ref == ref && k.mref == mask_ref,
               "wr_get_motion: reuse_template without a previous call on the same template, mask_ref, size, "
               "dtype, layer and precision");
    }
    ctx->c.tmpl.valid = 0;   // set again once this call's template levels are enqueued

    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Elem et = elem_of(dtype);

    // ---- geometry and tables of every level
    std::vector<LevelGeom> geo(layer + 1);
    std::vector<LkTile> tiles(layer + 1);
    std::vector<Plan2> down(layer + 1), up(layer + 1);
    Plan2 init_plan{};
    WR_ARG((prm->init_nx > 0) == (prm->init_ny > 0), "wr_get_motion: init_nx and init_ny must both be set or both 0");
    WR_ARG(!prm->init_cp || motion_init, "wr_get_motion: init_cp needs motion_init");
    std::vector<double> wa;
    std::vector<int> ia;
    size_t tmp_need = 0, ncp_max = 0, nblk_max = 0, ncpb_max = 0, nvox_prev = 0;
    for (int L = layer; L >= 0; --L) {
        LevelGeom& g = geo[L];
        g.nx = static_cast<int>(nx / std::ldexp(1.0, L));
        g.ny = static_cast<int>(ny / std::ldexp(1.0, L));
        g.nz = nz;
        WR_ARG(g.nx > 0 && g.ny > 0, "wr_get_motion: pyramid level " + std::to_string(L) + " has an empty axis");
        g.r = r;
        g.p = p;
        g.ncx = n_cp(g.nx, r);
        g.ncy = n_cp(g.ny, r);
        WR_ARG(g.ncx > 0 && g.ncy > 0,
               "wr_get_motion: pyramid level " + std::to_string(L) + " is smaller than one patch");
        g.extra_x = std::max(0, g.ncx * p - g.nx);
        g.extra_y = std::max(0, g.ncy * p - g.ny);
        g.z_ratio_d = prm->zRatio / std::ldexp(1.0, L);
        g.z_ratio = static_cast<float>(g.z_ratio_d);
        g.mov_range = prm->movRange;
        const int np_x = prm->tile_patches_level ? prm->tile_patches_level[L] : prm->tile_patches_x;
        WR_ARG(np_x >= 0, "wr_get_motion: negative tile width");
        tiles[L] = plan_tile(g, np_x);
        down[L] = make_plan(nx, g.nx, ny, g.ny, KERNEL_CUBIC, wa, ia);
        tmp_need = std::max(tmp_need, tmp_bytes(down[L], nz));
        if (L == layer && motion_init && !prm->init_cp && prm->init_nx > 0) {
            init_plan = make_plan(prm->init_nx, g.nx, prm->init_ny, g.ny, KERNEL_CUBIC, wa, ia);
            tmp_need = std::max(tmp_need, tmp_bytes(init_plan, nz));
        }
        if (L < layer) {
            up[L] = make_plan(geo[L + 1].nx, g.nx, geo[L + 1].ny, g.ny, KERNEL_TENT, wa, ia);
            tmp_need = std::max(tmp_need, tmp_bytes(up[L], nz));
            nvox_prev = std::max(nvox_prev, static_cast<size_t>(geo[L + 1].nx) * geo[L + 1].ny * nz);
        }
        const size_t ncp = static_cast<size_t>(g.ncx) * g.ncy * nz;
        ncp_max = std::max(ncp_max, ncp);
        nblk_max = std::max(nblk_max, static_cast<size_t>(tiles[L].blocks));
        ncpb_max = std::max(ncpb_max, (ncp + 63) / 64);   // neighbour-difference workgroups of >= 64 threads
    }
    if (prm->forced_records || prm->forced_updates) {
        WR_ARG(prm->forced_records && prm->forced_updates, "wr_get_motion: forced schedule needs both arrays");
        // a schedule the reference can produce: every record but possibly the
        // last is followed by an update, and the last iteration has none
        for (int L = 0; L <= layer; ++L) {
            const int fr = prm->forced_records[L], fu = prm->forced_updates[L];
            WR_ARG(fr >= 1 && fr <= iter && (fu == fr - 1 || (fu == fr && fr < iter)),
                   "wr_get_motion: forced schedule of level " + std::to_string(L) + " is not a reachable schedule");
        }
    }

    // ---- workspace
    const size_t n0 = static_cast<size_t>(nx) * ny * nz;
    WR_TRY(upload_tables(c, wa, ia, s));
    WR_TRY(grow(c.tmp, tmp_need));    // intermediate of the side stream (level images)
    WR_TRY(grow(c.tmp2, tmp_need));   // intermediate of the main stream (warm start, upsample)
    // the moving image (and mask_mov) of every level is stored padded in x, y
    // and z (padded_layout); the template is dense.  The LK gathers address
    // it with 32-bit byte offsets and 24-bit row products.
    const size_t padded0 = padded_floats(nx, ny, nz);
    WR_ARG(4.0 * padded0 < 4294967295.0, "wr_get_motion: padded moving volume above 4 GiB");
    WR_ARG(static_cast<double>(nz + kPadLo + kPadHi) * (ny + kPadLo + kPadHi) < 16777216.0,
           "wr_get_motion: more than 2^24 padded rows");
    // every level keeps its own images: they are built on the context's side
    // stream, ahead of the iterations that consume them
    std::vector<size_t> mov_off(layer + 1), ref_off(layer + 1), q_off(layer + 1);
    size_t mov_tot = 0, ref_tot = 0, q_tot = 0;
    for (int L = layer; L >= 0; --L) {
        mov_off[L] = mov_tot;
        mov_tot += (padded_floats(geo[L].nx, geo[L].ny, nz) + 63) & ~static_cast<size_t>(63);
        q_off[L] = q_tot;
        if (st != ST_F32)
            q_tot += (moving_records_bytes(geo[L].nx, geo[L].ny, nz, st) + 255) & ~static_cast<size_t>(255);
        ref_off[L] = ref_tot;
        ref_tot += (static_cast<size_t>(geo[L].nx) * geo[L].ny * nz + 255) & ~static_cast<size_t>(255);
    }
    WR_TRY(grow(c.ref_l, ref_tot * es));
    if (mask_ref) WR_TRY(grow(c.mref_l, ref_tot));
    if (!c.side) WR_HIP_CHECK(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    while (c.lvl_events.size() < static_cast<size_t>(layer + 2)) {
        hipEvent_t e;
        WR_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c.lvl_events.push_back(e);
    }
    // every per-frame buffer of a frame in one workspace; frame f's sits f * W
    // bytes after frame 0's (Frames)
    size_t W = 0;
    auto carve = [&W](size_t bytes) {
        const size_t o = W;
        W += (bytes + 255) & ~static_cast<size_t>(255);
        return o;
    };
    const size_t o_mov = carve(mov_tot * sizeof(float));
    const size_t o_movq = carve(q_tot);
    const size_t o_mmov = mask_mov ? carve(mov_tot * sizeof(float)) : 0;
    const size_t o_m0 = carve(3 * n0 * sizeof(float));
    const size_t o_prev = carve(3 * nvox_prev * sizeof(float));
    const size_t o_cp = carve(3 * ncp_max * sizeof(float));
    const size_t o_dnei = carve(3 * ncp_max * sizeof(float));
    const size_t o_sums = carve(9 * ncp_max * sizeof(float));
    const size_t o_err = carve(nblk_max * sizeof(double));
    const size_t o_pen = carve(ncpb_max * sizeof(double));
    const size_t o_md = carve(mdpart_slots(static_cast<long long>(ncp_max)) * sizeof(float));
    const size_t o_ctl = carve(sizeof(Ctl));
    const size_t o_log = carve(static_cast<size_t>(layer + 1) * iter * 4 * sizeof(double));
    const size_t o_cnt = carve(static_cast<size_t>(layer + 1) * 4 * sizeof(int));
    WR_TRY(grow(c.fws, W * static_cast<size_t>(fa.n)));
    char* const ws0 = c.fws.as<char>();
    auto at = [&](size_t off, int f) { return ws0 + static_cast<size_t>(f) * W + off; };
    Frames fr;
    fr.n = fa.n;
    fr.fs = static_cast<long long>(W);
    fr.ofs = fa.out_stride * static_cast<long long>(sizeof(float));
    fill_frames(reinterpret_cast<float*>(at(o_log, 0)), static_cast<size_t>(layer + 1) * iter * 8, 0.0f, fr, s);
    fill_frames(reinterpret_cast<float*>(at(o_cnt, 0)), static_cast<size_t>(layer + 1) * 4, 0.0f, fr, s);

    const bool timing = prm->timing != 0;
    if (timing) {
        const size_t need = static_cast<size_t>(layer + 1) * iter * 2;
        while (c.events.size() < need) {
            hipEvent_t e;
            WR_HIP_CHECK(hipEventCreate(&e));
            c.events.push_back(e);
        }
    }
    const double* tw = c.tab_w.as<double>();
    const int* ti = c.tab_i.as<int>();
    double* tmp = c.tmp.as<double>();
    double* tmp2 = c.tmp2.as<double>();
    float* m0 = reinterpret_cast<float*>(at(o_m0, 0));
    float* mprev = reinterpret_cast<float*>(at(o_prev, 0));
    Ctl* ctl = reinterpret_cast<Ctl*>(at(o_ctl, 0));
    const size_t in_es = dtype == WR_U16 ? 2 : 4;
    auto mov_of = [&](int f) {
        return static_cast<const void*>(static_cast<const char*>(mov) + static_cast<size_t>(f) * fa.mov_stride * in_es);
    };
    std::memset(&c.stats, 0, sizeof(c.stats));

    // ---- level images and masks of every level, on the side stream (coarsest
    // first); level L's iterations wait for event L.  The side stream's work
    // is enqueued two levels ahead of the iterations that need it, so neither
    // stream waits for the host.
    hipStream_t side = c.side;
    WR_HIP_CHECK(hipEventRecord(c.lvl_events[layer + 1], s));
    WR_HIP_CHECK(hipStreamWaitEvent(side, c.lvl_events[layer + 1], 0));
    std::vector<const void*> ref_lv(layer + 1);
    const OutMode om_level = store_mode(st);
    auto build_level = [&](int L) -> int {
        const LevelGeom& g = geo[L];
        const OutLayout padded = padded_layout(g.nx, g.ny);
        // the moving level image as float32, padded: what the float32 LK pass
        // gathers from; with 16-bit storage the pass reads its records instead,
        // made straight from the input where the level is the input (level 0)
        const bool direct = st != ST_F32 && down[L].tx.identity && down[L].ty.identity;
        for (int f = 0; f < fa.n; ++f) {
            float* mov_l = reinterpret_cast<float*>(at(o_mov, f)) + mov_off[L];
            char* movq = at(o_movq, f) + q_off[L];
            if (!direct) {
                resample2(mov_of(f), et, 1, 0, mov_l, OUT_F32, 1.0, down[L], tw, ti, nz, tmp, padded, side);
                launch_pad_volume(mov_l, 4, g.nx, g.ny, nz, side);
            }
            if (direct) launch_moving_records_input(mov_of(f), et, g.nx, g.ny, nz, st, movq, side);
            else if (st != ST_F32) launch_moving_records(mov_l, g.nx, g.ny, nz, st, movq, side);
        }
        if (L == 0 && dtype == WR_F32 && st == ST_F32) {
            ref_lv[L] = ref;
        } else {
            void* ref_l = c.ref_l.as<char>() + ref_off[L] * es;
            if (!reuse) resample2(ref, et, 1, 0, ref_l, om_level, 1.0, down[L], tw, ti, nz, tmp, dense(g.nx), side);
            ref_lv[L] = ref_l;
        }
        if (mask_ref && !reuse)
            resample2(mask_ref, E_F32, 1, 0, c.mref_l.as<uint8_t>() + ref_off[L], OUT_MASK, 1.0, down[L], tw, ti, nz,
                      tmp, dense(g.nx), side);
        if (mask_mov)
            for (int f = 0; f < fa.n; ++f) {
                float* mm = reinterpret_cast<float*>(at(o_mmov, f)) + mov_off[L];
                resample2(mask_mov + static_cast<size_t>(f) * fa.mmov_stride, E_F32, 1, 0, mm, OUT_F32, 1.0, down[L],
                          tw, ti, nz, tmp, padded, side);
                launch_pad_volume(mm, 4, g.nx, g.ny, nz, side);
            }
        WR_HIP_CHECK(hipEventRecord(c.lvl_events[L], side));
        return WR_OK;
    };
    for (int L = layer; L >= std::max(0, layer - 1); --L) WR_TRY(build_level(L));

    // level L (below the coarsest) forms its start field from the previous
    // level's CP field (launch_level_start) instead of upsampling a dense
    // end field written by launch_level_end
    auto start_from_cp = [&](int L) {
        return L < layer && !prm->resample_start &&
               level_start_fits(up[L].tx, up[L].ty, up[L].x_first, up[L].fu);
    };
    auto starts_at_cp = [&](int L) { return L == layer && motion_init && prm->init_cp; };

    // watch the levels' progress (see wait_progress) where the call waits for
    // its results anyway and the stream is not being captured into a graph
    bool observe = !fa.async && iter < 1024;
    if (observe) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        observe = hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone &&
                  ensure_hflags(c, fa.n) == WR_OK;
    }
    // the start field may come from another stream: the iterations wait for
    // it, the side stream's level images (forked above) do not
    if (prm->init_event) WR_HIP_CHECK(hipStreamWaitEvent(s, static_cast<hipEvent_t>(prm->init_event), 0));
    for (int L = layer; L >= 0; --L) {
        const LevelGeom& g = geo[L];
        const LkTile& t = tiles[L];
        const size_t nv = static_cast<size_t>(g.nx) * g.ny * nz;

        // initial dense field of the level, 3 planes (or, with init_cp, the CP field itself)
        const bool cp_start = starts_at_cp(L);
        float* const cp0 = reinterpret_cast<float*>(at(o_cp, 0));
        if (cp_start) {
            for (int f = 0; f < fa.n; ++f)
                WR_HIP_CHECK(hipMemcpyAsync(at(o_cp, f), motion_init + static_cast<size_t>(f) * fa.init_stride,
                                            3 * static_cast<size_t>(g.ncx) * g.ncy * nz * sizeof(float),
                                            hipMemcpyDeviceToDevice, s));
        } else if (L == layer) {
            if (motion_init) {
                const double div[3] = {1.0 * nx / g.nx, 1.0 * ny / g.ny, 1.0 * nz / g.nz};
                for (int f = 0; f < fa.n; ++f) {
                    float* m0f = reinterpret_cast<float*>(at(o_m0, f));
                    for (int q = 0; q < 3; ++q)
                        resample2(motion_init + static_cast<size_t>(f) * fa.init_stride, E_F32, 3, q, m0f + q * nv,
                                  OUT_F32, div[q], prm->init_nx > 0 ? init_plan : down[L], tw, ti, nz, tmp2,
                                  dense(g.nx), s);
                }
            } else {
                fill_frames(m0, 3 * nv, 0.0f, fr, s);
            }
        } else if (start_from_cp(L)) {
            launch_level_start(g, geo[L + 1], up[L].tx, up[L].ty, up[L].fu, tw, ti, ctl, cp0,
                               starts_at_cp(L + 1) ? 1 : 0, mprev, m0, fr, s);
            c.stats.start_from_cp[L] = 1;
        } else {
            const size_t np = static_cast<size_t>(geo[L + 1].nx) * geo[L + 1].ny * nz;
            for (int f = 0; f < fa.n; ++f) {
                const float* pf = reinterpret_cast<const float*>(at(o_prev, f));
                float* m0f = reinterpret_cast<float*>(at(o_m0, f));
                for (int q = 0; q < 3; ++q)
                    resample2(pf + q * np, E_F32, 1, 0, m0f + q * nv, OUT_F32, q < 2 ? 0.5 : 1.0, up[L], tw, ti, nz,
                              tmp2, dense(g.nx), s);
            }
        }
        if (!cp_start) launch_cp_init(g, m0, cp0, fr, s);
        const int frec = prm->forced_records ? prm->forced_records[L] : -1;
        const int fupd = prm->forced_updates ? prm->forced_updates[L] : -1;
        unsigned gen = 0;
        if (observe) {
            c.gen = (c.gen + 1) & kGenMask;
            if (c.gen == 0) c.gen = 1;
            gen = c.gen;
        }
        launch_ctl_reset(ctl, iter, frec, fupd, prm->tol, prm->smoothPenalty * p * p, static_cast<double>(nv), fr,
                         s, observe ? c.hflags_dev : nullptr, gen);

        if (L - 2 >= 0) WR_TRY(build_level(L - 2));

        WR_HIP_CHECK(hipStreamWaitEvent(s, c.lvl_events[L], 0));
        LkBuffers b;
        b.store = st;
        b.mov = reinterpret_cast<float*>(at(o_mov, 0)) + mov_off[L];
        b.movq = st != ST_F32 ? at(o_movq, 0) + q_off[L] : nullptr;
        b.ref = ref_lv[L];
        b.mref = mask_ref ? c.mref_l.as<uint8_t>() + ref_off[L] : nullptr;
        b.mmov = mask_mov ? reinterpret_cast<float*>(at(o_mmov, 0)) + mov_off[L] : nullptr;
        b.m0 = m0;
        b.cp = cp0;
        b.dnei = reinterpret_cast<float*>(at(o_dnei, 0));
        b.sums = reinterpret_cast<float*>(at(o_sums, 0));
        b.errpart = reinterpret_cast<double*>(at(o_err, 0));
        b.penpart = reinterpret_cast<double*>(at(o_pen, 0));
        b.ctl = ctl;
        b.log = reinterpret_cast<double*>(at(o_log, 0)) + static_cast<size_t>(L) * iter * 4;
        b.counts = reinterpret_cast<int*>(at(o_cnt, 0)) + L * 4;
        b.mdpart = reinterpret_cast<float*>(at(o_md, 0));
        b.fr = fr;
        b.events = timing ? &c.events[static_cast<size_t>(L) * iter * 2] : nullptr;
        b.cp_start = cp_start ? 1 : 0;
        b.maxdiff_gather = prm->maxdiff_gather ? 1 : 0;
        int launched = 0;
        for (int k = 0; k < iter; ++k) {
            if (observe && gen && k >= kAhead) {
                const int w = wait_progress(c, fa.n, gen, k - kAhead + 1);
                if (w == 1) break;        // every frame's level is done
                if (w < 0) gen = 0;       // no progress seen: enqueue the rest
            }
            WR_TRY(launch_iteration(g, t, b, k, prm->stop_kernels != 0, s));
            ++launched;
        }
        c.stats.lk_launches[L] = launched;
        c.stats.lk_blocks[L] = static_cast<int>(t.blocks);
        if (L == 0) {
            c.stats.tile_cols = t.cols;
            c.stats.tile_rows = t.rows;
        }
        if (L > 0 && start_from_cp(L - 1)) launch_level_keep(g, b, mprev, s);
        else if (L > 0) launch_level_end(g, b, L, iter, mprev, nullptr, nullptr, s);
        else if (fa.warp.src) launch_level_end_warp(g, b, motion_out, coords_out, fa.warp, s);
        else launch_level_end(g, b, L, iter, nullptr, motion_out, coords_out, s);
        WR_HIP_CHECK(hipGetLastError());
    }

    c.tmpl = key;   // every template level of this call is enqueued

    if (fa.async) return WR_OK;

    // ---- logs back to the host
    std::vector<double> hlog(static_cast<size_t>(layer + 1) * iter * 4 * fa.n);
    std::vector<int> hcnt(static_cast<size_t>(layer + 1) * 4 * fa.n);
    const size_t nlog = static_cast<size_t>(layer + 1) * iter * 4, ncnt = static_cast<size_t>(layer + 1) * 4;
    for (int f = 0; f < fa.n; ++f) {
        WR_HIP_CHECK(hipMemcpyAsync(hlog.data() + f * nlog, at(o_log, f), nlog * sizeof(double), hipMemcpyDeviceToHost,
                                    s));
        WR_HIP_CHECK(hipMemcpyAsync(hcnt.data() + f * ncnt, at(o_cnt, f), ncnt * sizeof(int), hipMemcpyDeviceToHost,
                                    s));
    }
    WR_HIP_CHECK(hipStreamSynchronize(s));
    if (timing) {
        for (int L = 0; L <= layer; ++L) {
            int n = 0;   // launches that did work for at least one frame
            for (int f = 0; f < fa.n; ++f) n = std::max(n, std::min(hcnt[f * ncnt + L * 4 + 0], iter));
            double tot = 0.0;
            for (int k = 0; k < n; ++k) {
                float ms = 0.0f;
                const size_t e = (static_cast<size_t>(L) * iter + k) * 2;
                WR_HIP_CHECK(hipEventElapsedTime(&ms, c.events[e], c.events[e + 1]));
                tot += ms;
            }
            c.stats.lk_exec[L] = n;
            c.stats.lk_ms[L] = tot;
        }
    }
    for (int L = 0; L <= layer; ++L) {
        int n64 = 0;
        for (int f = 0; f < fa.n; ++f) n64 += hcnt[f * ncnt + L * 4 + 3];
        c.stats.solve_fp64[L] = n64;
    }
    if (logs) {
        for (int f = 0; f < fa.n; ++f) {
            wr_log* log = logs + f;
            const double* hl = hlog.data() + f * nlog;
            const int* hc = hcnt.data() + f * ncnt;
            for (int L = 0; L <= layer; ++L) {
                log->records[L] = hc[L * 4 + 0];
                log->updates[L] = hc[L * 4 + 1];
                log->nans[L] = hc[L * 4 + 2];
                for (int k = 0; k < iter; ++k) {
                    const double* row = &hl[(static_cast<size_t>(L) * iter + k) * 4];
                    const size_t o = static_cast<size_t>(L) * iter + k;
                    if (log->diff_error) log->diff_error[o] = row[0];
                    if (log->penalty_error) log->penalty_error[o] = row[1];
                    if (log->current_error) log->current_error[o] = row[2];
                    if (log->max_diff_motion) log->max_diff_motion[o] = row[3];
                }
            }
            const int n = hc[0];
            log->final_error = n > 0 ? hl[static_cast<size_t>(n - 1) * 4 + 2] : NAN;
        }
    }
    return WR_OK;
}

}  // namespace

int wr_get_motion(wr_ctx* ctx, const void* mov, const void* ref, int dtype, int nx, int ny, int nz,
                  const float* mask_ref, const float* mask_mov, const float* motion_init, const wr_params* prm,
                  float* motion_out, float* coords_out, wr_log* log, void* stream) {
    return get_motion_frames(ctx, mov, ref, dtype, nx, ny, nz, mask_ref, mask_mov, motion_init, prm, motion_out,
                             coords_out, log, FrameArgs{}, stream);
}

int wr_get_motion_warp(wr_ctx* ctx, const void* mov, const void* ref, int dtype, int nx, int ny, int nz,
                       const float* mask_ref, const float* mask_mov, const float* motion_init, const wr_params* prm,
                       float* motion_out, const void* src, int src_dtype, float* warped_out, wr_log* log,
                       void* stream) {
    WR_ARG(src && warped_out, "wr_get_motion_warp: null argument");
    WR_ARG(src_dtype == WR_F32 || src_dtype == WR_U16, "wr_get_motion_warp: src_dtype must be WR_F32 or WR_U16");
    FrameArgs fa;
    fa.warp.src = src;
    fa.warp.src_t = elem_of(src_dtype);
    fa.warp.out = warped_out;
    return get_motion_frames(ctx, mov, ref, dtype, nx, ny, nz, mask_ref, mask_mov, motion_init, prm, motion_out,
                             nullptr, log, fa, stream);
}

int wr_get_motion_batch(wr_ctx* ctx, const void* mov, long long mov_stride, const void* ref, int dtype, int nx,
                        int ny, int nz, int frames, const float* mask_ref, const float* mask_mov,
                        long long mask_mov_stride, const float* motion_init, long long init_stride,
                        const wr_params* prm, float* motion_out, const void* warp_src, int warp_dtype,
                        long long warp_stride, float* warped_out, wr_log* logs, void* stream) {
    WR_ARG(frames >= 1, "wr_get_motion_batch: frames must be at least 1");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_get_motion: empty volume");
    const long long nvox = static_cast<long long>(nx) * ny * nz;
    WR_ARG(frames == 1 || mov_stride >= nvox, "wr_get_motion_batch: frames of mov overlap (mov_stride < nz*ny*nx)");
    WR_ARG(!mask_mov || frames == 1 || mask_mov_stride >= nvox,
           "wr_get_motion_batch: frames of mask_mov overlap (mask_mov_stride < nz*ny*nx)");
    WR_ARG(!motion_init || frames == 1 || init_stride > 0, "wr_get_motion_batch: init_stride must be positive");
    FrameArgs fa;
    fa.n = frames;
    fa.mov_stride = mov_stride;
    fa.mmov_stride = mask_mov_stride;
    fa.init_stride = init_stride;
    fa.out_stride = 3 * nvox;
    if (warp_src) {
        WR_ARG(warped_out, "wr_get_motion_batch: warp_src needs warped_out");
        WR_ARG(warp_dtype == WR_F32 || warp_dtype == WR_U16, "wr_get_motion_batch: warp_dtype must be WR_F32 or WR_U16");
        WR_ARG(frames == 1 || warp_stride >= nvox, "wr_get_motion_batch: frames of warp_src overlap");
        fa.warp.src = warp_src;
        fa.warp.src_t = elem_of(warp_dtype);
        fa.warp.src_fs = warp_stride * (warp_dtype == WR_U16 ? 2 : 4);
        fa.warp.out = warped_out;
        fa.warp.out_fs = nvox * static_cast<long long>(sizeof(float));
    }
    fa.async = logs == nullptr;
    WR_ARG(!(fa.async && prm && prm->timing), "wr_get_motion_batch: timing needs logs (the call then waits)");
    return get_motion_frames(ctx, mov, ref, dtype, nx, ny, nz, mask_ref, mask_mov, motion_init, prm, motion_out,
                             nullptr, logs, fa, stream);
}


// ------------------------------------------------------------------ masks

namespace {

struct MaskWs {
    SelCtl* c;
    unsigned* hist;
    double* part;
    long long* cpart;
    unsigned long long* ncomp;
    double* out;
};

static int mask_ws(Context& c, MaskWs& w) {
    const size_t o_hist = 256, o_part = o_hist + kHist16Bins * sizeof(unsigned);
    const size_t o_cpart = o_part + kStreamBlocks * sizeof(double);
    const size_t o_nc = o_cpart + kStreamBlocks * sizeof(long long);
    const size_t o_out = o_nc + 64;
    WR_TRY(grow(c.msel, o_out + 64));
    char* b = c.msel.as<char>();
    w.c = reinterpret_cast<SelCtl*>(b);
    w.hist = reinterpret_cast<unsigned*>(b + o_hist);
    w.part = reinterpret_cast<double*>(b + o_part);
    w.cpart = reinterpret_cast<long long*>(b + o_cpart);
    w.ncomp = reinterpret_cast<unsigned long long*>(b + o_nc);
    w.out = reinterpret_cast<double*>(b + o_out);
    return WR_OK;
}

}  // namespace

int wr_robust_mean_std(wr_ctx* ctx, const void* data, int dtype, long long n, double percentile, float* out,
                       void* stream) {
    WR_ARG(ctx && data && out, "wr_robust_mean_std: null argument");
    WR_ARG(n > 0, "wr_robust_mean_std: empty array");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_robust_mean_std: dtype must be WR_F32 or WR_U16");
    WR_ARG(percentile >= 0.0 && percentile <= 100.0, "Percentiles must be in the range [0, 100]");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    MaskWs w;
    WR_TRY(mask_ws(c, w));
    launch_robust_stats(data, elem_of(dtype), n, percentile, w.c, w.hist, w.part, w.cpart, s);
    WR_HIP_CHECK(hipGetLastError());
    SelCtl h;
    WR_HIP_CHECK(hipMemcpyAsync(&h, w.c, sizeof(h), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    out[0] = h.thr;
    out[1] = h.mu;
    out[2] = h.sigma;
    return WR_OK;
}

int wr_get_mask(wr_ctx* ctx, const void* vol, int dtype, int nx, int ny, int nz, double thres_factor,
                const float* mu_sigma, uint8_t* out, void* stream) {
    WR_ARG(ctx && vol && out, "wr_get_mask: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_get_mask: empty volume");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_get_mask: dtype must be WR_F32 or WR_U16");
    WR_ARG(nz <= 65535, "wr_get_mask: more than 65535 planes");
    WR_ARG(nx <= kMaskMaxRow, "wr_get_mask: x extent above 8192 voxels");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    MaskWs w;
    WR_TRY(mask_ws(c, w));
    const long long n = static_cast<long long>(nx) * ny * nz;
    if (mu_sigma)
        launch_set_musig(w.c, mu_sigma[0], mu_sigma[1], s);
    else
        launch_robust_stats(vol, elem_of(dtype), n, 95.0, w.c, w.hist, w.part, w.cpart, s);
    launch_getmask(vol, elem_of(dtype), nx, ny, nz, w.c, static_cast<float>(thres_factor), out, s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_bwareafilt3(wr_ctx* ctx, const uint8_t* in, int nx, int ny, int nz, double size_min, double size_max,
                   uint8_t* out_u8, float* out_f32, long long* n_components, void* stream) {
    WR_ARG(ctx && in && (out_u8 || out_f32), "wr_bwareafilt3: null argument");
    WR_ARG(nx > 0 && ny > 0 && nz > 0, "wr_bwareafilt3: empty volume");
    WR_ARG(static_cast<double>(nx) * ny * nz < 2147483647.0, "wr_bwareafilt3: volume exceeds 2^31 - 1 voxels");
    WR_ARG(reinterpret_cast<uintptr_t>(in) % 4 == 0 && reinterpret_cast<uintptr_t>(out_u8) % 4 == 0 &&
               reinterpret_cast<uintptr_t>(out_f32) % 16 == 0,
           "wr_bwareafilt3: masks must be 4-byte aligned (float32 output 16-byte aligned)");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    MaskWs w;
    WR_TRY(mask_ws(c, w));
    const size_t n = static_cast<size_t>(nx) * ny * nz;
    WR_TRY(grow(c.mlab, n * sizeof(int)));
    WR_TRY(grow(c.mcnt, n * sizeof(int)));
    launch_components(in, nx, ny, nz, size_min, size_max, c.mlab.as<int>(), c.mcnt.as<int>(), w.ncomp, out_u8,
                      out_f32, s);
    WR_HIP_CHECK(hipGetLastError());
    if (n_components) {
        unsigned long long h = 0;
        WR_HIP_CHECK(hipMemcpyAsync(&h, w.ncomp, sizeof(h), hipMemcpyDeviceToHost, s));
        WR_HIP_CHECK(hipStreamSynchronize(s));
        *n_components = static_cast<long long>(h);
    }
    return WR_OK;
}

int wr_mean_sq_diff(wr_ctx* ctx, const void* a, int dtype_a, const void* b, int dtype_b, long long n, double* out,
                    void* stream) {
    WR_ARG(ctx && a && b && out, "wr_mean_sq_diff: null argument");
    WR_ARG(n > 0, "wr_mean_sq_diff: empty arrays");
    for (int d : {dtype_a, dtype_b})
        WR_ARG(d == WR_F32 || d == WR_U16 || d == WR_F64, "wr_mean_sq_diff: unsupported dtype");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    MaskWs w;
    WR_TRY(mask_ws(c, w));
    launch_mean_sq_diff(a, elem_of(dtype_a), b, elem_of(dtype_b), n, w.part, w.out, s);
    WR_HIP_CHECK(hipGetLastError());
    WR_HIP_CHECK(hipMemcpyAsync(out, w.out, sizeof(double), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    return WR_OK;
}


// ---------------------------------------------------------------- harness

int wr_channel_transform(wr_ctx* ctx, const void* in, int dtype, long long n, double k, int method, int k_is_int,
                         float* out, void* stream) {
    WR_ARG(ctx && in && out, "wr_channel_transform: null argument");
    WR_ARG(n > 0, "wr_channel_transform: empty array");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_channel_transform: dtype must be WR_F32 or WR_U16");
    WR_ARG(method >= 0 && method <= 3, "Unknown method to process the image");
    WR_ARG(!k_is_int || (k >= 0.0 && k <= 65535.0 && k == std::floor(k)),
           "wr_channel_transform: integer k outside the uint16 range");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_transform(in, elem_of(dtype), n, k, method, k_is_int, out, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

namespace {
static int harness_small(Context& c, double** part, float** ms) {
    WR_TRY(grow(c.hsmall, 256 + kStreamBlocks * sizeof(double)));
    *ms = c.hsmall.as<float>();
    *part = reinterpret_cast<double*>(c.hsmall.as<char>() + 256);
    return WR_OK;
}
}  // namespace

int wr_mean_std(wr_ctx* ctx, const float* in, long long n, float* out, void* stream) {
    WR_ARG(ctx && in && out, "wr_mean_std: null argument");
    WR_ARG(n > 0, "wr_mean_std: empty array");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* part;
    float* ms;
    WR_TRY(harness_small(c, &part, &ms));
    launch_mean_std(in, n, part, ms, s);
    WR_HIP_CHECK(hipGetLastError());
    WR_HIP_CHECK(hipMemcpyAsync(out, ms, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    return WR_OK;
}

int wr_normalize_std(wr_ctx* ctx, const float* in, long long n, double mean, double std_, float* out,
                     void* stream) {
    WR_ARG(ctx && in && out, "wr_normalize_std: null argument");
    WR_ARG(n > 0, "wr_normalize_std: empty array");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* part;
    float* ms;
    WR_TRY(harness_small(c, &part, &ms));
    launch_mean_std(in, n, part, ms, s);
    launch_normalize(in, n, ms, static_cast<float>(mean), static_cast<float>(std_), out, s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_add_volumes(wr_ctx* ctx, const void* a, int dtype_a, const float* b, long long n, float* out,
                   void* stream) {
    WR_ARG(ctx && a && b && out, "wr_add_volumes: null argument");
    WR_ARG(n > 0, "wr_add_volumes: empty array");
    WR_ARG(dtype_a == WR_F32 || dtype_a == WR_U16, "wr_add_volumes: dtype must be WR_F32 or WR_U16");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_add(a, elem_of(dtype_a), b, n, out, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_frames_mean(wr_ctx* ctx, const void* const* frames, int count, int dtype, long long n, void* out,
                   void* stream) {
    WR_ARG(ctx && frames && out, "wr_frames_mean: null argument");
    WR_ARG(count > 0 && n > 0, "wr_frames_mean: no frames");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_frames_mean: dtype must be WR_F32 or WR_U16");
    for (int f = 0; f < count; ++f) WR_ARG(frames[f], "wr_frames_mean: null frame");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int f = 0; f < count; ++f) launch_accumulate(frames[f], elem_of(dtype), n, f == 0, out, s);
    launch_divide(out, elem_of(dtype), n, count, s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_frame_gram(wr_ctx* ctx, const void* slab, int dtype, int nt, int nzs, int ny, int nx, int ds, double* cc,
                  void* stream) {
    WR_ARG(ctx && slab && cc, "wr_frame_gram: null argument");
    WR_ARG(nt > 0 && nzs > 0 && ny > 0 && nx > 0 && ds > 0, "wr_frame_gram: empty shape");
    WR_ARG(dtype == WR_F32 || dtype == WR_U16, "wr_frame_gram: dtype must be WR_F32 or WR_U16");
    WR_ARG(nt <= 65535, "wr_frame_gram: more than 65535 frames");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long d = static_cast<long long>(nzs) * ((ny + ds - 1) / ds) * ((nx + ds - 1) / ds);
    const size_t nt2 = static_cast<size_t>(nt) * nt;
    const size_t m_bytes = (static_cast<size_t>(nt) * d * sizeof(float) + 255) / 256 * 256;
    const size_t mean_bytes = (nt * sizeof(float) + 255) / 256 * 256;
    WR_TRY(grow(c.hbig, m_bytes + mean_bytes + nt2 * sizeof(double)));
    float* m = c.hbig.as<float>();
    float* mean = reinterpret_cast<float*>(c.hbig.as<char>() + m_bytes);
    double* dcc = reinterpret_cast<double*>(c.hbig.as<char>() + m_bytes + mean_bytes);
    launch_frame_gram(slab, elem_of(dtype), nt, nzs, ny, nx, ds, m, mean, dcc, s);
    WR_HIP_CHECK(hipGetLastError());
    WR_HIP_CHECK(hipMemcpyAsync(cc, dcc, nt2 * sizeof(double), hipMemcpyDeviceToHost, s));
    WR_HIP_CHECK(hipStreamSynchronize(s));
    return WR_OK;
}


// ---------------------------------------------------------------- splats

int wr_project_planes_max(wr_ctx* ctx, const float* values, const float* coords, long long n, const float* target_z,
                          int nplanes, int xref, int yref, int downsample_xy, double z_window, int nearest,
                          int extra_radius, int fill, double fill_value, float* out, void* stream) {
    WR_ARG(ctx && (n == 0 || (values && coords)) && target_z && out, "wr_project_planes_max: null argument");
    WR_ARG(n >= 0 && nplanes > 0 && xref > 0 && yref > 0, "wr_project_planes_max: empty shape");
    WR_ARG(downsample_xy >= 1, "downsample_xy should be >= 1.");
    WR_ARG(extra_radius >= 0, "xy_extra_radius should be >= 0.");
    const int xout = (xref + downsample_xy - 1) / downsample_xy, yout = (yref + downsample_xy - 1) / downsample_xy;
    WR_ARG(static_cast<double>(nplanes) * xout * yout < 9.0e18, "wr_project_planes_max: output too large");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    WR_TRY(grow(c.splat, 512));
    unsigned* kmin = c.splat.as<unsigned>();
    float* shift = reinterpret_cast<float*>(c.splat.as<char>() + 256);
    launch_splat_max(values, coords, n, target_z, nplanes, xref, yref, xout, yout, static_cast<float>(z_window),
                     downsample_xy, nearest, extra_radius, fill, static_cast<float>(fill_value), kmin, shift, out,
                     static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_project_planes_weighted(wr_ctx* ctx, const float* values, const float* coords, long long n,
                               const float* target_z, int nplanes, int xref, int yref, int downsample_xy,
                               double z_window, double z_sigma, int mode, double eps, double fill_value, float* out,
                               float* weight_out, void* stream) {
    WR_ARG(ctx && (n == 0 || (values && coords)) && target_z && out && weight_out,
           "wr_project_planes_weighted: null argument");
    WR_ARG(n >= 0 && nplanes > 0 && xref > 0 && yref > 0, "wr_project_planes_weighted: empty shape");
    WR_ARG(downsample_xy >= 1, "downsample_xy should be >= 1.");
    WR_ARG(mode >= 0 && mode <= 2, "z_weight_mode should be 'hard', 'triangular', or 'gaussian'.");
    const int xout = (xref + downsample_xy - 1) / downsample_xy, yout = (yref + downsample_xy - 1) / downsample_xy;
    const size_t nout = static_cast<size_t>(nplanes) * xout * yout;
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    WR_TRY(grow(c.splat, 512 + nout * sizeof(float)));
    float* sum_val = reinterpret_cast<float*>(c.splat.as<char>() + 512);
    launch_splat_weighted(values, coords, n, target_z, nplanes, xref, yref, xout, yout, static_cast<float>(z_window),
                          static_cast<float>(z_sigma), downsample_xy, mode, static_cast<float>(eps),
                          static_cast<float>(fill_value), sum_val, weight_out, out, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

// ---------------------------------------------------------------- learned initialisation

}  // extern "C"

namespace {
bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int net_geom(int batch, int d, int h, int w, NetGeom* g) {
    WR_ARG(batch > 0 && d > 0 && h > 0 && w > 0, "net: empty shape");
    WR_ARG(static_cast<double>(batch) * d * h * w * 128.0 < 9.0e18, "net: volume too large");
    WR_ARG(static_cast<double>(d) * h * w / 256.0 < 2147483647.0, "net: volume too large for one launch");
    WR_ARG(batch <= 65535, "net: batch above 65535");
    g->d = d;
    g->h = h;
    g->w = w;
    return WR_OK;
}
}  // namespace

extern "C" {

int wr_net_conv3d(wr_ctx* ctx, const void* x, const void* w, const float* bias, void* y, int batch, int d, int h,
                  int wd, int cin, int cout, int leaky, double slope, void* stream) {
    WR_ARG(ctx && x && w && y, "wr_net_conv3d: null argument");
    WR_ARG(cin > 0 && cin % 8 == 0 && cout > 0 && cout % 16 == 0 && cout / 16 <= 65535,
           "wr_net_conv3d: cin must be a multiple of 8 and cout of 16");
    WR_ARG(al16(x) && al16(w) && al16(y), "wr_net_conv3d: x, w and y must be 16-byte aligned");
    // 32-bit element offsets within one batch item
    WR_ARG(static_cast<double>(d) * h * wd * std::max(cin, cout) < 2147483647.0, "wr_net_conv3d: volume too large");
    NetGeom g;
    WR_TRY(net_geom(batch, d, h, wd, &g));
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_conv_fwd(static_cast<const uint16_t*>(x), static_cast<const uint16_t*>(w), bias, static_cast<uint16_t*>(y),
                    g, batch, cin, cout, leaky, static_cast<float>(slope), static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_conv3d_bwd_data(wr_ctx* ctx, const void* dy, const void* w, void* dx, int batch, int d, int h, int wd,
                           int cin, int cout, void* stream) {
    WR_ARG(ctx && dy && w && dx, "wr_net_conv3d_bwd_data: null argument");
    WR_ARG(cin > 0 && cin % 16 == 0 && cout > 0 && cout % 16 == 0 && cin / 16 <= 65535,
           "wr_net_conv3d_bwd_data: cin and cout must be multiples of 16");
    WR_ARG(al16(dy) && al16(dx), "wr_net_conv3d_bwd_data: dy and dx must be 16-byte aligned");
    WR_ARG(static_cast<double>(d) * h * wd * std::max(cin, cout) < 2147483647.0,
           "wr_net_conv3d_bwd_data: volume too large");
    NetGeom g;
    WR_TRY(net_geom(batch, d, h, wd, &g));
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    const size_t nw = 27ull * cin * cout;
    WR_TRY(grow(c.netw, nw * sizeof(uint16_t)));
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint16_t* wf = c.netw.as<uint16_t>();
    launch_flip_weights(static_cast<const uint16_t*>(w), cin, cout, wf, s);
    launch_conv_fwd(static_cast<const uint16_t*>(dy), wf, nullptr, static_cast<uint16_t*>(dx), g, batch, cout, cin, 0,
                    0.0f, s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_conv3d_bwd_weight(wr_ctx* ctx, const void* dy, const void* x, float* dw, float* db, int batch, int d,
                             int h, int wd, int cin, int cout, void* stream) {
    WR_ARG(ctx && dy && x && dw, "wr_net_conv3d_bwd_weight: null argument");
    WR_ARG(cin > 0 && cin % 8 == 0 && cout > 0 && cout % 16 == 0 && cout / 16 <= 65535,
           "wr_net_conv3d_bwd_weight: cin must be a multiple of 8 and cout of 16");
    WR_ARG(wd % 8 == 0, "wr_net_conv3d_bwd_weight: w must be a multiple of 8");
    NetGeom g;
    WR_TRY(net_geom(batch, d, h, wd, &g));
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_conv_wgrad(static_cast<const uint16_t*>(dy), static_cast<const uint16_t*>(x), g, batch, cin, cout, dw, db,
                      static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_leaky_bwd(wr_ctx* ctx, void* grad, const void* y, long long n, double slope, void* stream) {
    WR_ARG(ctx && grad && y, "wr_net_leaky_bwd: null argument");
    WR_ARG(n > 0 && n % 8 == 0, "wr_net_leaky_bwd: n must be a positive multiple of 8");
    WR_ARG(al16(grad) && al16(y), "wr_net_leaky_bwd: buffers must be 16-byte aligned");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_lrelu_bwd(static_cast<uint16_t*>(grad), static_cast<const uint16_t*>(y), n, static_cast<float>(slope),
                     static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_down2(wr_ctx* ctx, const void* x, void* y, int batch, int d, int h, int wd, int c, double scale,
                 void* stream) {
    WR_ARG(ctx && x && y, "wr_net_down2: null argument");
    WR_ARG(c > 0 && c % 8 == 0, "wr_net_down2: c must be a multiple of 8");
    WR_ARG(d % 2 == 0 && h % 2 == 0 && wd % 2 == 0, "wr_net_down2: d, h and w must be even");
    WR_ARG(al16(x) && al16(y), "wr_net_down2: buffers must be 16-byte aligned");
    NetGeom g;
    WR_TRY(net_geom(batch, d, h, wd, &g));
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_down2(static_cast<const uint16_t*>(x), g, batch, c, static_cast<float>(scale), static_cast<uint16_t*>(y),
                 static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_up2(wr_ctx* ctx, const void* x, void* y, int batch, int d, int h, int wd, int c, double scale,
               int accumulate, void* stream) {
    WR_ARG(ctx && x && y, "wr_net_up2: null argument");
    WR_ARG(c > 0 && c % 8 == 0, "wr_net_up2: c must be a multiple of 8");
    WR_ARG(d % 2 == 0 && h % 2 == 0 && wd % 2 == 0, "wr_net_up2: d, h and w must be even");
    WR_ARG(al16(x) && al16(y), "wr_net_up2: buffers must be 16-byte aligned");
    NetGeom g;
    WR_TRY(net_geom(batch, d, h, wd, &g));
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_up2(static_cast<const uint16_t*>(x), g, batch, c, static_cast<float>(scale), accumulate,
               static_cast<uint16_t*>(y), static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_concat(wr_ctx* ctx, const void* a, int ca, const void* b, int cb, long long nvox, void* y, void* stream) {
    WR_ARG(ctx && a && b && y, "wr_net_concat: null argument");
    WR_ARG(ca > 0 && cb > 0 && ca % 8 == 0 && cb % 8 == 0 && nvox > 0, "wr_net_concat: bad channel counts");
    WR_ARG(al16(a) && al16(b) && al16(y), "wr_net_concat: buffers must be 16-byte aligned");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_concat(static_cast<const uint16_t*>(a), ca, static_cast<const uint16_t*>(b), cb, nvox,
                  static_cast<uint16_t*>(y), static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_upcat(wr_ctx* ctx, const void* x, int ca, const void* b, int cb, int batch, int d, int h, int wd, void* y,
                 void* stream) {
    WR_ARG(ctx && x && b && y, "wr_net_upcat: null argument");
    WR_ARG(ca > 0 && cb > 0 && ca % 8 == 0 && cb % 8 == 0, "wr_net_upcat: bad channel counts");
    WR_ARG(batch > 0 && d > 0 && h > 0 && wd > 0 && d % 2 == 0 && h % 2 == 0 && wd % 2 == 0,
           "wr_net_upcat: d, h and w must be positive and even");
    WR_ARG(al16(x) && al16(b) && al16(y), "wr_net_upcat: buffers must be 16-byte aligned");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    NetGeom g{};
    g.d = d;
    g.h = h;
    g.w = wd;
    launch_upcat(static_cast<const uint16_t*>(x), ca, static_cast<const uint16_t*>(b), cb, g, batch,
                 static_cast<uint16_t*>(y), static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_split(wr_ctx* ctx, const void* g, int ca, int cb, long long nvox, void* da, void* db, void* stream) {
    WR_ARG(ctx && g && da && db, "wr_net_split: null argument");
    WR_ARG(ca > 0 && cb > 0 && ca % 8 == 0 && cb % 8 == 0 && nvox > 0, "wr_net_split: bad channel counts");
    WR_ARG(al16(g) && al16(da) && al16(db), "wr_net_split: buffers must be 16-byte aligned");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_split(static_cast<const uint16_t*>(g), ca, cb, nvox, static_cast<uint16_t*>(da), static_cast<uint16_t*>(db),
                 static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_mse(wr_ctx* ctx, const void* p, int cp, const float* t, int ct, long long nvox, double pscale,
               double scale, void* dp, float* loss, void* stream) {
    WR_ARG(ctx && p && t && dp && loss, "wr_net_mse: null argument");
    WR_ARG(nvox > 0 && ct > 0 && ct <= cp, "wr_net_mse: bad sizes");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_mse(static_cast<const uint16_t*>(p), cp, t, ct, nvox, static_cast<float>(pscale),
               static_cast<float>(scale), static_cast<uint16_t*>(dp), loss, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_adam(wr_ctx* ctx, float* w, const float* g, float* m, float* v, void* w16, long long n, double lr,
                double beta1, double beta2, double eps, int step, void* stream) {
    WR_ARG(ctx && w && g && m && v, "wr_net_adam: null argument");
    WR_ARG(n > 0 && step >= 1, "wr_net_adam: n must be positive and step at least 1");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_adam(w, g, m, v, static_cast<uint16_t*>(w16), n, static_cast<float>(lr), static_cast<float>(beta1),
                static_cast<float>(beta2), static_cast<float>(eps), step, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_to_bf16(wr_ctx* ctx, const float* in, long long n, void* out, void* stream) {
    WR_ARG(ctx && in && out && n > 0, "wr_net_to_bf16: bad argument");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_to_bf16(in, n, static_cast<uint16_t*>(out), static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_input(wr_ctx* ctx, const float* mov, const float* ref, long long n, void* out, void* stream) {
    WR_ARG(ctx && mov && ref && out && n > 0, "wr_net_input: bad argument");
    WR_ARG(al16(out), "wr_net_input: out must be 16-byte aligned");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* part;
    float* ms;
    WR_TRY(harness_small(c, &part, &ms));
    launch_mean_std(ref, n, part, ms, s);
    launch_net_input(mov, ref, n, ms, static_cast<uint16_t*>(out), s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_output_cp(wr_ctx* ctx, const void* p, int cp, int nz, int h, int w, int hs, int ws, int q, int ncx,
                     int ncy, int r, int level, double scale, float* cp_out, void* stream) {
    WR_ARG(ctx && p && cp_out, "wr_net_output_cp: null argument");
    WR_ARG(cp >= 3 && nz > 0 && h > 0 && w > 0 && hs >= h && ws >= w && ncx > 0 && ncy > 0 && r >= 0,
           "wr_net_output_cp: bad sizes");
    WR_ARG(q >= 0 && q < 16 && level >= 0 && level < WR_MAX_LEVELS, "wr_net_output_cp: bad level");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_net_output_cp(static_cast<const uint16_t*>(p), cp, h, w, hs, ws, q, nz, ncx, ncy, r, level,
                         static_cast<float>(scale), cp_out, static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_input_padded(wr_ctx* ctx, const float* mov, const float* ref, int d, int h, int w, int dp, int hp, int wp,
                        void* out, void* stream) {
    WR_ARG(ctx && mov && ref && out, "wr_net_input_padded: null argument");
    WR_ARG(d > 0 && h > 0 && w > 0 && dp >= d && hp >= h && wp >= w, "wr_net_input_padded: bad sizes");
    WR_ARG(al16(out), "wr_net_input_padded: out must be 16-byte aligned");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* part;
    float* ms;
    WR_TRY(harness_small(c, &part, &ms));
    launch_mean_std(ref, static_cast<long long>(d) * h * w, part, ms, s);
    launch_net_input_padded(mov, ref, d, h, w, dp, hp, wp, ms, static_cast<uint16_t*>(out), s);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_input_padded_batch(wr_ctx* ctx, const float* movs, int nb, const float* ref, int d, int h, int w, int dp,
                              int hp, int wp, void* out, void* stream) {
    WR_ARG(ctx && movs && ref && out, "wr_net_input_padded_batch: null argument");
    WR_ARG(nb > 0 && nb <= 65535, "wr_net_input_padded_batch: frame count out of range");
    WR_ARG(d > 0 && h > 0 && w > 0 && dp >= d && hp >= h && wp >= w, "wr_net_input_padded_batch: bad sizes");
    WR_ARG(al16(out), "wr_net_input_padded_batch: out must be 16-byte aligned");
    Context& c = ctx->c;
    WR_HIP_CHECK(hipSetDevice(c.device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* part;
    float* ms;
    WR_TRY(harness_small(c, &part, &ms));
    launch_mean_std(ref, static_cast<long long>(d) * h * w, part, ms, s);
    launch_net_input_padded(movs, ref, d, h, w, dp, hp, wp, ms, static_cast<uint16_t*>(out), s, nb);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_output_cp_batch(wr_ctx* ctx, const void* p, int nb, int cp, int nz, int dp, int h, int w, int hs, int ws,
                           int q, int ncx, int ncy, int r, int level, double scale, float* cp_out, void* stream) {
    WR_ARG(ctx && p && cp_out, "wr_net_output_cp_batch: null argument");
    WR_ARG(nb > 0 && nb <= 65535, "wr_net_output_cp_batch: frame count out of range");
    WR_ARG(cp >= 3 && nz > 0 && dp >= nz && h > 0 && w > 0 && hs >= h && ws >= w && ncx > 0 && ncy > 0 && r >= 0,
           "wr_net_output_cp_batch: bad sizes");
    WR_ARG(q >= 0 && q < 16 && level >= 0 && level < WR_MAX_LEVELS, "wr_net_output_cp_batch: bad level");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_net_output_cp(static_cast<const uint16_t*>(p), cp, h, w, hs, ws, q, nz, ncx, ncy, r, level,
                         static_cast<float>(scale), cp_out, static_cast<hipStream_t>(stream), nb,
                         static_cast<long long>(dp) * hs * ws * cp);
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

int wr_net_output(wr_ctx* ctx, const void* p, int cp, long long n, double scale, float* field, void* stream) {
    WR_ARG(ctx && p && field && n > 0 && cp >= 3, "wr_net_output: bad argument");
    WR_HIP_CHECK(hipSetDevice(ctx->c.device));
    launch_net_output(static_cast<const uint16_t*>(p), cp, n, static_cast<float>(scale), field,
                      static_cast<hipStream_t>(stream));
    WR_HIP_CHECK(hipGetLastError());
    return WR_OK;
}

}  // extern "C"
