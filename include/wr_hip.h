/*
 * wr_hip.h -- C ABI of the MI355X (gfx950) pyramid Lucas-Kanade registration
 * library (libwrhip.so).
 *
 * The reference has no FFI for this path: its backend "plugin" is the module
 * pair getMotion / correctMotion in
 *   src/wholistic_registration/utils/calFlow3d_Wei_v1.py:295-673 (getMotion)
 *   src/wholistic_registration/utils/calFlow3d_Wei_v1.py:676-706 (correctMotion)
 * resolved by name from wbi_registration_3d
 * (src/wholistic_registration/utils/registration.py:320-326).  Each entry point
 * below replaces one of those functions; the Python module
 * wholistic_registration_amd.calFlow3d_Wei_v1 binds them with ctypes and keeps
 * the reference's Python signatures.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer on the context's device, in
 *     C order (Z, Y, X): x is the fastest axis.  This is the memory order of
 *     the (Z, Y, X) volumes the reference harness transposes into its (x, y, z)
 *     views (registration.py:251,313), so no transpose is needed.
 *   - Motion fields are (Z, Y, X, 3) float32, component c displacing the
 *     reference's axis c (0 = x, 1 = y, 2 = z).
 *   - Masks are float32 (Z, Y, X); a value > 0 after resampling means
 *     "excluded" (calFlow3d_Wei_v1.py:422-423,473-474).  NULL = no mask.
 *   - All work is enqueued on `stream` (NULL = the default stream).
 *     wr_get_motion returns after the stream has drained, because its return
 *     value includes host-side error logs.  wr_correct_motion is asynchronous.
 *   - Return value: WR_OK or a negative WR_E* code; wr_last_error() gives the
 *     message (thread-local).
 */
#ifndef WR_HIP_H
#define WR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WR_OK 0
#define WR_E_ARG -1      /* invalid argument (reference: ValueError)           */
#define WR_E_KEY -2      /* missing parameter (reference: KeyError)            */
#define WR_E_HIP -3      /* HIP runtime error                                   */
#define WR_E_MEM -4      /* device allocation failed                           */

#define WR_F32 0         /* input volume dtypes */
#define WR_U16 1
#define WR_F64 2         /* accepted by wr_imresize and wr_spline_coef */
#define WR_U8 3          /* byte masks: wr_spline_coef only */

#define WR_MAX_LEVELS 16

typedef struct wr_ctx wr_ctx;

/* Solver parameters: the `option` keys getMotion reads
 * (calFlow3d_Wei_v1.py:317-324,429,453,523). */
typedef struct wr_params {
    int layer;              /* option["layer"]: levels layer..0 are visited     */
    int iter;               /* option["iter"]: iterations per level             */
    int r;                  /* option["r"]: patch half-width, (2r+1)^2 x 1      */
    int timing;             /* 1 = time every LK pass with HIP events         */
    double zRatio;          /* option["zRatio"]                                 */
    double movRange;        /* option["movRange"]: per-update step cap          */
    double smoothPenalty;   /* option["smoothPenalty"] (already normalised)     */
    double tol;             /* option["tol"]: error-change stop tolerance       */
    /* Optional forced schedule (NULL = the data-dependent stop rules of
     * calFlow3d_Wei_v1.py:515-530,651-654).  Indexed by level 0..layer:
     * forced_records[L] error records, forced_updates[L] updates. */
    const int* forced_records;
    const int* forced_updates;
    int tile_patches_x;     /* 0 = autotuned; patches per workgroup along x   */
    int stop_kernels;       /* 0 = automatic; 1 = the stop rules always run in
                               their own single-workgroup kernels (testing)   */
    int precision;          /* storage of the pyramid level images the LK
                               iterations read: 0 = float32 (the reference's
                               values), 1 = bfloat16, 2 = float16 (moving
                               image and its central differences); every
                               interpolation and sum stays float32; the
                               per-patch solve is float64 with precision 0,
                               float32 with 1 and 2 except for systems whose
                               condition number exceeds 1e6 (float64) */
    int reuse_template;     /* 1 = `ref` and mask_ref hold the same data as in
                               the previous wr_get_motion on this context, with
                               the same ref and mask_ref pointers, size, dtype,
                               layer and precision: their pyramid levels are not
                               rebuilt (frames of a time series against one
                               template); a mismatch of those is an error */
    int init_nx;            /* with init_ny: 0, 0 = motion_init is (nz, ny, nx, 3); else
                               motion_init is (nz, init_ny, init_nx, 3), still
                               in full-resolution voxel units: the reference's
                               warm start imresizes option["motion"] of any
                               x-y size to the coarsest level and divides by
                               the full-to-level size ratio
                               (calFlow3d_Wei_v1.py:362-377) */
    int init_ny;
    const int* tile_patches_level;  /* NULL = tile_patches_x on every level; else
                               [layer + 1] entries, the patches per
                               workgroup along x of level L (0 = the
                               built-in rule): what the tile autotuner
                               picks per level */
    int init_cp;            /* 1 = motion_init is the coarsest level's control-
                               point field itself: 3 planes (x, y, z
                               components) of nz * ncy * ncx float32 in that
                               level's voxel units, ncx = len(range(r, nx_L - 1,
                               2r + 1)) (calFlow3d_Wei_v1.py:434-436); the level
                               starts from its bilinear interpolant with no
                               dense start field.  No counterpart in the
                               reference: the start of the learned
                               initialisation, which skips resampling a dense
                               field that the first update replaces anyway */
    int resample_start;     /* 0 = automatic; 1 = every level below the
                               coarsest starts from the resampler's upsample
                               of the previous level's dense end field,
                               written out first (testing: the automatic path
                               forms the same values from the control points) */
    int maxdiff_gather;     /* 0 = automatic; 1 = the |delta motion| maximum
                               of a level's first update always runs in the
                               gather kernel, not the LDS-staged one (testing) */
    void* init_event;       /* NULL, or a hipEvent_t recorded where motion_init
                               is produced (e.g. on another stream): the
                               iterations wait for it, the pyramid build on the
                               context's side stream does not, so the level
                               images are made while motion_init is still being
                               computed.  No counterpart in the reference. */
} wr_params;

/* Per-level iteration log, filled on the host when wr_get_motion returns
 * (the lists of error_log["layer_L"], calFlow3d_Wei_v1.py:329-335,505-507,643).
 * Arrays are [level][iter] with level 0..layer, row stride `iter`. */
typedef struct wr_log {
    int records[WR_MAX_LEVELS];   /* len(currentError) per level            */
    int updates[WR_MAX_LEVELS];   /* len(max_diff_motion) per level         */
    int nans[WR_MAX_LEVELS];      /* NaN components zeroed by the solve     */
    double* diff_error;           /* host, (layer+1)*iter, may be NULL      */
    double* penalty_error;
    double* current_error;
    double* max_diff_motion;
    double final_error;           /* currentError returned by getMotion     */
} wr_log;

/* Create / destroy a context bound to HIP device `device`.  The context owns
 * the solver workspace, grown on demand and reused across calls. */
wr_ctx* wr_create(int device);
void wr_destroy(wr_ctx* ctx);

/* Thread-local message of the last failed call. */
const char* wr_last_error(void);

/* Replaces getMotion(dat_mov, dat_ref, option)
 * (calFlow3d_Wei_v1.py:295-673).
 *   mov, ref      (Z,Y,X) volumes of dtype `dtype` (WR_F32 or WR_U16)
 *   mask_ref      (Z,Y,X) float32 or NULL   (option["mask_ref"])
 *   mask_mov      (Z,Y,X) float32 or NULL   (option["mask_mov"])
 *   motion_init   (Z,Y,X,3) float32 or NULL (option["motion"])
 *   motion_out    (Z,Y,X,3) float32, the level-0 motion field
 *   coords_out    (Z,Y,X,3) float32 or NULL: motion + voxel grid
 *   log           host struct or NULL */
int wr_get_motion(wr_ctx* ctx, const void* mov, const void* ref, int dtype,
                  int nx, int ny, int nz,
                  const float* mask_ref, const float* mask_mov,
                  const float* motion_init, const wr_params* params,
                  float* motion_out, float* coords_out, wr_log* log,
                  void* stream);

/* getMotion of `frames` moving volumes against one template, the frames of a
 * time series (the serial loop of registration.py:299-346, which calls
 * getMotion once per frame).  The frames go through one launch sequence:
 * every kernel of a level-iteration is one launch over all frames, each frame
 * with its own control-point field, stop rules and logs, so every frame's
 * result equals a wr_get_motion call on it alone, bit for bit.
 *   mov           frame f at mov + f * mov_stride elements, (Z,Y,X) each
 *   mask_mov      NULL, or frame f at mask_mov + f * mask_mov_stride floats
 *   motion_init   NULL, or frame f at motion_init + f * init_stride floats
 *                 (the layout params->init_nx / init_ny / init_cp select)
 *   motion_out    (frames, Z, Y, X, 3) float32
 *   warp_src      NULL, or frame f's volume to warp (correctMotion) at
 *                 warp_src + f * warp_stride elements of warp_dtype; the
 *                 warps go to warped_out, (frames, Z, Y, X) float32
 *   logs          `frames` host structs, or NULL: then the call returns once
 *                 its work is enqueued, without waiting for the stream (no
 *                 logs, no timing); a call with the geometry, frames and
 *                 buffers of an earlier call on the context allocates and
 *                 uploads nothing and can be captured in a HIP graph
 * ref, mask_ref, params and stream as for wr_get_motion (params->reuse_template
 * applies to the template).  Device memory: one workspace per frame, about
 * 21 bytes per voxel (11 more with 16-bit level storage). */
int wr_get_motion_batch(wr_ctx* ctx, const void* mov, long long mov_stride,
                        const void* ref, int dtype, int nx, int ny, int nz,
                        int frames, const float* mask_ref,
                        const float* mask_mov, long long mask_mov_stride,
                        const float* motion_init, long long init_stride,
                        const wr_params* params, float* motion_out,
                        const void* warp_src, int warp_dtype,
                        long long warp_stride, float* warped_out,
                        wr_log* logs, void* stream);

/* wr_get_motion followed by wr_correct_motion(src, motion_out) -- the
 * harness's getMotion / correctMotion pair (registration.py:320-325) -- with
 * the warp formed in the kernel that writes the level-0 field, so the field
 * is not read back.  `src` (Z,Y,X) of `src_dtype` (WR_F32 or WR_U16) is the
 * volume correctMotion warps (the moving volume, or another channel of it);
 * warped_out (Z,Y,X) float32.  Both outputs equal the two separate calls'
 * bit for bit. */
int wr_get_motion_warp(wr_ctx* ctx, const void* mov, const void* ref, int dtype,
                       int nx, int ny, int nz,
                       const float* mask_ref, const float* mask_mov,
                       const float* motion_init, const wr_params* params,
                       float* motion_out, const void* src, int src_dtype,
                       float* warped_out, wr_log* log, void* stream);

/* Replaces correctMotion(data_raw, motion_field)
 * (calFlow3d_Wei_v1.py:676-706): out(p) = trilinear(in, p + motion(p)),
 * clamp to edge, float32 out. */
int wr_correct_motion(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz,
                      const float* motion, float* out, void* stream);

/* Replaces getSmPnltNormFctr(dat_ref, option)
 * (src/wholistic_registration/utils/preprocess.py:42-66), called by the
 * harness before getMotion (registration.py:295):
 *   factor = (mean(((ref[x+1]-ref[x-1])/2)^2) + mean(((ref[y+1]-ref[y-1])/2)^2)) / 2
 * over the voxels whose mask_ref is 0 (interior x, resp. y, only), in the
 * input dtype's arithmetic (uint16 differences wrap modulo 2^16).
 *   ref       (Z,Y,X) volume, dtype WR_F32, WR_U16 or WR_F64
 *   mask_ref  (Z,Y,X) float32, nonzero = excluded, or NULL
 *   factor    host double; NaN when a direction has no unmasked voxel.
 * Synchronises the stream. */
int wr_smooth_penalty_factor(wr_ctx* ctx, const void* ref, int dtype, int nx, int ny, int nz,
                             const float* mask_ref, double* factor, void* stream);

/* Allocation-free variant of wr_smooth_penalty_factor.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and arithmetic, except that the factor goes to
 * DEVICE memory (`factor`, one double) and the call is asynchronous on
 * `stream` (no host synchronisation).  Its row partials live in a fixed
 * per-GPU scratch array of the library (WR_NOALLOC_MAX_ROWS rows of 4
 * doubles, 8 MiB, held until the process ends), resolved by wr_create:
 * ny * nz above WR_NOALLOC_MAX_ROWS is rejected with WR_E_ARG.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (several
 * streams or threads) race; calls in stream order do not. */
#define WR_NOALLOC_MAX_ROWS 262144
int wr_smooth_penalty_factor_noalloc(wr_ctx* ctx, const void* ref, int dtype, int nx, int ny, int nz,
                                     const float* mask_ref, double* factor, void* stream);

/* Pyramid level image as the reference builds it: imresize(vol,
 * output_shape=(int(nx/2^level), int(ny/2^level), nz)) (bicubic,
 * antialiased, imresize.py:148-191; calFlow3d_Wei_v1.py:341-351), stored as
 * float32.  out holds (nz, ny_L, nx_L). */
int wr_pyramid_level(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz,
                     int level, float* out, void* stream);

/* MATLAB-style resize of the x and y axes, z kept: imresize(vol,
 * output_shape=(nx_out, ny_out, nz), method) (imresize.py:148-191), float64
 * result in (nz, ny_out, nx_out).  method: 0 = bicubic, 1 = bilinear. */
int wr_imresize(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz,
                int nx_out, int ny_out, int method, double* out, void* stream);

/* ---- automatic masks and error metric of the registration harness
 * (wbi_registration_3d, src/wholistic_registration/utils/registration.py:
 * 259-260, 311-312, 328-329).  Volumes and masks are (Z,Y,X) device arrays. */

/* Replaces robust_mean_std(data, percentile)
 * (src/wholistic_registration/utils/preprocess.py:69-84) on float32 data (uint16
 * data is converted to float32, as getMask does before calling it):
 *   out[0] = numpy.percentile(data, percentile)   (linear method, float32)
 *   out[1] = mean of the values <= out[0]
 *   out[2] = std  of the values <= out[0]
 * out is host memory; synchronises the stream. */
int wr_robust_mean_std(wr_ctx* ctx, const void* data, int dtype, long long n, double percentile,
                       float* out, void* stream);

/* Replaces getMask(dat_mov, thres_factor)
 * (src/wholistic_registration/utils/mask.py:59-111): out = 1 where the
 * float32 z-score |(v - mu) / sigma| > thres_factor, after a 3x3 in-plane
 * binary opening and a 3x3 binary dilation; mu, sigma from robust_mean_std
 * at the 95th percentile, or from the host array mu_sigma[2] when not NULL.
 *   vol  (Z,Y,X) WR_F32 or WR_U16;   out (Z,Y,X) uint8 (0 / 1). */
int wr_get_mask(wr_ctx* ctx, const void* vol, int dtype, int nx, int ny, int nz, double thres_factor,
                const float* mu_sigma, uint8_t* out, void* stream);

/* Replaces bwareafilt3_wei(vol, size_range)
 * (src/wholistic_registration/utils/mask.py:8-56): keeps the 26-connected
 * components of the nonzero voxels of `in` whose voxel count lies in
 * [size_min, size_max].  Output as uint8 (out_u8) and/or float32 0 / 1
 * (out_f32, the form getMotion consumes); either may be NULL, not both.
 * n_components (host, may be NULL) receives the number of components, which
 * synchronises the stream. */
int wr_bwareafilt3(wr_ctx* ctx, const uint8_t* in, int nx, int ny, int nz, double size_min,
                   double size_max, uint8_t* out_u8, float* out_f32, long long* n_components,
                   void* stream);

/* Replaces np.mean((a - b) ** 2) of the harness's initial and eventual
 * errors (src/wholistic_registration/utils/registration.py:328-329) in
 * numpy's dtypes: uint16 - uint16 wraps modulo 2^16 (difference and square),
 * float32 steps are rounded in float32, anything with float64 is float64;
 * summed in float64.  out is host memory; synchronises the stream. */
int wr_mean_sq_diff(wr_ctx* ctx, const void* a, int dtype_a, const void* b, int dtype_b, long long n,
                    double* out, void* stream);

/* ---- channel arithmetic and reference picking of the harness */

/* Replaces transform(image, k, method)
 * (src/wholistic_registration/utils/registration.py:26-36):
 * out = k * f(in), f = identity (method 0), sqrt (1), log2(1 + x) (2),
 * log10(1 + x) (3), as float32.  uint16 input: 1 + x wraps modulo 2^16 and f
 * is evaluated in float32 (numpy's types for uint16 arrays); with method 0
 * an integer k (k_is_int = 1, 0 <= k <= 65535) multiplies modulo 2^16, a
 * float k in float64 (both stored as float32). */
int wr_channel_transform(wr_ctx* ctx, const void* in, int dtype, long long n, double k, int method,
                         int k_is_int, float* out, void* stream);

/* np.mean(image), np.std(image) of n float32 values (the statistics of
 * normalize_std, preprocess.py:288-293, and of the calcium reference,
 * registration.py:244-245): out[0] = mean, out[1] = std, host; synchronises
 * the stream. */
int wr_mean_std(wr_ctx* ctx, const float* in, long long n, float* out, void* stream);

/* Replaces normalize_std(mean, std, image)
 * (src/wholistic_registration/utils/preprocess.py:288-293):
 * out = ((in - mean(in)) / std(in)) * std + mean, float32. */
int wr_normalize_std(wr_ctx* ctx, const float* in, long long n, double mean, double std_, float* out,
                     void* stream);

/* dat_mem + dat_ca_tran (registration.py:306): out = float32(a) + b. */
int wr_add_volumes(wr_ctx* ctx, const void* a, int dtype_a, const float* b, long long n, float* out,
                   void* stream);

/* frames[top_indices].mean(axis=0) (reference.py:97 and registration.py:237):
 * the mean of `count` device volumes of n voxels, summed in the order given.
 * uint16 frames: out is float64 (exact sums); float32 frames: out is float32
 * (float32 running sum, then divided by count, as numpy). */
int wr_frames_mean(wr_ctx* ctx, const void* const* frames, int count, int dtype, long long n, void* out,
                   void* stream);

/* The correlation core of pick_initial_reference
 * (src/wholistic_registration/utils/reference.py:39-78): slab holds nt
 * frames' (nzs, ny, nx) slabs; each is sampled at [::ds, ::ds], converted to
 * float32, demeaned by its own mean, and cc (host, nt * nt) receives the Gram
 * matrix of the demeaned rows.  Synchronises the stream. */
int wr_frame_gram(wr_ctx* ctx, const void* slab, int dtype, int nt, int nzs, int ny, int nx, int ds,
                  double* cc, void* stream);

/* ---- fixed-plane projections of the cross-resolution path */

/* Replaces the max splat of project_coords_to_fixed_planes_gpu
 * (src/wholistic_registration/utils/calFlowCrossResolution.py:480-568 kernel,
 * :878-1141 host side) for given values: n samples (values[n], coords[n][3]
 * as reference x, y, z) are splatted onto the nplanes planes target_z
 * (device) with |z - z0| <= z_window: to the nearest pixel (nearest = 1) or
 * to the footprint floor(x) - r .. floor(x) + 1 + r (same in y), pixels mapped
 * to the output grid by integer division by downsample_xy.  out (device,
 * nplanes x xout x yout, y fastest, xout = ceil(xref / ds)) keeps each
 * pixel's maximum; with fill = 1 empty pixels get fill_value, else -inf. */
int wr_project_planes_max(wr_ctx* ctx, const float* values, const float* coords, long long n,
                          const float* target_z, int nplanes, int xref, int yref, int downsample_xy,
                          double z_window, int nearest, int extra_radius, int fill, double fill_value,
                          float* out, void* stream);

/* Replaces the weighted splat of project_coords_to_fixed_planes_weighted_gpu
 * (calFlowCrossResolution.py:571-669 kernel, :700-875 host side) for given
 * values: z weight mode 0 = hard, 1 = triangular, 2 = Gaussian (z_sigma <= 0:
 * z_window / 2) times bilinear weights at (x / ds, y / ds); samples with x
 * outside [0, xref - 1] or y outside [0, yref - 1] are dropped.
 * weight_out = sum of weights, out = sum(v * w) / sum(w) where sum(w) > eps,
 * else fill_value (both device, nplanes x xout x yout). */
int wr_project_planes_weighted(wr_ctx* ctx, const float* values, const float* coords, long long n,
                               const float* target_z, int nplanes, int xref, int yref,
                               int downsample_xy, double z_window, double z_sigma, int mode, double eps,
                               double fill_value, float* out, float* weight_out, void* stream);

/* Launch statistics of the last wr_get_motion on this context, for
 * measurement: LK passes launched per level, the LK kernel's tile shape and
 * grid size, and (when params->timing was 1) the number and summed HIP-event
 * duration of the passes that did work. */
typedef struct wr_stats {
    int lk_launches[WR_MAX_LEVELS];
    int tile_cols;                /* owned columns per workgroup            */
    int tile_rows;                /* owned rows per workgroup               */
    int lk_blocks[WR_MAX_LEVELS]; /* workgroups per LK pass                  */
    int lk_exec[WR_MAX_LEVELS];   /* LK passes that did work (timing = 1)   */
    double lk_ms[WR_MAX_LEVELS];  /* summed duration of those passes, ms    */
    int solve_fp64[WR_MAX_LEVELS];/* with 16-bit storage (float32 solves):
                                     systems solved in float64 because their
                                     condition number exceeded 1e6            */
    int start_from_cp[WR_MAX_LEVELS];/* 1 = the level's start field was formed
                                     from the previous level's control points
                                     (no dense end field written)             */
} wr_stats;
int wr_get_stats(wr_ctx* ctx, wr_stats* out);

/* ---- Learned initialisation.  No counterpart in the reference: these are
 * the layers of a 3-D U-Net whose predicted field enters getMotion through
 * its own warm start, option["motion"] (calFlow3d_Wei_v1.py:362-377;
 * wr_params.init_nx / init_ny take it at the network's x-y size).
 * Tensors are device, channels-last (batch, d, h, w, c) bfloat16 (uint16
 * bits) unless float32 is stated; weights (cout, 27, cin), tap
 * (kd * 3 + kh) * 3 + kw at offset (kd - 1, kh - 1, kw - 1); padding 1.
 * Every call is asynchronous on `stream`.  Buffers read or written 8
 * channels at a time must be 16-byte aligned. */

/* y = conv3x3x3(x, w) + bias (float32, nullable), then LeakyReLU(slope) when
 * leaky = 1; float32 accumulation (MFMA); cin % 8 == 0, cout % 16 == 0. */
int wr_net_conv3d(wr_ctx* ctx, const void* x, const void* w, const float* bias, void* y, int batch, int d, int h,
                  int wd, int cin, int cout, int leaky, double slope, void* stream);
/* dx (batch, d, h, w, cin) = gradient of the convolution's input for dy
 * (batch, d, h, w, cout); cin, cout multiples of 16. */
int wr_net_conv3d_bwd_data(wr_ctx* ctx, const void* dy, const void* w, void* dx, int batch, int d, int h, int wd,
                           int cin, int cout, void* stream);
/* dw (float32, cout x 27 x cin) += gradient of the weights, db (float32, cout,
 * nullable) += gradient of the bias; w % 8 == 0.  Float32 atomics: the
 * summation order varies between runs. */
int wr_net_conv3d_bwd_weight(wr_ctx* ctx, const void* dy, const void* x, float* dw, float* db, int batch, int d,
                             int h, int wd, int cin, int cout, void* stream);
/* grad *= (y > 0 ? 1 : slope) over n values (the LeakyReLU backward from its output). */
int wr_net_leaky_bwd(wr_ctx* ctx, void* grad, const void* y, long long n, double slope, void* stream);
/* y (batch, d/2, h/2, w/2, c) = scale * sum of the 2x2x2 children in x
 * (batch, d, h, w, c): average pooling (scale 1/8) and the backward of
 * nearest upsampling (scale 1). */
int wr_net_down2(wr_ctx* ctx, const void* x, void* y, int batch, int d, int h, int wd, int c, double scale,
                 void* stream);
/* y (batch, d, h, w, c) = (y +, when accumulate = 1) scale * the parent voxel
 * in x (batch, d/2, h/2, w/2, c): nearest upsampling (scale 1) and the
 * backward of average pooling (scale 1/8). */
int wr_net_up2(wr_ctx* ctx, const void* x, void* y, int batch, int d, int h, int wd, int c, double scale,
               int accumulate, void* stream);
/* y[v] = (a[v], b[v]) over nvox voxels; wr_net_split is its backward. */
int wr_net_concat(wr_ctx* ctx, const void* a, int ca, const void* b, int cb, long long nvox, void* y, void* stream);
/* y (batch, d, h, w, ca + cb) = (the nearest 2x upsampling of x (batch, d/2,
 * h/2, w/2, ca), b (batch, d, h, w, cb)): wr_net_up2 (scale 1) and
 * wr_net_concat in one pass, without the upsampled tensor. */
int wr_net_upcat(wr_ctx* ctx, const void* x, int ca, const void* b, int cb, int batch, int d, int h, int wd, void* y,
                 void* stream);
int wr_net_split(wr_ctx* ctx, const void* g, int ca, int cb, long long nvox, void* da, void* db, void* stream);
/* *loss (float32, device) += scale * sum over voxels and channels c < ct of
 * (pscale * p[v][c] - t[v][c])^2, t float32 (nvox, ct); dp (nvox, cp) = the
 * gradient with respect to p (zero in channels ct..cp-1). */
int wr_net_mse(wr_ctx* ctx, const void* p, int cp, const float* t, int ct, long long nvox, double pscale,
               double scale, void* dp, float* loss, void* stream);
/* One Adam step (Kingma & Ba 2015) on n float32 parameters w with gradient
 * g and moments m, v; step counts from 1; w16 (nullable) receives bfloat16
 * copies of w. */
int wr_net_adam(wr_ctx* ctx, float* w, const float* g, float* m, float* v, void* w16, long long n, double lr,
                double beta1, double beta2, double eps, int step, void* stream);
int wr_net_to_bf16(wr_ctx* ctx, const float* in, long long n, void* out, void* stream);
/* out (n, 8) = ((mov - mean) / std, (ref - mean) / std, 0, ...) with the
 * mean and standard deviation of ref's n float32 values. */
int wr_net_input(wr_ctx* ctx, const float* mov, const float* ref, long long n, void* out, void* stream);
/* The same input on a padded (dp, hp, wp) grid >= the level images' (d, h,
 * w): voxels past the volume repeat its nearest edge voxel; the statistics
 * are those of the d * h * w template voxels. */
int wr_net_input_padded(wr_ctx* ctx, const float* mov, const float* ref, int d, int h, int w, int dp, int hp, int wp,
                        void* out, void* stream);
/* wr_net_input_padded of nb frames in one call: movs (nb, d, h, w) float32
 * against the one template ref (its statistics once); out (nb, dp, hp, wp,
 * 8) bfloat16.  No counterpart in the reference (a learned-start helper). */
int wr_net_input_padded_batch(wr_ctx* ctx, const float* movs, int nb, const float* ref, int d, int h, int w, int dp,
                              int hp, int wp, void* out, void* stream);
/* field (n, 3) float32 = scale * p[:, 0:3] of p (n, cp). */
int wr_net_output(wr_ctx* ctx, const void* p, int cp, long long n, double scale, float* field, void* stream);
/* The network's field (p: nz planes of hs x ws voxels of cp channels, of
 * which h x w is pyramid level q of the full volume and the rest edge
 * padding; scale * channels 0..2 in full-resolution units) sampled
 * bilinearly at the control points of pyramid level `level` (ncx x ncy per
 * plane, half-width r), in that level's units: cp_out is the 3-plane field
 * wr_params.init_cp takes. */
int wr_net_output_cp(wr_ctx* ctx, const void* p, int cp, int nz, int h, int w, int hs, int ws, int q, int ncx,
                     int ncy, int r, int level, double scale, float* cp_out, void* stream);
/* wr_net_output_cp of nb frames in one call: p (nb, dp, hs, ws, cp) with dp
 * >= nz stored planes per frame (planes past nz are padding), cp_out (nb,
 * 3, nz, ncy, ncx).  No counterpart in the reference (a learned-start
 * helper). */
int wr_net_output_cp_batch(wr_ctx* ctx, const void* p, int nb, int cp, int nz, int dp, int h, int w, int hs, int ws,
                           int q, int ncx, int ncy, int r, int level, double scale, float* cp_out, void* stream);

/* ------------------------------------------------------------------------
 * Cross-resolution registration (getMotion_v2,
 * src/wholistic_registration/utils/calFlowCrossResolution.py:2470-2663).
 *
 * A moving volume on a coarse-z grid is registered to a reference volume on
 * its own (finer) grid: moving voxel p maps to reference coordinates
 * phase(p) = phase0(p) + motion(p).  Volumes are dense (z, y, x) device
 * arrays (the reference's (x, y, z) arrays transposed); fields (phase,
 * motion) are 3 planes (x, y, z components) of nz * ny * nx float32.  Level
 * images and residuals are float64, as in the reference.  The Python drop-in
 * (wholistic_registration_amd/calFlowCrossResolution.py) composes these
 * calls exactly as the reference composes its functions.
 * ------------------------------------------------------------------------ */

/* Cubic B-spline coefficients of a volume for map_coordinates(order=3,
 * mode="nearest") (the interpolator of generate_continuous_H_gpu, :463-477):
 * the volume padded by 12 replicated voxels on every side, then the cubic
 * B-spline prefilter along x, y and z in float64 (gain 6, pole sqrt(3) - 2,
 * half-sample symmetric initial conditions).  coef: (nz + 24) * (ny + 24) * (nx + 24)
 * doubles.  dtype: WR_F64, WR_F32 or WR_U8. */
int wr_spline_coef(wr_ctx* ctx, const void* in, int dtype, int nx, int ny, int nz, double* coef, void* stream);

/* Values of that interpolant at n points (n, 3) float32 (x, y, z), the z
 * coordinate divided by z_ratio first (generate_continuous_H_gpu's H and
 * apply_H_to_matrix_gpu, :463-477, :1144-1153); out: n doubles. */
int wr_spline_sample(wr_ctx* ctx, const double* coef, int nx, int ny, int nz, const float* pts, long long n,
                     double z_ratio, double* out, void* stream);

/* One pyramid layer's optimisation (optimize_layer_cross_resolution,
 * :2031-2237). */
typedef struct wr_xres_layer {
    int nx;                      /* moving level (x, y, z extents)                */
    int ny;
    int nz;
    int rx;                      /* reference level                               */
    int ry;
    int rz;
    int r;                       /* patch half-width                              */
    int iters;                   /* option["iter"]                                */
    double smooth_penalty_sum;   /* smoothPenalty * (2r + 1)^2                     */
    double mov_range;            /* option["movRange"]                             */
    double z_ratio_hr;           /* zRatio_HR / 2^layer                            */
    double tol;                  /* option["tol"]                                  */
} wr_xres_layer;

/* Iteration record of one optimisation.  reason: 0 last iteration, 1 error
 * increased repeatedly, 2 error change below tol, 3 motion update below tol,
 * 4 forced schedule.  The three arrays (host, `iters` entries each, may be
 * NULL) receive each iteration's diffError, penaltyError, currentError. */
typedef struct wr_xres_trace {
    int records;
    int updates;
    int reason;
    double* diff_error;
    double* penalty_error;
    double* current_error;
    double final_diff;           /* the final recompute (:2198-2222)               */
    double final_penalty;
    double final_error;
} wr_xres_trace;

/* mov_l (nz, ny, nx) and ref_l (rz, ry, rx) float64 level images; ref_coef /
 * mask_coef their spline coefficients (mask_coef: of the reference
 * exclusion mask, NULL = nothing excluded); mask_mov (nz, ny, nx) bytes,
 * nonzero = excluded (NULL = none); phase0 / motion_init 3-plane fields.
 * forced: NULL, or {records, updates} replacing the data-dependent stop tests.
 * Outputs: motion, phase (3 planes), sampled (reference at the phase),
 * residual (the masked 3x3 mean residual), valid (1 = voxel counts), and
 * first_residual (NULL or the residual of the first iteration, for the
 * 'highresidual' exclusion).  The host waits for each iteration's errors. */
int wr_xres_optimize(wr_ctx* ctx, const wr_xres_layer* layer, const double* mov_l, const double* ref_l,
                     const double* ref_coef, const double* mask_coef, const uint8_t* mask_mov,
                     const float* phase0, const float* motion_init, const int* forced, float* motion,
                     float* phase, double* sampled, double* residual, uint8_t* valid, double* first_residual,
                     wr_xres_trace* trace, void* stream);

/* compute_valid_mask_on_moving_grid (:2006-2028) at a given phase: valid
 * (nullable) and its count (host). */
int wr_xres_valid(wr_ctx* ctx, const float* phase, int nx, int ny, int nz, const double* mask_coef, int rx,
                  int ry, int rz, const uint8_t* mask_mov, uint8_t* valid, long long* count, void* stream);

/* Wrong-region detection of one layer (correct_wrong_regions_one_layer
 * :2306-2351 with get_local_error_on_control_points :1190-1222, metric 0
 * 'mse' or 1 'mae',
 * build_bad_region_mask_from_cp_error :1570-1787, the 'highresidual'
 * selection :2329-2347 when mode = 1 (first_residual and q used), and
 * & valid): bad (nz, ny, nx) bytes; *n_regions = kept components. */
int wr_xres_bad_regions(wr_ctx* ctx, const double* residual, const uint8_t* valid, const double* first_residual,
                        int nx, int ny, int nz, int r, int metric, double mad_threshold, int min_size, int mode,
                        double q, uint8_t* bad, int* n_regions, void* stream);

/* build_reference_trap_mask_from_bad_moving_fast_roi (:1413-1567) united with
 * the reference level's exclusion mask: corrected (rz, ry, rx) bytes =
 * mask_ref_l (NULL = none) | trap; *trap_count = voxels of the trap. */
int wr_xres_trap_mask(wr_ctx* ctx, const uint8_t* bad, const float* phase, int nx, int ny, int nz,
                      const double* ref_l, int rx, int ry, int rz, const uint8_t* mask_ref_l, double z_ratio_ref,
                      double expand_radius_xy, double sigma, double intensity_k, uint8_t* corrected,
                      long long* trap_count, void* stream);

/* FindInitZ_stack_global_fixed_spacing (:370-460): the reference z of each of
 * nk moving slices spaced delta * direction apart, chosen by the summed
 * weighted ZNCC (Hann window when use_hann) of the slices' Sobel gradient
 * magnitudes (raw values when use_gradient = 0) against the nz reference
 * slices.  mov (nk, w, h) and ref (nz, w, h) float32 device stacks (the
 * reference's (h, w, k) arrays transposed); z_init (host, nk) the chosen
 * z indices truncated to uint16 as the reference returns them; scores_out
 * (host, nk * nz, nullable) the ZNCC matrix; best_score / best_z0 (host,
 * nullable).  WR_E_ARG with the reference's message when no start fits. */
int wr_find_init_z(wr_ctx* ctx, const float* mov, int nk, const float* ref, int nz, int h, int w, double delta,
                   int use_gradient, int use_hann, int direction, double eps, uint16_t* z_init, float* scores_out,
                   double* best_score, int* best_z0, void* stream);

/* Number of nonzero bytes of m (int(cp.sum(valid_mask)), :2375-2378). */
int wr_count_u8(wr_ctx* ctx, const uint8_t* m, long long n, long long* count, void* stream);

/* Element-wise helpers of the layer set-up (:2540-2558, :1917-1992):
 * out[i] = float32(in[i] * mul / div) (the level's start fields);
 * mask[i] = in[i] > thr (resample_exclude_mask); the identity phase
 * (x, y, z * z_mul) of a level. */
int wr_xres_scale(wr_ctx* ctx, const double* in, long long n, double mul, double div, float* out, void* stream);
int wr_xres_threshold(wr_ctx* ctx, const double* in, long long n, double thr, uint8_t* out, void* stream);
int wr_xres_phase_identity(wr_ctx* ctx, int nx, int ny, int nz, double z_ratio, double z_ratio_hr, float* phase,
                           void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WR_HIP_H */
