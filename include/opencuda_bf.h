/*
 * opencuda_bf.h -- C ABI of the MI355X bilateral filter (libopencuda_bf.so).
 *
 * Plain C: pointers, sizes and status codes only.  Every call is
 * stream-ordered and asynchronous on the HIP stream passed as `stream`
 * (a hipStream_t, NULL = the null stream) unless it says otherwise.  Status
 * codes are the ErrorCode.h values (0 = NO_ERROR, negative = error); a
 * failing HIP runtime call is reported as CUDA_ERROR (-7) and its text is
 * kept for bf_last_error().  The library never allocates or frees caller
 * images.
 *
 * The C++ object model (Image.h, Template.h, TemplateFactory.h,
 * BilateralFilter.h) is implemented on top of these entry points and keeps
 * the reference's class API; FFI callers (ctypes, cgo, JNI) bind this file.
 * Paths in citations are relative to the LitLeo/OpenCUDA tree.
 */
#ifndef OPENCUDA_BF_H
#define OPENCUDA_BF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Filter modes.
 * BF_MODE_REF: the reference kernel as written (okano_3_0/BilateralFilter.cu
 *   :71-199): each 4-row group reads the window of its first row
 *   (:119-120,:137,:154,:171) and runs (2r+1)^2+1 taps (:78,:113).  This is
 *   what BilateralFilter::doFilter computes.
 * BF_MODE_STD: each pixel reads its own (2r+1)^2 window.  No reference
 *   counterpart; the learned surrogate approximates this mode. */
#define BF_MODE_REF 0
#define BF_MODE_STD 1

/* Largest radius (okano_3_0/BilateralFilter.h:26, DEF_FILTER_RANGE). */
#define BF_MAX_RADIUS 20

/* A ROI sub-image in device memory: the view ImageBasicOp::roiSubImage
 * builds (okano_3_0/Image.h:218-258): data points at the ROI origin,
 * width/height are the ROI size, pitch is the row stride in bytes. */
typedef struct bf_image_view {
    unsigned char *data;
    int32_t width;
    int32_t height;
    int64_t pitch;
} bf_image_view;

/* Device-resident filter tables: the spatial template (offsets + weights)
 * and the 256-entry range table.  Opaque. */
typedef struct bf_lut_s *bf_lut;

/* Spatial ("GAUSS") template of edge `size` (odd, >= 3): (dy, dx) offsets in
 * ring order and their weights exp((float)(-(dy*dy+dx*dx)) / (2*s*s)), with
 * (0, 0) first at weight 1.  Replaces _gaussCreateTemplate
 * (okano_3_0/TemplateFactory.cu:1092-1168; legality :1057-1076).
 * offsets holds 2*size*size ints, weights size*size floats.  Returns the
 * point count, or INVALID_DATA. */
int bf_gauss_template(float sigma, int32_t size, int32_t *offsets, float *weights);

/* Range ("EUCLIDE") table: weights[d] = exp((float)(-(d*d)) / (2*s*s)),
 * d = 0..size-1.  Replaces _euclideCreateTemplate
 * (okano_3_0/TemplateFactory.cu:1323-1373; legality :1293-1310).
 * Returns size, or INVALID_DATA. */
int bf_euclid_template(float sigma, int32_t size, float *weights);

/* Upload a spatial template and a range table to the current device.
 * Replaces TemplateBasicOp::copyToCurrentDevice(gaussian / euclid)
 * (okano_3_0/BilateralFilter.cu:280-289, okano_3_0/Template.cu:385-463) and
 * the TemplateCuda pair the kernel receives (:331-333).
 *   offsets: 2*count ints, pairs (dy, dx) -- the first coordinate is the
 *            row offset, as the kernel reads it (:116-120);
 *   weights: count spatial weights (attachedData);
 *   range:   256 range weights (the EUCLIDE template's attachedData).
 * Host pointers; the data is copied.  Synchronous. */
int bf_lut_create(const int32_t *offsets, const float *weights, int32_t count,
                  const float *range, bf_lut *out);
int bf_lut_destroy(bf_lut lut);
/* Number of distinct spatial weights among the first ntaps taps (rows of
 * the fused weight table the kernels build in LDS); <0 on error. */
int bf_lut_rows(bf_lut lut, int32_t ntaps);

/* Exact FP32 filter of one ROI.  Replaces _initTexture + the grid set-up +
 * the _bilateralFilterKer launch of BilateralFilter::doFilter(in, out)
 * (okano_3_0/BilateralFilter.cu:38-68, :71-199, :322-333).
 *   in:  input ROI view; its width/height are the clamp extent of the
 *        sampler (the texture binds the untrimmed input ROI, :61-64);
 *   out: output ROI view; the filtered size is min(in, out) per axis
 *        (:312-320), other output pixels are left untouched;
 *   radius 1..20 (REF needs the template to hold (2r+1)^2+1 taps; r = 20
 *        in REF mode would read past the 41x41 template and is rejected
 *        with INVALID_DATA).
 * Output u8 = trunc(t / s) (clamped to 0..255, NaN -> 0), bit-exact to the
 * host transcription.  In and out must not overlap. */
int bf_exact(const bf_image_view *in, const bf_image_view *out, bf_lut lut,
             int32_t radius, int32_t mode, void *stream);

/* In-place filtering with `repeat` iterations.  Replaces
 * BilateralFilter::doFilter(Image *inoutimg) (okano_3_0/BilateralFilter.cu
 * :202-254), whose launches read and write one buffer concurrently; here
 * every iteration reads a snapshot (ping-pong through `work`), so the
 * result is deterministic and equals `repeat` out-of-place passes.
 *   work: device scratch of at least width*height bytes, or NULL to let
 *         the call allocate (and free) it. */
int bf_exact_inplace(const bf_image_view *img, bf_lut lut, int32_t radius,
                     int32_t mode, int32_t repeat, unsigned char *work,
                     void *stream);

/* ---------------------------------------------------------------------
 * Batched exact filter: many (image, ROI, radius, tables) jobs in one
 * persistent launch.  No reference counterpart (the reference launches one
 * grid per image); per image the output is bit-exact to bf_exact.
 */
typedef struct bf_job {
    bf_image_view in;    /* input ROI view (clamp extent)               */
    bf_image_view out;   /* output ROI view; filtered size = min(in,out) */
    bf_lut lut;
    int32_t radius;
    int32_t mode;
} bf_job;

typedef struct bf_batch_s *bf_batch;

/* Tile configuration ids accepted by bf_batch_create / bf_exact_config.
 * bf_tune_count() reports how many exist; bf_tune_describe() names one:
 * tile width and height in pixels, threads per workgroup, and the adjacent
 * columns each lane filters (1 or 2; each lane owns 4 rows of them). */
int bf_tune_count(void);
int bf_tune_describe(int32_t config, int32_t *tile_w, int32_t *tile_h,
                     int32_t *threads, int32_t *cols_per_lane);

/* Plan a batch: validates every job, orders tiles longest-first, uploads
 * the job table.  config < 0 picks the built-in default per radius.
 * Synchronous. */
int bf_batch_create(const bf_job *jobs, int32_t njobs, int32_t config,
                    bf_batch *out);
int bf_batch_destroy(bf_batch batch);
/* Run the whole batch as one persistent launch (plus a small reset
 * kernel).  Safe to capture into a hipGraph: it records no memset. */
int bf_batch_run(bf_batch batch, void *stream);
/* Capture bf_batch_run into a hipGraph owned by the batch (first call) and
 * launch it. */
int bf_batch_run_graph(bf_batch batch, void *stream);
/* Pixels the batch filters (sum of min(in,out) areas). */
int64_t bf_batch_pixels(bf_batch batch);
/* The tile configuration the batch runs with (the one asked for, or the
 * one bf_batch_create chose for config < 0). */
int bf_batch_config(bf_batch batch);
/* 1 when every job uses the canonical GAUSS-factory template at a radius
 * with an unrolled kernel (3, 5, 7, 9, 15) and the batch runs those
 * kernels (tile configs 0 and 1), 0 when it runs the tap-list kernel,
 * < 0 on error.  Same output bits either way. */
int bf_batch_unrolled(bf_batch batch);

/* Autotuner (no reference counterpart).  bf_tune_radius times every tile
 * configuration (tile shape and columns per lane, bf_tune_describe) for
 * (radius, mode) on the current device -- a synthetic width x height image,
 * sigma_space = radius/1.5, sigma_range = 30, run as a one-job batch with
 * that configuration (so the unrolled kernels are timed where a batch would
 * use them), `reps` timed launches after one warm-up -- and installs the
 * fastest in a process-wide table.  bf_exact and bf_batch_create (config < 0) use that
 * table: a single image takes the tuned configuration of its radius; a
 * batch whose radii are all tuned takes the configuration with the least
 * estimated total time (sum over jobs of pixels x measured cost).
 *   best_config: receives the winner; ms_per_mpix (bf_tune_count() floats,
 *   may be NULL): measured cost of each configuration, < 0 if it could not
 *   run.  Synchronous. */
int bf_tune_radius(int32_t radius, int32_t mode, int32_t width, int32_t height,
                   int32_t reps, int32_t *best_config, float *ms_per_mpix);
/* Install / read a table entry (e.g. from a saved tuning run):
 * ms_per_mpix as above.  bf_tune_get returns the tuned configuration, or
 * -1 if (radius, mode) has not been tuned. */
int bf_tune_set(int32_t radius, int32_t mode, const float *ms_per_mpix);
int bf_tune_get(int32_t radius, int32_t mode, float *ms_per_mpix);
/* Forget every tuned entry (back to the built-in defaults). */
int bf_tune_reset(void);

/* bf_exact with an explicit tile configuration (the autotuner's probe). */
int bf_exact_config(const bf_image_view *in, const bf_image_view *out,
                    bf_lut lut, int32_t radius, int32_t mode, int32_t config,
                    void *stream);

/* ---------------------------------------------------------------------
 * Learned surrogate (BF16 MFMA conv net, FP32 accumulate) with per-tile
 * exact fallback.  No reference counterpart.
 *
 * Network: 6 conv3x3 layers, 32 channels, ReLU between layers, inputs
 * (image/255, sigma_range/100), output = image/255 + conv6(...).
 */
typedef struct bf_sur_s *bf_sur;

/* Parameter count of the network (weights + biases). */
int64_t bf_sur_param_count(void);
/* Create a network; weights from `seed` (He init) in FP32 master copies,
 * BF16 working copies, zeroed Adam moments.  Synchronous. */
int bf_sur_create(uint64_t seed, bf_sur *out);
int bf_sur_destroy(bf_sur sur);
/* Copy FP32 master parameters to/from host (`params` has
 * bf_sur_param_count() floats; layout documented in surrogate.hip).
 * set also refreshes the BF16 copies.  Synchronous. */
int bf_sur_get_params(bf_sur sur, float *params);
int bf_sur_set_params(bf_sur sur, const float *params);

/* One training step on a batch of n u8 crops (h x w, dense, device):
 * forward, MSE loss against target/255, backward, Adam (lr, beta1, beta2,
 * eps; bias-corrected, step counter kept in the net).  sigma_range: n
 * floats on the device.  The loss leaves out a 6-pixel band along each
 * crop side, except the sides that lie on their image's edge: edges (n
 * bytes on the device, or NULL = no such side) has bit 0 top, 1 bottom,
 * 2 left, 3 right.  loss_out: device float (mean loss), may be NULL. */
int bf_sur_train_step(bf_sur sur, const unsigned char *x, const unsigned char *y,
                      const float *sigma_range, const unsigned char *edges,
                      int32_t n, int32_t h, int32_t w,
                      float lr, float beta1, float beta2, float eps,
                      float *loss_out, void *stream);

/* Loss of bf_sur_train_step and bf_sur_forward_backward: the MSE plus,
 * per pixel, weight * max(|pred - y| - start_lsb / 255, 0)^2 (over the same
 * pixels and count).  weight 0 (the default of a new net) is the plain MSE. */
int bf_sur_set_loss_tail(bf_sur sur, float weight, float start_lsb);

/* Forward only, FP32 output (n x h x w, device) of the same crops: the
 * gradient-check hook.  grads (optional, device, param_count floats):
 * receives d(loss)/d(params) of the loss bf_sur_train_step minimises
 * against y (y may be NULL when grads is NULL; edges as there). */
int bf_sur_forward_backward(bf_sur sur, const unsigned char *x,
                            const unsigned char *y, const float *sigma_range,
                            const unsigned char *edges,
                            int32_t n, int32_t h, int32_t w, float *pred,
                            float *grads, float *loss_out, void *stream);

/* Surrogate inference of one ROI into out (u8, rounded, clamped). */
int bf_sur_infer(bf_sur sur, const bf_image_view *in, const bf_image_view *out,
                 float sigma_range, void *stream);

/* Surrogate + check + fallback for one ROI:
 *   1. surrogate inference into out;
 *   2. per 64x64 tile, 1/64 of the pixels (one per 8x8 cell, at a fixed
 *      pseudo-random position) are recomputed exactly (mode, radius, lut)
 *      and compared with the surrogate;
 *   3. every tile with a sampled |diff| > tol_lsb is recomputed exactly.
 * tiles_total / tiles_fallback (host pointers, may be NULL) receive the
 * counts after the call completes (the call synchronises the stream only
 * when one of them is non-NULL). */
int bf_sur_filter(bf_sur sur, const bf_image_view *in, const bf_image_view *out,
                  bf_lut lut, int32_t radius, int32_t mode, float sigma_range,
                  int32_t tol_lsb, int32_t *tiles_total, int32_t *tiles_fallback,
                  void *stream);

/* bf_sur_filter without host synchronisation, safe to capture into a
 * hipGraph once the network's workspace has grown to the largest image
 * (run each size once outside the capture): adds the number of tiles that
 * fell back to *fallback_accum and the number of tiles to *tiles_accum
 * (device ints, either may be NULL). */
int bf_sur_filter_async(bf_sur sur, const bf_image_view *in, const bf_image_view *out,
                        bf_lut lut, int32_t radius, int32_t mode, float sigma_range,
                        int32_t tol_lsb, int32_t *fallback_accum, int32_t *tiles_accum,
                        void *stream);

/* bf_sur_filter_async without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, hosts that forbid
 * hipMalloc).  The fallback tile count and list live in a file-scope device
 * array of the library (one per device, looked up on the first call on that
 * device and held until the process ends) instead of the network's
 * workspace, so the call never grows a buffer.  Limit: 65536 tiles of
 * 64x64, i.e. filtered images up to 16384 x 16384 pixels; larger ones are
 * rejected with INVALID_DATA before anything is launched.  Calls on one
 * device share that array: calls that overlap in time on one device (other
 * streams, other threads) race and must be serialised by the caller.  The
 * first call on a device should run outside graph capture (it looks up the
 * array's address).  Output and counts as bf_sur_filter_async. */
int bf_sur_filter_noalloc(bf_sur sur, const bf_image_view *in, const bf_image_view *out,
                          bf_lut lut, int32_t radius, int32_t mode, float sigma_range,
                          int32_t tol_lsb, int32_t *fallback_accum, int32_t *tiles_accum,
                          void *stream);

/* ---------------------------------------------------------------------
 * Misc.
 */
/* Text of the last error on this thread ("" if none). */
const char *bf_last_error(void);
/* Library version (major*10000 + minor*100 + patch). */
int bf_version(void);

/* Deterministic synthetic test image (no reference counterpart): linear
 * gradient 0..200 along x, `nshapes` rectangles and `nshapes` discs
 * (interleaved, sizes smin..smax, grey 0..255), Gaussian noise of
 * noise_sigma (Box-Muller in double), rounded half-to-even, clipped to
 * 0..255.  Host memory, dense (pitch = width).  Synchronous. */
int bf_synth_image(unsigned char *dst, int32_t width, int32_t height,
                   int32_t nshapes, int32_t smin, int32_t smax,
                   double noise_sigma, uint64_t seed);

/* Host-side shape templates (no device work): the BOX, CIRCLE and ARC
 * vendors of okano_3_0/TemplateFactory.cu:340-402, :501-666, :771-936 (what
 * TemplateFactory::getTemplate pools for those shapes).  Return the point
 * count (or a negative status); with points or weights NULL only the count
 * is returned.  points = (x, y) pairs. */
int bf_box_template(int32_t size_x, int32_t size_y, int32_t *points, float *weights);
int bf_circle_template(int32_t size, int32_t *points, float *weights);
int bf_arc_template(int32_t size_x, int32_t size_y, int32_t *points, float *weights);

/* ------------------------------------------------------------------
 * The stages around the filter in okano_3_0/SimpleRegionDetect.cu:287-312
 * (detectRegion: ... -> LocalCluster -> BilateralFilter -> ConnectRegion).
 * ------------------------------------------------------------------ */

/* LocalCluster's parameters (okano_3_0/LocalCluster.h members gapThred,
 * diffeThred, proBlack, proWhite, pntRange, pntCount). */
typedef struct bf_local_cluster_params {
    int32_t pnt_range;      /* points walked per direction */
    int32_t pnt_count;      /* directional averages kept (of 8) */
    uint8_t gap_thred;      /* max |centre - walked pixel| */
    uint8_t diffe_thred;    /* max difference of the two side sums */
    uint8_t pro_black;      /* pixels <= pro_black are copied */
    uint8_t pro_white;      /* pixels >= pro_white are copied */
} bf_local_cluster_params;

/* Replaces LocalCluster::localCluster's kernel launch
 * (okano_3_0/LocalCluster.cu:45-153, launch :246-270).  `in` is the ROI
 * view; [alloc_lo, alloc_hi) are the byte offsets from in->data that may be
 * read (the whole allocation holding the ROI): the reference's walks read
 * up to 2 (pnt_range - 1) pixels from a pixel by linear address and reads
 * outside the allocation are clamped into it.  alloc_lo = alloc_hi = 0
 * means the ROI's own bytes.  Output size = min of the two views.  Async on
 * `stream`. */
int bf_local_cluster(const bf_image_view *in, int64_t alloc_lo, int64_t alloc_hi,
                     const bf_image_view *out, const bf_local_cluster_params *params,
                     void *stream);

/* Replaces ConnectRegion::connectRegion (okano_3_0/ConnectRegion.cu:789-1043,
 * which the reference calls with min_gray 10, max_gray 250): 8-connected
 * regions of pixels in [min_gray, max_gray], area filter [min_area,
 * max_area], regions numbered in raster order of their first pixel, u8
 * output = number + 1 (0 = none).  Output size = min of the two views.
 * Allocates 12 bytes per pixel of stream-ordered scratch.  Async. */
int bf_connect_region(const bf_image_view *in, const bf_image_view *out, int32_t min_area,
                      int32_t max_area, int32_t min_gray, int32_t max_gray, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OPENCUDA_BF_H */
