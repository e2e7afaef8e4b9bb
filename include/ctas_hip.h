/*
 * ctas_hip.h - C ABI of the MI355X filtered-backprojection engine for CTAS.
 *
 * One handle per reconstruction geometry, mirroring the reference's CTrec
 * (src/common/ct.h:235-298). All functions return 0 on success and a
 * negative CTAS_E_* code on failure; ctas_last_error() then holds a
 * thread-local message. Host pointers are row-major float32: sinograms
 * [S][N][w] (N angles x w detector pixels per slice), slices [S][w][w]
 * (row j, column i). Functions with a _dev suffix take device pointers and a
 * hipStream_t passed as void* (NULL = the handle's own stream) and return
 * without synchronising. A handle is used by one thread at a time, as a
 * CTrec is; clones share nothing mutable.
 */
#ifndef CTAS_HIP_H
#define CTAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ctas_fbp ctas_fbp;

/* error codes */
#define CTAS_OK 0
#define CTAS_E_ARG -1      /* bad argument (reference: throw_error)          */
#define CTAS_E_SHAPE -2    /* array shape does not match the handle         */
#define CTAS_E_CENTER -3   /* |c| >= w/2 (lct.cpp:614-616, 651-653)          */
#define CTAS_E_DATA -4     /* deAbs: maximum not positive (physical.world.cpp:94) */
#define CTAS_E_DEVICE -5   /* HIP runtime error or no device                */
#define CTAS_E_NOMEM -6    /* device allocation failed                      */
#define CTAS_E_STATE -7    /* call order (e.g. repeat before sino)           */
#define CTAS_E_AXIS -8     /* axis search failed (lct.cpp:1136-1210 throws)  */
#define CTAS_E_IO -9       /* a reader or writer callback reported failure   */
#define CTAS_E_RANGE -10   /* FP16 storage: a filtered sample beyond 65504, or
                            * beyond 16376 where the mirrored CTAS kernel adds
                            * four samples in FP16 (no reference counterpart: the
                            * reference has no FP16 storage)                 */

/* contrast codes = Contrast::Cntype (physical.world.h:22-27) */
#define CTAS_CONTRAST_ABS 0
#define CTAS_CONTRAST_REF 1
#define CTAS_CONTRAST_PHS 2
#define CTAS_CONTRAST_FLT 3

/* filter codes = Filter::Ftype (ct.h:95-146) */
#define CTAS_FILTER_NONE 0
#define CTAS_FILTER_RAMP 1
#define CTAS_FILTER_BARLETT 2
#define CTAS_FILTER_WELCH 3
#define CTAS_FILTER_PARZEN 4
#define CTAS_FILTER_HANN 5
#define CTAS_FILTER_HAMMING 6
#define CTAS_FILTER_BLACKMAN 7
#define CTAS_FILTER_LANCKZOS 8
#define CTAS_FILTER_KAISER 9
#define CTAS_FILTER_GAUSS 10

/* flags for ctas_fbp_create */
#define CTAS_INTERP_CTAS 0x0   /* floor sampling in the reference's float arithmetic (parity) */
#define CTAS_INTERP_LINEAR 0x1 /* linear interpolation between bin centres           */
#define CTAS_INTERP_MASK 0x3
#define CTAS_STORE_FP32 0x00   /* filtered sinogram stored as float32             */
#define CTAS_STORE_FP16 0x10   /* ... as IEEE half, accumulated in float32         */
#define CTAS_STORE_BF16 0x20   /* ... as bfloat16, accumulated in float32          */
#define CTAS_STORE_MASK 0x30

/* Thread-local message of the last failure on this thread. */
const char *ctas_last_error(void);
/* Number of visible HIP devices (0 when none). */
int ctas_device_count(int *n);
/* Library version string. */
const char *ctas_version(void);

/* ---- geometry helpers (no device needed) ---------------------------- */
/* zidth = 2^ceil(log2(2w-1)) (lct.cpp:490) */
int ctas_zidth(int w);
/* arcO / angO (lct.cpp:470-484): arc or step in degrees -> radians */
float ctas_arc_rad(float anarc, int points);
float ctas_step_rad(float anarc, int points);
/* Filter::fill + CTrec gain/REF rule (lct.cpp:122-209, 505-508): W[Z] */
int ctas_filter_window(int filter, float param, int w, int contrast, float *W);
/* Filter::fromName (lct.cpp:71-86) and _conversion "NAME[:param]"
 * (lct.cpp:236-248): name -> code and parameter. */
int ctas_filter_parse(const char *spec, int *filter, float *param);
/* Contrast(const std::string&) (physical.world.cpp:19-27). */
int ctas_contrast_parse(const char *name, int *contrast);
/* The float angle table of lct.cpp:512-517 ({sin, cos} per angle). */
int ctas_angles(int n, float anarc, float *sincos);

/* ---- reconstruction handle (CTrec) ------------------------------------ */
/* CTrec(sinoshape=(n_angles, w), contrast, anarc, filter)
 * (lct.cpp:487-526). Throws there for w < 2 or n_angles == 0. */
int ctas_fbp_create(ctas_fbp **h, int w, int n_angles, int contrast, int filter,
                    float filter_param, float anarc, int flags);
/* CTrec(CTrec&) (lct.cpp:529-545): same geometry, own device buffers. */
int ctas_fbp_clone(const ctas_fbp *src, ctas_fbp **h);
/* ~CTrec (lct.cpp:549-555) */
int ctas_fbp_destroy(ctas_fbp *h);
int ctas_fbp_info(const ctas_fbp *h, int *w, int *n_angles, int *zidth, int *flags);
/* CTrec::recShape (ct.h:272-274): the slice rows left after cropping
 * ceil(|center|) rows at the top and at the bottom, and all w columns. */
int ctas_fbp_rec_shape(const ctas_fbp *h, float center, int *rows, int *cols);
/* CTrec::recCrop (ct.h:276-279): that crop as the half-open row range
 * [row_begin, row_end) of the w x w slice; columns are not cropped. */
int ctas_fbp_rec_crop(const ctas_fbp *h, float center, int *row_begin, int *row_end);
/* CTrec::setPhysics (lct.cpp:558-566): pixel size and wavelength in m. */
int ctas_fbp_set_physics(ctas_fbp *h, float pixel_size, float lambda);
int ctas_fbp_get_reccof(const ctas_fbp *h, float *recCof);
/* Replace / read the {sin, cos} float table (2*N floats). */
int ctas_fbp_set_angles(ctas_fbp *h, const float *sincos);
int ctas_fbp_get_angles(const ctas_fbp *h, float *sincos);
int ctas_fbp_get_window(const ctas_fbp *h, float *W);

/* CTrec::prepare_sino on S host slices, in place (lct.cpp:569-597). */
int ctas_fbp_prepare(ctas_fbp *h, float *sino, int S);
/* CTrec::sino (lct.cpp:600-609): prepare in place and keep the filtered
 * sinogram on the device for ctas_fbp_repeat / ctas_axis_scores. */
int ctas_fbp_sino(ctas_fbp *h, float *sino);
/* CTrec::repeat (lct.cpp:612-646) for n_centers centres at once:
 * out[n_centers][w][w]. */
int ctas_fbp_repeat(ctas_fbp *h, float *out, const float *centers, int n_centers);
/* CTrec::reconstruct (lct.cpp:649-685) for S slices, one centre per slice:
 * sino[S][N][w] is filtered in place, out[S][w][w]. */
int ctas_fbp_reconstruct(ctas_fbp *h, float *sino, float *out, const float *centers, int S);

/* Device-resident variants (inputs already in HBM; the input sinogram is not
 * modified; filtered rows go to the handle's storage-dtype buffer). */
int ctas_fbp_reconstruct_dev(ctas_fbp *h, const float *d_sino, float *d_out,
                             const float *centers, int S, void *stream);
/* Backproject the device-resident filtered sinogram kept by ctas_fbp_sino
 * (or by the last _dev reconstruct, slice `slice`) at n centres. */
int ctas_fbp_repeat_dev(ctas_fbp *h, int slice, float *d_out, const float *centers,
                        int n_centers, void *stream);
/* The two stages of reconstruct_dev on their own: prepare S device slices
 * into the handle's filtered buffer (CTrec::prepare_sino, lct.cpp:569-597),
 * and backproject those S filtered slices, centre i for slice i (the kernel
 * launch of ForCLdev::reconstruct, lct.cpp:406-409). */
int ctas_fbp_prepare_dev(ctas_fbp *h, const float *d_sino, int S, void *stream);
int ctas_fbp_backproject_dev(ctas_fbp *h, float *d_out, const float *centers, int S, void *stream);
/* Capture reconstruct_dev(d_sino, d_out, centers, S) into a hipGraph owned
 * by the handle, then replay it. No memset nodes are recorded. */
int ctas_fbp_graph_capture(ctas_fbp *h, const float *d_sino, float *d_out,
                           const float *centers, int S, void *stream);
int ctas_fbp_graph_launch(ctas_fbp *h, void *stream);
/* Last per-slice status of a _dev call: 0 ok, CTAS_E_DATA (deAbs, ABS
 * contrast) or CTAS_E_RANGE (FP16 storage of PHS, REF or FLT data);
 * synchronises the handle's stream. */
int ctas_fbp_status(ctas_fbp *h, int S, int *status);

/* Tile choice for this handle's width class: rows x cols of output per
 * workgroup and angles staged per LDS pass; 0 = let the autotuner decide. */
int ctas_fbp_set_tiles(ctas_fbp *h, int tile_kind, int angles_per_pass);
int ctas_fbp_get_tiles(const ctas_fbp *h, int *tile_kind, int *angles_per_pass);
/* Time the candidate tiles on the current device for this handle's shape
 * and keep the fastest (cached per width class in the process). */
int ctas_fbp_autotune(ctas_fbp *h, int S);

/* ---- uneven batches in one launch ------------------------------------- */
typedef struct {
  int w, n_angles;        /* shape of this sinogram                       */
  float center;           /* centre offset c                              */
  const float *d_sino;    /* device [N][w] float32 (unfiltered)           */
  float *d_out;           /* device [w][w] float32                        */
} ctas_batch_item;
/* Filters and backprojects n items of different widths and angle counts
 * with one persistent backprojection launch (contrast, filter, arc and
 * flags shared). */
int ctas_fbp_batch_uneven(const ctas_batch_item *items, int n, int contrast, int filter,
                          float filter_param, float anarc, int flags, void *stream);

/* ---- pipelined reconstruction of a series --------------------------------
 * The loop of `ctas ct` over its sinograms (RecInThread, ct.cpp:156-236:
 * read, ring filter, reconstruct, write) as a pipeline: `read` fills
 * sinogram `index` ([N][w] float32) on a reader thread, batches of `batch`
 * slices are uploaded from pinned buffers, ring-filtered (`ring_box` > 0,
 * RingFilter::apply) and reconstructed on the device at centers[index], and
 * `write` receives slice `index` ([w][w]) on a writer thread. Reading batch
 * k+1, computing batch k and writing batch k-1 overlap. With io_threads = 1
 * the callbacks run in index order on one reader and one writer thread;
 * with io_threads > 1 the slices of a batch are read (and written) by that
 * many threads at once, in no fixed order, so the callbacks must be safe
 * to call concurrently for different indices. Callbacks return 0 on
 * success; any other value stops the pipeline and the call returns
 * CTAS_E_IO. The filtered sinograms are not returned. */
typedef int (*ctas_read_fn)(void *user, long index, float *sino);
typedef int (*ctas_write_fn)(void *user, long index, const float *slice);
int ctas_fbp_pipeline(ctas_fbp *h, long n_slices, int batch, const float *centers, int ring_box,
                      ctas_read_fn read, ctas_write_fn write, void *user, int io_threads);

/* ---- ring filter (RingFilter::apply, lct.cpp:295-327) ----------------- */
int ctas_ring_filter(float *sino, int N, int w, int box);
int ctas_ring_filter_dev(float *d_sino, int S, int N, int w, int box, void *stream);
/* No counterpart in the reference. ctas_ring_filter_dev for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * loops): the same kernels, with the S*w column means in a fixed device
 * array of 1,048,576 floats (4 MiB per GPU, held until the process ends)
 * instead of a per-call allocation, and no synchronisation (the work is
 * enqueued on `stream`). S*w above 1,048,576 is rejected with CTAS_E_ARG.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other threads) race. */
int ctas_ring_filter_dev_noalloc(float *d_sino, int S, int N, int w, int box, void *stream);

/* ---- tomosynthesis (ts_add, ct.h:317 / lct.cpp:739-787) --------------- */
/* Filter::fill(filt, n) (lct.cpp:122-209): the window of an n-point line,
 * without the CTrec gain. */
int ctas_filter_fill(int filter, float param, int n, float *W);
/* The window of ts_add's row filter as a transform of *Z points (Z = n for
 * a power of two n >= 2, else zidth(n)); W (optional) receives Z values. */
int ctas_ts_window(int filter, float param, int n, int *Z, float *W);
/* ts_add(projection, result, Filter(filter, param), center, contrast,
 * angle, plane): projection (rows x pixels, row-major) is replaced by its
 * filtered rows (after -log for ABS, running sums for REF), as in the
 * reference; result (result_rows x result_pixels) accumulates the projection
 * resampled for the plane. Angle in radians. Fails, like the reference's
 * size check, unless result has the projection's shape. */
int ctas_ts_add(float *projection, int rows, int pixels, float *result, int result_rows,
                int result_pixels, int filter, float filter_param, float center, int contrast,
                float angle, int plane);
/* Device plan of ts_add for projections of rows x pixels (pixels <= 8192;
 * 4096 unless a power of two): the window and twiddles of the row filter. */
typedef struct ctas_ts ctas_ts;
int ctas_ts_create(ctas_ts **h, int rows, int pixels, int contrast, int filter, float filter_param);
int ctas_ts_destroy(ctas_ts *h);
int ctas_ts_info(const ctas_ts *h, int *rows, int *pixels, int *zidth);
/* n_proj successive ts_add calls on device arrays, enqueued on `stream`:
 * d_proj holds the projections back to back ([n_proj][rows][pixels]),
 * angles[n_proj] (host, radians) their angles; d_result accumulates them in
 * order; d_filtered (optional, may equal d_proj) receives the filtered
 * projections. */
int ctas_ts_add_dev(ctas_ts *h, const float *d_proj, int n_proj, const float *angles, float center,
                    int plane, float *d_filtered, float *d_result, void *stream);
/* n_proj successive ts_add calls on host arrays: projections
 * ([n_proj][rows][pixels], left unchanged) are streamed to the device in
 * chunks and added in order to result (host, [rows][pixels]); the ts tool's
 * loop over its projection list (ts.cpp:212-217). */
int ctas_ts_accumulate(ctas_ts *h, const float *projections, int n_proj, const float *angles,
                       float center, int plane, float *result);

/* ---- denoiser (Denoiser, matrix.world.cpp:952-1037; `ax -n/-N`) -------- */
/* Denoiser(shape, rad, threshold).proc(image) without a mask: the disc mean
 * of radius |rad| (offsets inside the image only) replaces each pixel
 * always (threshold 0) or where the pixel is 0 or deviates from it by more
 * than threshold (absolute, > 0) or -threshold relative to the mean (< 0).
 * rad 0 leaves the image unchanged. */
int ctas_denoise(float *image, int rows, int cols, int rad, float threshold);
/* The same on device arrays (d_in and d_out distinct), on `stream`. */
int ctas_denoise_dev(const float *d_in, float *d_out, int rows, int cols, int rad, float threshold,
                     void *stream);

/* ---- rotation axis (lct.cpp:804-1326, axis.cpp:116-208) -------------- */
/* AverageGradient::apply for n trial axis positions axisR (pixels from the
 * left edge) on the sinogram kept by ctas_fbp_sino of an ABS/RAMP handle. */
int ctas_axis_scores(ctas_fbp *h, const float *axisR, int n, float *scores);
/* raxis(sino, anarc, maxDev, algo) (lct.cpp:1133-1212): axis position. */
int ctas_raxis(const float *sino, int N, int w, float anarc, float maxDev, int algo,
               float *result);
/* raxis(proj0, proj180, maxDev) (lct.cpp:1217-1294). */
int ctas_raxis_pair(const float *p0, const float *p180, int rows, int w, float maxDev,
                    float *result);

/* ---- learned axis finder (no counterpart in the reference) ------------
 * A BF16 CNN (4 conv 3x3 + 2 FC, leaky ReLU, 245,554 parameters) on the
 * matrix cores that estimates the centre offset of a sinogram and a confidence; it
 * fronts the sweep of trial reconstructions of raxis (lct.cpp:1133-1212,
 * AverageGradient lct.cpp:804-917). Input: [2][180][256] bf16 (the -log
 * sinogram averaged over 180 angle cells x 256 detector bins, standardised,
 * and its copy turned by 180 degrees: detector and angle order reversed). Output per sample: {offset in detector
 * bins of the input, confidence logit}. */
typedef struct ctas_cnn ctas_cnn;
int ctas_cnn_create(ctas_cnn **net, uint64_t seed);
int ctas_cnn_destroy(ctas_cnn *net);
/* Parameters: w1 b1 w2 b2 w3 b3 w4 b4 wf1 bf1 wf2 bf2 back to back (conv
 * weights [cout][cin][3][3], fully connected [out][in]), float32. */
int ctas_cnn_num_params(const ctas_cnn *net, int *count);
int ctas_cnn_get_params(ctas_cnn *net, float *params);
int ctas_cnn_set_params(ctas_cnn *net, const float *params);
/* Gradients of the last ctas_cnn_gradients_dev / train_step call. */
int ctas_cnn_get_grads(ctas_cnn *net, float *grads);
int ctas_cnn_save(const ctas_cnn *net, const char *path);
int ctas_cnn_load(ctas_cnn *net, const char *path);
/* The network's _dev calls read inputs made ready on `stream` (NULL: the
 * null stream) and return when their results are complete.
 * d_x [B][2][180][256] bf16 -> d_out [B][2] float32. */
int ctas_cnn_forward_dev(ctas_cnn *net, const uint16_t *d_x, int B, float *d_out, void *stream);
/* d_target [B][2] = {offset (bins), tolerance (bins)}: loss = mean of
 * Huber(offset error, 0.25) + 0.2 BCE(logit, |error| < tolerance). */
int ctas_cnn_gradients_dev(ctas_cnn *net, const uint16_t *d_x, const float *d_target, int B,
                           float *loss, void *stream);
/* gradients + one Adam step (beta 0.9 / 0.999, eps 1e-8) at rate lr, with
 * the gradient scaled to global norm clip_norm when larger (clip_norm <= 0:
 * no clipping); grad_norm (optional) gets the norm before clipping. */
int ctas_cnn_train_step_dev(ctas_cnn *net, const uint16_t *d_x, const float *d_target, int B,
                            float lr, float clip_norm, float *loss, float *grad_norm,
                            void *stream);
/* Network input of one device sinogram [N][w] (transmission). */
int ctas_cnn_features_dev(const float *d_sino, int N, int w, uint16_t *d_x, void *stream);
/* Training sample `index` of stream `seed` (w, N, offset c, photons N0 and
 * up to 60 ellipses (x0, y0, a, b, phi, mu)); returns the ellipse count. */
int ctas_axis_sample(uint64_t seed, uint64_t index, int *w, int *n_angles, double *c, double *N0,
                     double *ell, int *n_ell);
/* Samples first .. first+B-1 of `seed` projected on the GPU and reduced to
 * network inputs d_x and targets d_target ({c/bin, 1 px/bin}). */
int ctas_cnn_make_batch(uint64_t seed, uint64_t first, int B, uint16_t *d_x, float *d_target,
                        float *offsets_px, int *widths_px, void *stream);
/* Estimate (px) and confidence in [0, 1] for one sinogram. */
int ctas_cnn_predict_dev(ctas_cnn *net, const float *d_sino, int n_angles, int w, float *offset,
                         float *confidence, void *stream);
int ctas_cnn_predict(ctas_cnn *net, const float *sino, int n_angles, int w, float *offset,
                     float *confidence);
/* Sweep of trial centres -half_range + k*step (AverageGradient score,
 * argmin) on a device sinogram of h's shape (h: ABS contrast, RAMP filter).
 * With net (and confidence >= min_confidence) only a window around the
 * estimate is scored (3 points, widened one point at a time until its
 * minimum is interior; past 21 points the whole grid).
 * info[4]: trials, whole grid swept (0/1), estimate (milli-px),
 * confidence (per mille). */
int ctas_axis_find(ctas_fbp *h, ctas_cnn *net, const float *d_sino, float half_range, float step,
                   float min_confidence, float *offset, int *info);
/* The same windowed search from a given start offset `start` (px) instead
 * of the network's estimate (e.g. the 0/180 registration's seed, or 0): the
 * search the network fronts, without it. info as above (estimate = start). */
int ctas_axis_find_from(ctas_fbp *h, const float *d_sino, float half_range, float step, float start,
                        float *offset, int *info);
/* raxis algorithm 1 (lct.cpp:1133-1212) with its 8 px sweep replaced by the
 * network's estimate (improveMe from +-4 px around it); falls back to
 * ctas_raxis(algo 1) below min_confidence. */
int ctas_raxis_cnn(ctas_cnn *net, const float *sino, int N, int w, float anarc, float maxDev,
                   float min_confidence, float *result, int *used_sweep);
/* The network's bf16 GEMM on row-major operands (verification):
 * C[M][N] = A[M][K] B[K][N]. */
int ctas_cnn_gemm_dev(const uint16_t *d_a, const uint16_t *d_b, float *d_c, int M, int N, int K,
                      void *stream);

/* ---- synthetic ellipse phantoms ---------------------------------------- */
/* n_ell random ellipses (x0, y0, a, b, phi, mu per pixel) for width w: centres
 * uniform in the inner 80% of the disc, semi-axes uniform in
 * [5, 200]*axis_scale px, phi uniform in [0, pi), densities uniform in
 * [-0.5, 1] scaled by 2/w per pixel. Seeded (SplitMix64). */
int ctas_phantom_ellipses(int n_ell, int w, double axis_scale, uint64_t seed, double *ell);
/* Analytic line integrals at bin centres (FP64 on the device); with N0 > 0
 * the noisy transmission counts/N0; gain (optional, [w]) multiplies it. */
int ctas_phantom_sino_dev(int N, int w, double c, int n_ell, const double *ell, double N0,
                          uint64_t seed, const float *d_gain, float *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
