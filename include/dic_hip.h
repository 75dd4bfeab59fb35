/*
 * dic_hip.h -- C-ABI of the MI355X (gfx950) subset-correlation library.
 *
 * This is the drop-in boundary for Traditional-DIC's subset IC-GN hot path:
 * the B-spline prefilter and sampler (src/interpolation/bspline.cpp), the
 * reference-gradient cache (src/subset/seed/reliability_propagation.cpp:84-112)
 * and the per-POI IC-GN solver family (src/subset/solver/icgn.cpp).  The C++
 * host layer (dic_hip.hpp) mirrors the reference's SubsetSolver /
 * BSplineImagePreprocessor / BSplineInterpolator surface on top of it.
 *
 * Conventions
 *   - Every entry point is extern "C", never throws, and returns an int status
 *     (DIC_OK = 0).  The message of the last failure on a context is returned
 *     by dic_last_error().  Numerical outcomes per POI are NOT errors: they are
 *     reported in dic_result.status exactly as the reference reports
 *     SolverStatus in Displacement2D.
 *   - Images are row-major float32, index y*width + x (dic::Image::data(),
 *     include/dic/core/image.hpp:58-81).
 *   - Handles own device memory; a context's calls run in order on its stream.
 *     One context per host thread.
 */
#ifndef DIC_HIP_H
#define DIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes of the C-ABI calls ---- */
enum {
    DIC_OK = 0,
    DIC_ERR_INVALID_ARGUMENT = 1,
    DIC_ERR_HIP = 2,
    DIC_ERR_UNSUPPORTED = 3,
    DIC_ERR_BUDGET = 4,
    DIC_ERR_NO_DEVICE = 5
};

/* ---- per-POI solver status: SolverStatus order, include/dic/core/types.hpp:35-40 ---- */
enum { DIC_SUCCESS = 0, DIC_NOT_CONVERGED = 1, DIC_INVALID_INPUT = 2, DIC_NUMERICAL_FAILURE = 3 };

/* ---- enumerations mirroring the reference config ---- */
enum { DIC_PRECISION_FP64 = 0, DIC_PRECISION_MIXED = 1 };    /* plain FP64 / FP32-compensated + FP64 solve */
enum { DIC_OBJECTIVE_ZNSSD = 0, DIC_OBJECTIVE_SSD = 1 };     /* CorrelationCriterionKind */
enum { DIC_OPTIMIZER_ICGN = 0, DIC_OPTIMIZER_FGN = 1 };      /* SubsetOptimizationMethod */
enum { DIC_GRADIENT_CENTRAL = 0, DIC_GRADIENT_BSPLINE_HALF = 1 };

typedef struct dic_ctx dic_ctx;
typedef struct dic_image dic_image;   /* device float32 W x H image */
typedef struct dic_grad dic_grad;     /* device reference-gradient planes gx, gy (H x W) */
typedef struct dic_coeffs dic_coeffs; /* device B-spline coefficients, padded (H+2b) x (W+2b) */

/* SubsetConfig fields used by the solver (include/dic/subset/subset_config.hpp:33-50). */
typedef struct dic_icgn_params {
    int32_t subset_radius;        /* subset_radius */
    int32_t max_iterations;       /* max_iterations */
    double convergence_threshold; /* convergence_threshold on ||dp||_2 */
    int32_t shape_order;          /* 1 = FirstOrder, 2 = SecondOrder (or use_second_order) */
    int32_t objective;            /* DIC_OBJECTIVE_* */
    int32_t optimizer;            /* DIC_OPTIMIZER_*: IC-GN, or forward Gauss-Newton (ForwardGaussNewtonSolver) */
    int32_t precision;            /* DIC_PRECISION_*; must match the coefficient/gradient handles */
    int32_t masked;               /* 1: solve_with_mask semantics (roi required) */
    int32_t min_valid_samples;    /* masked only */
    double min_valid_sample_ratio;/* masked only */
} dic_icgn_params;

/* InitialDisplacement (include/dic/initialization/initializer.hpp:33-44). */
typedef struct dic_initial {
    double u, v, du_dx, du_dy, dv_dx, dv_dy, confidence;
} dic_initial;

/* Displacement2D (include/dic/core/result.hpp:32-51) + executed iterations. */
typedef struct dic_result {
    double x, y;
    double p[12]; /* u v du_dx du_dy dv_dx dv_dy d2u_dx2 d2u_dxdy d2u_dy2 d2v_dx2 d2v_dxdy d2v_dy2 */
    double correlation;
    int32_t status;
    int32_t valid;
    int32_t iterations;
    int32_t reserved;
} dic_result;

/* ---- context ---- */
/* New boundary (the reference has no device): owns the HIP stream and tracks
 * the HBM bytes its handles hold against hbm_budget (0 = unlimited). */
int dic_ctx_create(int device, uint64_t hbm_budget_bytes, dic_ctx **out);
void dic_ctx_destroy(dic_ctx *ctx);
const char *dic_last_error(const dic_ctx *ctx);
/* Run subsequent work on a caller-owned hipStream_t (NULL restores the context's own). */
int dic_ctx_set_stream(dic_ctx *ctx, void *hip_stream);
void *dic_ctx_stream(dic_ctx *ctx);
int dic_ctx_synchronize(dic_ctx *ctx);
uint64_t dic_ctx_bytes_in_use(const dic_ctx *ctx);
int dic_device_count(int *count);
const char *dic_version(void);

/* ---- plain device buffers and events (new boundary; callers without a
 * framework of their own use these to keep inputs resident in HBM and to
 * time work on the context's stream) ---- */
int dic_buffer_alloc(dic_ctx *ctx, size_t bytes, void **device_ptr);
void dic_buffer_free(dic_ctx *ctx, void *device_ptr);
int dic_buffer_upload(dic_ctx *ctx, void *device_dst, const void *host_src, size_t bytes);
int dic_buffer_download(dic_ctx *ctx, void *host_dst, const void *device_src, size_t bytes);
/* Device-to-device copy on the context's stream (not synchronised). */
int dic_buffer_copy(dic_ctx *ctx, void *device_dst, const void *device_src, size_t bytes);
/* Records a new event on the context's stream; free with dic_event_destroy. */
int dic_event_record(dic_ctx *ctx, void **event);
int dic_event_elapsed_ms(void *start, void *end, float *ms);
void dic_event_destroy(void *event);

/* ---- images: replaces dic::Image(width, height, data) (src/core/image.cpp:128-139) ---- */
int dic_image_upload(dic_ctx *ctx, const float *host, int width, int height, dic_image **out);
/* Borrow device memory the caller keeps alive (no copy). */
int dic_image_wrap_device(dic_ctx *ctx, const float *device_ptr, int width, int height, dic_image **out);
/* Copies the W x H pixels to host (Image::data(), image.hpp:58-81). */
int dic_image_download(dic_ctx *ctx, const dic_image *img, float *host);
void dic_image_free(dic_image *img);

/* ---- reference gradients ----
 * mode DIC_GRADIENT_CENTRAL replaces build_reference_gradient_cache
 *   (src/subset/seed/reliability_propagation.cpp:84-112; the pipeline convention);
 * mode DIC_GRADIENT_BSPLINE_HALF replaces the gradient_x/y planes of
 *   BSplineImagePreprocessor::compute (src/interpolation/bspline.cpp:282-294;
 *   used by standalone ICGNSolver::solve, icgn.cpp:441-443). */
int dic_ref_gradient(dic_ctx *ctx, const dic_image *ref, int mode, int degree, int border, int precision,
                     dic_grad **out);
/* Recompute central-difference gradients of ref into an existing handle of the same size. */
int dic_ref_gradient_into(dic_ctx *ctx, const dic_image *ref, dic_grad *out);
int dic_grad_download(dic_ctx *ctx, const dic_grad *g, double *gx_host, double *gy_host);
void dic_grad_free(dic_grad *g);

/* ---- B-spline prefilter: replaces BSplineImagePreprocessor::compute / compute_lazy
 * coefficient formation (src/interpolation/bspline.cpp:270-318, :369-392).
 * degree in {1,3,5}; border is raised to 3 (:261-268); exact = use_exact_prefilter.
 * precision selects FP64 or FP32 storage (arithmetic is FP64 either way). ---- */
int dic_bspline_prefilter(dic_ctx *ctx, const dic_image *img, int degree, int border, int exact, int precision,
                          dic_coeffs **out);
/* Recompute the coefficients of img into an existing handle of the same size
 * (its degree, border, exact flag and precision); no allocation after the
 * first call on a context.  New boundary: lets a resident handle follow a
 * sequence of deformed frames. */
int dic_bspline_prefilter_into(dic_ctx *ctx, const dic_image *img, dic_coeffs *out);
int dic_coeffs_info(const dic_coeffs *c, int *width, int *height, int *degree, int *border, int *precision);
/* Copies the padded coefficient image to host as FP64, row-major (H+2b) x (W+2b). */
int dic_coeffs_download(dic_ctx *ctx, const dic_coeffs *c, double *host);
void dic_coeffs_free(dic_coeffs *c);

/* ---- sampling hook: replaces BSplineInterpolator::value / gradient
 * (src/interpolation/bspline.cpp:458-500).  xy: 2n host doubles; val: n; grad: 2n
 * (either output may be NULL). ---- */
int dic_bspline_sample(dic_ctx *ctx, const dic_coeffs *c, const double *xy, int n, double *val, double *grad);

/* ---- IC-GN: replaces ICGNSolver::solve_with_interpolators / solve_with_mask
 * (src/subset/solver/icgn.cpp:384-432 -> :446-1912), one POI per entry.
 * With params->optimizer = DIC_OPTIMIZER_FGN the same entry points replace
 * ForwardGaussNewtonSolver::solve_with_interpolators / solve_with_mask
 * (src/subset/solver/forward_gauss_newton.cpp:117-175 -> :188-1941); grad is
 * then unused and may be NULL.
 * Largest subset: one POI's workgroup holds a 3.4 KiB solver block in LDS,
 * plus one double per sample for FP64 ZNSSD IC-GN (two-pass statistics).
 * On a 160 KiB-LDS device (MI355X) that is radius <= 79 for FP64 ZNSSD
 * IC-GN; beyond it the call returns DIC_ERR_UNSUPPORTED before any launch.
 * The reference has no such limit; nor do the other solves here.
 * Host-pointer form: copies in, runs, copies out, synchronises. ---- */
int dic_icgn_solve_batch(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                         const dic_coeffs *def, const uint8_t *roi_host, const double *poi_xy,
                         const dic_initial *init, int n, dic_result *out);
/* Device-pointer form: all arrays already in HBM; enqueued on `stream`
 * (NULL = the context's stream), no synchronisation. */
int dic_icgn_solve_batch_device(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref,
                                const dic_grad *grad, const dic_coeffs *def, const uint8_t *roi_device,
                                const double *poi_xy_device, const dic_initial *init_device, int n,
                                dic_result *out_device, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  dic_icgn_solve_batch_device with the same arguments,
 * kernels and results, but no allocation of any kind: the circular-subset
 * offset table, which the other entry allocates on the first use of each
 * radius, is rebuilt at every call by a small kernel on `stream` into a
 * static device array of DIC_NOALLOC_MAX_SAMPLES entries (128 KiB, held by
 * every GPU the library runs on until the process ends), so subset radii up
 * to DIC_NOALLOC_MAX_RADIUS are accepted and larger ones are refused with
 * DIC_ERR_UNSUPPORTED.  Each GPU has one such array: calls that overlap on
 * one GPU (two streams or two host threads) race on it. */
#define DIC_NOALLOC_MAX_SAMPLES 32768
#define DIC_NOALLOC_MAX_RADIUS 102
int dic_icgn_solve_batch_device_noalloc(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref,
                                        const dic_grad *grad, const dic_coeffs *def, const uint8_t *roi_device,
                                        const double *poi_xy_device, const dic_initial *init_device, int n,
                                        dic_result *out_device, void *stream);

/* ---- Batched solving: POIs of many image pairs, with uneven subset radii and
 * shape orders, in one launch captured in a hipGraph.  No counterpart in the
 * reference, which solves pair after pair and POI after POI (the callers
 * ReliabilityPropagation::propagate, reliability_propagation.cpp:233-354, and
 * the per-pair workflow loop); each POI's arithmetic is the same as
 * dic_icgn_solve_batch_device's (with the item's optimizer), bit for bit;
 * FGN items run one POI per workgroup whatever the mapping.  All coefficient images of a
 * batch share one B-spline degree and one precision; every array of an item
 * is a device pointer.  The item arrays must stay valid while the batch is
 * used; a batch is replayed with dic_batch_run as often as wanted (the inputs
 * may change between runs, the array addresses may not). ---- */
typedef struct dic_batch_item {
    const dic_image *reference;
    const dic_grad *gradient;
    const dic_coeffs *deformed;
    const double *poi_xy;         /* 2n doubles (device) */
    const dic_initial *init;      /* n (device) */
    dic_result *out;              /* n (device) */
    int32_t n;
    int32_t subset_radius;
    int32_t shape_order;          /* 1 or 2 */
    int32_t max_iterations;
    double convergence_threshold;
    int32_t optimizer;            /* DIC_OPTIMIZER_*: IC-GN (gradient required) or FGN (gradient unused) */
    int32_t reserved;             /* 0 */
} dic_batch_item;

enum { DIC_BATCH_GRID = 0, DIC_BATCH_PERSISTENT = 1, DIC_BATCH_TILE = 2 };
enum { DIC_BATCH_ORDER_NATURAL = 0, DIC_BATCH_ORDER_COST = 1, DIC_BATCH_ORDER_INTERLEAVE = 2, DIC_BATCH_ORDER_XCD = 3 };

/* Launch mapping chosen by dic_batch_autotune (or set by hand).  A batch is
 * one launch per shape order present (second order first), replayed as one
 * hipGraph.
 *   mapping DIC_BATCH_GRID: one workgroup (one 64-lane wave) per POI;
 *   mapping DIC_BATCH_PERSISTENT: waves_per_cu x CUs workgroups take `chunk`
 *     POIs at a time from a queue counter until the queue is empty;
 *   mapping DIC_BATCH_TILE: first-order IC-GN items of subset radius <= 24
 *     in a mixed-precision batch (those whose four neighbouring subsets can
 *     share one tile, the per-pair launch's rule) run four neighbouring POIs
 *     of one item per workgroup (one per wave) over a tile of the
 *     coefficients staged in LDS around their starting subsets (workgroups
 *     dealt over the 8 XCDs in contiguous runs); there the POIs of an item
 *     keep their natural order and `order` only ranks the items (by cost
 *     under COST).  All other items run as GRID in the given order.
 *   order: NATURAL = item after item; COST = items by descending estimated
 *     cost per POI (subset samples x shape order); INTERLEAVE = round robin
 *     over the items in runs of 256 POIs; XCD = natural order with each
 *     item's POIs dealt so that each of the 8 XCDs gets one contiguous eighth
 *     of them (grid mapping). */
typedef struct dic_batch_tuning {
    int32_t mapping;
    int32_t waves_per_cu;
    int32_t chunk;
    int32_t order;
} dic_batch_tuning;

typedef struct dic_batch dic_batch;

int dic_batch_create(dic_ctx *ctx, const dic_batch_item *items, int n_items, int objective, dic_batch **out);
int dic_batch_set_tuning(dic_batch *b, const dic_batch_tuning *t);
int dic_batch_get_tuning(const dic_batch *b, dic_batch_tuning *t);
/* Total POIs and the hipGraph's launch count so far. */
int dic_batch_info(const dic_batch *b, int64_t *total_pois, int64_t *runs);
/* Replays the batch's hipGraph on `stream` (NULL = the context's stream). */
int dic_batch_run(dic_batch *b, void *stream);
/* Times every candidate (NULL = the built-in sweep) `reps` times on the
 * context's stream, keeps the fastest (median) and reports it; the outputs
 * hold the last run's (identical) results. */
int dic_batch_autotune(dic_batch *b, const dic_batch_tuning *candidates, int n_candidates, int reps,
                       dic_batch_tuning *best, float *best_ms);
void dic_batch_destroy(dic_batch *b);

/* ---- Pinned host memory: frames handed to the server from pinned memory are
 * uploaded by the copy engines while the GPU solves earlier frames. ---- */
int dic_host_alloc(dic_ctx *ctx, size_t bytes, void **ptr);
void dic_host_free(void *ptr);

/* ---- Serving: a request loop with continuous batching.  No counterpart in the
 * reference (whose workflows correlate one frame pair at a time,
 * python/traditional_dic/workflows/subset_2d.py); every frame's arithmetic is
 * the IC-GN of dic_icgn_solve_batch_device.
 *
 * A request is an incremental frame sequence against one reference frame:
 * frame f is correlated with the reference at the request's POIs, starting
 * from the request's initial guesses (f < K) or from frame f-K's results
 * (u, v, gradients; confidence = correlation), K = guess_lag, or from their
 * extrapolation along the sequence (guess_extrapolate).  dic_server_submit
 * returns at once; a worker thread admits pending requests in arrival order
 * while the number of active requests stays below max_active and the device
 * bytes they hold (reference, gradients, 3K frame, coefficient, guess and
 * result buffers, POIs, the request's guesses) stay within
 * hbm_budget.  Each tick is one launch over the next frame of every active
 * request and, while the tick holds fewer than tick_pois POIs, up to K - 1
 * more of each; a request that arrives while a tick runs joins the next tick,
 * and a finished request leaves and frees its slot.  Three ticks are in
 * flight: uploads and prefilters of the next ticks (copy engines, prep
 * stream) and the copy-out of the previous one overlap the solve of the
 * current one.  Each POI's result depends only on its own inputs and the
 * frames its guesses come from, which K and guess_extrapolate fix, so a request's results
 * are the same served alone or in any mix.  The caller keeps every array of
 * a request valid until dic_server_wait reports it done. ---- */
#define DIC_SERVER_MAX_GUESS_LAG 16
#define DIC_SERVER_DEFAULT_TICK_POIS (1 << 21)
typedef struct dic_server_config {
    uint64_t hbm_budget;          /* device bytes held by admitted requests (0 = no cap) */
    int32_t max_active;           /* requests served at once */
    int32_t max_width, max_height; /* largest frame accepted */
    int32_t subset_radius;
    int32_t shape_order;          /* 1 or 2 */
    int32_t max_iterations;
    double convergence_threshold;
    int32_t degree;               /* B-spline degree 1, 3 or 5 */
    int32_t precision;            /* DIC_PRECISION_* */
    int32_t objective;            /* DIC_OBJECTIVE_* */
    int32_t optimizer;            /* DIC_OPTIMIZER_*: IC-GN, or FGN (ForwardGaussNewtonSolver) */
    int32_t guess_lag;            /* K in [0, DIC_SERVER_MAX_GUESS_LAG] (0 means 1): frame f starts from
                                   * frame f - K's results, frames below K from the request's guesses;
                                   * up to K frames of one request share a tick.  The reference's own
                                   * sequence workflow correlates every frame against the reference
                                   * (python/traditional_dic/workflows/subset_2d.py:152-227). */
    int32_t tick_pois;            /* a tick takes the next frame of every active request, then more
                                   * frames (up to K per request, arrival order) while it holds fewer
                                   * POIs than this (0: DIC_SERVER_DEFAULT_TICK_POIS) */
    int32_t guess_extrapolate;    /* 1: frames f >= 2K start from the linear extrapolation
                                   * 2 r(f-K) - r(f-2K) of u, v and the gradients where both earlier
                                   * results are Success (r(f-K) otherwise); 0: from r(f-K) */
    int32_t reserved;
} dic_server_config;

typedef struct dic_request {
    const float *reference;       /* width x height, host */
    int32_t width, height;
    int32_t n_frames;
    int32_t n;                    /* POIs */
    const float *const *frames;   /* n_frames host pointers, width x height each */
    const double *poi_xy;         /* 2n, host */
    const dic_initial *init;      /* n, host: guesses for frame 0 */
    dic_result *results;          /* n_frames x n, host, written as frames finish */
    double *frame_done_s;         /* n_frames completion times (dic_server_now clock), or NULL */
} dic_request;

typedef struct dic_server_stats {
    int64_t submitted, completed, frames_done, pois_done, ticks;
    int32_t active, pending, max_active_seen, reserved;
    uint64_t bytes_held, peak_bytes_held;
} dic_server_stats;

typedef struct dic_server dic_server;

int dic_server_create(dic_ctx *ctx, const dic_server_config *cfg, dic_server **out);
/* Queues a request; *id identifies it; *arrival_s (may be NULL) is its arrival time. */
int dic_server_submit(dic_server *s, const dic_request *req, int64_t *id, double *arrival_s);
/* Blocks until request id is done (DIC_OK), the server failed, or timeout_s
 * (< 0: no limit) passed (DIC_ERR_UNSUPPORTED). */
int dic_server_wait(dic_server *s, int64_t id, double timeout_s);
int dic_server_frames_done(dic_server *s, int64_t id, int *frames);
int dic_server_stats_get(dic_server *s, dic_server_stats *st);
/* Seconds since dic_server_create on the clock of arrival and completion times. */
double dic_server_now(const dic_server *s);
const char *dic_server_error(const dic_server *s);
/* Finishes all submitted requests, then stops the worker and frees everything. */
void dic_server_destroy(dic_server *s);

/* ---- Reliability-guided propagation over the POI grid: replaces
 * ReliabilityPropagation::propagate (src/subset/seed/reliability_propagation.cpp:130-394).
 * Grid: step = spacing + 1, ceil(W / step) x ceil(H / step) points, row-major.
 * Each seed (aligned to its nearest grid point and re-solved there when it
 * moved, :206-262) starts a walk over 4-neighbours with the parent's affine
 * extrapolation as guess (:314-327) and the reference's acceptance test
 * (Success, valid, correlation < max_znssd, |du|, |dv| < step from the guess,
 * finite; :359-368); rejected, unvisited and out-of-ROI points keep the
 * default record (status InvalidInput, valid 0, x and y set).  The reference
 * pops one point at a time from a heap ordered by correlation:
 *   DIC_PROPAGATE_HEAP     -- that order exactly: the host replays the heap
 *                             walk (the same std::priority_queue operations)
 *                             and takes each solve from speculative launches
 *                             over the heap's points in pop order, so every
 *                             point is solved from the parent the reference
 *                             would use and the valid set is the reference's
 *                             at any max_znssd;
 *   DIC_PROPAGATE_FRONTIER -- every frontier (all points accepted in the
 *                             previous wave) is expanded at once, best
 *                             correlation first, and its new neighbours are
 *                             solved in one launch: fewer, larger launches,
 *                             but which parent seeds a neighbour, and hence
 *                             which points a strict max_znssd rejects, can
 *                             differ from the reference's order.
 * params->masked plays truncate_roi_subsets (solve_with_mask).  roi is host
 * memory, W x H. */
enum { DIC_PROPAGATE_HEAP = 0, DIC_PROPAGATE_FRONTIER = 1 };
typedef struct dic_propagation_params {
    int32_t spacing;      /* propagation_spacing (>= 0) */
    int32_t order;        /* DIC_PROPAGATE_* */
    double max_znssd;     /* propagation_max_znssd */
} dic_propagation_params;

typedef struct dic_seed {
    double x, y;          /* PropagationSeed::point */
    dic_initial displacement; /* u, v, gradients; confidence = correlation */
} dic_seed;

/* grid_out must hold ceil(W/step) * ceil(H/step) records; *grid_w, *grid_h,
 * *points_computed (may be NULL) receive the grid size and the count. */
int dic_propagate(dic_ctx *ctx, const dic_icgn_params *params, const dic_propagation_params *prop,
                  const dic_image *ref, const dic_grad *grad, const dic_coeffs *def, const uint8_t *roi_host,
                  const dic_seed *seeds, int n_seeds, dic_result *grid_out, int *grid_w, int *grid_h,
                  int *points_computed);

/* ---- Seeding: the step before propagation ----
 * Integer-pixel search: replaces IntegerSearchInitializer
 * (src/initialization/integer_search.cpp:303-645).  Per point, the ZNCC of a
 * circular subset (radius subset_radius, FP64, the reference's two-pass
 * mean/norm in its summation order, no fused multiply-add) is evaluated at
 * integer shifts and the first strict maximum in row-major scan order wins:
 *   DIC_SEARCH_GLOBAL  (estimate / estimate_with_mask): with pyramid_enabled,
 *     every shift that keeps the subset inside the image on a grid of step
 *     pyramid_scale with the subsampled subset, then the full subset within
 *     max(pyramid_refinement_radius, ceil(1.5 r), step) of the coarse winner;
 *     without it, every shift at step 1.  With search_radius > 0 the window
 *     [-search_radius, search_radius]^2 is searched too and preferred when its
 *     ZNCC + 0.02 >= the global one (:267-280).
 *   DIC_SEARCH_AROUND_PRIOR (estimate[_with_mask]_around_displacement): the
 *     window of radius search_radius around llround(prior).
 * masked = 1 restricts the subset samples to the ROI (the _with_mask variants).
 * Where the reference's unmasked around-prior form reads outside the image
 * (it throws), the point is reported invalid. */
enum { DIC_SEARCH_GLOBAL = 0, DIC_SEARCH_AROUND_PRIOR = 1 };

typedef struct dic_integer_search_params { /* SeedIntegerSearchConfig, seed_config.hpp:33-40 */
    int32_t subset_radius;             /* >= 1 */
    int32_t search_radius;             /* >= 0 */
    int32_t pyramid_enabled;
    int32_t pyramid_scale;             /* >= 1 */
    int32_t pyramid_refinement_radius; /* >= 1 */
    int32_t mode;                      /* DIC_SEARCH_* */
    int32_t masked;                    /* 1: roi required */
    int32_t reserved;
} dic_integer_search_params;

/* InitialDisplacement with every field (initializer.hpp:33-44). */
typedef struct dic_estimate {
    double u, v, du_dx, du_dy, dv_dx, dv_dy, confidence;
    int32_t valid, reserved;
    double zncc, znssd;
} dic_estimate;

/* points_xy: 2n host doubles; prior_uv: 2n (AROUND_PRIOR only, else NULL);
 * roi_host: W x H bytes (masked only); out: n host records.  ref and def must
 * have the same size (else every point is invalid, as in the reference). */
int dic_integer_search(dic_ctx *ctx, const dic_integer_search_params *params, const dic_image *ref,
                       const dic_image *def, const uint8_t *roi_host, const double *points_xy,
                       const double *prior_uv, int n, dic_estimate *out);

/* Seed selection: replaces SeedSelector::generate_candidates and
 * select_best_seed (src/subset/seed/seed_selector.cpp:146-316) on the
 * integer-search path, with SubsetInitializer's sub-pixel refinement
 * (src/initialization/subset_initializer.cpp:128-264).  Candidates are the
 * rounded k-means centres of (a deterministic sample of) the ROI points that
 * keep the subset margin, or, where a centre fails the ROI / margin / texture
 * test, the most textured of the 50 sampled points nearest to it (equally
 * distant points are taken in sampled order; the reference's std::sort leaves
 * that order unspecified).  Each candidate is searched (GLOBAL, masked when
 * truncate_roi_subsets) and refined by IC-GN on the non-prefiltered spline of
 * the deformed image with central-difference reference gradients (or by
 * forward Gauss-Newton on the prefiltered spline), all in FP64.  The SIFT
 * prior (OpenCV) is not part of this library.  seed_count <= 2048. */
enum { DIC_QUALITY_ZNCC = 0, DIC_QUALITY_ZNSSD = 1, DIC_QUALITY_SSD = 2 };

typedef struct dic_seed_params {
    dic_integer_search_params search;  /* mode and masked are set by the selector */
    int32_t subpixel_enabled;          /* SeedSubpixelRefinementConfig, seed_config.hpp:42-50 */
    int32_t subpixel_shape_order;      /* 1 or 2 */
    int32_t subpixel_optimizer;        /* DIC_OPTIMIZER_* */
    int32_t subpixel_objective;        /* DIC_OBJECTIVE_* */
    int32_t subpixel_radius;
    int32_t subpixel_max_iterations;
    double subpixel_convergence_threshold;
    int32_t seed_count;                /* SeedSelectionConfig, seed_config.hpp:58-69 */
    int32_t quality_metric;            /* DIC_QUALITY_* */
    double min_zncc, max_znssd, max_ssd, min_displacement_norm, min_texture_std;
    int32_t kmeans_iterations;
    int32_t kmeans_sample_limit;
    int32_t truncate_roi_subsets;      /* SubsetConfig, subset_config.hpp:33-50 */
    int32_t min_valid_samples;
    double min_valid_sample_ratio;
    int32_t degree;                    /* image_precompute: spline degree and border */
    int32_t border;
} dic_seed_params;

/* SeedEvaluation (seed_selector.hpp:20-27). */
typedef struct dic_seed_evaluation {
    double x, y;
    dic_result displacement;           /* Displacement2D; iterations unused */
    double quality, displacement_norm;
    int32_t quality_passed, valid;
} dic_seed_evaluation;

/* Candidate points (host xy pairs, at most capacity) into xy_out, count in *n_out. */
int dic_seed_candidates(dic_ctx *ctx, const dic_seed_params *params, const dic_image *ref,
                        const uint8_t *roi_host, double *xy_out, int capacity, int *n_out);
/* Evaluations of every candidate (at most capacity) and the index of the best
 * seed in *best_index (-1: SeedSelectionResult::found == false). */
int dic_select_seed(dic_ctx *ctx, const dic_seed_params *params, const dic_image *ref, const dic_image *def,
                    const uint8_t *roi_host, dic_seed_evaluation *evaluations, int capacity, int *n_out,
                    int *best_index);

/* mirror_pad_image (src/subset/padding.cpp:11-59): a new (W+2p) x (H+2p) image
 * whose border repeats the image mirrored about its edges (edge pixels doubled). */
int dic_image_mirror_pad(dic_ctx *ctx, const dic_image *img, int pad, dic_image **out);

#ifdef __cplusplus
}
#endif
#endif
