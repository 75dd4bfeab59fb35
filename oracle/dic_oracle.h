/*
 * dic_oracle.h -- CPU FP64 restatement of Traditional-DIC's subset-correlation
 * hot path (B-spline prefilter + sampling, reference gradients, IC-GN solvers).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product library links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Traditional-DIC source tree).  Arithmetic is FP64 throughout, with the
 * reference's serial summation order (dy-outer / dx-inner subset loops,
 * two-pass mean/norm) so that its rounding behaviour is close to the
 * reference's; bit identity with Eigen is not claimed.
 */
#ifndef DIC_ORACLE_H
#define DIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SolverStatus enum order: include/dic/core/types.hpp:35-40 */
enum { ORC_SUCCESS = 0, ORC_NOT_CONVERGED = 1, ORC_INVALID_INPUT = 2, ORC_NUMERICAL_FAILURE = 3 };

/* Displacement2D (include/dic/core/result.hpp:32-51) plus the executed
 * iteration count.  Layout is shared with the HIP library's dic_result. */
typedef struct orc_result {
    double x, y;
    double p[12]; /* u v du_dx du_dy dv_dx dv_dy d2u_dx2 d2u_dxdy d2u_dy2 d2v_dx2 d2v_dxdy d2v_dy2 */
    double correlation;
    int32_t status;
    int32_t valid;
    int32_t iterations;
    int32_t reserved;
} orc_result;

/* InitialDisplacement (include/dic/initialization/initializer.hpp:33-44):
 * the six first-order parameters and the confidence used on failure exits. */
typedef struct orc_initial {
    double u, v, du_dx, du_dy, dv_dx, dv_dy, confidence;
} orc_initial;

/* SubsetConfig fields read by ICGNSolver (include/dic/subset/subset_config.hpp:33-50). */
typedef struct orc_config {
    int32_t subset_radius;
    int32_t max_iterations;
    double convergence_threshold;
    int32_t second_order;      /* shape_function == SecondOrder || use_second_order */
    int32_t objective;         /* 0 = ZNSSD, 1 = SSD */
    int32_t min_valid_samples; /* masked variants only */
    int32_t reserved;
    double min_valid_sample_ratio;
} orc_config;

/* BSplinePrecomputedImage (include/dic/interpolation/bspline.hpp:69-81), row-major. */
typedef struct orc_precomputed {
    int32_t width, height, degree, border, n;
    int32_t has_blocks;
    double qk[36];     /* n x n, row = derivative order */
    double *coeffs;    /* (height + 2*border) x (width + 2*border), row-major */
    double *blocks;    /* width*height blocks of n*n (row = dy power, col = dx power), or NULL */
    double *grad_x;    /* height x width, B-spline derivative at (+0.5,+0.5) */
    double *grad_y;
} orc_precomputed;

/* --- B-spline (src/interpolation/bspline.cpp) --- */
double orc_basis(double x, int derivative_order, int degree);              /* :342-367 */
void orc_build_qk(int degree, double *qk_out);                             /* :320-340 */
int orc_symmetric_index(int value, int size);                              /* :78-89 */
/* Circular deconvolution of one periodic line by the sampled B-spline kernel
 * (:177-214), restated as the exact periodic recursive inverse. */
void orc_deconvolve_periodic(double *line, int length, int degree);
/* form_coefficients (:369-392): pad + rows + columns.  out is
 * (H+2b) x (W+2b) with b = max(3, border). Returns b. */
int orc_form_coefficients(const float *image, int width, int height, int degree,
                          int border, int exact, double *out);
/* compute (:270-300) when with_blocks != 0, else compute_lazy (:302-318). */
int orc_precompute(const float *image, int width, int height, int degree, int border,
                   int exact, int with_blocks, orc_precomputed *out);
void orc_precomputed_free(orc_precomputed *pre);
/* Threads of the OpenMP regions without an explicit thread count (the
 * per-pixel block build of orc_precompute, orc_value_batch) started from the
 * calling thread; n <= 0 restores the OpenMP default. */
void orc_set_threads(int n);
double orc_value(const orc_precomputed *pre, double x, double y);                 /* :458-474 */
void orc_gradient(const orc_precomputed *pre, double x, double y, double *gxgy);  /* :476-500 */
void orc_value_batch(const orc_precomputed *pre, const double *xy, int n, double *val, double *grad);

/* --- reference gradients --- */
/* build_reference_gradient_cache: src/subset/seed/reliability_propagation.cpp:84-112 */
void orc_reference_gradient_cache(const float *image, int width, int height, double *gx, double *gy);

/* --- dense small-matrix kernels (Eigen 3.4.0 semantics) --- */
/* LDLT with diagonal pivoting; returns info (0 ok), *is_positive */
int orc_ldlt_factor(double *a, int n, int *transpositions, int *is_positive);
void orc_ldlt_solve(const double *a, int n, const int *transpositions, const double *rhs, double *x);
/* FullPivLU invertibility + inverse of a 6x6 matrix; returns 1 if invertible */
int orc_fullpivlu_inverse6(const double *m, double *inv);

/* --- warp updates (src/subset/solver/icgn.cpp) --- */
void orc_ic_affine_update(const double *p, const double *dp, double *out);        /* :83-120 */
void orc_second_order_warp_matrix(const double *p, double *g);                    /* :196-261 */
void orc_ic_second_order_update(const double *p, const double *dp, double *out);  /* :289-357 */

/* --- IC-GN solver (src/subset/solver/icgn.cpp:361-1912) --- */
/* grad_x/grad_y: H x W planes used for the reference gradient
 * (reference_gradient_at, :55-71).  roi may be NULL (unmasked variants);
 * when non-NULL the masked variants run. */
void orc_icgn_solve(const orc_config *cfg,
                    const float *reference, const float *deformed_unused,
                    int width, int height,
                    const double *grad_x, const double *grad_y,
                    const orc_precomputed *deformed,
                    const uint8_t *roi,
                    double px, double py, const orc_initial *init, orc_result *out);
/* Batch over POIs; threads <= 0 means all OpenMP threads. */
void orc_icgn_solve_batch(const orc_config *cfg, const float *reference,
                          int width, int height,
                          const double *grad_x, const double *grad_y,
                          const orc_precomputed *deformed, const uint8_t *roi,
                          const double *poi_xy, const orc_initial *init, int n,
                          orc_result *out, int threads);

/* --- Forward-compositional Gauss-Newton (src/subset/solver/forward_gauss_newton.cpp) --- */
/* compose_warp, :64-76 (first order) and :1025-1053 (second order) */
void orc_fgn_compose(const double *p, const double *dp, int P, double *out);
/* One POI.  Uses only the deformed image's interpolant (value and gradient);
 * the reference image gives f.  roi may be NULL (unmasked variants). */
void orc_fgn_solve(const orc_config *cfg, const float *reference, int width, int height,
                   const orc_precomputed *deformed, const uint8_t *roi, double px, double py,
                   const orc_initial *init, orc_result *out);
void orc_fgn_solve_batch(const orc_config *cfg, const float *reference, int width, int height,
                         const orc_precomputed *deformed, const uint8_t *roi, const double *poi_xy,
                         const orc_initial *init, int n, orc_result *out, int threads);

/* --- Reliability-guided propagation (src/subset/seed/reliability_propagation.cpp:130-394) --- */
/* seeds: n x 9 doubles (x, y, u, v, du_dx, du_dy, dv_dx, dv_dy, correlation).
 * grid: ceil(W / step) x ceil(H / step) results, step = spacing + 1.  The solver
 * is IC-GN (optimizer 0) or forward Gauss-Newton (1); truncate_roi selects the
 * masked solves.  Returns points_computed. */
int orc_propagate(const orc_config *cfg, int optimizer, int spacing, double max_znssd, int truncate_roi,
                  const float *reference, int width, int height, const double *grad_x, const double *grad_y,
                  const orc_precomputed *deformed, const uint8_t *roi, const double *seeds, int n_seeds,
                  orc_result *grid);

/* POI grid of ReliabilityPropagation::propagate (reliability_propagation.cpp:154-176). */
int orc_poi_grid(int width, int height, int spacing, double *xy_out /* may be NULL */);

/* --- Seeding: integer-pixel ZNCC search (src/initialization/integer_search.cpp)
 * and seed selection (src/subset/seed/seed_selector.cpp,
 * src/initialization/subset_initializer.cpp).  Layouts are shared with the HIP
 * library's dic_integer_search_params / dic_estimate / dic_seed_params /
 * dic_seed_evaluation. --- */
typedef struct orc_search_params {
    int32_t subset_radius;             /* SeedIntegerSearchConfig (seed_config.hpp:33-40) */
    int32_t search_radius;
    int32_t pyramid_enabled;
    int32_t pyramid_scale;
    int32_t pyramid_refinement_radius;
    int32_t around_prior;              /* 1: estimate[_with_mask]_around_displacement */
    int32_t masked;                    /* 1: the _with_mask variants (roi required) */
    int32_t reserved;
} orc_search_params;

/* InitialDisplacement with every field (initializer.hpp:33-44). */
typedef struct orc_estimate {
    double u, v, du_dx, du_dy, dv_dx, dv_dy, confidence;
    int32_t valid, reserved;
    double zncc, znssd;
} orc_estimate;

typedef struct orc_seed_params {
    orc_search_params search;          /* around_prior / masked are set by the selector */
    int32_t subpixel_enabled;          /* SeedSubpixelRefinementConfig (seed_config.hpp:42-50) */
    int32_t subpixel_shape_order;      /* 1 or 2 */
    int32_t subpixel_optimizer;        /* 0 IC-GN, 1 forward Gauss-Newton */
    int32_t subpixel_objective;        /* 0 ZNSSD, 1 SSD */
    int32_t subpixel_radius;
    int32_t subpixel_max_iterations;
    double subpixel_convergence_threshold;
    int32_t seed_count;                /* SeedSelectionConfig (seed_config.hpp:58-69) */
    int32_t quality_metric;            /* 0 ZNCC, 1 ZNSSD, 2 SSD */
    double min_zncc, max_znssd, max_ssd, min_displacement_norm, min_texture_std;
    int32_t kmeans_iterations;
    int32_t kmeans_sample_limit;
    int32_t truncate_roi_subsets;      /* SubsetConfig (subset_config.hpp:33-50) */
    int32_t min_valid_samples;
    double min_valid_sample_ratio;
    int32_t degree;                    /* image_precompute of the seed interpolant */
    int32_t border;
} orc_seed_params;

/* SeedEvaluation (seed_selector.hpp:20-27). */
typedef struct orc_seed_eval {
    double x, y;
    orc_result displacement;
    double quality, displacement_norm;
    int32_t quality_passed, valid;
} orc_seed_eval;

/* IntegerSearchInitializer::estimate / estimate_with_mask /
 * estimate[_with_mask]_around_displacement (integer_search.cpp:303-645) for n
 * points.  prior_uv: 2n (around_prior only).  roi: W x H (masked only). */
void orc_integer_search(const orc_search_params *p, const float *reference, const float *deformed, int width,
                        int height, const uint8_t *roi, const double *xy, const double *prior_uv, int n,
                        orc_estimate *out, int threads);
/* ZNCC of one circular subset at integer shift (u, v) (circular_subset_zncc,
 * :127-165, full circle at sample step 1); -inf where the reference does. */
double orc_subset_zncc(const float *reference, const float *deformed, int width, int height, int radius, int cx,
                       int cy, int u, int v);
/* SeedSelector::local_texture_std (seed_selector.cpp:421-446). */
double orc_texture_std(const float *reference, int width, int height, int x, int y, int radius);
/* deterministic_sample index map (seed_selector.cpp:34-48): idx_out[limit'] for n points. */
int orc_deterministic_sample(int n, int limit, int *idx_out);
/* initialize_centers + run_kmeans (seed_selector.cpp:50-110) over integer points (xy pairs). */
int orc_kmeans(const int *points, int n, int requested, int iterations, double *centers_out);
/* SeedSelector::generate_candidates (seed_selector.cpp:208-316); returns the count (<= cap). */
int orc_seed_candidates(const orc_seed_params *p, const float *reference, int width, int height,
                        const uint8_t *roi, double *xy_out, int cap);
/* SeedSelector::select_best_seed (seed_selector.cpp:146-206): evaluations of every
 * candidate into out (<= cap), *n_out; returns the index of the best seed or -1. */
int orc_select_seed(const orc_seed_params *p, const float *reference, const float *deformed, int width, int height,
                    const uint8_t *roi, orc_seed_eval *out, int cap, int *n_out);

#ifdef __cplusplus
}
#endif
#endif
