// dic_hip.hpp -- C++17 host layer over the C-ABI (dic_hip.h), mirroring the
// surface Traditional-DIC's C++ core calls on this path:
//   dic::SubsetSolver / ICGNSolver      include/dic/subset/solver/subset_solver.hpp:37-67,
//                                       include/dic/subset/solver/icgn.hpp:35-124
//   ForwardGaussNewtonSolver            include/dic/subset/solver/forward_gauss_newton.hpp:11-107
//   BSplineImagePreprocessor::compute   include/dic/interpolation/bspline.hpp:84-100
//   BSplineInterpolator::value/gradient include/dic/interpolation/bspline.hpp:102-130
//   IntegerSearchInitializer            include/dic/initialization/integer_search.hpp:32-71
//   SeedSelector::select_best_seed      include/dic/subset/seed/seed_selector.hpp:20-64
//   Displacement2D, InitialDisplacement, SolverStatus, SubsetConfig, the seed configs
// Types are Eigen-free (points are plain doubles); each method keeps the
// reference's argument meaning and reports per-POI outcomes in the result's
// status exactly as the reference does.  Errors of the C-ABI (bad handles,
// HIP failures) surface as std::runtime_error, like the reference's
// exceptions for invalid configuration.
#ifndef DIC_HIP_HPP
#define DIC_HIP_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dic_hip.h"

namespace dic_hip {

enum class SolverStatus { Success = 0, NotConverged = 1, InvalidInput = 2, NumericalFailure = 3 };
enum class Precision { FP64 = DIC_PRECISION_FP64, Mixed = DIC_PRECISION_MIXED };

struct Displacement2D {
    double x{0.0}, y{0.0}, u{0.0}, v{0.0};
    double du_dx{0.0}, du_dy{0.0}, dv_dx{0.0}, dv_dy{0.0};
    double d2u_dx2{0.0}, d2u_dxdy{0.0}, d2u_dy2{0.0}, d2v_dx2{0.0}, d2v_dxdy{0.0}, d2v_dy2{0.0};
    double correlation{0.0};
    SolverStatus status{SolverStatus::InvalidInput};
    bool valid{false};
    int iterations{0};
};

struct InitialDisplacement {
    double u{0.0}, v{0.0}, du_dx{0.0}, du_dy{0.0}, dv_dx{0.0}, dv_dy{0.0};
    double confidence{0.0};
};

struct BSplinePrecomputeConfig {
    int degree{5};
    int border{3};
    bool use_exact_prefilter{true};
};

// include/dic/initialization/seed_config.hpp:33-69
struct SeedIntegerSearchConfig {
    int subset_radius{10};
    int search_radius{30};
    bool pyramid_enabled{true};
    int pyramid_scale{4};
    int pyramid_refinement_radius{4};
};

struct SeedSubpixelRefinementConfig {
    bool enabled{true};
    bool second_order{false};         // shape_function == SecondOrder
    bool forward_gauss_newton{false}; // optimizer == ForwardGaussNewton
    bool ssd{false};                  // objective == SSD
    int subset_radius{15};
    int max_iterations{30};
    double convergence_threshold{1e-3};
};

enum class SeedQualityMetric { ZNCC = DIC_QUALITY_ZNCC, ZNSSD = DIC_QUALITY_ZNSSD, SSD = DIC_QUALITY_SSD };

struct SeedSelectionConfig {
    int seed_count{256};
    SeedQualityMetric quality_metric{SeedQualityMetric::ZNSSD};
    double min_zncc{0.7};
    double max_znssd{0.2};
    double max_ssd{0.05};
    double min_displacement_norm{0.0};
    double min_texture_std{0.0};
    int kmeans_iterations{20};
    int kmeans_sample_limit{20000};
};

struct SubsetConfig {
    int subset_radius{15};
    int max_iterations{30};
    double convergence_threshold{1e-3};
    bool use_second_order{false};
    bool second_order_shape{false}; // shape_function == SecondOrder
    bool ssd{false};                // objective == SSD (else ZNSSD)
    bool forward_gauss_newton{false}; // optimization == FGN (else IC-GN)
    bool truncate_roi_subsets{false};
    double min_valid_sample_ratio{0.5};
    int min_valid_samples{12};
    BSplinePrecomputeConfig image_precompute{};
    SeedIntegerSearchConfig seed_integer_search{};
    SeedSubpixelRefinementConfig seed_subpixel{};
    SeedSelectionConfig seed_selection{};
};

// InitialDisplacement with valid / zncc / znssd (the integer search's result).
struct InitialEstimate {
    double u{0.0}, v{0.0}, du_dx{0.0}, du_dy{0.0}, dv_dx{0.0}, dv_dy{0.0};
    double confidence{0.0};
    bool valid{false};
    double zncc{0.0}, znssd{0.0};
};

inline void check(dic_ctx *ctx, int rc)
{
    if (rc != DIC_OK) throw std::runtime_error(std::string("dic_hip: ") + dic_last_error(ctx));
}

class Context {
public:
    explicit Context(int device = 0, uint64_t hbm_budget = 0)
    {
        int rc = dic_ctx_create(device, hbm_budget, &ctx_);
        if (rc != DIC_OK) throw std::runtime_error("dic_hip: no usable HIP device (code " + std::to_string(rc) + ")");
    }
    ~Context() { dic_ctx_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    dic_ctx *get() const { return ctx_; }

private:
    dic_ctx *ctx_ = nullptr;
};

// dic::Image: row-major float32, resident on the device.
class Image {
public:
    Image(Context &ctx, int width, int height, const float *data) : ctx_(ctx.get())
    {
        check(ctx_, dic_image_upload(ctx_, data, width, height, &img_));
        width_ = width;
        height_ = height;
    }
    ~Image() { dic_image_free(img_); }
    Image(const Image &) = delete;
    Image &operator=(const Image &) = delete;
    int width() const { return width_; }
    int height() const { return height_; }
    const dic_image *get() const { return img_; }

private:
    dic_ctx *ctx_;
    dic_image *img_ = nullptr;
    int width_ = 0, height_ = 0;
};

// BSplinePrecomputedImage equivalent: padded coefficient image in HBM.
class BSplineImage {
public:
    BSplineImage(Context &ctx, const Image &img, BSplinePrecomputeConfig cfg = {}, Precision p = Precision::FP64)
        : ctx_(ctx.get())
    {
        check(ctx_, dic_bspline_prefilter(ctx_, img.get(), cfg.degree, cfg.border, cfg.use_exact_prefilter ? 1 : 0,
                                          (int)p, &c_));
    }
    ~BSplineImage() { dic_coeffs_free(c_); }
    BSplineImage(const BSplineImage &) = delete;
    BSplineImage &operator=(const BSplineImage &) = delete;
    const dic_coeffs *get() const { return c_; }

private:
    dic_ctx *ctx_;
    dic_coeffs *c_ = nullptr;
};

// BSplineInterpolator::value / gradient over a device coefficient image.
class BSplineInterpolator {
public:
    BSplineInterpolator(Context &ctx, const BSplineImage &img) : ctx_(ctx.get()), img_(&img) {}
    double value(double x, double y) const
    {
        double xy[2] = {x, y}, v = 0.0;
        check(ctx_, dic_bspline_sample(ctx_, img_->get(), xy, 1, &v, nullptr));
        return v;
    }
    void gradient(double x, double y, double *gx, double *gy) const
    {
        double xy[2] = {x, y}, g[2] = {0.0, 0.0};
        check(ctx_, dic_bspline_sample(ctx_, img_->get(), xy, 1, nullptr, g));
        *gx = g[0];
        *gy = g[1];
    }
    std::vector<double> values(const std::vector<double> &xy) const
    {
        std::vector<double> v(xy.size() / 2);
        check(ctx_, dic_bspline_sample(ctx_, img_->get(), xy.data(), (int)v.size(), v.data(), nullptr));
        return v;
    }

private:
    dic_ctx *ctx_;
    const BSplineImage *img_;
};

// dic::Mask: W x H bytes, nonzero = valid (src/core/mask.cpp:95-98).
struct Mask {
    int width = 0, height = 0;
    std::vector<uint8_t> valid; // row-major, width * height
};

namespace detail {
// What a prepared pair was built from and how: the solver checks it against
// the images and configuration of every call that hands the pair over.
struct PairKey {
    const void *reference = nullptr, *deformed = nullptr;
    int degree = 0, border = 0, exact = 0, precision = 0, gradient_mode = 0;
};
inline void check_pair(const PairKey &have, const PairKey &want)
{
    if (have.reference != want.reference || have.deformed != want.deformed)
        throw std::invalid_argument("dic_hip: the prepared pair was built from other images than the ones passed");
    if (have.degree != want.degree || have.border != want.border || have.exact != want.exact)
        throw std::invalid_argument("dic_hip: the prepared pair's B-spline settings differ from the solver's");
    if (have.precision != want.precision || have.gradient_mode != want.gradient_mode)
        throw std::invalid_argument("dic_hip: the prepared pair's precision or gradient mode differs from the solver's");
}
} // namespace detail

// The interpolators of one (reference, deformed) pair: the reference-gradient
// cache and the deformed image's coefficients, resident on the device
// (reliability_propagation.cpp:182-185 builds the same two objects).  It
// remembers which Image objects it was built from; both must outlive it.
class PreparedPair {
public:
    PreparedPair(Context &ctx, const Image &reference, const Image &deformed, const BSplinePrecomputeConfig &ip,
                 Precision p, int gradient_mode = DIC_GRADIENT_CENTRAL)
        : ctx_(ctx.get())
    {
        key_ = {&reference, &deformed, ip.degree, ip.border, ip.use_exact_prefilter ? 1 : 0, (int)p, gradient_mode};
        check(ctx_, dic_ref_gradient(ctx_, reference.get(), gradient_mode, ip.degree, ip.border, (int)p, &grad_));
        const int rc = dic_bspline_prefilter(ctx_, deformed.get(), ip.degree, ip.border, key_.exact, (int)p, &coeffs_);
        if (rc != DIC_OK) {
            dic_grad_free(grad_);
            check(ctx_, rc);
        }
    }
    ~PreparedPair()
    {
        dic_grad_free(grad_);
        dic_coeffs_free(coeffs_);
    }
    PreparedPair(PreparedPair &&o) noexcept : ctx_(o.ctx_), key_(o.key_), grad_(o.grad_), coeffs_(o.coeffs_)
    {
        o.grad_ = nullptr;
        o.coeffs_ = nullptr;
    }
    PreparedPair(const PreparedPair &) = delete;
    PreparedPair &operator=(const PreparedPair &) = delete;
    PreparedPair &operator=(PreparedPair &&) = delete;
    const detail::PairKey &key() const { return key_; }
    const dic_grad *gradient() const { return grad_; }
    const dic_coeffs *coefficients() const { return coeffs_; }

private:
    dic_ctx *ctx_;
    detail::PairKey key_;
    dic_grad *grad_ = nullptr;
    dic_coeffs *coeffs_ = nullptr;
};

// ICGNSolver: solve / solve_with_interpolators / solve_with_mask with the
// reference's arguments, plus solve_batch over many POIs (the form the GPU
// wants).  solve() prepares the pair for the call, as the reference's
// standalone solve() builds its interpolators (icgn.cpp:363-382); the other
// forms take a PreparedPair and throw std::invalid_argument when it was built
// from other Image objects or settings.  Images of different sizes, or an
// ROI of another size, give InvalidInput records with the initial u, v
// (icgn.cpp:464-470, :648-652), as in the reference.
class ICGNSolver {
public:
    explicit ICGNSolver(Context &ctx, SubsetConfig cfg = {}, Precision p = Precision::FP64,
                        int gradient_mode = DIC_GRADIENT_CENTRAL)
        : ctx_(&ctx), cfg_(cfg), precision_(p), gradient_mode_(gradient_mode)
    {
    }

    PreparedPair prepare(const Image &reference, const Image &deformed) const
    {
        return PreparedPair(*ctx_, reference, deformed, cfg_.image_precompute, precision_, gradient_mode_);
    }

    std::vector<Displacement2D> solve_batch(const Image &reference, const Image &deformed,
                                            const std::vector<double> &poi_xy,
                                            const std::vector<InitialDisplacement> &initial,
                                            const PreparedPair *pair = nullptr, const Mask *roi = nullptr) const
    {
        const int n = (int)initial.size();
        if ((int)poi_xy.size() != 2 * n) throw std::invalid_argument("dic_hip: one (x, y) per initial guess");
        if (reference.width() != deformed.width() || reference.height() != deformed.height() ||
            (roi && (roi->width != reference.width() || roi->height != reference.height() ||
                     roi->valid.size() != (size_t)roi->width * (size_t)roi->height)))
            return invalid(poi_xy, initial);
        if (!pair) {
            const PreparedPair own = prepare(reference, deformed);
            return run(reference, own, poi_xy, initial, roi);
        }
        const auto &ip = cfg_.image_precompute;
        detail::check_pair(pair->key(), {&reference, &deformed, ip.degree, ip.border, ip.use_exact_prefilter ? 1 : 0,
                                         (int)precision_, gradient_mode_});
        return run(reference, *pair, poi_xy, initial, roi);
    }

    Displacement2D solve(const Image &reference, const Image &deformed, double x, double y,
                         const InitialDisplacement &initial) const
    {
        return solve_batch(reference, deformed, {x, y}, {initial})[0];
    }

    Displacement2D solve_with_interpolators(const Image &reference, const Image &deformed, double x, double y,
                                            const InitialDisplacement &initial, const PreparedPair &pair) const
    {
        return solve_batch(reference, deformed, {x, y}, {initial}, &pair)[0];
    }

    Displacement2D solve_with_mask(const Image &reference, const Image &deformed, const Mask &roi, double x, double y,
                                   const InitialDisplacement &initial, const PreparedPair &pair) const
    {
        return solve_batch(reference, deformed, {x, y}, {initial}, &pair, &roi)[0];
    }

private:
    static std::vector<Displacement2D> invalid(const std::vector<double> &poi_xy,
                                               const std::vector<InitialDisplacement> &initial)
    {
        std::vector<Displacement2D> out(initial.size());
        for (size_t i = 0; i < initial.size(); ++i) {
            out[i].x = poi_xy[2 * i];
            out[i].y = poi_xy[2 * i + 1];
            out[i].u = initial[i].u;
            out[i].v = initial[i].v;
            out[i].correlation = initial[i].confidence;
            out[i].status = SolverStatus::InvalidInput;
            out[i].valid = false;
        }
        return out;
    }

    std::vector<Displacement2D> run(const Image &reference, const PreparedPair &pair, const std::vector<double> &poi_xy,
                                    const std::vector<InitialDisplacement> &initial, const Mask *roi) const
    {
        const int n = (int)initial.size();
        dic_icgn_params p{};
        p.subset_radius = cfg_.subset_radius;
        p.max_iterations = cfg_.max_iterations;
        p.convergence_threshold = cfg_.convergence_threshold;
        p.shape_order = (cfg_.second_order_shape || cfg_.use_second_order) ? 2 : 1;
        p.objective = cfg_.ssd ? DIC_OBJECTIVE_SSD : DIC_OBJECTIVE_ZNSSD;
        p.optimizer = cfg_.forward_gauss_newton ? DIC_OPTIMIZER_FGN : DIC_OPTIMIZER_ICGN;
        p.precision = (int)precision_;
        p.masked = roi ? 1 : 0;
        p.min_valid_samples = cfg_.min_valid_samples;
        p.min_valid_sample_ratio = cfg_.min_valid_sample_ratio;
        std::vector<dic_initial> ini(n);
        for (int i = 0; i < n; ++i)
            ini[i] = {initial[i].u, initial[i].v, initial[i].du_dx, initial[i].du_dy,
                      initial[i].dv_dx, initial[i].dv_dy, initial[i].confidence};
        std::vector<dic_result> raw(n);
        check(ctx_->get(), dic_icgn_solve_batch(ctx_->get(), &p, reference.get(), pair.gradient(), pair.coefficients(),
                                                roi ? roi->valid.data() : nullptr, poi_xy.data(), ini.data(), n,
                                                raw.data()));
        std::vector<Displacement2D> out(n);
        for (int i = 0; i < n; ++i) {
            const dic_result &r = raw[i];
            Displacement2D &d = out[i];
            d.x = r.x;
            d.y = r.y;
            d.u = r.p[0];
            d.v = r.p[1];
            d.du_dx = r.p[2];
            d.du_dy = r.p[3];
            d.dv_dx = r.p[4];
            d.dv_dy = r.p[5];
            d.d2u_dx2 = r.p[6];
            d.d2u_dxdy = r.p[7];
            d.d2u_dy2 = r.p[8];
            d.d2v_dx2 = r.p[9];
            d.d2v_dxdy = r.p[10];
            d.d2v_dy2 = r.p[11];
            d.correlation = r.correlation;
            d.status = (SolverStatus)r.status;
            d.valid = r.valid != 0;
            d.iterations = r.iterations;
        }
        return out;
    }

    Context *ctx_;
    SubsetConfig cfg_;
    Precision precision_;
    int gradient_mode_;
};

// ForwardGaussNewtonSolver: the same surface; the deformed image's B-spline
// gradient drives the steepest descent (the reference gradients go unused).
class ForwardGaussNewtonSolver : public ICGNSolver {
public:
    explicit ForwardGaussNewtonSolver(Context &ctx, SubsetConfig cfg = {}, Precision p = Precision::FP64)
        : ICGNSolver(ctx, with_fgn(cfg), p)
    {
    }

private:
    static SubsetConfig with_fgn(SubsetConfig c)
    {
        c.forward_gauss_newton = true;
        return c;
    }
};

inline dic_integer_search_params search_params(const SeedIntegerSearchConfig &c, int mode, bool masked)
{
    dic_integer_search_params p{};
    p.subset_radius = c.subset_radius < 1 ? 1 : c.subset_radius; // integer_search.cpp:293-301
    p.search_radius = c.search_radius < 0 ? 0 : c.search_radius;
    p.pyramid_enabled = c.pyramid_enabled ? 1 : 0;
    p.pyramid_scale = c.pyramid_scale < 1 ? 1 : c.pyramid_scale;
    p.pyramid_refinement_radius = c.pyramid_refinement_radius < 1 ? 1 : c.pyramid_refinement_radius;
    p.mode = mode;
    p.masked = masked ? 1 : 0;
    return p;
}

// IntegerSearchInitializer: estimate / estimate_with_mask /
// estimate[_with_mask]_around_displacement over many points at once.
class IntegerSearchInitializer {
public:
    explicit IntegerSearchInitializer(Context &ctx, SeedIntegerSearchConfig cfg = {}) : ctx_(ctx.get()), cfg_(cfg) {}

    std::vector<InitialEstimate> estimate(const Image &reference, const Image &deformed,
                                          const std::vector<double> &points_xy, const uint8_t *roi = nullptr) const
    {
        return run(reference, deformed, points_xy, nullptr, roi, DIC_SEARCH_GLOBAL);
    }
    std::vector<InitialEstimate> estimate_around_displacement(const Image &reference, const Image &deformed,
                                                              const std::vector<double> &points_xy,
                                                              const std::vector<double> &prior_uv,
                                                              const uint8_t *roi = nullptr) const
    {
        if (prior_uv.size() != points_xy.size()) throw std::invalid_argument("dic_hip: one prior (u, v) per point");
        return run(reference, deformed, points_xy, prior_uv.data(), roi, DIC_SEARCH_AROUND_PRIOR);
    }

private:
    std::vector<InitialEstimate> run(const Image &reference, const Image &deformed, const std::vector<double> &xy,
                                     const double *prior, const uint8_t *roi, int mode) const
    {
        const int n = (int)(xy.size() / 2);
        const dic_integer_search_params p = search_params(cfg_, mode, roi != nullptr);
        std::vector<dic_estimate> raw(n);
        check(ctx_, dic_integer_search(ctx_, &p, reference.get(), deformed.get(), roi, xy.data(), prior, n,
                                       raw.data()));
        std::vector<InitialEstimate> out(n);
        for (int i = 0; i < n; ++i)
            out[i] = {raw[i].u,          raw[i].v,     raw[i].du_dx,    raw[i].du_dy,
                      raw[i].dv_dx,      raw[i].dv_dy, raw[i].confidence, raw[i].valid != 0,
                      raw[i].zncc,       raw[i].znssd};
        return out;
    }
    dic_ctx *ctx_;
    SeedIntegerSearchConfig cfg_;
};

struct SeedEvaluation {
    double x{0.0}, y{0.0};
    Displacement2D displacement{};
    double quality{0.0};
    double displacement_norm{0.0};
    bool quality_passed{false};
    bool valid{false};
};

struct SeedSelectionResult {
    std::vector<SeedEvaluation> candidates;
    SeedEvaluation best_seed{};
    bool found{false};
};

// SeedSelector (integer-search path): generate_candidates / select_best_seed.
class SeedSelector {
public:
    explicit SeedSelector(Context &ctx, SubsetConfig cfg = {}) : ctx_(ctx.get()), cfg_(cfg) {}

    dic_seed_params params() const
    {
        dic_seed_params p{};
        p.search = search_params(cfg_.seed_integer_search, DIC_SEARCH_GLOBAL, false);
        const auto &s = cfg_.seed_subpixel;
        p.subpixel_enabled = s.enabled ? 1 : 0;
        p.subpixel_shape_order = s.second_order ? 2 : 1;
        p.subpixel_optimizer = s.forward_gauss_newton ? DIC_OPTIMIZER_FGN : DIC_OPTIMIZER_ICGN;
        p.subpixel_objective = s.ssd ? DIC_OBJECTIVE_SSD : DIC_OBJECTIVE_ZNSSD;
        p.subpixel_radius = s.subset_radius < 1 ? 1 : s.subset_radius;
        p.subpixel_max_iterations = s.max_iterations < 0 ? 0 : s.max_iterations;
        p.subpixel_convergence_threshold = s.convergence_threshold;
        const auto &q = cfg_.seed_selection;
        p.seed_count = q.seed_count < 1 ? 1 : q.seed_count; // seed_selector.cpp:138-144
        p.quality_metric = (int)q.quality_metric;
        p.min_zncc = q.min_zncc;
        p.max_znssd = q.max_znssd;
        p.max_ssd = q.max_ssd;
        p.min_displacement_norm = q.min_displacement_norm;
        p.min_texture_std = q.min_texture_std;
        p.kmeans_iterations = q.kmeans_iterations < 0 ? 0 : q.kmeans_iterations;
        p.kmeans_sample_limit = q.kmeans_sample_limit < 1 ? 1 : q.kmeans_sample_limit;
        p.truncate_roi_subsets = cfg_.truncate_roi_subsets ? 1 : 0;
        p.min_valid_samples = cfg_.min_valid_samples;
        p.min_valid_sample_ratio = cfg_.min_valid_sample_ratio;
        p.degree = cfg_.image_precompute.degree;
        p.border = cfg_.image_precompute.border;
        return p;
    }

    std::vector<double> generate_candidates(const Image &reference, const uint8_t *roi) const
    {
        const dic_seed_params p = params();
        std::vector<double> xy(2 * (size_t)p.seed_count);
        int n = 0;
        check(ctx_, dic_seed_candidates(ctx_, &p, reference.get(), roi, xy.data(), p.seed_count, &n));
        xy.resize(2 * (size_t)n);
        return xy;
    }

    SeedSelectionResult select_best_seed(const Image &reference, const Image &deformed, const uint8_t *roi) const
    {
        const dic_seed_params p = params();
        std::vector<dic_seed_evaluation> raw((size_t)p.seed_count);
        int n = 0, best = -1;
        check(ctx_, dic_select_seed(ctx_, &p, reference.get(), deformed.get(), roi, raw.data(), p.seed_count, &n,
                                    &best));
        SeedSelectionResult out;
        out.candidates.resize(n);
        for (int i = 0; i < n; ++i) {
            const dic_seed_evaluation &e = raw[i];
            SeedEvaluation &s = out.candidates[i];
            s.x = e.x;
            s.y = e.y;
            const dic_result &r = e.displacement;
            Displacement2D &d = s.displacement;
            d.x = r.x;
            d.y = r.y;
            d.u = r.p[0];
            d.v = r.p[1];
            d.du_dx = r.p[2];
            d.du_dy = r.p[3];
            d.dv_dx = r.p[4];
            d.dv_dy = r.p[5];
            d.correlation = r.correlation;
            d.status = (SolverStatus)r.status;
            d.valid = r.valid != 0;
            s.quality = e.quality;
            s.displacement_norm = e.displacement_norm;
            s.quality_passed = e.quality_passed != 0;
            s.valid = e.valid != 0;
        }
        if (best >= 0) {
            out.best_seed = out.candidates[best];
            out.found = true;
        }
        return out;
    }

private:
    dic_ctx *ctx_;
    SubsetConfig cfg_;
};

} // namespace dic_hip

#endif
