// dic_hip.hpp -- C++17 host layer over the C-ABI (dic_hip.h), mirroring the
// surface Traditional-DIC's C++ core calls on this path:
//   dic::SubsetSolver / ICGNSolver      include/dic/subset/solver/subset_solver.hpp:37-67,
//                                       include/dic/subset/solver/icgn.hpp:35-124
//   ForwardGaussNewtonSolver            include/dic/subset/solver/forward_gauss_newton.hpp:11-107
//   BSplineImagePreprocessor::compute   include/dic/interpolation/bspline.hpp:84-100
//   BSplineInterpolator::value/gradient include/dic/interpolation/bspline.hpp:102-130
//   Displacement2D, InitialDisplacement, SolverStatus, SubsetConfig
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

struct SubsetConfig {
    int subset_radius{15};
    int max_iterations{30};
    double convergence_threshold{1e-3};
    bool use_second_order{false};
    bool second_order_shape{false}; // shape_function == SecondOrder
    bool ssd{false};                // objective == SSD (else ZNSSD)
    bool forward_gauss_newton{false}; // optimization == FGN (else IC-GN)
    double min_valid_sample_ratio{0.5};
    int min_valid_samples{12};
    BSplinePrecomputeConfig image_precompute{};
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

// ICGNSolver: solve / solve_with_interpolators / solve_with_mask, plus
// solve_batch over many POIs (the form the GPU wants).
class ICGNSolver {
public:
    explicit ICGNSolver(Context &ctx, SubsetConfig cfg = {}, Precision p = Precision::FP64)
        : ctx_(ctx.get()), cfg_(cfg), precision_(p)
    {
    }

    // Prepares the pair: pipeline reference gradients (central difference,
    // reliability_propagation.cpp:84-112) and the deformed coefficient image.
    void prepare(const Image &reference, const Image &deformed)
    {
        release();
        const auto &ip = cfg_.image_precompute;
        check(ctx_, dic_ref_gradient(ctx_, reference.get(), DIC_GRADIENT_CENTRAL, ip.degree, ip.border,
                                     (int)precision_, &grad_));
        check(ctx_, dic_bspline_prefilter(ctx_, deformed.get(), ip.degree, ip.border,
                                          ip.use_exact_prefilter ? 1 : 0, (int)precision_, &coeffs_));
        ref_ = &reference;
    }

    std::vector<Displacement2D> solve_batch(const std::vector<double> &poi_xy,
                                            const std::vector<InitialDisplacement> &initial,
                                            const uint8_t *roi = nullptr) const
    {
        if (!ref_) throw std::runtime_error("dic_hip: prepare() the image pair first");
        const int n = (int)initial.size();
        if ((int)poi_xy.size() != 2 * n) throw std::invalid_argument("dic_hip: one (x, y) per initial guess");
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
        check(ctx_, dic_icgn_solve_batch(ctx_, &p, ref_->get(), grad_, coeffs_, roi, poi_xy.data(), ini.data(), n,
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

    Displacement2D solve_with_interpolators(double x, double y, const InitialDisplacement &initial) const
    {
        return solve_batch({x, y}, {initial})[0];
    }

    Displacement2D solve_with_mask(const uint8_t *roi, double x, double y, const InitialDisplacement &initial) const
    {
        return solve_batch({x, y}, {initial}, roi)[0];
    }

    ~ICGNSolver() { release(); }
    ICGNSolver(const ICGNSolver &) = delete;
    ICGNSolver &operator=(const ICGNSolver &) = delete;

private:
    void release()
    {
        dic_grad_free(grad_);
        dic_coeffs_free(coeffs_);
        grad_ = nullptr;
        coeffs_ = nullptr;
        ref_ = nullptr;
    }
    dic_ctx *ctx_;
    SubsetConfig cfg_;
    Precision precision_;
    const Image *ref_ = nullptr;
    dic_grad *grad_ = nullptr;
    dic_coeffs *coeffs_ = nullptr;
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

} // namespace dic_hip

#endif
