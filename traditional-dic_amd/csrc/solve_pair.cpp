// solve_pair.cpp -- C++ host example over dic_hip.hpp: builds an analytic
// texture pair with a known sub-pixel shift, solves a POI grid with the
// HIP IC-GN and reports the worst error.  Exit codes: 0 ok, 1 accuracy
// failure, 3 no usable HIP device.
#include <cmath>
#include <cstdio>
#include <vector>

#include "dic_hip.hpp"

static float texture(double x, double y)
{
    return (float)(0.5 + 0.2 * std::sin(0.31 * x + 0.7 * std::cos(0.05 * y)) * std::cos(0.23 * y) +
                   0.1 * std::sin(0.17 * x * 0.9 + 0.29 * y));
}

int main()
{
    const int W = 256, H = 256;
    const double su = 0.4, sv = -0.3;
    std::vector<float> ref((size_t)W * H), def((size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            ref[(size_t)y * W + x] = texture(x, y);
            def[(size_t)y * W + x] = texture(x - su, y - sv); // def(x) = ref(x - s)
        }
    try {
        dic_hip::Context ctx(0);
        dic_hip::Image r(ctx, W, H, ref.data()), d(ctx, W, H, def.data());
        dic_hip::SubsetConfig cfg;
        cfg.subset_radius = 15;
        cfg.image_precompute.degree = 5;
        dic_hip::ICGNSolver solver(ctx, cfg, dic_hip::Precision::FP64);
        solver.prepare(r, d);
        std::vector<double> xy;
        std::vector<dic_hip::InitialDisplacement> ini;
        for (int y = 40; y <= 216; y += 16)
            for (int x = 40; x <= 216; x += 16) {
                xy.push_back(x);
                xy.push_back(y);
                ini.push_back({});
            }
        auto out = solver.solve_batch(xy, ini);
        double worst = 0.0;
        int bad = 0;
        for (const auto &o : out) {
            if (o.status != dic_hip::SolverStatus::Success) ++bad;
            worst = std::fmax(worst, std::fmax(std::fabs(o.u - su), std::fabs(o.v - sv)));
        }
        dic_hip::BSplineImage bi(ctx, d, cfg.image_precompute);
        dic_hip::BSplineInterpolator interp(ctx, bi);
        const double identity = std::fabs(interp.value(100.0, 60.0) - def[(size_t)60 * W + 100]);
        std::printf("solve_pair: %zu POIs, %d not Success, max |error| = %.3e px, identity %.1e\n", out.size(), bad,
                    worst, identity);
        return (bad == 0 && worst < 1e-2 && identity < 1e-6) ? 0 : 1;
    } catch (const std::exception &e) {
        std::printf("solve_pair: %s\n", e.what());
        return 3;
    }
}
