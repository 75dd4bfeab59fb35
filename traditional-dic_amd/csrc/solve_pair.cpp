// solve_pair.cpp -- C++ host example over dic_hip.hpp: builds one reference
// texture and two deformed frames with known sub-pixel shifts, solves a POI
// grid of each pair with the HIP IC-GN (per-call preparation, then one
// PreparedPair per pair reused over a second call), checks that a pair handed
// over with the other frame's images is refused, and reports the worst error.
// Exit codes: 0 ok, 1 accuracy failure, 2 the pair guard did not fire,
// 3 no usable HIP device.
#include <cmath>
#include <cstdio>
#include <vector>

#include "dic_hip.hpp"

static float texture(double x, double y)
{
    return (float)(0.5 + 0.2 * std::sin(0.31 * x + 0.7 * std::cos(0.05 * y)) * std::cos(0.23 * y) +
                   0.1 * std::sin(0.17 * x * 0.9 + 0.29 * y));
}

static std::vector<float> render(int W, int H, double su, double sv)
{
    std::vector<float> img((size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) img[(size_t)y * W + x] = texture(x - su, y - sv); // img(x) = ref(x - s)
    return img;
}

static double worst_error(const std::vector<dic_hip::Displacement2D> &out, double su, double sv, int &bad)
{
    double worst = 0.0;
    for (const auto &o : out) {
        if (o.status != dic_hip::SolverStatus::Success) ++bad;
        worst = std::fmax(worst, std::fmax(std::fabs(o.u - su), std::fabs(o.v - sv)));
    }
    return worst;
}

int main()
{
    const int W = 256, H = 256;
    const double shift[2][2] = {{0.4, -0.3}, {-1.35, 0.8}};
    const std::vector<float> ref = render(W, H, 0.0, 0.0);
    const std::vector<float> def0 = render(W, H, shift[0][0], shift[0][1]);
    const std::vector<float> def1 = render(W, H, shift[1][0], shift[1][1]);
    try {
        dic_hip::Context ctx(0);
        dic_hip::Image r(ctx, W, H, ref.data()), d0(ctx, W, H, def0.data()), d1(ctx, W, H, def1.data());
        const dic_hip::Image *frames[2] = {&d0, &d1};
        dic_hip::SubsetConfig cfg;
        cfg.subset_radius = 15;
        cfg.image_precompute.degree = 5;
        dic_hip::ICGNSolver solver(ctx, cfg, dic_hip::Precision::FP64);
        std::vector<double> xy;
        std::vector<dic_hip::InitialDisplacement> ini;
        for (int y = 40; y <= 216; y += 16)
            for (int x = 40; x <= 216; x += 16) {
                xy.push_back(x);
                xy.push_back(y);
                dic_hip::InitialDisplacement g{};
                ini.push_back(g);
            }
        double worst = 0.0;
        int bad = 0, same = 1;
        for (int k = 0; k < 2; ++k) { // per-call preparation, frame after frame
            for (auto &g : ini) {
                g.u = std::round(shift[k][0]);
                g.v = std::round(shift[k][1]);
            }
            const auto once = solver.solve_batch(r, *frames[k], xy, ini);
            worst = std::fmax(worst, worst_error(once, shift[k][0], shift[k][1], bad));
            const dic_hip::PreparedPair pair = solver.prepare(r, *frames[k]);
            const auto reused = solver.solve_batch(r, *frames[k], xy, ini, &pair);
            for (size_t i = 0; i < once.size(); ++i) same = same && reused[i].u == once[i].u && reused[i].v == once[i].v;
            const auto one = solver.solve_with_interpolators(r, *frames[k], xy[0], xy[1], ini[0], pair);
            same = same && one.u == once[0].u;
        }
        int guarded = 0;
        {
            const dic_hip::PreparedPair pair0 = solver.prepare(r, d0);
            try {
                solver.solve_with_interpolators(r, d1, xy[0], xy[1], ini[0], pair0);
            } catch (const std::invalid_argument &) {
                guarded = 1;
            }
        }
        dic_hip::BSplineImage bi(ctx, d0, cfg.image_precompute);
        dic_hip::BSplineInterpolator interp(ctx, bi);
        const double identity = std::fabs(interp.value(100.0, 60.0) - def0[(size_t)60 * W + 100]);
        std::printf("solve_pair: 2 pairs x %zu POIs, %d not Success, max |error| = %.3e px, reuse identical %d, "
                    "guard %d, identity %.1e\n", ini.size(), bad, worst, same, guarded, identity);
        if (!guarded) return 2;
        return (bad == 0 && same && worst < 1e-2 && identity < 1e-6) ? 0 : 1;
    } catch (const std::exception &e) {
        std::printf("solve_pair: %s\n", e.what());
        return 3;
    }
}
