// propagate.hip -- reliability-guided propagation over the POI grid
// (dic_propagate): the reference's seed alignment, neighbour rules, guess
// extrapolation and acceptance test.  Two walk orders: the reference's heap
// order (speculative launches, replayed on the host) and frontier waves
// (each frontier solved in one launch).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "dic_internal.hpp"

using namespace dic;

namespace {

struct Item { // reliability_propagation.cpp:44-53
    double corr;
    int gx, gy;
    double u, v, ux, uy, vx, vy;
};

struct Cand {
    int idx, nx, ny;
    double iu, iv; // extrapolated guess
    const Item *parent;
};

bool finite6(const dic_result &r)
{
    for (int j = 0; j < 6; ++j)
        if (!std::isfinite(r.p[j])) return false;
    return std::isfinite(r.correlation);
}

// The walk's launches are small, so copy latency counts: a launch packs its
// POIs and guesses into one block of coherent pinned host memory, which the
// kernels read in place (one POI record each), and the kernels write their
// results into the same block; no copy is queued around the solve.  The ROI
// stays in device memory.
struct DeviceScratch {
    dic_ctx *ctx = nullptr;
    uint8_t *roi = nullptr;
    size_t roi_bytes = 0;
    unsigned char *pin = nullptr;     // host view: n POI positions (16 B each), n guesses, then the results
    unsigned char *pin_dev = nullptr; // the same block as the device sees it
    size_t pin_cap = 0;               // POIs the block holds
    static constexpr size_t kInBytes = 16 + sizeof(dic_initial);
    ~DeviceScratch()
    {
        ctx_free(ctx, roi, roi_bytes);
        if (pin) (void)hipHostFree(pin);
    }
    hipError_t pinned(size_t n)
    {
        if (n <= pin_cap) return hipSuccess;
        const size_t want = std::max(n, 2 * pin_cap + 256);
        if (pin) (void)hipHostFree(pin);
        pin = pin_dev = nullptr;
        pin_cap = 0;
        hipError_t e = hipHostMalloc((void **)&pin, (kInBytes + sizeof(dic_result)) * want,
                                     hipHostMallocCoherent | hipHostMallocMapped);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&pin_dev, pin, 0);
        if (e == hipSuccess) pin_cap = want;
        return e;
    }
};

// Launches of at most this many IC-GN POIs give each POI a workgroup of
// several waves: the walk's launches are latency-bound (a few POIs each).
constexpr int kCoopMaxPois = 256;

// Whether the cooperative kernel's LDS (its larger static block beside the
// per-sample buffer) fits the device for this subset; where it does not
// (FP64 ZNSSD at the largest admitted radii), launches keep one wave per POI.
bool coop_fits(dic_ctx *ctx, const dic_icgn_params *params)
{
    const int *table = nullptr;
    int count = 0;
    if (get_subset_table(ctx, params->subset_radius, &table, &count) != DIC_OK) return false;
    return ctx->max_lds == 0 ||
           coop_static_lds() + solve_dynamic_lds(false, params->precision, params->objective, count) <= ctx->max_lds;
}

// one launch over n POIs (host arrays in, host results out)
int solve_host(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
               const dic_coeffs *def, DeviceScratch &d, const std::vector<double> &xy,
               const std::vector<dic_initial> &ini, std::vector<dic_result> &res)
{
    const int n = (int)ini.size();
    res.resize(n);
    if (n == 0) return DIC_OK;
    hipStream_t s = ctx->stream;
    int rc = check_hip(ctx, d.pinned((size_t)n), "pinned staging");
    if (rc != DIC_OK) return rc;
    const size_t xy_bytes = 16 * (size_t)n;
    std::memcpy(d.pin, xy.data(), xy_bytes);
    std::memcpy(d.pin + xy_bytes, ini.data(), sizeof(dic_initial) * (size_t)n);
    const double *dxy = reinterpret_cast<const double *>(d.pin_dev);
    const dic_initial *dini = reinterpret_cast<const dic_initial *>(d.pin_dev + xy_bytes);
    const size_t res_off = DeviceScratch::kInBytes * d.pin_cap;
    dic_result *pres = reinterpret_cast<dic_result *>(d.pin + res_off);
    dic_result *dres = reinterpret_cast<dic_result *>(d.pin_dev + res_off);
    const int shape = params->optimizer == DIC_OPTIMIZER_ICGN && n <= kCoopMaxPois && coop_fits(ctx, params)
                          ? kShapeCoop
                          : host_tile_hint(params, def, xy.data(), ini.data(), n);
    rc = run_solve(ctx, params, ref, grad, def, params->masked ? d.roi : nullptr, dxy, dini, n, dres, s, shape);
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(s), "propagation sync");
    if (rc == DIC_OK) std::memcpy(res.data(), pres, sizeof(dic_result) * (size_t)n);
    return rc;
}

// Heap order (reliability_propagation.cpp:277-389): the reference pops one
// point at a time from a std::priority_queue ordered by correlation and
// solves the popped point's unvisited neighbours with guesses extrapolated
// from it; whether a neighbour is accepted, and so what is pushed next,
// depends on that solve.  Here the host replays exactly that walk, with the
// same container operations (std::push_heap / std::pop_heap under the
// reference's comparison), and takes every solve from a cache keyed by
// (neighbour, parent): a solve's result depends on nothing but the
// neighbour's position and the parent's parameters.  When the point on top
// of the heap needs a solve the cache lacks, one launch solves, for the
// heap's points in their pop order, every neighbour that is not yet visited
// nor taken by an earlier point of that order; points accepted after the
// launch can claim neighbours first, so some of these solves go unused, and
// the next launch covers what the replay then needs.
struct HeapWalk {
    dic_ctx *ctx;
    const dic_icgn_params *params;
    const dic_image *ref;
    const dic_grad *grad;
    const dic_coeffs *def;
    DeviceScratch &d;
    const uint8_t *roi;
    int W, H, step, gw, gh, radius;
    double max_znssd;
    std::vector<uint8_t> &calculated;
    dic_result *grid_out;
    int &computed;
    std::unordered_map<long long, dic_result> cache;
    std::vector<int> taken; // launch number that took a neighbour in speculate()
    int launches = 0;
    std::vector<double> xy;
    std::vector<dic_initial> ini;
    std::vector<dic_result> res;

    static bool later(const Item &a, const Item &b) { return a.corr > b.corr; } // QueueItem::operator<
    long long key(int nidx, const Item &p) const { return (long long)nidx * gw * gh + (p.gy * gw + p.gx); }

    // neighbour q of p (up, right, down, left) as grid index, or -1 when the
    // bounds test of :296-303 skips it; roi_ok / margin_ok: the tests of :300-311
    int neighbour(const Item &p, int q, int &nx, int &ny, bool &usable) const
    {
        const int px = p.gx * step, py = p.gy * step;
        const int nb[4][2] = {{px, py - step}, {px + step, py}, {px, py + step}, {px - step, py}};
        nx = nb[q][0];
        ny = nb[q][1];
        if (!(nx >= 0 && ny >= 0 && nx < W && ny < H)) return -1;
        const int gnx = nx / step, gny = ny / step;
        if (gnx < 0 || gny < 0 || gnx >= gw || gny >= gh) return -1;
        usable = roi[(size_t)ny * W + nx] &&
                 !(nx - radius < 0 || ny - radius < 0 || nx + radius >= W || ny + radius >= H);
        return gny * gw + gnx;
    }
    dic_initial guess(const Item &p, int nx, int ny) const
    {
        const double ddx = (double)(nx - p.gx * step), ddy = (double)(ny - p.gy * step);
        return {p.u + p.ux * ddx + p.uy * ddy, p.v + p.vx * ddx + p.vy * ddy, p.ux, p.uy, p.vx, p.vy, 0.0};
    }
    bool needs_launch(const Item &p) const
    {
        for (int q = 0; q < 4; ++q) {
            int nx, ny;
            bool usable = false;
            const int nidx = neighbour(p, q, nx, ny, usable);
            if (nidx >= 0 && !calculated[nidx] && usable && !cache.count(key(nidx, p))) return true;
        }
        return false;
    }
    // one launch: the neighbours of the heap's points in their pop order
    int speculate(const std::vector<Item> &heap)
    {
        constexpr size_t kMaxLaunch = 8192;
        std::vector<Item> h(heap);
        if (taken.empty()) taken.assign((size_t)gw * gh, 0);
        ++launches;
        xy.clear();
        ini.clear();
        std::vector<long long> keys;
        while (!h.empty() && ini.size() + 4 <= kMaxLaunch) {
            std::pop_heap(h.begin(), h.end(), later);
            const Item p = h.back();
            h.pop_back();
            for (int q = 0; q < 4; ++q) {
                int nx, ny;
                bool usable = false;
                const int nidx = neighbour(p, q, nx, ny, usable);
                if (nidx < 0 || calculated[nidx] || !usable) continue;
                if (taken[nidx] == launches) continue;
                taken[nidx] = launches;
                const long long k = key(nidx, p);
                if (cache.count(k)) continue;
                keys.push_back(k);
                xy.push_back((double)nx);
                xy.push_back((double)ny);
                ini.push_back(guess(p, nx, ny));
            }
        }
        const int rc = solve_host(ctx, params, ref, grad, def, d, xy, ini, res);
        if (rc != DIC_OK) return rc;
        for (size_t c = 0; c < keys.size(); ++c) cache.emplace(keys[c], res[c]);
        return DIC_OK;
    }
    int walk(const Item &seed)
    {
        std::vector<Item> heap{seed};
        while (!heap.empty()) {
            if (needs_launch(heap.front())) {
                const int rc = speculate(heap);
                if (rc != DIC_OK) return rc;
            }
            std::pop_heap(heap.begin(), heap.end(), later);
            const Item p = heap.back();
            heap.pop_back();
            for (int q = 0; q < 4; ++q) {
                int nx, ny;
                bool usable = false;
                const int nidx = neighbour(p, q, nx, ny, usable);
                if (nidx < 0 || calculated[nidx]) continue;
                calculated[nidx] = 1; // marked even when outside the ROI or the margin (:300-311, :357)
                if (!usable) continue;
                const dic_initial g = guess(p, nx, ny);
                const dic_result &r = cache.at(key(nidx, p));
                const bool accepted = r.valid && r.status == DIC_SUCCESS && r.correlation < max_znssd &&
                                      std::fabs(g.u - r.p[0]) < (double)step &&
                                      std::fabs(g.v - r.p[1]) < (double)step && finite6(r);
                if (!accepted) continue;
                grid_out[nidx] = r;
                grid_out[nidx].x = (double)nx;
                grid_out[nidx].y = (double)ny;
                ++computed;
                heap.push_back({r.correlation, nx / step, ny / step, r.p[0], r.p[1], r.p[2], r.p[3], r.p[4], r.p[5]});
                std::push_heap(heap.begin(), heap.end(), later);
            }
        }
        return DIC_OK;
    }
};

} // namespace

extern "C" int dic_propagate(dic_ctx *ctx, const dic_icgn_params *params, const dic_propagation_params *prop,
                             const dic_image *ref, const dic_grad *grad, const dic_coeffs *def,
                             const uint8_t *roi, const dic_seed *seeds, int n_seeds, dic_result *grid_out,
                             int *grid_w, int *grid_h, int *points_computed)
{
    if (!ctx || !prop || !grid_out || n_seeds < 0 || (n_seeds > 0 && !seeds)) return DIC_ERR_INVALID_ARGUMENT;
    int rc = validate_solve(ctx, params, ref, grad, def, 0);
    if (rc != DIC_OK) return rc;
    const int W = ref->width, H = ref->height;
    const int step = std::max(prop->spacing, 0) + 1;
    const int gw = (W + step - 1) / step, gh = (H + step - 1) / step, total = gw * gh;
    if (grid_w) *grid_w = gw;
    if (grid_h) *grid_h = gh;
    if (points_computed) *points_computed = 0;
    for (int i = 0; i < total; ++i) { // defaults (:167-171)
        dic_result &r = grid_out[i];
        r = dic_result{};
        r.x = (double)((i % gw) * step);
        r.y = (double)((i / gw) * step);
        r.status = DIC_INVALID_INPUT;
    }
    // input validation (:137-149): empty ROI or no seeds gives the empty result
    if (!roi || W != def->width || H != def->height || params->subset_radius < 1 || n_seeds == 0) return DIC_OK;
    const int radius = params->subset_radius;
    const double max_znssd = prop->max_znssd;

    DeviceScratch d;
    d.ctx = ctx;
    if (params->masked) {
        d.roi_bytes = (size_t)W * H;
        rc = check_hip(ctx, ctx_malloc(ctx, (void **)&d.roi, d.roi_bytes), "roi buffer");
        if (rc == DIC_OK)
            rc = check_hip(ctx, hipMemcpy(d.roi, roi, d.roi_bytes, hipMemcpyHostToDevice), "roi upload");
    }
    if (rc != DIC_OK) return rc;

    auto valid_px = [&](int x, int y) { return x >= 0 && y >= 0 && x < W && y < H && roi[(size_t)y * W + x]; };
    std::vector<uint8_t> calculated((size_t)total, 0);
    std::vector<int> claimed((size_t)total, -1);
    std::vector<double> xy;
    std::vector<dic_initial> ini;
    std::vector<dic_result> res;
    int computed = 0;

    for (int k = 0; k < n_seeds; ++k) {
        const dic_seed &sd = seeds[k];
        const int sxr = (int)std::lround(sd.x), syr = (int)std::lround(sd.y);
        if (!valid_px(sxr, syr)) continue;
        const int gxs = std::clamp((int)std::llround((double)sxr / step), 0, gw - 1);
        const int gys = std::clamp((int)std::llround((double)syr / step), 0, gh - 1);
        const int sx = gxs * step, sy = gys * step;
        if (!valid_px(sx, sy)) continue;
        const int sidx = gys * gw + gxs;
        if (calculated[sidx]) continue;
        dic_result al{};
        al.p[0] = sd.displacement.u;
        al.p[1] = sd.displacement.v;
        al.p[2] = sd.displacement.du_dx;
        al.p[3] = sd.displacement.du_dy;
        al.p[4] = sd.displacement.dv_dx;
        al.p[5] = sd.displacement.dv_dy;
        al.correlation = sd.displacement.confidence;
        if (sx != sxr || sy != syr) { // re-solve the seed on the grid point (:233-261)
            xy.assign({(double)sx, (double)sy});
            ini.assign(1, sd.displacement);
            rc = solve_host(ctx, params, ref, grad, def, d, xy, ini, res);
            if (rc != DIC_OK) return rc;
            al = res[0];
            if (!al.valid || al.status != DIC_SUCCESS || !finite6(al)) continue;
        }
        calculated[sidx] = 1;
        al.x = (double)sx;
        al.y = (double)sy;
        al.valid = 1;
        al.status = DIC_SUCCESS;
        grid_out[sidx] = al;
        ++computed;
        const Item seed_item{std::max(0.0, al.correlation), gxs, gys, al.p[0], al.p[1], al.p[2], al.p[3], al.p[4],
                             al.p[5]};
        if (prop->order == DIC_PROPAGATE_HEAP) {
            HeapWalk hw{ctx, params, ref, grad, def, d, roi, W, H, step, gw, gh, radius, max_znssd, calculated,
                        grid_out, computed, {}, {}, 0, {}, {}, {}};
            rc = hw.walk(seed_item);
            if (rc != DIC_OK) return rc;
            continue;
        }
        std::vector<Item> frontier{seed_item};
        while (!frontier.empty()) {
            // best (lowest correlation) parents claim their neighbours first
            std::stable_sort(frontier.begin(), frontier.end(),
                             [](const Item &a, const Item &b) { return a.corr < b.corr; });
            std::vector<Cand> cands;
            for (const Item &it : frontier) {
                const int px = it.gx * step, py = it.gy * step;
                const int nb[4][2] = {{px, py - step}, {px + step, py}, {px, py + step}, {px - step, py}};
                for (int q = 0; q < 4; ++q) {
                    const int nx = nb[q][0], ny = nb[q][1];
                    if (!(nx >= 0 && ny >= 0 && nx < W && ny < H)) continue;
                    const int gnx = nx / step, gny = ny / step;
                    if (gnx < 0 || gny < 0 || gnx >= gw || gny >= gh) continue;
                    const int nidx = gny * gw + gnx;
                    if (calculated[nidx] || claimed[nidx] >= 0) continue;
                    if (!roi[(size_t)ny * W + nx]) { // :300-304
                        calculated[nidx] = 1;
                        continue;
                    }
                    if (nx - radius < 0 || ny - radius < 0 || nx + radius >= W || ny + radius >= H) { // :306-311
                        calculated[nidx] = 1;
                        continue;
                    }
                    const double ddx = (double)(nx - px), ddy = (double)(ny - py);
                    claimed[nidx] = (int)cands.size();
                    cands.push_back({nidx, nx, ny, it.u + it.ux * ddx + it.uy * ddy, it.v + it.vx * ddx + it.vy * ddy,
                                     &it});
                }
            }
            if (cands.empty()) break;
            xy.resize(2 * cands.size());
            ini.resize(cands.size());
            for (size_t c = 0; c < cands.size(); ++c) {
                const Cand &cd = cands[c];
                xy[2 * c] = (double)cd.nx;
                xy[2 * c + 1] = (double)cd.ny;
                ini[c] = {cd.iu, cd.iv, cd.parent->ux, cd.parent->uy, cd.parent->vx, cd.parent->vy, 0.0};
            }
            rc = solve_host(ctx, params, ref, grad, def, d, xy, ini, res);
            if (rc != DIC_OK) return rc;
            std::vector<Item> next;
            for (size_t c = 0; c < cands.size(); ++c) {
                const Cand &cd = cands[c];
                const dic_result &r = res[c];
                calculated[cd.idx] = 1;
                claimed[cd.idx] = -1;
                const bool accepted = r.valid && r.status == DIC_SUCCESS && r.correlation < max_znssd &&
                                      std::fabs(cd.iu - r.p[0]) < (double)step &&
                                      std::fabs(cd.iv - r.p[1]) < (double)step && finite6(r);
                if (!accepted) continue;
                grid_out[cd.idx] = r;
                grid_out[cd.idx].x = (double)cd.nx;
                grid_out[cd.idx].y = (double)cd.ny;
                ++computed;
                next.push_back({r.correlation, cd.nx / step, cd.ny / step, r.p[0], r.p[1], r.p[2], r.p[3], r.p[4],
                                r.p[5]});
            }
            frontier.swap(next);
        }
    }
    if (points_computed) *points_computed = computed;
    return DIC_OK;
}
