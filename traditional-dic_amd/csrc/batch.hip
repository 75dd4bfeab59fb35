// batch.hip -- host side of batched solving (dic_batch_*): descriptor and
// dispatch-order tables in HBM, the hipGraph that replays the launch, and the
// autotuner that times launch mappings on the caller's own batch.
#include <algorithm>
#include <map>
#include <cmath>
#include <new>
#include <numeric>
#include <tuple>
#include <vector>

#include "dic_internal.hpp"

using namespace dic;

struct dic_batch {
    dic_ctx *ctx = nullptr;
    int objective = 0, precision = 0, degree = 0;
    int cus = 0;
    std::vector<BatchPair> pairs;  // host copy of the descriptors
    std::vector<double> cost;      // estimated cost per POI of each descriptor
    BatchPair *d_pairs = nullptr;
    int2 *d_items = nullptr;
    unsigned *d_counter = nullptr;
    int64_t total = 0;
    int max_nfull = 0;
    dic_batch_tuning tuning{DIC_BATCH_GRID, 8, 1, DIC_BATCH_ORDER_COST};
    int items_key = -1;            // order (and tile layout) the device item table holds
    std::vector<BatchGroup> groups; // launch groups of the current item table
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t runs = 0;
    // recorded after every graph launch on the stream it ran on (possibly a
    // caller's): the item table and the graph are not touched before it
    hipEvent_t last_run = nullptr;
    bool run_recorded = false;
};

namespace {

constexpr int kInterleaveRun = 256; // POIs per turn of DIC_BATCH_ORDER_INTERLEAVE
constexpr int kXcds = 8;            // accelerator complex dies of an MI355X

// Waits for the last replay, whatever stream it was launched on, so that its
// kernels no longer read the item table or belong to the graph.
void wait_last_run(dic_batch *b)
{
    if (b->run_recorded) (void)hipEventSynchronize(b->last_run);
    b->run_recorded = false;
}

void drop_graph(dic_batch *b)
{
    wait_last_run(b);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int valid_tuning(dic_ctx *ctx, const dic_batch_tuning &t)
{
    if (t.mapping != DIC_BATCH_GRID && t.mapping != DIC_BATCH_PERSISTENT && t.mapping != DIC_BATCH_TILE)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "batch mapping must be GRID, PERSISTENT or TILE");
    if (t.order < DIC_BATCH_ORDER_NATURAL || t.order > DIC_BATCH_ORDER_XCD)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown batch order");
    if (t.mapping == DIC_BATCH_PERSISTENT && (t.waves_per_cu < 1 || t.waves_per_cu > 64 || t.chunk < 1 ||
                                              t.chunk > 4096))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "persistent mapping needs 1 <= waves_per_cu <= 64 and "
                                                        "1 <= chunk <= 4096");
    return DIC_OK;
}

// Whether descriptor p runs the tile kernel under `mapping`: under TILE, the
// first-order FP32 IC-GN descriptors whose four neighbouring subsets can share
// one tile (r <= 24, the per-pair launch's rule); the others keep one POI per
// workgroup.
bool tile_item(const dic_batch *b, const BatchPair &p, int mapping)
{
    return mapping == DIC_BATCH_TILE && !p.a.fgn && p.shape_order == 1 && b->precision != DIC_PRECISION_FP64 &&
           grouped_radius(p.a.radius);
}

// Launch groups: descriptors of one optimizer and shape order (FGN, then
// second order first), tile descriptors apart; where the solve keeps
// per-sample values in LDS (FP64 ZNSSD IC-GN, FGN) also of one LDS size class
// (power of two of the subset's bytes), so a small subset is not limited by
// the LDS the largest one needs.
std::vector<std::vector<int>> make_groups(const dic_batch *b, int mapping)
{
    // (-fgn, -order, tile, -class) -> descriptors
    std::map<std::tuple<int, int, int, int>, std::vector<int>> m;
    for (int p = 0; p < (int)b->pairs.size(); ++p) {
        const IcgnArgs &a = b->pairs[p].a;
        const size_t bytes = solve_dynamic_lds(a.fgn != 0, b->precision, b->objective, a.nfull);
        int cls = 0;
        if (bytes > 0)
            while ((size_t(1) << cls) < bytes) ++cls;
        m[{-a.fgn, -b->pairs[p].shape_order, tile_item(b, b->pairs[p], mapping) ? 1 : 0, -cls}].push_back(p);
    }
    std::vector<std::vector<int>> g;
    for (auto &kv : m) g.push_back(kv.second);
    return g;
}

// (descriptor, POI) list, group after group, each group in the requested
// order; a tile group keeps each descriptor's POIs in their natural order
// (four consecutive items are neighbouring POIs; the kernel deals the quads
// over the XCDs) and ranks the descriptors by cost under ORDER_COST only
std::vector<int2> make_items(const dic_batch *b, int order, int mapping, std::vector<BatchGroup> &groups)
{
    const int np = (int)b->pairs.size();
    std::vector<int2> items;
    items.reserve((size_t)b->total);
    groups.clear();
    for (std::vector<int> ids : make_groups(b, mapping)) {
        const size_t start = items.size();
        BatchGroup G{b->pairs[ids[0]].shape_order, 0, (int)start, 0, b->pairs[ids[0]].a.fgn,
                     tile_item(b, b->pairs[ids[0]], mapping) ? 1 : 0};
        for (int p : ids) G.max_nfull = std::max(G.max_nfull, b->pairs[p].a.nfull);
        if (G.tile) {
            if (order == DIC_BATCH_ORDER_COST)
                std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) { return b->cost[x] > b->cost[y]; });
            for (int p : ids)
                for (int i = 0; i < b->pairs[p].a.n; ++i) items.push_back(make_int2(p, i));
            G.count = (int)(items.size() - start);
            if (G.count > 0) groups.push_back(G);
            continue;
        }
        if (order == DIC_BATCH_ORDER_COST)
            std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) { return b->cost[x] > b->cost[y]; });
        if (order == DIC_BATCH_ORDER_INTERLEAVE) {
            std::vector<int> next(np, 0);
            bool left = true;
            while (left) {
                left = false;
                for (int p : ids) {
                    const int n = b->pairs[p].a.n;
                    const int end = std::min(n, next[p] + kInterleaveRun);
                    for (int i = next[p]; i < end; ++i) items.push_back(make_int2(p, i));
                    next[p] = end;
                    left = left || end < n;
                }
            }
        } else {
            for (int p : ids)
                for (int i = 0; i < b->pairs[p].a.n; ++i) items.push_back(make_int2(p, i));
        }
        if (order == DIC_BATCH_ORDER_XCD) {
            // workgroup k runs on XCD k % 8: deal each descriptor's POIs so that
            // every XCD takes one contiguous eighth of them (neighbouring
            // subsets then share that XCD's L2), descriptors in natural order
            size_t at = start;
            for (int p : ids) {
                const int n = b->pairs[p].a.n, per = (n + kXcds - 1) / kXcds;
                std::vector<int2> sw;
                sw.reserve((size_t)n);
                for (int j = 0; j < per; ++j)
                    for (int x = 0; x < kXcds; ++x) {
                        const int poi = x * per + j;
                        if (poi < n) sw.push_back(make_int2(p, poi));
                    }
                std::copy(sw.begin(), sw.end(), items.begin() + at);
                at += sw.size();
            }
        }
        G.count = (int)(items.size() - start);
        if (G.count > 0) groups.push_back(G);
    }
    return items;
}

int prepare(dic_batch *b)
{
    dic_ctx *ctx = b->ctx;
    const int key = b->tuning.order * 4 + b->tuning.mapping;
    if (b->items_key != key) {
        const std::vector<int2> items = make_items(b, b->tuning.order, b->tuning.mapping, b->groups);
        if (b->groups.size() > (size_t)kMaxBatchGroups)
            return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "too many launch groups in one batch");
        if (!items.empty()) {
            wait_last_run(b);
            int rc = check_hip(ctx, hipMemcpyAsync(b->d_items, items.data(), sizeof(int2) * items.size(),
                                                   hipMemcpyHostToDevice, ctx->stream),
                               "batch item upload");
            if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "batch item upload");
            if (rc != DIC_OK) return rc;
        }
        b->items_key = key;
    }
    if (b->exec) return DIC_OK;
    BatchLaunch L;
    L.objective = b->objective;
    L.degree = b->degree;
    L.precision = b->precision;
    L.pairs = b->d_pairs;
    L.items = b->d_items;
    L.groups = b->groups.data();
    L.ngroups = (int)b->groups.size();
    L.persistent = b->tuning.mapping == DIC_BATCH_PERSISTENT;
    L.grid = std::max(1, b->tuning.waves_per_cu * b->cus);
    L.chunk = std::max(1, b->tuning.chunk);
    L.counter = b->d_counter;
    int rc = check_hip(ctx, hipGraphCreate(&b->graph, 0), "batch graph");
    if (rc == DIC_OK) rc = check_hip(ctx, add_batch_nodes(L, b->graph), "batch graph nodes");
    if (rc == DIC_OK) rc = check_hip(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                                     "batch graph instantiate");
    if (rc != DIC_OK) drop_graph(b);
    return rc;
}

void release(dic_batch *b)
{
    drop_graph(b);
    ctx_free(b->ctx, b->d_pairs, sizeof(BatchPair) * b->pairs.size());
    ctx_free(b->ctx, b->d_items, sizeof(int2) * (size_t)b->total);
    ctx_free(b->ctx, b->d_counter, sizeof(unsigned) * kMaxBatchGroups);
    if (b->last_run) (void)hipEventDestroy(b->last_run);
    b->last_run = nullptr;
}

} // namespace

extern "C" {

int dic_batch_create(dic_ctx *ctx, const dic_batch_item *items, int n_items, int objective, dic_batch **out)
{
    if (!ctx || !out || n_items < 0 || (n_items > 0 && !items)) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (objective != DIC_OBJECTIVE_ZNSSD && objective != DIC_OBJECTIVE_SSD)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown objective");
    dic_batch *b = new (std::nothrow) dic_batch;
    if (!b) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "out of host memory");
    b->ctx = ctx;
    b->objective = objective;
    int rc = check_hip(ctx, hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
                       "device attribute");
    if (rc == DIC_OK)
        rc = check_hip(ctx, hipEventCreateWithFlags(&b->last_run, hipEventDisableTiming), "batch event");
    for (int k = 0; k < n_items && rc == DIC_OK; ++k) {
        const dic_batch_item &it = items[k];
        if (it.optimizer != DIC_OPTIMIZER_ICGN && it.optimizer != DIC_OPTIMIZER_FGN) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "batch item " + std::to_string(k) + ": unknown optimizer");
            break;
        }
        const bool fgn = it.optimizer == DIC_OPTIMIZER_FGN;
        if (!it.reference || (!fgn && !it.gradient) || !it.deformed || it.n < 0) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "batch item " + std::to_string(k) + ": null handle");
            break;
        }
        if (it.n > 0 && (!it.poi_xy || !it.init || !it.out)) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "batch item " + std::to_string(k) + ": null array");
            break;
        }
        if (it.shape_order != 1 && it.shape_order != 2) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "batch item " + std::to_string(k) + ": shape_order");
            break;
        }
        if (it.subset_radius < 1) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "batch item " + std::to_string(k) + ": subset_radius");
            break;
        }
        const dic_coeffs *c = it.deformed;
        if (k == 0) {
            b->precision = c->precision;
            b->degree = c->degree;
        }
        if (c->precision != b->precision || (it.gradient && it.gradient->precision != b->precision) ||
            c->degree != b->degree) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT,
                           "batch item " + std::to_string(k) + ": degree or precision differs from item 0");
            break;
        }
        if (it.gradient && (it.gradient->width != it.reference->width || it.gradient->height != it.reference->height)) {
            rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT,
                           "batch item " + std::to_string(k) + ": gradient planes do not match the reference");
            break;
        }
        const int *table = nullptr;
        int count = 0;
        rc = get_subset_table(ctx, it.subset_radius, &table, &count);
        if (rc == DIC_OK) rc = check_solve_lds(ctx, fgn, c->precision, objective, it.subset_radius, count);
        if (rc != DIC_OK) break;
        BatchPair bp{};
        IcgnArgs &a = bp.a;
        // size mismatch: every POI of the item reports InvalidInput (icgn.cpp:464-470)
        const bool same = it.reference->width == c->width && it.reference->height == c->height;
        a.W = same ? c->width : 0;
        a.H = same ? c->height : 0;
        a.pitch = c->pitch;
        a.b = c->border;
        a.ref = it.reference->data;
        a.gx = it.gradient ? it.gradient->gx : nullptr;
        a.gy = it.gradient ? it.gradient->gy : nullptr;
        a.coeffs = c->data;
        a.roi = nullptr;
        a.table = table;
        a.nfull = count;
        a.coeff_bytes = (int)c->bytes;
        a.radius = it.subset_radius;
        a.max_it = it.max_iterations;
        a.thr = it.convergence_threshold;
        a.min_needed = it.shape_order == 2 ? 12 : 6;
        a.fgn = fgn ? 1 : 0;
        a.xy = it.poi_xy;
        a.init = it.init;
        a.out = it.out;
        a.n = it.n;
        bp.shape_order = it.shape_order;
        b->pairs.push_back(bp);
        b->cost.push_back((double)count * (it.shape_order == 2 ? 1.7 : 1.0));
        b->total += it.n;
        b->max_nfull = std::max(b->max_nfull, count);
    }
    if (rc == DIC_OK && b->total > 0x7fffffff)
        rc = set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "a batch holds at most 2^31 - 1 POIs");
    if (rc == DIC_OK) {
        rc = check_hip(ctx, ctx_malloc(ctx, (void **)&b->d_pairs, sizeof(BatchPair) * b->pairs.size()), "batch alloc");
        if (rc == DIC_OK)
            rc = check_hip(ctx, ctx_malloc(ctx, (void **)&b->d_items, sizeof(int2) * (size_t)b->total), "batch alloc");
        if (rc == DIC_OK) rc = check_hip(ctx, ctx_malloc(ctx, (void **)&b->d_counter, sizeof(unsigned) * kMaxBatchGroups), "batch alloc");
        if (rc == DIC_OK && !b->pairs.empty())
            rc = check_hip(ctx, hipMemcpy(b->d_pairs, b->pairs.data(), sizeof(BatchPair) * b->pairs.size(),
                                          hipMemcpyHostToDevice),
                           "batch descriptor upload");
    }
    if (rc != DIC_OK) {
        release(b);
        delete b;
        return rc;
    }
    *out = b;
    return DIC_OK;
}

int dic_batch_set_tuning(dic_batch *b, const dic_batch_tuning *t)
{
    if (!b || !t) return DIC_ERR_INVALID_ARGUMENT;
    int rc = valid_tuning(b->ctx, *t);
    if (rc != DIC_OK) return rc;
    const dic_batch_tuning &o = b->tuning;
    // the graph holds the grid size, chunk and the two groups' item counts
    if (o.mapping != t->mapping || o.waves_per_cu != t->waves_per_cu || o.chunk != t->chunk || o.order != t->order)
        drop_graph(b);
    b->tuning = *t;
    return DIC_OK;
}

int dic_batch_get_tuning(const dic_batch *b, dic_batch_tuning *t)
{
    if (!b || !t) return DIC_ERR_INVALID_ARGUMENT;
    *t = b->tuning;
    return DIC_OK;
}

int dic_batch_info(const dic_batch *b, int64_t *total_pois, int64_t *runs)
{
    if (!b) return DIC_ERR_INVALID_ARGUMENT;
    if (total_pois) *total_pois = b->total;
    if (runs) *runs = b->runs;
    return DIC_OK;
}

int dic_batch_run(dic_batch *b, void *stream)
{
    if (!b) return DIC_ERR_INVALID_ARGUMENT;
    if (b->total == 0) return DIC_OK;
    int rc = prepare(b);
    if (rc != DIC_OK) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
    rc = check_hip(b->ctx, hipGraphLaunch(b->exec, s), "batch graph launch");
    if (rc == DIC_OK) rc = check_hip(b->ctx, hipEventRecord(b->last_run, s), "batch event");
    if (rc == DIC_OK) {
        b->run_recorded = true;
        ++b->runs;
    }
    return rc;
}

int dic_batch_autotune(dic_batch *b, const dic_batch_tuning *candidates, int n_candidates, int reps,
                       dic_batch_tuning *best, float *best_ms)
{
    if (!b || reps < 1 || n_candidates < 0 || (n_candidates > 0 && !candidates)) return DIC_ERR_INVALID_ARGUMENT;
    dic_ctx *ctx = b->ctx;
    std::vector<dic_batch_tuning> cand;
    if (n_candidates > 0) {
        cand.assign(candidates, candidates + n_candidates);
    } else {
        for (int order : {DIC_BATCH_ORDER_NATURAL, DIC_BATCH_ORDER_COST, DIC_BATCH_ORDER_INTERLEAVE,
                          DIC_BATCH_ORDER_XCD})
            cand.push_back({DIC_BATCH_GRID, 0, 1, order});
        for (int wpc : {4, 8, 16})
            for (int chunk : {1, 4, 16})
                for (int order : {DIC_BATCH_ORDER_NATURAL, DIC_BATCH_ORDER_COST})
                    cand.push_back({DIC_BATCH_PERSISTENT, wpc, chunk, order});
        bool first_fp32 = false;
        for (const BatchPair &p : b->pairs)
            first_fp32 = first_fp32 || tile_item(b, p, DIC_BATCH_TILE);
        if (first_fp32)
            for (int order : {DIC_BATCH_ORDER_NATURAL, DIC_BATCH_ORDER_COST, DIC_BATCH_ORDER_XCD})
                cand.push_back({DIC_BATCH_TILE, 0, 1, order});
    }
    for (const auto &c : cand) {
        int rc = valid_tuning(ctx, c);
        if (rc != DIC_OK) return rc;
    }
    hipEvent_t e0, e1;
    int rc = check_hip(ctx, hipEventCreate(&e0), "event");
    if (rc != DIC_OK) return rc;
    rc = check_hip(ctx, hipEventCreate(&e1), "event");
    if (rc != DIC_OK) {
        (void)hipEventDestroy(e0);
        return rc;
    }
    float best_t = 0.0f;
    dic_batch_tuning best_c = b->tuning;
    bool have = false;
    for (const auto &c : cand) {
        rc = dic_batch_set_tuning(b, &c);
        if (rc == DIC_OK) rc = dic_batch_run(b, nullptr); // warm: builds the graph and item order
        std::vector<float> t;
        for (int r = 0; r < reps && rc == DIC_OK; ++r) {
            rc = check_hip(ctx, hipEventRecord(e0, ctx->stream), "event");
            if (rc == DIC_OK) rc = dic_batch_run(b, nullptr);
            if (rc == DIC_OK) rc = check_hip(ctx, hipEventRecord(e1, ctx->stream), "event");
            if (rc == DIC_OK) rc = check_hip(ctx, hipEventSynchronize(e1), "event");
            float ms = 0.0f;
            if (rc == DIC_OK) rc = check_hip(ctx, hipEventElapsedTime(&ms, e0, e1), "event");
            t.push_back(ms);
        }
        if (rc != DIC_OK) break;
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (!have || med < best_t) {
            best_t = med;
            best_c = c;
            have = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != DIC_OK) return rc;
    rc = dic_batch_set_tuning(b, &best_c);
    if (rc == DIC_OK) rc = prepare(b);
    if (rc != DIC_OK) return rc;
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    return DIC_OK;
}

void dic_batch_destroy(dic_batch *b)
{
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    release(b); // waits for the last replay on whichever stream it ran
    delete b;
}

} // extern "C"
