// server.hip -- the serving loop (dic_server_*): admission by HBM budget and
// slots, continuous batching of up to K frames per active request per tick
// (K = the guess lag: frame f starts from frame f - K's results, frames below
// K from the request's guesses), three ticks in flight over two streams
// (prep: frame uploads and result downloads; compute: reference gradients,
// prefilters, initial guesses, IC-GN).  A tick's uploads run on the copy
// engines while the previous tick's solve holds the CUs: the kernels that
// prepare a tick run on the compute stream, so that no upload waits behind a
// kernel that cannot get a CU, and a tick's downloads are queued after the
// next tick's uploads.  Two streams, not three: a process gets a few hardware
// queues (GPU_MAX_HW_QUEUES, 4 by default), and streams beyond them share
// one, in whose order their work then runs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "dic_internal.hpp"

using namespace dic;

namespace {

size_t elem_bytes(int precision) { return precision == DIC_PRECISION_FP64 ? sizeof(double) : sizeof(float); }
size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Device and pinned buffers of one request shape (W, H, n); pooled for reuse
// so the loop never calls hipFree (which would wait for the whole device).
// Per-frame buffers form rings of kDepth * K slots, frame f in slot
// f mod (kDepth * K): the kDepth ticks in flight hold at most kDepth * K
// consecutive frames of a request, frame f's guess source f - K lies in an
// earlier tick on the same compute stream, and a slot is reused only by a
// tick built after the one that used it was finalized.
constexpr int kDepth = 3; // ticks in flight
struct Slab {
    int W = 0, H = 0, n = 0;
    size_t bytes = 0;       // device bytes
    char *dev = nullptr;    // one allocation, carved below
    float *ref = nullptr;
    std::vector<float *> frame;         // kDepth * K rings (uploaded frames)
    void *gx = nullptr, *gy = nullptr;
    size_t coef_bytes = 0;
    double *xy = nullptr;
    dic_initial *init0 = nullptr;       // the request's guesses (frames below K)
    std::vector<void *> coef;           // kDepth * K rings
    std::vector<dic_initial *> init;    // guesses of frame f from frame f - K
    std::vector<dic_result *> out;
    std::vector<dic_result *> pinned;   // host, pinned
};

struct Req {
    int64_t id = 0;
    dic_request r{};
    double arrival = 0.0;
    Slab *slab = nullptr;
    int next = 0;       // next frame to put into a tick
    int done = 0;       // frames whose results are on the host
    bool gradients = false; // reference gradients formed (compute stream)
    bool finished = false;
    int rc = DIC_OK;    // a request that could not be served
};

struct Tick {
    int slot = 0;
    std::vector<std::pair<Req *, int>> frames;
    hipEvent_t copied = nullptr;
    bool downloads_queued = false;
};

} // namespace

struct dic_server {
    dic_ctx *ctx = nullptr;
    dic_server_config cfg{};
    int border = 3, pitch = 0, rows_max = 0;
    int lag = 1;   // K
    bool extrapolate = false; // guesses 2 r(f-K) - r(f-2K)
    int ring = kDepth; // kDepth * K
    long tick_pois = 0;   // extra frames join a tick while it holds fewer POIs
    int copy_threads = 1; // threads of the result copy-out
    const int *table = nullptr;
    int nfull = 0;
    hipStream_t prep = nullptr, compute = nullptr;
    hipEvent_t prep_done[kDepth] = {}, solved[kDepth] = {};
    double *tmp_a = nullptr, *tmp_b = nullptr; // prefilter temporaries (compute stream)
    size_t tmp_bytes = 0;
    BatchPair *h_pairs[kDepth] = {}, *d_pairs[kDepth] = {};
    int *h_off[kDepth] = {}, *d_off[kDepth] = {};

    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int failed = DIC_OK;
    std::string error;
    std::deque<Req *> pending;
    std::vector<Req *> active;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::multimap<std::tuple<int, int, int>, Slab *> pool; // idle slabs by (W, H, n)
    std::vector<Slab *> slabs;                             // every slab, for teardown
    int64_t next_id = 1;
    uint64_t held = 0, peak = 0; // device bytes of active + pooled slabs
    dic_server_stats st{};
    std::chrono::steady_clock::time_point t0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    size_t max_pairs() const { return (size_t)cfg.max_active * (size_t)lag; } // frames in one tick
};

namespace {

size_t slab_device_bytes(const dic_server *s, int W, int H, int n)
{
    const int b = s->border;
    const size_t px = (size_t)W * H, cpx = (size_t)(W + 2 * b) * (H + 2 * b);
    const size_t e = elem_bytes(s->cfg.precision), R = (size_t)s->ring;
    return align256(px * sizeof(float)) + 2 * align256(px * e) + align256(16 * (size_t)n) +
           align256(sizeof(dic_initial) * (size_t)n) +
           R * (align256(px * sizeof(float)) + align256(cpx * e) + align256(sizeof(dic_initial) * (size_t)n) +
                align256(sizeof(dic_result) * (size_t)n));
}

int fail(dic_server *s, int code, const std::string &msg)
{
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed == DIC_OK) {
        s->failed = code;
        s->error = msg;
    }
    s->cv_done.notify_all();
    return code;
}

int hipcheck(dic_server *s, hipError_t e, const char *what)
{
    if (e == hipSuccess) return DIC_OK;
    (void)hipGetLastError(); // consumed, as check_hip does
    return fail(s, DIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void free_slab(dic_server *s, Slab *sl)
{
    if (sl->dev) ctx_free(s->ctx, sl->dev, sl->bytes);
    for (auto *p : sl->pinned)
        if (p) (void)hipHostFree(p);
    delete sl;
}

// a slab for (W, H, n): from the pool, or newly allocated (bytes already
// reserved against the budget by the caller)
Slab *take_slab(dic_server *s, int W, int H, int n)
{
    auto it = s->pool.find({W, H, n});
    if (it != s->pool.end()) {
        Slab *sl = it->second;
        s->pool.erase(it);
        return sl;
    }
    Slab *sl = new (std::nothrow) Slab;
    if (!sl) return nullptr;
    sl->W = W;
    sl->H = H;
    sl->n = n;
    sl->bytes = slab_device_bytes(s, W, H, n);
    if (ctx_malloc(s->ctx, (void **)&sl->dev, sl->bytes) != hipSuccess) {
        delete sl;
        return nullptr;
    }
    const int b = s->border;
    const size_t px = (size_t)W * H, cpx = (size_t)(W + 2 * b) * (H + 2 * b);
    const size_t e = elem_bytes(s->cfg.precision);
    char *p = sl->dev;
    auto carve = [&](size_t nb) {
        char *q = p;
        p += align256(nb);
        return q;
    };
    sl->ref = (float *)carve(px * sizeof(float));
    sl->gx = carve(px * e);
    sl->gy = carve(px * e);
    sl->coef_bytes = cpx * e;
    sl->xy = (double *)carve(16 * (size_t)n);
    sl->init0 = (dic_initial *)carve(sizeof(dic_initial) * (size_t)n);
    sl->frame.assign(s->ring, nullptr);
    sl->coef.assign(s->ring, nullptr);
    sl->init.assign(s->ring, nullptr);
    sl->out.assign(s->ring, nullptr);
    sl->pinned.assign(s->ring, nullptr);
    for (int k = 0; k < s->ring; ++k) {
        sl->frame[k] = (float *)carve(px * sizeof(float));
        sl->coef[k] = carve(cpx * e);
        sl->init[k] = (dic_initial *)carve(sizeof(dic_initial) * (size_t)n);
        sl->out[k] = (dic_result *)carve(sizeof(dic_result) * (size_t)n);
    }
    for (int k = 0; k < s->ring; ++k)
        if (hipHostMalloc((void **)&sl->pinned[k], sizeof(dic_result) * (size_t)std::max(n, 1), 0) != hipSuccess) {
            free_slab(s, sl);
            return nullptr;
        }
    s->slabs.push_back(sl);
    return sl;
}

// Admission (worker thread, lock held): arrival order, slots and budget.
// Idle pooled slabs of other shapes are released when the budget needs it.
void admit(dic_server *s, std::vector<Req *> &admitted)
{
    while (!s->pending.empty() && (int)s->active.size() < s->cfg.max_active) {
        Req *q = s->pending.front();
        const int W = q->r.width, H = q->r.height, n = q->r.n;
        const bool pooled = s->pool.count({W, H, n}) > 0;
        const uint64_t need = pooled ? 0 : slab_device_bytes(s, W, H, n);
        if (s->cfg.hbm_budget != 0 && s->held + need > s->cfg.hbm_budget) {
            // release idle slabs (other shapes) until it fits or none are left
            while (!s->pool.empty() && s->held + need > s->cfg.hbm_budget) {
                auto it = s->pool.begin();
                Slab *sl = it->second;
                s->pool.erase(it);
                s->held -= sl->bytes;
                s->slabs.erase(std::find(s->slabs.begin(), s->slabs.end(), sl));
                free_slab(s, sl);
            }
            if (s->held + need > s->cfg.hbm_budget) {
                if (s->active.empty()) { // can never fit: refuse this request only
                    s->pending.pop_front();
                    q->finished = true;
                    q->rc = DIC_ERR_BUDGET;
                    s->cv_done.notify_all();
                    continue;
                }
                break; // wait for active requests to leave
            }
        }
        s->pending.pop_front();
        Slab *sl = take_slab(s, W, H, n);
        if (!sl) {
            q->finished = true;
            if (s->failed == DIC_OK) {
                s->failed = DIC_ERR_HIP;
                s->error = "device or pinned allocation failed for request " + std::to_string(q->id);
            }
            s->cv_done.notify_all();
            continue;
        }
        if (!pooled) s->held += sl->bytes;
        s->peak = std::max(s->peak, s->held);
        q->slab = sl;
        s->active.push_back(q);
        s->st.max_active_seen = std::max<int32_t>(s->st.max_active_seen, (int32_t)s->active.size());
        admitted.push_back(q);
    }
}

// Uploads of an admitted request (prep stream): reference, POIs, frame-0
// guesses.  Its reference gradients are formed on the compute stream before
// its first solve.
int setup_request(dic_server *s, Req *q)
{
    Slab *sl = q->slab;
    const size_t px = (size_t)q->r.width * q->r.height;
    int rc = hipcheck(s, hipMemcpyAsync(sl->ref, q->r.reference, px * sizeof(float), hipMemcpyHostToDevice, s->prep),
                      "reference upload");
    if (rc == DIC_OK && q->r.n > 0) {
        rc = hipcheck(s, hipMemcpyAsync(sl->xy, q->r.poi_xy, 16 * (size_t)q->r.n, hipMemcpyHostToDevice, s->prep),
                      "POI upload");
        if (rc == DIC_OK)
            rc = hipcheck(s, hipMemcpyAsync(sl->init0, q->r.init, sizeof(dic_initial) * (size_t)q->r.n,
                                            hipMemcpyHostToDevice, s->prep),
                          "initial-guess upload");
    }
    return rc;
}

int build_tick(dic_server *s, Tick &t, int slot)
{
    t.slot = slot;
    t.frames.clear();
    // the next frame of every active request, then (arrival order) up to K - 1
    // more consecutive frames of each while the tick holds fewer than
    // tick_pois POIs.  At most K frames of a request share a tick, so frame
    // f's source f - K lies in an earlier one, already ordered before this
    // one on the compute stream; which tick a frame joins changes no result.
    long pois = 0;
    std::vector<char> open(s->active.size(), 0); // took a frame in the previous pass
    for (size_t i = 0; i < s->active.size(); ++i) {
        Req *q = s->active[i];
        if (q->next < q->r.n_frames) {
            t.frames.push_back({q, q->next++});
            pois += q->r.n;
            open[i] = 1;
        }
    }
    for (int j = 1; j < s->lag && pois < s->tick_pois; ++j)
        for (size_t i = 0; i < s->active.size(); ++i) {
            Req *q = s->active[i];
            if (!open[i]) continue;
            if (pois >= s->tick_pois || q->next >= q->r.n_frames) {
                open[i] = 0;
                continue;
            }
            t.frames.push_back({q, q->next++});
            pois += q->r.n;
        }
    if (t.frames.empty()) return DIC_OK;
    // prep: frame uploads into the frames' ring slots (the slot's previous
    // frame was prefiltered by a tick already finalized)
    for (auto &fr : t.frames) {
        Req *q = fr.first;
        const int f = fr.second;
        const size_t px = (size_t)q->r.width * q->r.height;
        int rc = hipcheck(s, hipMemcpyAsync(q->slab->frame[f % s->ring], q->r.frames[f], px * sizeof(float),
                                            hipMemcpyHostToDevice, s->prep),
                          "frame upload");
        if (rc != DIC_OK) return rc;
    }
    int rc = hipcheck(s, hipEventRecord(s->prep_done[slot], s->prep), "event");
    if (rc == DIC_OK) rc = hipcheck(s, hipStreamWaitEvent(s->compute, s->prep_done[slot], 0), "event");
    if (rc != DIC_OK) return rc;
    // compute: reference gradients of requests new to this tick, prefilters
    for (auto &fr : t.frames) {
        Req *q = fr.first;
        Slab *sl = q->slab;
        const int k = fr.second % s->ring;
        if (!q->gradients) {
            rc = hipcheck(s, launch_gradient_central(sl->ref, q->r.width, q->r.height, s->cfg.precision, sl->gx,
                                                     sl->gy, s->compute),
                          "reference gradients");
            if (rc != DIC_OK) return rc;
            q->gradients = true;
        }
        rc = hipcheck(s, launch_prefilter(sl->frame[k], q->r.width, q->r.height, s->cfg.degree, s->border, 1,
                                          s->cfg.precision, sl->coef[k], s->tmp_a, s->tmp_b, s->compute),
                      "frame prefilter");
        if (rc != DIC_OK) return rc;
    }
    // compute: guesses from frame f - K, then one launch over all frames
    BatchPair *hp = s->h_pairs[slot];
    int *ho = s->h_off[slot];
    int total = 0;
    for (size_t k = 0; k < t.frames.size(); ++k) {
        Req *q = t.frames[k].first;
        Slab *sl = q->slab;
        const int f = t.frames[k].second, r = f % s->ring;
        const dic_initial *guess = sl->init0;
        if (f >= s->lag) {
            // frame f - 2K's result slot is reused only by frame f + K, whose
            // solve follows this tick on the compute stream
            if (s->extrapolate && f >= 2 * s->lag)
                rc = hipcheck(s, launch_init_extrapolated(sl->out[(f - s->lag) % s->ring],
                                                          sl->out[(f - 2 * s->lag) % s->ring], sl->init[r], q->r.n,
                                                          s->compute),
                              "initial guesses extrapolated from earlier frames");
            else
                rc = hipcheck(s, launch_init_from_results(sl->out[(f - s->lag) % s->ring], sl->init[r], q->r.n,
                                                          s->compute),
                              "initial guesses from an earlier frame");
            if (rc != DIC_OK) return rc;
            guess = sl->init[r];
        }
        BatchPair bp{};
        IcgnArgs &a = bp.a;
        a.W = q->r.width;
        a.H = q->r.height;
        a.pitch = q->r.width + 2 * s->border;
        a.b = s->border;
        a.ref = sl->ref;
        a.gx = sl->gx;
        a.gy = sl->gy;
        a.coeffs = sl->coef[r];
        a.roi = nullptr;
        a.table = s->table;
        a.nfull = s->nfull;
        a.coeff_bytes = (int)sl->coef_bytes;
        a.radius = s->cfg.subset_radius;
        a.max_it = s->cfg.max_iterations;
        a.thr = s->cfg.convergence_threshold;
        a.min_needed = s->cfg.shape_order == 2 ? 12 : 6;
        a.fgn = s->cfg.optimizer == DIC_OPTIMIZER_FGN ? 1 : 0;
        a.xy = sl->xy;
        a.init = guess;
        a.out = sl->out[r];
        a.n = q->r.n;
        bp.shape_order = s->cfg.shape_order;
        hp[k] = bp;
        ho[k] = total;
        total += q->r.n;
    }
    const int np = (int)t.frames.size();
    ho[np] = total;
    rc = hipcheck(s, hipMemcpyAsync(s->d_pairs[slot], hp, sizeof(BatchPair) * np, hipMemcpyHostToDevice, s->compute),
                  "tick descriptors");
    if (rc == DIC_OK)
        rc = hipcheck(s, hipMemcpyAsync(s->d_off[slot], ho, sizeof(int) * (np + 1), hipMemcpyHostToDevice, s->compute),
                      "tick offsets");
    RangeLaunch L;
    L.objective = s->cfg.objective;
    L.degree = s->cfg.degree;
    L.precision = s->cfg.precision;
    L.shape_order = s->cfg.shape_order;
    L.pairs = s->d_pairs[slot];
    L.offsets = s->d_off[slot];
    L.npairs = np;
    L.total = total;
    L.max_nfull = s->nfull;
    L.fgn = s->cfg.optimizer == DIC_OPTIMIZER_FGN ? 1 : 0;
    if (rc == DIC_OK) rc = hipcheck(s, launch_range(L, s->compute), "tick launch");
    if (rc == DIC_OK) rc = hipcheck(s, hipEventRecord(s->solved[slot], s->compute), "event");
    t.downloads_queued = false;
    return rc;
}

// The tick's result downloads (prep stream, after its solve).  Queued after
// the next tick's uploads: a download waits for its solve, and the prep
// stream runs in order, so a download queued first would hold the next
// tick's uploads (and so its solve) until this tick's solve has ended.
int queue_downloads(dic_server *s, Tick &t)
{
    if (t.downloads_queued) return DIC_OK;
    t.downloads_queued = true;
    int rc = hipcheck(s, hipStreamWaitEvent(s->prep, s->solved[t.slot], 0), "event");
    for (auto &fr : t.frames) {
        if (rc != DIC_OK) break;
        Slab *sl = fr.first->slab;
        const int f = fr.second;
        rc = hipcheck(s, hipMemcpyAsync(sl->pinned[f % s->ring], sl->out[f % s->ring],
                                        sizeof(dic_result) * fr.first->r.n, hipMemcpyDeviceToHost, s->prep),
                      "result download");
    }
    if (rc == DIC_OK) rc = hipcheck(s, hipEventRecord(t.copied, s->prep), "event");
    return rc;
}

// The tick's results, from the pinned staging slots into the callers' arrays
// (worker thread, no lock: only the worker touches slabs, and a request's
// arrays stay valid until it is reported done).  Writes into freshly
// allocated host arrays fault their pages in and run at a few GB/s per core,
// so a large tick is split in pieces over up to copy_threads threads.
void copy_results(dic_server *s, const Tick &t)
{
    struct Piece {
        char *dst;
        const char *src;
        size_t bytes;
    };
    constexpr size_t kPiece = (size_t)8 << 20;
    std::vector<Piece> pieces;
    for (auto &fr : t.frames) {
        const Req *q = fr.first;
        const int f = fr.second;
        const size_t n = (size_t)q->r.n, bytes = sizeof(dic_result) * n;
        char *dst = reinterpret_cast<char *>(q->r.results + (size_t)f * n);
        const char *src = reinterpret_cast<const char *>(q->slab->pinned[f % s->ring]);
        for (size_t o = 0; o < bytes; o += kPiece) pieces.push_back({dst + o, src + o, std::min(kPiece, bytes - o)});
    }
    std::atomic<size_t> next{0};
    auto body = [&] {
        for (size_t i; (i = next.fetch_add(1)) < pieces.size();) std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
    };
    std::vector<std::thread> helpers;
    const size_t want = std::min<size_t>((size_t)s->copy_threads, pieces.size());
    for (size_t k = 1; k < want; ++k) {
        try {
            helpers.emplace_back(body);
        } catch (...) {
            break; // fewer helpers; the calling thread copies the rest
        }
    }
    body();
    for (auto &h : helpers) h.join();
}

void finalize_tick(dic_server *s, Tick &t)
{
    if (queue_downloads(s, t) != DIC_OK) return;
    if (hipcheck(s, hipEventSynchronize(t.copied), "tick completion") != DIC_OK) return;
    copy_results(s, t);
    const double now = s->now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &fr : t.frames) {
        Req *q = fr.first;
        const int f = fr.second, n = q->r.n;
        if (q->r.frame_done_s) q->r.frame_done_s[f] = now;
        q->done = f + 1;
        s->st.frames_done += 1;
        s->st.pois_done += n;
        if (q->done == q->r.n_frames) {
            q->finished = true;
            s->st.completed += 1;
            s->active.erase(std::find(s->active.begin(), s->active.end(), q));
            s->pool.insert({{q->slab->W, q->slab->H, q->slab->n}, q->slab});
            q->slab = nullptr;
        }
    }
    s->st.ticks += 1;
    t.frames.clear();
    s->cv_done.notify_all();
}

void worker_main(dic_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    Tick ticks[kDepth];
    for (auto &t : ticks)
        if (hipcheck(s, hipEventCreateWithFlags(&t.copied, hipEventDisableTiming), "event") != DIC_OK) return;
    std::deque<int> inflight; // tick slots in launch order
    int next_slot = 0;
    for (;;) {
        std::vector<Req *> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            admit(s, admitted);
            const bool work = !admitted.empty() || !s->active.empty() || !inflight.empty();
            if (!work) {
                if (s->stop && s->pending.empty()) break;
                s->cv_work.wait(lk);
                continue;
            }
        }
        bool ok = true;
        for (Req *q : admitted) ok = ok && setup_request(s, q) == DIC_OK;
        if (!ok) break;
        // a new tick while fewer than kDepth are in flight and some frame is waiting
        bool waiting = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Req *q : s->active) waiting = waiting || q->next < q->r.n_frames;
        }
        if (waiting && inflight.size() < (size_t)kDepth) {
            const int slot = next_slot;
            next_slot = (next_slot + 1) % kDepth;
            if (build_tick(s, ticks[slot], slot) != DIC_OK) break;
            if (!ticks[slot].frames.empty()) {
                // the previous tick's downloads now follow this tick's uploads
                if (!inflight.empty() && queue_downloads(s, ticks[inflight.back()]) != DIC_OK) break;
                inflight.push_back(slot);
            }
            continue;
        }
        if (!inflight.empty()) {
            finalize_tick(s, ticks[inflight.front()]);
            inflight.pop_front();
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->failed != DIC_OK) break;
            continue;
        }
    }
    for (int slot : inflight) {
        if (!ticks[slot].downloads_queued) continue; // never queued: nothing recorded on its event
        (void)hipEventSynchronize(ticks[slot].copied);
    }
    for (auto &t : ticks) (void)hipEventDestroy(t.copied);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed != DIC_OK) { // fail every unfinished request
        for (auto &kv : s->reqs) kv.second->finished = true;
        s->pending.clear();
        s->active.clear();
    }
    s->cv_done.notify_all();
}

void destroy_resources(dic_server *s)
{
    if (s->prep) (void)hipStreamSynchronize(s->prep);
    if (s->compute) (void)hipStreamSynchronize(s->compute);
    for (Slab *sl : s->slabs) free_slab(s, sl);
    s->slabs.clear();
    s->pool.clear();
    for (int k = 0; k < kDepth; ++k) {
        if (s->h_pairs[k]) (void)hipHostFree(s->h_pairs[k]);
        if (s->h_off[k]) (void)hipHostFree(s->h_off[k]);
        if (s->d_pairs[k]) ctx_free(s->ctx, s->d_pairs[k], sizeof(BatchPair) * s->max_pairs());
        if (s->d_off[k]) ctx_free(s->ctx, s->d_off[k], sizeof(int) * (s->max_pairs() + 1));
        if (s->prep_done[k]) (void)hipEventDestroy(s->prep_done[k]);
        if (s->solved[k]) (void)hipEventDestroy(s->solved[k]);
    }
    if (s->tmp_a) ctx_free(s->ctx, s->tmp_a, s->tmp_bytes);
    if (s->tmp_b) ctx_free(s->ctx, s->tmp_b, s->tmp_bytes);
    for (hipStream_t st : {s->prep, s->compute})
        if (st) (void)hipStreamDestroy(st);
}

} // namespace

extern "C" {

int dic_host_alloc(dic_ctx *ctx, size_t bytes, void **ptr)
{
    if (!ctx || !ptr) return DIC_ERR_INVALID_ARGUMENT;
    *ptr = nullptr;
    return check_hip(ctx, hipHostMalloc(ptr, bytes ? bytes : 1, 0), "pinned host allocation");
}

void dic_host_free(void *ptr)
{
    if (ptr) (void)hipHostFree(ptr);
}

int dic_server_create(dic_ctx *ctx, const dic_server_config *cfg, dic_server **out)
{
    if (!ctx || !cfg || !out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const dic_server_config &c = *cfg;
    if (c.max_active < 1 || c.max_active > 4096 || c.max_width < 1 || c.max_height < 1)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: max_active in [1, 4096] and a positive max frame size");
    if (c.shape_order != 1 && c.shape_order != 2)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: shape_order must be 1 or 2");
    if (c.subset_radius < 1) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: subset_radius must be >= 1");
    if (c.degree != 1 && c.degree != 3 && c.degree != 5)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: degree must be 1, 3 or 5");
    if (c.precision != DIC_PRECISION_FP64 && c.precision != DIC_PRECISION_MIXED)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: unknown precision");
    if (c.objective != DIC_OBJECTIVE_ZNSSD && c.objective != DIC_OBJECTIVE_SSD)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: unknown objective");
    if (c.optimizer != DIC_OPTIMIZER_ICGN && c.optimizer != DIC_OPTIMIZER_FGN)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: unknown optimizer");
    if (c.tick_pois < 0) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "server: tick_pois must be >= 0");
    if (c.guess_lag < 0 || c.guess_lag > DIC_SERVER_MAX_GUESS_LAG)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT,
                         "server: guess_lag must be in [0, " + std::to_string(DIC_SERVER_MAX_GUESS_LAG) + "]");
    const int b = 3;
    int rc0 = check_prefilter_size(ctx, c.max_width, c.max_height, b, c.precision, "server: largest frame");
    if (rc0 != DIC_OK) return rc0;
    dic_server *s = new (std::nothrow) dic_server;
    if (!s) return set_error(ctx, DIC_ERR_HIP, "host allocation failed");
    s->ctx = ctx;
    s->cfg = c;
    s->border = b;
    s->lag = c.guess_lag < 1 ? 1 : c.guess_lag;
    s->extrapolate = c.guess_extrapolate != 0;
    s->ring = kDepth * s->lag;
    s->tick_pois = c.tick_pois > 0 ? c.tick_pois : DIC_SERVER_DEFAULT_TICK_POIS;
    s->copy_threads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
    s->t0 = std::chrono::steady_clock::now();
    int rc = get_subset_table(ctx, c.subset_radius, &s->table, &s->nfull);
    if (rc == DIC_OK)
        rc = check_solve_lds(ctx, c.optimizer == DIC_OPTIMIZER_FGN, c.precision, c.objective, c.subset_radius,
                             s->nfull);
    auto hc = [&](hipError_t e, const char *w) {
        if (rc == DIC_OK) rc = check_hip(ctx, e, w);
    };
    hc(hipSetDevice(ctx->device), "set device");
    hc(hipStreamCreateWithFlags(&s->prep, hipStreamNonBlocking), "stream");
    hc(hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking), "stream");
    s->tmp_bytes = sizeof(double) * (size_t)(c.max_width + 2 * b) * (c.max_height + 2 * b);
    hc(ctx_malloc(ctx, (void **)&s->tmp_a, s->tmp_bytes), "prefilter temporaries");
    hc(ctx_malloc(ctx, (void **)&s->tmp_b, s->tmp_bytes), "prefilter temporaries");
    for (int k = 0; k < kDepth; ++k) {
        hc(hipHostMalloc((void **)&s->h_pairs[k], sizeof(BatchPair) * s->max_pairs(), 0), "pinned descriptors");
        hc(hipHostMalloc((void **)&s->h_off[k], sizeof(int) * (s->max_pairs() + 1), 0), "pinned offsets");
        hc(ctx_malloc(ctx, (void **)&s->d_pairs[k], sizeof(BatchPair) * s->max_pairs()), "descriptors");
        hc(ctx_malloc(ctx, (void **)&s->d_off[k], sizeof(int) * (s->max_pairs() + 1)), "offsets");
        hc(hipEventCreateWithFlags(&s->prep_done[k], hipEventDisableTiming), "event");
        hc(hipEventCreateWithFlags(&s->solved[k], hipEventDisableTiming), "event");
    }
    if (rc != DIC_OK) {
        destroy_resources(s);
        delete s;
        return rc;
    }
    try {
        s->worker = std::thread(worker_main, s);
    } catch (...) {
        destroy_resources(s);
        delete s;
        return set_error(ctx, DIC_ERR_HIP, "server: could not start the worker thread");
    }
    *out = s;
    return DIC_OK;
}

int dic_server_submit(dic_server *s, const dic_request *req, int64_t *id, double *arrival_s)
{
    if (!s || !req) return DIC_ERR_INVALID_ARGUMENT;
    const dic_request &r = *req;
    if (!r.reference || r.width < 1 || r.height < 1 || r.n_frames < 1 || r.n < 0 || !r.frames || !r.results ||
        (r.n > 0 && (!r.poi_xy || !r.init)))
        return DIC_ERR_INVALID_ARGUMENT;
    if (r.width > s->cfg.max_width || r.height > s->cfg.max_height) return DIC_ERR_INVALID_ARGUMENT;
    for (int f = 0; f < r.n_frames; ++f)
        if (!r.frames[f]) return DIC_ERR_INVALID_ARGUMENT;
    auto q = std::make_unique<Req>();
    q->r = r;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed != DIC_OK) return s->failed;
    if (s->stop) return DIC_ERR_INVALID_ARGUMENT;
    q->id = s->next_id++;
    q->arrival = s->now();
    if (id) *id = q->id;
    if (arrival_s) *arrival_s = q->arrival;
    s->pending.push_back(q.get());
    s->reqs[q->id] = std::move(q);
    s->st.submitted += 1;
    s->cv_work.notify_all();
    return DIC_OK;
}

int dic_server_wait(dic_server *s, int64_t id, double timeout_s)
{
    if (!s) return DIC_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return DIC_ERR_INVALID_ARGUMENT;
    Req *q = it->second.get();
    auto done = [&] { return q->finished || s->failed != DIC_OK; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        return DIC_ERR_UNSUPPORTED;
    }
    if (q->done == q->r.n_frames) return DIC_OK;
    if (q->rc != DIC_OK) return q->rc;
    return s->failed != DIC_OK ? s->failed : DIC_ERR_HIP;
}

int dic_server_frames_done(dic_server *s, int64_t id, int *frames)
{
    if (!s || !frames) return DIC_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return DIC_ERR_INVALID_ARGUMENT;
    *frames = it->second->done;
    return DIC_OK;
}

int dic_server_stats_get(dic_server *s, dic_server_stats *st)
{
    if (!s || !st) return DIC_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->st;
    st->active = (int32_t)s->active.size();
    st->pending = (int32_t)s->pending.size();
    st->bytes_held = s->held;
    st->peak_bytes_held = s->peak;
    return DIC_OK;
}

double dic_server_now(const dic_server *s) { return s ? s->now() : 0.0; }

const char *dic_server_error(const dic_server *s)
{
    if (!s) return "null server";
    return s->error.c_str();
}

void dic_server_destroy(dic_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    destroy_resources(s);
    delete s;
}

} // extern "C"
