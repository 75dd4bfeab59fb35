// seed.hip -- seeding on the GPU: the integer-pixel ZNCC search of
// IntegerSearchInitializer (src/initialization/integer_search.cpp), the
// candidate generation and selection of SeedSelector
// (src/subset/seed/seed_selector.cpp) with SubsetInitializer's sub-pixel
// refinement (src/initialization/subset_initializer.cpp), and mirror padding
// (src/subset/padding.cpp).
//
// The search is data-parallel over (point, shift): one thread evaluates the
// ZNCC of one integer shift, a workgroup reduces 256 shifts to the first
// maximum in scan order, and a per-point pass reduces the workgroups and sets
// up the next stage (coarse grid -> refinement + local window).  Arithmetic
// is FP64 in the reference's summation order; fused multiply-add is off in
// this file so every mean, norm and correlation is the same double as the
// reference's (a default x86-64 build has no FMA contraction).
#include <algorithm>
#include <climits>
#include <cmath>
#include <new>
#include <vector>

#include "dic_internal.hpp"

#pragma clang fp contract(off)

using namespace dic;

namespace {

constexpr double kSearchEps = 1e-12;    // kEpsilon, integer_search.cpp:10
constexpr double kLocalMargin = 0.02;   // kLocalPreferenceZnccMargin, integer_search.cpp:11
constexpr int kSearchBlock = 256;
constexpr int kMaxSeeds = 2048;         // k-means centres and their sums held in LDS (56 KiB)
constexpr int kNearest = 50;            // seed_selector.cpp:300

// One block of shifts: u = u0 + iu * step, v = v0 + iv * step, row-major.
struct Region {
    int u0, v0, nu, nv, step;
    int coarse; // 1: the subsampled subset (coarse grid), 0: the full subset
};

struct SearchState {
    int cx, cy, valid, coarse_ok;
    int n_full, n_coarse;
    int stage_global_ok; // the multigrid search produced a finite coarse winner
    int pad;
    double mean_full, norm_full, mean_coarse, norm_coarse;
    int min_u, max_u, min_v, max_v;
    Region task[2]; // [0] coarse grid or refinement, [1] local / prior window
};

struct Partial {
    double z;
    int idx;
    int pad;
};

struct SearchArgs {
    const float *ref, *def;
    const uint8_t *roi; // masked only
    int W, H;
    int r, sr, coarse_step, refine_radius;
    int masked, around_prior;
    const double *xy, *prior;
    SearchState *st;
    Partial *part;
    int npoints, max_chunks;
    dic_estimate *out;
};

__device__ __forceinline__ int isqrt_floor(int m)
{
    int w = (int)sqrt((double)m);
    while (w * w > m) --w;
    while ((w + 1) * (w + 1) <= m) ++w;
    return w;
}

// dx offsets (from -r) of the circle's samples in row dy at sample step ss:
// dx = k * ss - r for k in [k0, k1], the samples with dx^2 + dy^2 <= r^2 in
// the reference's order (collect_circular_reference_samples, :62-79).
__device__ __forceinline__ void row_range(int r, int dy, int ss, int &k0, int &k1)
{
    const int w = isqrt_floor(r * r - dy * dy);
    k0 = (r - w + ss - 1) / ss;
    k1 = (r + w) / ss;
}

__device__ __forceinline__ bool roi_at(const uint8_t *roi, int W, int H, int x, int y)
{
    return x >= 0 && y >= 0 && x < W && y < H && roi[(size_t)y * W + x] != 0;
}

// Sample count, mean and norm of the reference subset (sample_mean /
// sample_norm, :107-125).  Returns false where an unmasked sample leaves the
// image (the reference's Image::at throws there).
template <bool MASKED>
__device__ bool subset_stats(const float *ref, const uint8_t *roi, int W, int H, int cx, int cy, int r, int ss,
                             int &n, double &mean, double &norm)
{
    n = 0;
    double s = 0.0;
    for (int dy = -r; dy <= r; dy += ss) {
        int k0, k1;
        row_range(r, dy, ss, k0, k1);
        const int y = cy + dy;
        for (int k = k0; k <= k1; ++k) {
            const int x = cx - r + k * ss;
            if (MASKED) {
                if (!roi_at(roi, W, H, x, y)) continue;
            } else if (x < 0 || y < 0 || x >= W || y >= H) {
                return false;
            }
            s += (double)ref[(size_t)y * W + x];
            ++n;
        }
    }
    mean = n > 0 ? s / (double)n : 0.0;
    double q = 0.0;
    for (int dy = -r; dy <= r; dy += ss) {
        int k0, k1;
        row_range(r, dy, ss, k0, k1);
        const int y = cy + dy;
        for (int k = k0; k <= k1; ++k) {
            const int x = cx - r + k * ss;
            if (MASKED && !roi_at(roi, W, H, x, y)) continue;
            const double d = (double)ref[(size_t)y * W + x] - mean;
            q += d * d;
        }
    }
    norm = q;
    return true;
}

// circular_subset_zncc (:127-165) at shift (u, v); the caller has checked
// that the shifted subset lies inside the deformed image.
template <bool MASKED>
__device__ double shift_zncc(const float *__restrict__ ref, const float *__restrict__ def,
                             const uint8_t *__restrict__ roi, int W, int H, int cx, int cy, int u, int v, int r,
                             int ss, int n, double rm, double rn)
{
    if (n == 0 || rn <= kSearchEps) return -INFINITY;
    double dsum = 0.0;
    for (int dy = -r; dy <= r; dy += ss) {
        int k0, k1;
        row_range(r, dy, ss, k0, k1);
        const int yr = cy + dy;
        const float *drow = def + (size_t)(yr + v) * W + (cx + u - r);
        for (int k = k0; k <= k1; ++k) {
            if (MASKED && !roi_at(roi, W, H, cx - r + k * ss, yr)) continue;
            dsum += (double)drow[k * ss];
        }
    }
    const double dm = dsum / (double)n;
    double num = 0.0, dn = 0.0;
    for (int dy = -r; dy <= r; dy += ss) {
        int k0, k1;
        row_range(r, dy, ss, k0, k1);
        const int yr = cy + dy;
        const float *drow = def + (size_t)(yr + v) * W + (cx + u - r);
        const float *rrow = ref + (size_t)yr * W + (cx - r);
        for (int k = k0; k <= k1; ++k) {
            if (MASKED && !roi_at(roi, W, H, cx - r + k * ss, yr)) continue;
            const double rd = (double)rrow[k * ss] - rm;
            const double dd = (double)drow[k * ss] - dm;
            num += rd * dd;
            dn += dd * dd;
        }
    }
    if (dn <= kSearchEps) return -INFINITY;
    return num / sqrt(rn * dn);
}

__device__ __forceinline__ int min_required(int full_count) // min_required_samples, :49-52
{
    const int h = (int)ceil(0.5 * (double)full_count);
    return h > 12 ? h : 12;
}

__device__ int circle_count(int r, int ss) // full_circle_sample_count, :35-47
{
    int c = 0;
    for (int dy = -r; dy <= r; dy += ss) {
        int k0, k1;
        row_range(r, dy, ss, k0, k1);
        c += k1 - k0 + 1;
    }
    return c;
}

__device__ __forceinline__ Region empty_region()
{
    Region g;
    g.u0 = g.v0 = g.nu = g.nv = 0;
    g.step = 1;
    g.coarse = 0;
    return g;
}

// [lo, hi] x [vlo, vhi] at step 1 (empty when lo > hi or vlo > vhi)
__device__ __forceinline__ Region window(int lo, int hi, int vlo, int vhi)
{
    Region g = empty_region();
    if (lo <= hi && vlo <= vhi) {
        g.u0 = lo;
        g.v0 = vlo;
        g.nu = hi - lo + 1;
        g.nv = vhi - vlo + 1;
    }
    return g;
}

template <bool MASKED>
__device__ void prep_point(const SearchArgs &a, int i)
{
    SearchState s = {};
    s.task[0] = empty_region();
    s.task[1] = empty_region();
    const int r = a.r;
    const double px = a.xy[2 * i], py = a.xy[2 * i + 1];
    // a point that is not a finite image coordinate is invalid (its integer
    // conversion would be undefined)
    if (!(fabs(px) < 1e9 && fabs(py) < 1e9)) {
        a.st[i] = s;
        return;
    }
    if (a.around_prior) {
        // estimate_around_displacement :519-579, _with_mask_ :581-645
        const double pu = a.prior[2 * i], pv = a.prior[2 * i + 1];
        if (fabs(pu) < 1e9 && fabs(pv) < 1e9) {
            s.cx = (int)llround(px);
            s.cy = (int)llround(py);
            const bool inside = subset_stats<MASKED>(a.ref, a.roi, a.W, a.H, s.cx, s.cy, r, 1, s.n_full,
                                                     s.mean_full, s.norm_full);
            if (inside && s.n_full >= min_required(circle_count(r, 1)) && s.norm_full > kSearchEps) {
                s.valid = 1;
                s.min_u = r - s.cx;
                s.max_u = a.W - 1 - r - s.cx;
                s.min_v = r - s.cy;
                s.max_v = a.H - 1 - r - s.cy;
                const int iu = (int)llround(pu), iv = (int)llround(pv);
                s.task[1] = window(max(s.min_u, iu - a.sr), min(s.max_u, iu + a.sr), max(s.min_v, iv - a.sr),
                                   min(s.max_v, iv + a.sr));
            }
        }
        a.st[i] = s;
        return;
    }
    s.cx = (int)round(px);
    s.cy = (int)round(py);
    const int cs = a.coarse_step;
    if (!MASKED) {
        // estimate_with_interpolators :318-403
        const bool inb = s.cx - r >= 0 && s.cy - r >= 0 && s.cx + r < a.W && s.cy + r < a.H;
        if (inb) {
            subset_stats<false>(a.ref, a.roi, a.W, a.H, s.cx, s.cy, r, 1, s.n_full, s.mean_full, s.norm_full);
            if (s.n_full > 0 && s.norm_full > kSearchEps) {
                s.valid = 1;
                subset_stats<false>(a.ref, a.roi, a.W, a.H, s.cx, s.cy, r, cs, s.n_coarse, s.mean_coarse,
                                    s.norm_coarse);
                s.coarse_ok = s.n_coarse > 0 && s.norm_coarse > kSearchEps;
            }
        }
    } else {
        // estimate_with_mask_interpolators :425-517
        if (roi_at(a.roi, a.W, a.H, s.cx, s.cy)) {
            subset_stats<true>(a.ref, a.roi, a.W, a.H, s.cx, s.cy, r, 1, s.n_full, s.mean_full, s.norm_full);
            if (s.n_full >= min_required(circle_count(r, 1)) && s.norm_full > kSearchEps) {
                s.valid = 1;
                subset_stats<true>(a.ref, a.roi, a.W, a.H, s.cx, s.cy, r, cs, s.n_coarse, s.mean_coarse,
                                   s.norm_coarse);
                s.coarse_ok = s.n_coarse >= min_required(circle_count(r, cs)) && s.norm_coarse > kSearchEps;
            }
        }
    }
    if (s.valid) {
        s.min_u = r - s.cx;
        s.max_u = a.W - 1 - r - s.cx;
        s.min_v = r - s.cy;
        s.max_v = a.H - 1 - r - s.cy;
        // search_ncorr_multigrid :208-265, coarse pass
        if (s.coarse_ok && s.min_u <= s.max_u && s.min_v <= s.max_v) {
            Region g = empty_region();
            g.u0 = s.min_u;
            g.v0 = s.min_v;
            g.nu = (s.max_u - s.min_u) / cs + 1;
            g.nv = (s.max_v - s.min_v) / cs + 1;
            g.step = cs;
            g.coarse = 1;
            s.task[0] = g;
        }
    }
    a.st[i] = s;
}

__global__ void search_prep_kernel(SearchArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npoints) return;
    if (a.masked)
        prep_point<true>(a, i);
    else
        prep_point<false>(a, i);
}

__device__ __forceinline__ void better_of(double &z, int &idx, double oz, int oi)
{
    if (oz > z || (oz == z && oi < idx)) {
        z = oz;
        idx = oi;
    }
}

// Block-wide first maximum (largest z, then smallest scan index).
__device__ void block_best(double &z, int &idx)
{
    __shared__ double sz[kSearchBlock / 64];
    __shared__ int si[kSearchBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double oz = __shfl_xor(z, off);
        const int oi = __shfl_xor(idx, off);
        better_of(z, idx, oz, oi);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sz[wave] = z;
        si[wave] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) better_of(z, idx, sz[w], si[w]);
    __syncthreads();
}

// One thread per shift of task (task0 + blockIdx.z) of point blockIdx.y.
// Consecutive logical chunks run on one XCD (workgroups are dealt to the 8
// XCDs round robin), so an XCD's L2 holds one band of the deformed image.
template <bool MASKED>
__global__ __launch_bounds__(kSearchBlock) void search_shift_kernel(SearchArgs a, int task0)
{
    const int point = blockIdx.y;
    const int task = task0 + blockIdx.z;
    const int per = gridDim.x >> 3;
    const int chunk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const SearchState &s = a.st[point];
    const Region g = s.task[task];
    const int total = g.nu * g.nv;
    if (!s.valid || chunk * kSearchBlock >= total) return;
    const int i = chunk * kSearchBlock + threadIdx.x;
    double z = -INFINITY;
    int idx = INT_MAX;
    if (i < total) {
        const int iv = i / g.nu, iu = i - iv * g.nu;
        const int u = g.u0 + iu * g.step, v = g.v0 + iv * g.step;
        const int x = s.cx + u, y = s.cy + v, r = a.r;
        if (x - r >= 0 && y - r >= 0 && x + r < a.W && y + r < a.H) { // subset_in_bounds, :54-60
            const double q = g.coarse ? shift_zncc<MASKED>(a.ref, a.def, a.roi, a.W, a.H, s.cx, s.cy, u, v, r,
                                                           a.coarse_step, s.n_coarse, s.mean_coarse, s.norm_coarse)
                                      : shift_zncc<MASKED>(a.ref, a.def, a.roi, a.W, a.H, s.cx, s.cy, u, v, r, 1,
                                                           s.n_full, s.mean_full, s.norm_full);
            if (q > -INFINITY) { // a NaN never wins the reference's strict comparison
                z = q;
                idx = i;
            }
        }
    }
    block_best(z, idx);
    if (threadIdx.x == 0) {
        Partial p;
        p.z = z;
        p.idx = idx;
        p.pad = 0;
        a.part[((size_t)task * a.npoints + point) * a.max_chunks + chunk] = p;
    }
}

// First maximum over the chunks of one task of one point (whole block).
__device__ void reduce_task(const SearchArgs &a, const Region &g, int task, int point, double &z, int &idx)
{
    const int chunks = (g.nu * g.nv + kSearchBlock - 1) / kSearchBlock;
    z = -INFINITY;
    idx = INT_MAX;
    const Partial *p = a.part + ((size_t)task * a.npoints + point) * a.max_chunks;
    for (int c = threadIdx.x; c < chunks; c += blockDim.x) better_of(z, idx, p[c].z, p[c].idx);
    block_best(z, idx);
}

__device__ __forceinline__ void shift_of(const Region &g, int idx, int &u, int &v)
{
    const int iv = idx / g.nu, iu = idx - iv * g.nu;
    u = g.u0 + iu * g.step;
    v = g.v0 + iv * g.step;
}

__device__ double znssd_from_zncc(double z) // :22-28
{
    if (!isfinite(z)) return INFINITY;
    const double c = z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
    const double q = 2.0 * (1.0 - c);
    return q > 0.0 ? q : 0.0;
}

// stage 1: coarse winner -> refinement window + local window
// stage 2: refinement / local winners -> estimate (prefer_local_when_ambiguous, :267-280)
__global__ __launch_bounds__(kSearchBlock) void search_stage_kernel(SearchArgs a, int stage)
{
    const int point = blockIdx.x;
    SearchState &s = a.st[point];
    if (stage == 1) {
        if (!s.valid) return;
        double z;
        int idx;
        reduce_task(a, s.task[0], 0, point, z, idx);
        if (threadIdx.x == 0) {
            Region refine = empty_region();
            if (s.task[0].nu > 0 && z > -INFINITY) {
                int u, v;
                shift_of(s.task[0], idx, u, v);
                const int rr = max(a.refine_radius, a.coarse_step);
                refine = window(max(s.min_u, u - rr), min(s.max_u, u + rr), max(s.min_v, v - rr),
                                min(s.max_v, v + rr));
                s.stage_global_ok = 1;
            }
            s.task[0] = refine;
            s.task[1] = a.sr > 0 ? window(-a.sr, a.sr, -a.sr, a.sr) : empty_region();
        }
        return;
    }
    double zg = -INFINITY, zl = -INFINITY;
    int ig = INT_MAX, il = INT_MAX;
    if (s.valid) {
        reduce_task(a, s.task[0], 0, point, zg, ig);
        reduce_task(a, s.task[1], 1, point, zl, il);
    }
    if (threadIdx.x != 0) return;
    dic_estimate e = {};
    if (s.valid) {
        bool have = false;
        int u = 0, v = 0;
        double z = -INFINITY;
        if (a.around_prior) {
            if (zl > -INFINITY) {
                shift_of(s.task[1], il, u, v);
                z = zl;
                have = true;
            }
        } else {
            const bool g_ok = s.stage_global_ok && zg > -INFINITY;
            const bool l_ok = a.sr > 0 && zl > -INFINITY;
            if (l_ok && (!g_ok || zl + kLocalMargin >= zg)) {
                shift_of(s.task[1], il, u, v);
                z = zl;
                have = true;
            } else if (g_ok) {
                shift_of(s.task[0], ig, u, v);
                z = zg;
                have = true;
            }
        }
        if (have) {
            e.u = (double)u;
            e.v = (double)v;
            e.confidence = znssd_from_zncc(z);
            e.valid = 1;
            e.zncc = z;
            e.znssd = e.confidence;
        }
    }
    a.out[point] = e;
}

// ---------------------------------------------------------------- candidates

// Points of the ROI whose margin box fits the image (the bounds part of
// is_candidate_margin_valid, and the truncate branch's loop bounds, :220-240).
__device__ __forceinline__ bool point_flag(const uint8_t *roi, int W, int H, int x, int y, int m)
{
    return roi[(size_t)y * W + x] != 0 && (m < 0 || (x - m >= 0 && y - m >= 0 && x + m < W && y + m < H));
}

__global__ void count_rows_kernel(const uint8_t *roi, int W, int H, int m, int *cnt_margin, int *cnt_all)
{
    const int y = blockIdx.x;
    int a = 0, b = 0;
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
        a += point_flag(roi, W, H, x, y, m);
        b += point_flag(roi, W, H, x, y, -1);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    __shared__ int sa[16], sb[16];
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        sb[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            a += sa[w];
            b += sb[w];
        }
        cnt_margin[y] = a;
        cnt_all[y] = b;
    }
}

// Exclusive prefix sums of two length-H arrays in one 1024-thread block
// (offsets has H + 1 entries; the last is the total).
__global__ __launch_bounds__(1024) void scan_rows_kernel(const int *c0, const int *c1, int H, int *o0, int *o1)
{
    __shared__ long long s0[1024], s1[1024];
    const int t = threadIdx.x;
    const int seg = (H + 1023) / 1024;
    const int lo = min(H, t * seg), hi = min(H, lo + seg);
    long long a = 0, b = 0;
    for (int y = lo; y < hi; ++y) {
        a += c0[y];
        b += c1[y];
    }
    s0[t] = a;
    s1[t] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long x0 = t >= off ? s0[t - off] : 0, x1 = t >= off ? s1[t - off] : 0;
        __syncthreads();
        s0[t] += x0;
        s1[t] += x1;
        __syncthreads();
    }
    a = s0[t] - a;
    b = s1[t] - b;
    for (int y = lo; y < hi; ++y) {
        o0[y] = (int)a;
        o1[y] = (int)b;
        a += c0[y];
        b += c1[y];
    }
    if (t == 1023) {
        o0[H] = (int)s0[1023];
        o1[H] = (int)s1[1023];
    }
}

// Row-major compaction of the flagged points of row blockIdx.x.
__global__ __launch_bounds__(256) void compact_rows_kernel(const uint8_t *roi, int W, int H, int m,
                                                           const int *offsets, int2 *pts)
{
    const int y = blockIdx.x;
    __shared__ int wave_tot[4];
    int base = offsets[y];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int x0 = 0; x0 < W; x0 += 256) {
        const int x = x0 + threadIdx.x;
        const bool f = x < W && point_flag(roi, W, H, x, y, m);
        const unsigned long long bal = __ballot(f);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wave_tot[wave] = __popcll(bal);
        __syncthreads();
        int wbase = 0, tot = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) wbase += wave_tot[w];
            tot += wave_tot[w];
        }
        if (f) pts[base + wbase + before] = make_int2(x, y);
        base += tot;
        __syncthreads();
    }
}

// deterministic_sample (seed_selector.cpp:34-48).  With limit == 1 the
// reference's index is llround(0 * inf), past the end on x86-64, clamped to
// the last point.
__global__ void sample_points_kernel(const int2 *pts, int n, int limit, int ns, int2 *sp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    int j = i;
    if (ns < n) {
        if (limit == 1) {
            j = n - 1;
        } else {
            const double step = (double)(n - 1) / (double)(limit - 1);
            const long long k = llround((double)i * step);
            j = k > n - 1 ? n - 1 : (int)k;
        }
    }
    sp[i] = pts[j];
}

// SeedSelector::local_texture_std (:421-446)
__device__ double texture_std(const float *ref, int W, int H, int x, int y, int r)
{
    if (x - r < 0 || y - r < 0 || x + r >= W || y + r >= H) return 0.0;
    double sum = 0.0, sq = 0.0;
    int count = 0;
    for (int yy = y - r; yy <= y + r; ++yy)
        for (int xx = x - r; xx <= x + r; ++xx) {
            const double v = (double)ref[(size_t)yy * W + xx];
            sum += v;
            sq += v * v;
            ++count;
        }
    if (count == 0) return 0.0;
    const double mean = sum / (double)count;
    const double var = sq / (double)count - mean * mean;
    return sqrt(var > 0.0 ? var : 0.0);
}

struct SeedGeom {
    int W, H;
    int radius;      // max(integer subset radius, sub-pixel subset radius)
    int tex_radius;  // integer subset radius
    int search;      // integer search radius
    int truncate, min_valid_samples;
    double min_valid_sample_ratio;
    double min_texture;
};

// is_candidate_margin_valid (:385-419)
__device__ bool margin_valid(const SeedGeom &g, const uint8_t *roi, int x, int y)
{
    const int radius = g.radius;
    const int margin = radius + (g.truncate ? g.search : 0);
    if (x - margin < 0 || y - margin < 0 || x + margin >= g.W || y + margin >= g.H) return false;
    if (!g.truncate) return true;
    int valid = 0, full = 0;
    for (int yy = y - radius; yy <= y + radius; ++yy)
        for (int xx = x - radius; xx <= x + radius; ++xx) {
            const int dx = xx - x, dy = yy - y;
            if (dx * dx + dy * dy <= radius * radius) {
                ++full;
                valid += roi_at(roi, g.W, g.H, xx, yy);
            }
        }
    const int need = max(g.min_valid_samples, (int)ceil(g.min_valid_sample_ratio * (double)full));
    return valid >= max(6, need);
}

// requested == 1 (:259-276): texture of every margin-valid sampled point, -1 elsewhere
__global__ void texture_points_kernel(SeedGeom g, const float *ref, const uint8_t *roi, const int2 *sp, int ns,
                                      double *tex)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int2 p = sp[i];
    tex[i] = margin_valid(g, roi, p.x, p.y) ? texture_std(ref, g.W, g.H, p.x, p.y, g.tex_radius) : -1.0;
}

// first maximum of tex (entries < 0 excluded) -> out_xy[0..1], *count
__global__ __launch_bounds__(kSearchBlock) void best_texture_kernel(const double *tex, const int2 *sp, int ns,
                                                                    double *out_xy, int *count)
{
    double z = -INFINITY;
    int idx = INT_MAX;
    for (int i = threadIdx.x; i < ns; i += blockDim.x)
        if (tex[i] >= 0.0) better_of(z, idx, tex[i], i);
    block_best(z, idx);
    if (threadIdx.x == 0) {
        if (idx != INT_MAX) {
            out_xy[0] = (double)sp[idx].x;
            out_xy[1] = (double)sp[idx].y;
            *count = 1;
        } else {
            *count = 0;
        }
    }
}

// initialize_centers (:50-68)
__global__ void kmeans_init_kernel(const int2 *sp, int ns, int k, double2 *centers, unsigned long long *sums,
                                   int *counts)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const double step = k == 1 ? 0.0 : (double)(ns - 1) / (double)(k - 1);
    long long j = llround((double)i * step);
    if (j > ns - 1) j = ns - 1;
    centers[i] = make_double2((double)sp[j].x, (double)sp[j].y);
    sums[2 * i] = 0;
    sums[2 * i + 1] = 0;
    counts[i] = 0;
}

__device__ __forceinline__ double dist2(double x0, double y0, double x1, double y1) // :27-32
{
    const double dx = x0 - x1, dy = y0 - y1;
    return dx * dx + dy * dy;
}

// run_kmeans (:70-110), assignment: nearest centre (first on ties).  The
// coordinates are integers, so integer sums are the reference's double sums
// exactly, in any order: each workgroup sums its kKmeansPoints points per
// centre in LDS (int32 holds 1024 coordinates below 2^16) and adds them to the
// global 64-bit sums once, instead of every point contending for a few
// addresses.
constexpr int kKmeansPoints = 1024;

__global__ __launch_bounds__(256) void kmeans_assign_kernel(const int2 *sp, int ns, const double2 *centers, int k,
                                                            unsigned long long *sums, int *counts)
{
    extern __shared__ double2 c_lds[];
    int *acc = (int *)(c_lds + k); // per centre: sum x, sum y, count
    for (int i = threadIdx.x; i < k; i += blockDim.x) c_lds[i] = centers[i];
    for (int i = threadIdx.x; i < 3 * k; i += blockDim.x) acc[i] = 0;
    __syncthreads();
    const int first = blockIdx.x * kKmeansPoints;
    const int last = min(ns, first + kKmeansPoints);
    for (int p = first + threadIdx.x; p < last; p += blockDim.x) {
        const int2 q = sp[p];
        const double px = (double)q.x, py = (double)q.y;
        int best = 0;
        double bd = INFINITY;
        for (int i = 0; i < k; ++i) {
            const double d = dist2(px, py, c_lds[i].x, c_lds[i].y);
            if (d < bd) {
                bd = d;
                best = i;
            }
        }
        atomicAdd(&acc[3 * best], q.x);
        atomicAdd(&acc[3 * best + 1], q.y);
        atomicAdd(&acc[3 * best + 2], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        const int c = acc[3 * i + 2];
        if (c == 0) continue;
        atomicAdd(&sums[2 * i], (unsigned long long)acc[3 * i]);
        atomicAdd(&sums[2 * i + 1], (unsigned long long)acc[3 * i + 1]);
        atomicAdd(&counts[i], c);
    }
}

__global__ void kmeans_update_kernel(double2 *centers, int k, unsigned long long *sums, int *counts)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    if (counts[i] > 0)
        centers[i] = make_double2((double)sums[2 * i] / (double)counts[i], (double)sums[2 * i + 1] / (double)counts[i]);
    sums[2 * i] = 0;
    sums[2 * i + 1] = 0;
    counts[i] = 0;
}

// Candidate of centre blockIdx.x (:278-313): the rounded centre when it passes
// the ROI, margin and texture tests, else the most textured of the kNearest
// sampled points nearest to the centre (distance, then sampled order).
__global__ __launch_bounds__(kSearchBlock) void center_candidate_kernel(SeedGeom g, const float *ref,
                                                                        const uint8_t *roi, const double2 *centers,
                                                                        const int2 *sp, int ns, double *out_xy)
{
    const int c = blockIdx.x;
    const double cx = centers[c].x, cy = centers[c].y;
    __shared__ int ok;
    __shared__ int sel[kNearest];
    __shared__ double tex[kNearest];
    if (threadIdx.x == 0) {
        const int rx = (int)llround(cx), ry = (int)llround(cy);
        ok = roi_at(roi, g.W, g.H, rx, ry) && margin_valid(g, roi, rx, ry) &&
             texture_std(ref, g.W, g.H, rx, ry, g.tex_radius) >= g.min_texture;
        if (ok) {
            out_xy[2 * c] = (double)rx;
            out_xy[2 * c + 1] = (double)ry;
        }
    }
    __syncthreads();
    if (ok) return;
    const int lim = min(kNearest, ns);
    double last_d = -1.0;
    int last_i = -1;
    for (int round = 0; round < lim; ++round) {
        double z = -INFINITY; // maximise -distance: the nearest, then the smallest index
        int idx = INT_MAX;
        for (int i = threadIdx.x; i < ns; i += blockDim.x) {
            const double d = dist2((double)sp[i].x, (double)sp[i].y, cx, cy);
            if (d > last_d || (d == last_d && i > last_i)) better_of(z, idx, -d, i);
        }
        block_best(z, idx);
        __shared__ double bd;
        __shared__ int bi;
        if (threadIdx.x == 0) {
            bd = -z;
            bi = idx;
            sel[round] = idx;
        }
        __syncthreads();
        last_d = bd;
        last_i = bi;
        __syncthreads();
    }
    if ((int)threadIdx.x < lim) {
        const int2 p = sp[sel[threadIdx.x]];
        tex[threadIdx.x] = texture_std(ref, g.W, g.H, p.x, p.y, g.tex_radius);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double best = -1.0;
        int bi = sel[0];
        for (int i = 0; i < lim; ++i)
            if (tex[i] > best) {
                best = tex[i];
                bi = sel[i];
            }
        out_xy[2 * c] = (double)sp[bi].x;
        out_xy[2 * c + 1] = (double)sp[bi].y;
    }
}

__device__ int mirror_index(int i, int size) // padding.cpp:11-24
{
    if (size <= 1) return 0;
    while (i < 0 || i >= size) i = i < 0 ? -i - 1 : 2 * size - i - 1;
    return i;
}

__global__ void mirror_pad_kernel(const float *src, int W, int H, int pad, float *dst)
{
    const int PW = W + 2 * pad, PH = H + 2 * pad;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)PW * PH) return;
    const int y = (int)(i / PW), x = (int)(i - (long long)y * PW);
    dst[i] = src[(size_t)mirror_index(y - pad, H) * W + mirror_index(x - pad, W)];
}

// ---------------------------------------------------------------- host side

// Device memory owned for the duration of one call.
struct DevBuf {
    dic_ctx *ctx = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf()
    {
        if (p) ctx_free(ctx, p, bytes);
    }
    int alloc(dic_ctx *c, size_t b, const char *what)
    {
        ctx = c;
        bytes = b ? b : 1;
        const hipError_t e = ctx_malloc(c, &p, bytes);
        if (e == hipSuccess) return DIC_OK;
        p = nullptr;
        if (e == hipErrorOutOfMemory && c->budget)
            return set_error(c, DIC_ERR_BUDGET, std::string(what) + ": HBM budget exceeded");
        return check_hip(c, e, what);
    }
    template <class T> T *as() const { return (T *)p; }
};

int launch_check(dic_ctx *ctx, const char *what) { return check_hip(ctx, hipGetLastError(), what); }

int check_search_params(dic_ctx *ctx, const dic_integer_search_params *p)
{
    if (p->subset_radius < 1 || p->subset_radius > 500)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "integer search subset_radius must be in [1, 500]");
    if (p->search_radius < 0) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "search_radius must be >= 0");
    if (p->pyramid_scale < 1 || p->pyramid_refinement_radius < 1)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "pyramid_scale and pyramid_refinement_radius must be >= 1");
    if (p->mode != DIC_SEARCH_GLOBAL && p->mode != DIC_SEARCH_AROUND_PRIOR)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown integer search mode");
    return DIC_OK;
}

// The search over n points whose coordinates (and priors) are in HBM; the
// estimates are written to out_dev.  roi_dev is required when masked.
int run_integer_search(dic_ctx *ctx, const dic_integer_search_params *p, const dic_image *ref,
                       const dic_image *def, const uint8_t *roi_dev, const double *xy_dev, const double *prior_dev,
                       int n, dic_estimate *out_dev)
{
    hipStream_t s = ctx->stream;
    if (n == 0) return DIC_OK;
    const int W = ref->width, H = ref->height;
    SearchArgs a = {};
    a.ref = ref->data;
    a.def = def->data;
    a.roi = roi_dev;
    a.W = W;
    a.H = H;
    a.r = p->subset_radius;
    a.sr = p->search_radius;
    a.coarse_step = p->pyramid_enabled ? p->pyramid_scale : 1;
    a.refine_radius = std::max(p->pyramid_refinement_radius, (int)std::ceil(1.5 * (double)p->subset_radius));
    a.masked = p->masked ? 1 : 0;
    a.around_prior = p->mode == DIC_SEARCH_AROUND_PRIOR;
    a.xy = xy_dev;
    a.prior = prior_dev;
    a.npoints = n;
    a.out = out_dev;
    if (n > 65535) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "at most 65535 points per integer search");
    // shifts per task: the coarse grid spans the whole image; the windows are bounded by their radii
    const long long span_u = std::max(0, W - 2 * a.r), span_v = std::max(0, H - 2 * a.r);
    const long long coarse = ((span_u + a.coarse_step - 1) / a.coarse_step) * ((span_v + a.coarse_step - 1) / a.coarse_step);
    const long long rr = std::max(a.refine_radius, a.coarse_step);
    const long long refine = (2 * rr + 1) * (2 * rr + 1), local = (2LL * a.sr + 1) * (2LL * a.sr + 1);
    const long long most = a.around_prior ? local : std::max(coarse, std::max(refine, local));
    const long long chunks = (most + kSearchBlock - 1) / kSearchBlock;
    if (chunks > (1LL << 24)) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "integer search region too large");
    a.max_chunks = (int)std::max(1LL, chunks);
    DevBuf st, part;
    int rc = st.alloc(ctx, sizeof(SearchState) * (size_t)n, "search state");
    if (rc == DIC_OK) rc = part.alloc(ctx, sizeof(Partial) * 2 * (size_t)n * a.max_chunks, "search partials");
    if (rc != DIC_OK) return rc;
    a.st = st.as<SearchState>();
    a.part = part.as<Partial>();
    search_prep_kernel<<<(n + 63) / 64, 64, 0, s>>>(a);
    if ((rc = launch_check(ctx, "search prep")) != DIC_OK) return rc;
    auto shifts = [&](int task0, int ntasks, long long count) {
        const long long c = (count + kSearchBlock - 1) / kSearchBlock;
        const unsigned gx = (unsigned)(((std::max(1LL, c) + 7) / 8) * 8); // a multiple of the 8 XCDs
        dim3 grid(gx, (unsigned)n, (unsigned)ntasks);
        if (a.masked)
            search_shift_kernel<true><<<grid, kSearchBlock, 0, s>>>(a, task0);
        else
            search_shift_kernel<false><<<grid, kSearchBlock, 0, s>>>(a, task0);
        return launch_check(ctx, "search shifts");
    };
    if (!a.around_prior) {
        if ((rc = shifts(0, 1, coarse)) != DIC_OK) return rc;
        search_stage_kernel<<<n, kSearchBlock, 0, s>>>(a, 1);
        if ((rc = launch_check(ctx, "search stage")) != DIC_OK) return rc;
        if ((rc = shifts(0, 2, std::max(refine, local))) != DIC_OK) return rc;
    } else {
        if ((rc = shifts(1, 1, local)) != DIC_OK) return rc;
    }
    search_stage_kernel<<<n, kSearchBlock, 0, s>>>(a, 2);
    if ((rc = launch_check(ctx, "search final")) != DIC_OK) return rc;
    // the temporaries are freed on return: wait for the kernels that use them
    return check_hip(ctx, hipStreamSynchronize(s), "integer search");
}

int check_seed_params(dic_ctx *ctx, const dic_seed_params *p)
{
    dic_integer_search_params q = p->search;
    q.mode = DIC_SEARCH_GLOBAL;
    int rc = check_search_params(ctx, &q);
    if (rc != DIC_OK) return rc;
    if (p->seed_count < 1 || p->seed_count > kMaxSeeds)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "seed_count must be in [1, 2048]");
    if (p->subpixel_radius < 1) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "subpixel subset_radius must be >= 1");
    if (p->kmeans_sample_limit < 1) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "kmeans_sample_limit must be >= 1");
    if (p->subpixel_shape_order != 1 && p->subpixel_shape_order != 2)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "subpixel shape order must be 1 or 2");
    if (p->degree != 1 && p->degree != 3 && p->degree != 5)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "degree must be 1, 3 or 5");
    if (p->quality_metric < DIC_QUALITY_ZNCC || p->quality_metric > DIC_QUALITY_SSD)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown quality metric");
    return DIC_OK;
}

int upload(dic_ctx *ctx, DevBuf &b, const void *host, size_t bytes, const char *what)
{
    int rc = b.alloc(ctx, bytes, what);
    if (rc == DIC_OK) rc = check_hip(ctx, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream), what);
    return rc;
}

// generate_candidates (:208-316) into xy_dev (room for seed_count pairs); *count on the host.
int run_candidates(dic_ctx *ctx, const dic_seed_params *p, const dic_image *ref, const uint8_t *roi_dev,
                   double *xy_dev, int *count)
{
    hipStream_t s = ctx->stream;
    const int W = ref->width, H = ref->height;
    *count = 0;
    SeedGeom g;
    g.W = W;
    g.H = H;
    g.radius = std::max(p->search.subset_radius, p->subpixel_radius);
    g.tex_radius = p->search.subset_radius;
    g.search = p->search.search_radius;
    g.truncate = p->truncate_roi_subsets ? 1 : 0;
    g.min_valid_samples = p->min_valid_samples;
    g.min_valid_sample_ratio = p->min_valid_sample_ratio;
    g.min_texture = p->min_texture_std;
    const int m = g.radius + (g.truncate ? g.search : 0);
    DevBuf rows;
    int rc = rows.alloc(ctx, sizeof(int) * (4 * (size_t)H + 2), "row counts");
    if (rc != DIC_OK) return rc;
    int *cnt_m = rows.as<int>(), *cnt_a = cnt_m + H, *off_m = cnt_a + H, *off_a = off_m + H + 1;
    count_rows_kernel<<<H, 256, 0, s>>>(roi_dev, W, H, m, cnt_m, cnt_a);
    scan_rows_kernel<<<1, 1024, 0, s>>>(cnt_m, cnt_a, H, off_m, off_a);
    if ((rc = launch_check(ctx, "roi point count")) != DIC_OK) return rc;
    int tot[2] = {0, 0};
    rc = check_hip(ctx, hipMemcpyAsync(&tot[0], off_m + H, sizeof(int), hipMemcpyDeviceToHost, s), "count");
    if (rc == DIC_OK) rc = check_hip(ctx, hipMemcpyAsync(&tot[1], off_a + H, sizeof(int), hipMemcpyDeviceToHost, s), "count");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(s), "count");
    if (rc != DIC_OK) return rc;
    const bool use_margin = tot[0] > 0; // else every ROI point (:242-251)
    const int n = use_margin ? tot[0] : tot[1];
    if (n == 0) return DIC_OK;
    DevBuf pts;
    if ((rc = pts.alloc(ctx, sizeof(int2) * (size_t)n, "roi points")) != DIC_OK) return rc;
    compact_rows_kernel<<<H, 256, 0, s>>>(roi_dev, W, H, use_margin ? m : -1, use_margin ? off_m : off_a,
                                          pts.as<int2>());
    if ((rc = launch_check(ctx, "roi point compaction")) != DIC_OK) return rc;
    const int limit = p->kmeans_sample_limit;
    const int ns = n <= limit ? n : limit;
    DevBuf spb;
    if ((rc = spb.alloc(ctx, sizeof(int2) * (size_t)ns, "sampled points")) != DIC_OK) return rc;
    int2 *sp = spb.as<int2>();
    sample_points_kernel<<<(ns + 255) / 256, 256, 0, s>>>(pts.as<int2>(), n, limit, ns, sp);
    if ((rc = launch_check(ctx, "point sampling")) != DIC_OK) return rc;
    if (p->seed_count == 1) {
        DevBuf tex, cnt;
        if ((rc = tex.alloc(ctx, sizeof(double) * (size_t)ns, "texture")) != DIC_OK) return rc;
        if ((rc = cnt.alloc(ctx, sizeof(int), "count")) != DIC_OK) return rc;
        texture_points_kernel<<<(ns + 255) / 256, 256, 0, s>>>(g, ref->data, roi_dev, sp, ns, tex.as<double>());
        best_texture_kernel<<<1, kSearchBlock, 0, s>>>(tex.as<double>(), sp, ns, xy_dev, cnt.as<int>());
        if ((rc = launch_check(ctx, "texture candidate")) != DIC_OK) return rc;
        rc = check_hip(ctx, hipMemcpyAsync(count, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s), "count");
        if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(s), "texture candidate");
        return rc;
    }
    const int k = std::min(p->seed_count, ns);
    DevBuf cen, sums, cnts;
    if ((rc = cen.alloc(ctx, sizeof(double2) * (size_t)k, "centres")) != DIC_OK) return rc;
    if ((rc = sums.alloc(ctx, sizeof(unsigned long long) * 2 * (size_t)k, "centre sums")) != DIC_OK) return rc;
    if ((rc = cnts.alloc(ctx, sizeof(int) * (size_t)k, "centre counts")) != DIC_OK) return rc;
    kmeans_init_kernel<<<(k + 255) / 256, 256, 0, s>>>(sp, ns, k, cen.as<double2>(), sums.as<unsigned long long>(),
                                                      cnts.as<int>());
    for (int it = 0; it < std::max(0, p->kmeans_iterations); ++it) {
        kmeans_assign_kernel<<<(ns + kKmeansPoints - 1) / kKmeansPoints, 256, (sizeof(double2) + 12) * (size_t)k,
                               s>>>(sp, ns, cen.as<double2>(), k, sums.as<unsigned long long>(), cnts.as<int>());
        kmeans_update_kernel<<<(k + 255) / 256, 256, 0, s>>>(cen.as<double2>(), k, sums.as<unsigned long long>(),
                                                            cnts.as<int>());
    }
    if ((rc = launch_check(ctx, "k-means")) != DIC_OK) return rc;
    center_candidate_kernel<<<k, kSearchBlock, 0, s>>>(g, ref->data, roi_dev, cen.as<double2>(), sp, ns, xy_dev);
    if ((rc = launch_check(ctx, "centre candidates")) != DIC_OK) return rc;
    rc = check_hip(ctx, hipStreamSynchronize(s), "candidates");
    if (rc == DIC_OK) *count = k;
    return rc;
}

int check_pair(dic_ctx *ctx, const dic_image *ref, const dic_image *def)
{
    if (!ref || !ref->data || ref->width <= 0 || ref->height <= 0)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "reference image required");
    if (def && !def->data) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "deformed image has no data");
    return DIC_OK;
}

double zncc_from_znssd(double q) // subset_initializer.cpp:13-19
{
    if (!std::isfinite(q)) return -1.0;
    const double z = 1.0 - 0.5 * q;
    return z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
}

bool quality_passes(const dic_seed_params *p, double q) // seed_selector.cpp:448-460
{
    if (p->quality_metric == DIC_QUALITY_ZNCC) return q >= p->min_zncc;
    if (p->quality_metric == DIC_QUALITY_SSD) return q <= p->max_ssd;
    return q <= p->max_znssd;
}

bool seed_is_better(const dic_seed_evaluation &a, const dic_seed_evaluation &cur, int metric) // :112-136
{
    const double eps = 1.0e-9;
    if (!cur.valid) return true;
    if (metric == DIC_QUALITY_ZNCC) {
        if (a.quality > cur.quality + eps) return true;
        if (a.quality < cur.quality - eps) return false;
    } else {
        if (a.quality < cur.quality - eps) return true;
        if (a.quality > cur.quality + eps) return false;
    }
    return a.displacement_norm > cur.displacement_norm;
}

} // namespace

extern "C" int dic_integer_search(dic_ctx *ctx, const dic_integer_search_params *params, const dic_image *ref,
                                  const dic_image *def, const uint8_t *roi, const double *xy, const double *prior,
                                  int n, dic_estimate *out)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    if (!params || n < 0 || (n > 0 && (!xy || !out)))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null argument or negative point count");
    int rc = check_search_params(ctx, params);
    if (rc == DIC_OK) rc = check_pair(ctx, ref, def);
    if (rc != DIC_OK) return rc;
    if (!def) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "deformed image required");
    if (params->mode == DIC_SEARCH_AROUND_PRIOR && n > 0 && !prior)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "around-prior search needs prior_uv");
    if (params->masked && !roi) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "masked search needs the roi");
    if (n == 0) return DIC_OK;
    if (ref->width != def->width || ref->height != def->height) { // :309-314
        for (int i = 0; i < n; ++i) out[i] = dic_estimate{};
        return DIC_OK;
    }
    const size_t px = (size_t)ref->width * ref->height;
    DevBuf xyb, prb, roib, outb;
    rc = upload(ctx, xyb, xy, sizeof(double) * 2 * (size_t)n, "point upload");
    if (rc == DIC_OK && prior && params->mode == DIC_SEARCH_AROUND_PRIOR)
        rc = upload(ctx, prb, prior, sizeof(double) * 2 * (size_t)n, "prior upload");
    if (rc == DIC_OK && params->masked) rc = upload(ctx, roib, roi, px, "roi upload");
    if (rc == DIC_OK) rc = outb.alloc(ctx, sizeof(dic_estimate) * (size_t)n, "estimates");
    if (rc == DIC_OK)
        rc = run_integer_search(ctx, params, ref, def, roib.as<uint8_t>(), xyb.as<double>(), prb.as<double>(), n,
                                outb.as<dic_estimate>());
    if (rc == DIC_OK)
        rc = check_hip(ctx, hipMemcpyAsync(out, outb.p, sizeof(dic_estimate) * (size_t)n, hipMemcpyDeviceToHost,
                                           ctx->stream), "estimate download");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "integer search");
    return rc;
}

extern "C" int dic_seed_candidates(dic_ctx *ctx, const dic_seed_params *params, const dic_image *ref,
                                   const uint8_t *roi, double *xy_out, int capacity, int *n_out)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    if (!params || !roi || !n_out || capacity < 0 || (capacity > 0 && !xy_out))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null argument");
    *n_out = 0;
    int rc = check_seed_params(ctx, params);
    if (rc == DIC_OK) rc = check_pair(ctx, ref, nullptr);
    if (rc != DIC_OK) return rc;
    DevBuf roib, xyb;
    rc = upload(ctx, roib, roi, (size_t)ref->width * ref->height, "roi upload");
    if (rc == DIC_OK) rc = xyb.alloc(ctx, sizeof(double) * 2 * (size_t)params->seed_count, "candidates");
    int k = 0;
    if (rc == DIC_OK) rc = run_candidates(ctx, params, ref, roib.as<uint8_t>(), xyb.as<double>(), &k);
    k = std::min(k, capacity);
    if (rc == DIC_OK && k > 0)
        rc = check_hip(ctx, hipMemcpyAsync(xy_out, xyb.p, sizeof(double) * 2 * (size_t)k, hipMemcpyDeviceToHost,
                                           ctx->stream), "candidate download");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "candidate download");
    if (rc == DIC_OK) *n_out = k;
    return rc;
}

extern "C" int dic_select_seed(dic_ctx *ctx, const dic_seed_params *params, const dic_image *ref,
                               const dic_image *def, const uint8_t *roi, dic_seed_evaluation *evals, int capacity,
                               int *n_out, int *best_index)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    if (!params || !roi || !n_out || !best_index || capacity < 0 || (capacity > 0 && !evals))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null argument");
    *n_out = 0;
    *best_index = -1;
    int rc = check_seed_params(ctx, params);
    if (rc == DIC_OK) rc = check_pair(ctx, ref, def);
    if (rc != DIC_OK) return rc;
    if (!def) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "deformed image required");
    if (ref->width != def->width || ref->height != def->height) return DIC_OK; // :151-157: not found
    hipStream_t s = ctx->stream;
    const int W = ref->width, H = ref->height;
    DevBuf roib, xyb, estb;
    rc = upload(ctx, roib, roi, (size_t)W * H, "roi upload");
    if (rc == DIC_OK) rc = xyb.alloc(ctx, sizeof(double) * 2 * (size_t)params->seed_count, "candidates");
    int k = 0;
    if (rc == DIC_OK) rc = run_candidates(ctx, params, ref, roib.as<uint8_t>(), xyb.as<double>(), &k);
    if (rc != DIC_OK) return rc;
    k = std::min(k, capacity);
    if (k == 0) return DIC_OK;
    dic_integer_search_params sp = params->search;
    sp.mode = DIC_SEARCH_GLOBAL;
    sp.masked = params->truncate_roi_subsets ? 1 : 0;
    if ((rc = estb.alloc(ctx, sizeof(dic_estimate) * (size_t)k, "estimates")) != DIC_OK) return rc;
    rc = run_integer_search(ctx, &sp, ref, def, roib.as<uint8_t>(), xyb.as<double>(), nullptr, k,
                            estb.as<dic_estimate>());
    std::vector<double> xy(2 * (size_t)k);
    std::vector<dic_estimate> est(k);
    if (rc == DIC_OK)
        rc = check_hip(ctx, hipMemcpyAsync(xy.data(), xyb.p, sizeof(double) * 2 * (size_t)k, hipMemcpyDeviceToHost, s),
                       "candidate download");
    if (rc == DIC_OK)
        rc = check_hip(ctx, hipMemcpyAsync(est.data(), estb.p, sizeof(dic_estimate) * (size_t)k, hipMemcpyDeviceToHost,
                                           s), "estimate download");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(s), "integer search");
    if (rc != DIC_OK) return rc;
    if (params->subpixel_enabled) {
        // refine_initial_with[_mask]_interpolators (subset_initializer.cpp:171-264)
        std::vector<int> which;
        std::vector<double> pxy;
        std::vector<dic_initial> ini;
        for (int i = 0; i < k; ++i)
            if (est[i].valid) {
                which.push_back(i);
                pxy.push_back(xy[2 * i]);
                pxy.push_back(xy[2 * i + 1]);
                ini.push_back(dic_initial{est[i].u, est[i].v, est[i].du_dx, est[i].du_dy, est[i].dv_dx, est[i].dv_dy,
                                          est[i].confidence});
            }
        if (!which.empty()) {
            const bool fgn = params->subpixel_optimizer == DIC_OPTIMIZER_FGN;
            dic_icgn_params ip = {};
            ip.subset_radius = params->subpixel_radius;
            ip.max_iterations = std::max(0, params->subpixel_max_iterations);
            ip.convergence_threshold = params->subpixel_convergence_threshold;
            ip.shape_order = params->subpixel_shape_order;
            ip.objective = params->subpixel_objective;
            ip.optimizer = fgn ? DIC_OPTIMIZER_FGN : DIC_OPTIMIZER_ICGN;
            ip.precision = DIC_PRECISION_FP64;
            ip.masked = (!fgn && params->truncate_roi_subsets) ? 1 : 0; // FGN's seed path is unmasked (:212-214)
            ip.min_valid_samples = params->min_valid_samples;
            ip.min_valid_sample_ratio = params->min_valid_sample_ratio;
            dic_coeffs *co = nullptr;
            dic_grad *gr = nullptr;
            // IC-GN: compute_lazy with use_exact_prefilter = false (seed_selector.cpp:163-173), reference
            // gradients by central differences (icgn.cpp:42-71); FGN: the configured spline (forward_gauss_newton.cpp:98)
            rc = dic_bspline_prefilter(ctx, def, params->degree, params->border, fgn ? 1 : 0, DIC_PRECISION_FP64, &co);
            if (rc == DIC_OK && !fgn)
                rc = dic_ref_gradient(ctx, ref, DIC_GRADIENT_CENTRAL, params->degree, params->border,
                                      DIC_PRECISION_FP64, &gr);
            std::vector<dic_result> res(which.size());
            if (rc == DIC_OK)
                rc = dic_icgn_solve_batch(ctx, &ip, ref, gr, co, ip.masked ? roi : nullptr, pxy.data(), ini.data(),
                                          (int)which.size(), res.data());
            dic_grad_free(gr);
            dic_coeffs_free(co);
            if (rc != DIC_OK) return rc;
            for (size_t j = 0; j < which.size(); ++j) {
                const dic_result &r = res[j];
                if (!(r.valid && r.status == DIC_SUCCESS)) continue;
                dic_estimate &e = est[which[j]];
                e.u = r.p[0];
                e.v = r.p[1];
                e.du_dx = r.p[2];
                e.du_dy = r.p[3];
                e.dv_dx = r.p[4];
                e.dv_dy = r.p[5];
                e.confidence = r.correlation;
                e.zncc = zncc_from_znssd(r.correlation);
                e.znssd = r.correlation;
            }
        }
    }
    // evaluate_candidate (:466-539) and the selection loop (:188-201)
    int best = -1;
    for (int i = 0; i < k; ++i) {
        dic_seed_evaluation &e = evals[i];
        e = dic_seed_evaluation{};
        e.x = xy[2 * i];
        e.y = xy[2 * i + 1];
        e.displacement.status = DIC_INVALID_INPUT;
        const dic_estimate &q = est[i];
        if (q.valid && std::isfinite(q.u) && std::isfinite(q.v) && std::isfinite(q.confidence)) {
            e.displacement.x = e.x;
            e.displacement.y = e.y;
            e.displacement.p[0] = q.u;
            e.displacement.p[1] = q.v;
            e.displacement.p[2] = q.du_dx;
            e.displacement.p[3] = q.du_dy;
            e.displacement.p[4] = q.dv_dx;
            e.displacement.p[5] = q.dv_dy;
            e.displacement.correlation = q.confidence;
            e.displacement.status = DIC_SUCCESS;
            e.displacement.valid = 1;
            e.quality = params->quality_metric == DIC_QUALITY_ZNCC ? q.zncc : q.znssd;
            e.displacement_norm = std::hypot(q.u, q.v);
            e.quality_passed = quality_passes(params, e.quality);
            e.valid = 1;
        }
        if (e.valid && e.quality_passed && e.displacement_norm >= params->min_displacement_norm)
            if (best < 0 || seed_is_better(e, evals[best], params->quality_metric)) best = i;
    }
    *n_out = k;
    *best_index = best;
    return DIC_OK;
}

extern "C" int dic_image_mirror_pad(dic_ctx *ctx, const dic_image *img, int pad, dic_image **out)
{
    if (!ctx || !out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!img || !img->data) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "image required");
    if (pad < 0) pad = 0;
    const int PW = img->width + 2 * pad, PH = img->height + 2 * pad;
    dic_image *o = new (std::nothrow) dic_image();
    if (!o) return set_error(ctx, DIC_ERR_HIP, "host allocation failed");
    o->ctx = ctx;
    o->width = PW;
    o->height = PH;
    o->bytes = sizeof(float) * (size_t)PW * PH;
    o->owned = true;
    hipError_t e = ctx_malloc(ctx, (void **)&o->data, o->bytes);
    if (e != hipSuccess) {
        delete o;
        return e == hipErrorOutOfMemory && ctx->budget ? set_error(ctx, DIC_ERR_BUDGET, "padded image: HBM budget exceeded")
                                                       : check_hip(ctx, e, "padded image allocation");
    }
    const long long total = (long long)PW * PH;
    mirror_pad_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(img->data, img->width, img->height,
                                                                                pad, o->data);
    int rc = launch_check(ctx, "mirror pad");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "mirror pad");
    if (rc != DIC_OK) {
        dic_image_free(o);
        return rc;
    }
    *out = o;
    return DIC_OK;
}
