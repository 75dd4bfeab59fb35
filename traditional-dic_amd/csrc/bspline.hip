// bspline.hip -- B-spline coefficient prefilter, reference gradients and the
// sampling test hook for gfx950.
//
// Prefilter (replaces form_coefficients + deconvolve_rows/columns,
// src/interpolation/bspline.cpp:177-244, :369-392).  The reference pads the
// image by b = max(3, border) with half-sample-symmetric indices and divides
// the DFT of every padded line by the DFT of the sampled kernel
// [beta(0), beta(1), beta(2), 0, ..., beta(2), beta(1)], i.e. it applies the
// exact inverse of a circulant matrix on a line of length L = N + 2b.  That
// inverse factors into first-order causal/anticausal recursions per pole z
// (|z| < 1) times the gain prod (1 - z)^2.  On a periodic line the recursion
// y[k] = s[k] + z y[k-1] started from y = 0 at index k0 - T reaches the exact
// periodic solution to within |z|^T; T is chosen so that every pole is below
// 1e-18, which makes each output chunk independent: every thread starts T
// samples early and needs no scan or carry exchange.
//
// Layout in HBM: padded coefficients, row-major, pitch = W + 2b, FP64
// (plain) or FP32 (mixed) storage.  Arithmetic is FP64 in both cases.
//   rows pass:    one workgroup per padded row, the line lives in LDS
//                 (coalesced global load/store, recursion in LDS);
//   column pass:  one thread per (column, 64-row chunk), so each step of
//                 the recursion is a coalesced row segment load.
#include <cmath>

#include "bspline_device.hpp"
#include "dic_internal.hpp"

namespace dic {
namespace {

__device__ __forceinline__ int sym_index(int v, int size)
{
    const int period = 2 * size;
    int w = v % period;
    if (w < 0) w += period;
    return w < size ? w : period - 1 - w;
}

__device__ __forceinline__ int clamp_index(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ int wrap(int k, int L)
{
    k %= L;
    return k < 0 ? k + L : k;
}
// the next and the previous index of a period-L sequence (one compare, no
// division per step of the recursive filters)
__device__ __forceinline__ int wrap_next(int kk, int L) { return kk + 1 == L ? 0 : kk + 1; }
__device__ __forceinline__ int wrap_prev(int kk, int L) { return kk == 0 ? L - 1 : kk - 1; }

struct PoleArgs {
    int count;
    double z0, z1, gain;
    int warmup;
};

PoleArgs make_poles(int degree)
{
    PoleArgs p{0, 0.0, 0.0, 1.0, 0};
    if (degree == 3) {
        p.count = 1;
        p.z0 = std::sqrt(3.0) - 2.0;
    } else if (degree == 5) {
        const double w1 = -13.0 + std::sqrt(105.0), w2 = -13.0 - std::sqrt(105.0);
        p.count = 2;
        p.z0 = 0.5 * (w1 + std::sqrt(w1 * w1 - 4.0));
        p.z1 = 0.5 * (w2 + std::sqrt(w2 * w2 - 4.0));
    }
    double g = 1.0;
    int T = 0;
    const double zz[2] = {p.z0, p.z1};
    for (int i = 0; i < p.count; ++i) {
        g *= (1.0 - zz[i]) * (1.0 - zz[i]);
        T += (int)std::ceil(std::log(1e-18) / std::log(std::fabs(zz[i])));
    }
    p.gain = g;
    p.warmup = T + 2;
    return p;
}

template <typename OUT>
__global__ void pad_kernel(const float *__restrict__ img, int W, int H, int b, int symmetric, OUT *__restrict__ out)
{
    const int PW = W + 2 * b, PH = H + 2 * b;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= PW || y >= PH) return;
    const int sx = symmetric ? sym_index(x - b, W) : clamp_index(x - b, 0, W - 1);
    const int sy = symmetric ? sym_index(y - b, H) : clamp_index(y - b, 0, H - 1);
    out[(size_t)y * PW + x] = (OUT)img[(size_t)sy * W + sx];
}

// One recursion of the FP64 pass: y1 = v + z0*y1, y2 = y1 + z1*y2 (z1 = 0 for
// a single pole), returning the last stage.
struct PoleChain {
    double y1 = 0.0, y2 = 0.0;
    __device__ __forceinline__ double step(double v, const PoleArgs &p)
    {
        y1 = v + p.z0 * y1;
        y2 = y1 + p.z1 * y2;
        return (p.count == 2) ? y2 : y1;
    }
};

// samples per block of loads issued ahead of the serial chain
constexpr int kBlock = 16;

// Runs one recursion over n_warm + n_out consecutive samples of a period-L
// sequence, starting at periodic index kk and stepping forward (FWD) or
// backward; the first n_warm steps only warm the chain, the rest store their
// result through st(index, value).  Whole blocks of kBlock samples issue
// every load before the block's chain (the same operations in the same order
// as one sample at a time); the remainder goes one sample at a time.
template <bool FWD, typename Load, typename Store>
__device__ __forceinline__ void sweep(int kk, int L, int n_warm, int n_out, const PoleArgs &p, Load ld, Store st)
{
    PoleChain c;
    auto adv = [L](int q) { return FWD ? wrap_next(q, L) : wrap_prev(q, L); };
    int left = n_warm;
    for (; left >= kBlock; left -= kBlock) {
        double v[kBlock];
#pragma unroll
        for (int i = 0; i < kBlock; ++i) {
            v[i] = ld(kk);
            kk = adv(kk);
        }
#pragma unroll
        for (int i = 0; i < kBlock; ++i) c.step(v[i], p);
    }
    for (; left > 0; --left) {
        c.step(ld(kk), p);
        kk = adv(kk);
    }
    left = n_out;
    for (; left >= kBlock; left -= kBlock) {
        double v[kBlock];
        int at[kBlock];
#pragma unroll
        for (int i = 0; i < kBlock; ++i) {
            v[i] = ld(kk);
            at[i] = kk;
            kk = adv(kk);
        }
#pragma unroll
        for (int i = 0; i < kBlock; ++i) st(at[i], c.step(v[i], p));
    }
    for (; left > 0; --left) {
        st(kk, c.step(ld(kk), p));
        kk = adv(kk);
    }
}

// One workgroup per padded row; the symmetric-padded line is built in LDS,
// filtered by both recursions, and written as FP64.
__global__ void prefilter_rows_kernel(const float *__restrict__ img, int W, int H, int b, PoleArgs poles,
                                      double *__restrict__ out)
{
    extern __shared__ double lds[];
    const int PW = W + 2 * b;
    double *line = lds;
    double *buf = lds + PW;
    const int y = blockIdx.x;
    const int sy = sym_index(y - b, H);
    for (int x = threadIdx.x; x < PW; x += blockDim.x) line[x] = (double)img[(size_t)sy * W + sym_index(x - b, W)];
    __syncthreads();

    const int chunk = (PW + blockDim.x - 1) / blockDim.x;
    const int c0 = threadIdx.x * chunk;
    const int c1 = min(c0 + chunk, PW);
    const int T = poles.warmup;
    if (c0 < c1)
        sweep<true>(wrap(c0 - T, PW), PW, T, c1 - c0, poles, [&](int q) { return line[q]; },
                    [&](int q, double v) { buf[q] = v; });
    __syncthreads();
    if (c0 < c1)
        sweep<false>(wrap(c1 - 1 + T, PW), PW, T, c1 - c0, poles, [&](int q) { return buf[q]; },
                     [&](int q, double v) { line[q] = poles.gain * v; });
    __syncthreads();
    for (int x = threadIdx.x; x < PW; x += blockDim.x) out[(size_t)y * PW + x] = line[x];
}

constexpr int kColChunk = 64;

// one thread per (column, kColChunk-row chunk), each starting its chain
// poles.warmup rows early
__global__ void prefilter_cols_causal_kernel(const double *__restrict__ in, int PW, int PH, PoleArgs poles,
                                             double *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= PW) return;
    const int r0 = blockIdx.y * kColChunk;
    const int r1 = min(r0 + kColChunk, PH);
    sweep<true>(wrap(r0 - poles.warmup, PH), PH, poles.warmup, r1 - r0, poles,
                [&](int q) { return in[(size_t)q * PW + x]; }, [&](int q, double v) { out[(size_t)q * PW + x] = v; });
}

__global__ void prefilter_cols_anticausal_kernel(const double *__restrict__ in, int PW, int PH, PoleArgs poles,
                                                 double *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= PW) return;
    const int r0 = blockIdx.y * kColChunk;
    const int r1 = min(r0 + kColChunk, PH);
    sweep<false>(wrap(r1 - 1 + poles.warmup, PH), PH, poles.warmup, r1 - r0, poles,
                 [&](int q) { return in[(size_t)q * PW + x]; },
                 [&](int q, double v) { out[(size_t)q * PW + x] = poles.gain * v; });
}

// ---- FP32 storage (the mixed path): two passes with an FP32 intermediate ----
// The recursions run in FP64; every chunk starts kF32Warm steps early, which is
// exact to |z|^28 (6e-11 for the quintic's larger pole, 1e-16 for the cubic
// pole), below the FP32 rounding of the stored coefficients.  Both passes load
// a thread's whole window (warm-up included) before the recursion starts, so
// the serial chain waits on no memory.  A single pole has z1 = 0, and then the
// second stage (y2 = y1 + 0 * y2) passes y1 through unchanged.  Each padded
// pixel moves 16 bytes of HBM (image read, intermediate write and read,
// coefficient write) against 8 compulsory.
constexpr int kF32Warm = 28;
constexpr int kF32ColChunk = 32;

// Rows: one workgroup per padded row, one thread per CH samples.  The row
// lives in LDS as floats in an extended index e = j + kF32Warm, j the
// period-PW sample index (a line with the kF32Warm samples before sample 0 and
// CH after the last, the causal intermediate with CH + kF32Warm after it), so
// every window is a run of consecutive e with no wrap test.  Slot e + e / CH
// keeps one spare float per CH entries: with CH even the lanes' windows start
// an odd stride apart, and lanes reading the same step hit distinct banks.
template <int CH>
__device__ __forceinline__ int ext_slot(int e)
{
    return e + e / CH;
}

template <int CH>
__host__ __device__ constexpr int f32_line_len(int PW)
{
    return PW + CH + kF32Warm;
}
template <int CH>
__host__ __device__ constexpr int f32_buf_len(int PW)
{
    return PW + CH + 2 * kF32Warm;
}
template <int CH>
__host__ __device__ constexpr int f32_slots(int len)
{
    return len + len / CH + 1;
}

template <int CH>
__global__ __launch_bounds__(1024) void prefilter_rows_f32_kernel(const float *__restrict__ img, int W, int H,
                                                                  int b, PoleArgs poles, float *__restrict__ out)
{
    extern __shared__ float lf[];
    const int PW = W + 2 * b;
    const int LL = f32_line_len<CH>(PW), BL = f32_buf_len<CH>(PW);
    float *line = lf;
    float *buf = lf + f32_slots<CH>(LL);
    const int y = blockIdx.x;
    const int sy = sym_index(y - b, H);
    const float *row = img + (size_t)sy * W;
    for (int e = threadIdx.x; e < LL; e += blockDim.x) {
        int j = e - kF32Warm;
        if ((unsigned)j >= (unsigned)PW) j = wrap(j, PW);
        int x = j - b;
        if ((unsigned)x >= (unsigned)W) x = sym_index(x, W);
        line[ext_slot<CH>(e)] = row[x];
    }
    __syncthreads();
    const int t = threadIdx.x;
    const int c0 = t * CH;
    const int n = min(CH, PW - c0); // samples this thread owns (none past the row)
    const int base = t * (CH + 1);  // ext_slot<CH>(c0) for a window starting at e = c0
    const double z0 = poles.z0, z1 = poles.z1;
    if (n > 0) {
        // causal: line e in [c0, c0 + CH + kF32Warm) is sample c0 - kF32Warm + i
        float v[CH + kF32Warm];
#pragma unroll
        for (int i = 0; i < CH + kF32Warm; ++i) v[i] = line[base + i + i / CH];
        double y1 = 0.0, y2 = 0.0;
#pragma unroll
        for (int i = 0; i < CH + kF32Warm; ++i) {
            y1 = (double)v[i] + z0 * y1;
            y2 = y1 + z1 * y2;
            if (i >= kF32Warm && i - kF32Warm < n) buf[base + i + i / CH] = (float)y2; // e = c0 + i
        }
    }
    __syncthreads();
    // the causal values past the row's end, for the anticausal warm-up
    for (int e = PW + kF32Warm + t; e < BL; e += blockDim.x)
        buf[ext_slot<CH>(e)] = buf[ext_slot<CH>(wrap(e - kF32Warm, PW) + kF32Warm)];
    __syncthreads();
    if (n > 0) {
        // anticausal: buf e = c0 + kF32Warm + i is sample c0 + i
        float v[CH + kF32Warm];
#pragma unroll
        for (int i = 0; i < CH + kF32Warm; ++i) v[i] = buf[base + (i + kF32Warm) + (i + kF32Warm) / CH];
        double w1 = 0.0, w2 = 0.0;
#pragma unroll
        for (int i = CH + kF32Warm - 1; i >= 0; --i) {
            w1 = (double)v[i] + z0 * w1;
            w2 = w1 + z1 * w2;
            if (i < n) line[base + (i + kF32Warm) + (i + kF32Warm) / CH] = (float)(poles.gain * w2);
        }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < PW; x += blockDim.x) out[(size_t)y * PW + x] = line[ext_slot<CH>(x + kF32Warm)];
}

// Columns: one thread per (column, kF32ColChunk-row chunk) loads the rows
// from kF32Warm above the chunk to 2 kF32Warm below its start + chunk, runs the
// causal recursion down and the anticausal one back up; every load and store
// is a coalesced row segment across the wave.  The last chunk of a column
// runs past the final row into the period's start, as the periodic recursion
// does, and stores only its own rows.
__global__ __launch_bounds__(64) void prefilter_cols_f32_kernel(const float *__restrict__ in, int PW, int PH,
                                                                 PoleArgs poles, float *__restrict__ out)
{
    constexpr int CC = kF32ColChunk, L = CC + 2 * kF32Warm;
    const int x = blockIdx.x * 64 + threadIdx.x;
    if (x >= PW) return;
    const int r0 = blockIdx.y * CC;
    const int n = min(CC, PH - r0);
    // kk: the period-PH row index of r0 - kF32Warm + i (the image may have
    // fewer rows than the warm-up: the start takes a true modulo, every step wraps)
    int kk = wrap(r0 - kF32Warm, PH);
    float v[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        v[i] = in[(size_t)kk * PW + x];
        kk = wrap_next(kk, PH);
    }
    const double z0 = poles.z0, z1 = poles.z1;
    double y1 = 0.0, y2 = 0.0;
    float c[CC + kF32Warm]; // causal values of rows r0 .. r0 + CC + kF32Warm - 1
#pragma unroll
    for (int i = 0; i < L; ++i) {
        y1 = (double)v[i] + z0 * y1;
        y2 = y1 + z1 * y2;
        if (i >= kF32Warm) c[i - kF32Warm] = (float)y2;
    }
    double w1 = 0.0, w2 = 0.0;
#pragma unroll
    for (int i = CC + kF32Warm - 1; i >= 0; --i) {
        w1 = (double)c[i] + z0 * w1;
        w2 = w1 + z1 * w2;
        if (i < n) out[(size_t)(r0 + i) * PW + x] = (float)(poles.gain * w2);
    }
}

// Samples per row thread: the candidate that makes the fewest wave-steps per
// row (waves per row x window length), at most 1024 threads a row.
int f32_row_chunk(int PW)
{
    static const int cand[] = {8, 16, 24, 32};
    int best = 32;
    long best_cost = -1;
    for (int ch : cand) {
        const int threads = (PW + ch - 1) / ch;
        if (threads > 1024) continue;
        const long cost = (long)((threads + 63) / 64) * (ch + kF32Warm);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = ch;
        }
    }
    return best;
}

template <int CH>
size_t f32_row_lds_bytes(int PW)
{
    return sizeof(float) * (size_t)(f32_slots<CH>(f32_line_len<CH>(PW)) + f32_slots<CH>(f32_buf_len<CH>(PW)));
}

template <int CH>
hipError_t launch_rows_f32(const float *img, int W, int H, int b, const PoleArgs &poles, float *mid, hipStream_t s)
{
    const int PW = W + 2 * b, PH = H + 2 * b;
    const size_t lds = f32_row_lds_bytes<CH>(PW);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)prefilter_rows_f32_kernel<CH>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int threads = ((PW + CH - 1) / CH + 63) / 64 * 64;
    hipLaunchKernelGGL(prefilter_rows_f32_kernel<CH>, dim3(PH), dim3(threads), lds, s, img, W, H, b, poles, mid);
    return hipGetLastError();
}

// Pipeline reference gradient (reliability_propagation.cpp:84-112): central
// difference, one-sided at the border, evaluated in FP64 from float pixels.
template <typename R>
__global__ void gradient_central_kernel(const float *__restrict__ img, int W, int H, R *__restrict__ gx,
                                        R *__restrict__ gy)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    const int xm = max(0, x - 1), xp = min(W - 1, x + 1);
    const int ym = max(0, y - 1), yp = min(H - 1, y + 1);
    const double xden = (double)max(1, xp - xm), yden = (double)max(1, yp - ym);
    const size_t i = (size_t)y * W + x;
    gx[i] = (R)(((double)img[(size_t)y * W + xp] - (double)img[(size_t)y * W + xm]) / xden);
    gy[i] = (R)(((double)img[(size_t)yp * W + x] - (double)img[(size_t)ym * W + x]) / yden);
}

// gradient_x/y planes of BSplineImagePreprocessor::compute
// (bspline.cpp:289-294): derivative of the local polynomial at (0.5, 0.5).
template <int D, typename CS, typename R>
__global__ void gradient_bspline_kernel(const CS *__restrict__ coeffs, int W, int H, int pitch, int b,
                                        R *__restrict__ gx, R *__restrict__ gy)
{
    using S = Spline<D>;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    double w[S::n], dw[S::n];
    S::weights(0.5, w);
    S::dweights(0.5, dw);
    const CS *base = coeffs + (size_t)(y + b - S::off) * pitch + (x + b - S::off);
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int j = 0; j < S::n; ++j) {
        double rw = 0.0, rdw = 0.0;
#pragma unroll
        for (int k = 0; k < S::n; ++k) {
            const double c = (double)base[(size_t)j * pitch + k];
            rw += w[k] * c;
            rdw += dw[k] * c;
        }
        sx += w[j] * rdw;
        sy += dw[j] * rw;
    }
    gx[(size_t)y * W + x] = (R)sx;
    gy[(size_t)y * W + x] = (R)sy;
}

// BSplineInterpolator::value / gradient (bspline.cpp:458-500) at arbitrary
// points: floor clamped to the image, polynomial extrapolated beyond.
template <int D, typename CS>
__global__ void sample_kernel(const CS *__restrict__ coeffs, int W, int H, int pitch, int b,
                              const double *__restrict__ xy, int n, double *__restrict__ val,
                              double *__restrict__ grad)
{
    using S = Spline<D>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = xy[2 * i], y = xy[2 * i + 1];
    const int ix = clamp_index((int)floor(x), 0, W - 1);
    const int iy = clamp_index((int)floor(y), 0, H - 1);
    const double tx = x - ix, ty = y - iy;
    double wx[S::n], wy[S::n], dwx[S::n], dwy[S::n];
    S::weights(tx, wx);
    S::weights(ty, wy);
    S::dweights(tx, dwx);
    S::dweights(ty, dwy);
    const CS *base = coeffs + (size_t)(iy + b - S::off) * pitch + (ix + b - S::off);
    double v = 0.0, gxs = 0.0, gys = 0.0;
#pragma unroll
    for (int j = 0; j < S::n; ++j) {
        double rw = 0.0, rdw = 0.0;
#pragma unroll
        for (int k = 0; k < S::n; ++k) {
            const double c = (double)base[(size_t)j * pitch + k];
            rw += wx[k] * c;
            rdw += dwx[k] * c;
        }
        v += wy[j] * rw;
        gxs += wy[j] * rdw;
        gys += dwy[j] * rw;
    }
    if (val) val[i] = v;
    if (grad) {
        grad[2 * i] = gxs;
        grad[2 * i + 1] = gys;
    }
}

} // namespace

size_t prefilter_row_lds(int PW, int precision)
{
    if (precision == DIC_PRECISION_FP64) return sizeof(double) * 2 * (size_t)PW; // two lines of doubles
    switch (f32_row_chunk(PW)) {                                                   // two extended lines of floats
    case 8: return f32_row_lds_bytes<8>(PW);
    case 16: return f32_row_lds_bytes<16>(PW);
    case 24: return f32_row_lds_bytes<24>(PW);
    default: return f32_row_lds_bytes<32>(PW);
    }
}

hipError_t launch_prefilter(const float *img, int W, int H, int degree, int border, int exact, int precision,
                            void *out, double *tmp_a, double *tmp_b, hipStream_t s)
{
    const int b = border < 3 ? 3 : border;
    const int PW = W + 2 * b, PH = H + 2 * b;
    const PoleArgs poles = make_poles(degree);
    const bool fp64 = precision == DIC_PRECISION_FP64;
    if (!exact || poles.count == 0) {
        dim3 grid((PW + 255) / 256, PH);
        if (fp64)
            hipLaunchKernelGGL(pad_kernel<double>, grid, dim3(256), 0, s, img, W, H, b, exact ? 1 : 0, (double *)out);
        else
            hipLaunchKernelGGL(pad_kernel<float>, grid, dim3(256), 0, s, img, W, H, b, exact ? 1 : 0, (float *)out);
        return hipGetLastError();
    }
    hipError_t e = hipSuccess;
    if (!fp64) {
        float *mid = reinterpret_cast<float *>(tmp_a);
        switch (f32_row_chunk(PW)) {
        case 8: e = launch_rows_f32<8>(img, W, H, b, poles, mid, s); break;
        case 16: e = launch_rows_f32<16>(img, W, H, b, poles, mid, s); break;
        case 24: e = launch_rows_f32<24>(img, W, H, b, poles, mid, s); break;
        default: e = launch_rows_f32<32>(img, W, H, b, poles, mid, s); break;
        }
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(prefilter_cols_f32_kernel, dim3((PW + 63) / 64, (PH + kF32ColChunk - 1) / kF32ColChunk),
                           dim3(64), 0, s, mid, PW, PH, poles, (float *)out);
        return hipGetLastError();
    }
    const size_t lds = sizeof(double) * 2 * (size_t)PW;
    if (lds > 64 * 1024) {
        e = hipFuncSetAttribute((const void *)prefilter_rows_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(prefilter_rows_kernel, dim3(PH), dim3(256), lds, s, img, W, H, b, poles, tmp_a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 grid((PW + 63) / 64, (PH + kColChunk - 1) / kColChunk);
    hipLaunchKernelGGL(prefilter_cols_causal_kernel, grid, dim3(64), 0, s, tmp_a, PW, PH, poles, tmp_b);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(prefilter_cols_anticausal_kernel, grid, dim3(64), 0, s, tmp_b, PW, PH, poles, (double *)out);
    return hipGetLastError();
}

hipError_t launch_gradient_central(const float *img, int W, int H, int precision, void *gx, void *gy, hipStream_t s)
{
    dim3 grid((W + 255) / 256, H);
    if (precision == DIC_PRECISION_FP64)
        hipLaunchKernelGGL(gradient_central_kernel<double>, grid, dim3(256), 0, s, img, W, H, (double *)gx,
                           (double *)gy);
    else
        hipLaunchKernelGGL(gradient_central_kernel<float>, grid, dim3(256), 0, s, img, W, H, (float *)gx,
                           (float *)gy);
    return hipGetLastError();
}

template <int D>
static hipError_t launch_gradient_bspline_d(const dic_coeffs *c, int precision, void *gx, void *gy, hipStream_t s)
{
    dim3 grid((c->width + 255) / 256, c->height);
    const bool cs64 = c->precision == DIC_PRECISION_FP64;
    const bool r64 = precision == DIC_PRECISION_FP64;
    if (cs64 && r64)
        hipLaunchKernelGGL((gradient_bspline_kernel<D, double, double>), grid, dim3(256), 0, s,
                           (const double *)c->data, c->width, c->height, c->pitch, c->border, (double *)gx,
                           (double *)gy);
    else if (cs64)
        hipLaunchKernelGGL((gradient_bspline_kernel<D, double, float>), grid, dim3(256), 0, s,
                           (const double *)c->data, c->width, c->height, c->pitch, c->border, (float *)gx,
                           (float *)gy);
    else if (r64)
        hipLaunchKernelGGL((gradient_bspline_kernel<D, float, double>), grid, dim3(256), 0, s,
                           (const float *)c->data, c->width, c->height, c->pitch, c->border, (double *)gx,
                           (double *)gy);
    else
        hipLaunchKernelGGL((gradient_bspline_kernel<D, float, float>), grid, dim3(256), 0, s, (const float *)c->data,
                           c->width, c->height, c->pitch, c->border, (float *)gx, (float *)gy);
    return hipGetLastError();
}

hipError_t launch_gradient_bspline(const dic_coeffs *c, int precision, void *gx, void *gy, hipStream_t s)
{
    switch (c->degree) {
    case 1: return launch_gradient_bspline_d<1>(c, precision, gx, gy, s);
    case 3: return launch_gradient_bspline_d<3>(c, precision, gx, gy, s);
    case 5: return launch_gradient_bspline_d<5>(c, precision, gx, gy, s);
    default: return hipErrorInvalidValue;
    }
}

template <int D>
static hipError_t launch_sample_d(const dic_coeffs *c, const double *xy, int n, double *val, double *grad,
                                  hipStream_t s)
{
    dim3 grid((n + 255) / 256);
    if (c->precision == DIC_PRECISION_FP64)
        hipLaunchKernelGGL((sample_kernel<D, double>), grid, dim3(256), 0, s, (const double *)c->data, c->width,
                           c->height, c->pitch, c->border, xy, n, val, grad);
    else
        hipLaunchKernelGGL((sample_kernel<D, float>), grid, dim3(256), 0, s, (const float *)c->data, c->width,
                           c->height, c->pitch, c->border, xy, n, val, grad);
    return hipGetLastError();
}

hipError_t launch_sample(const dic_coeffs *c, const double *xy, int n, double *val, double *grad, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    switch (c->degree) {
    case 1: return launch_sample_d<1>(c, xy, n, val, grad, s);
    case 3: return launch_sample_d<3>(c, xy, n, val, grad, s);
    case 5: return launch_sample_d<5>(c, xy, n, val, grad, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace dic
