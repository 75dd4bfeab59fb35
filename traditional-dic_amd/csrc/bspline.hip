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
    if (c0 < c1) {
        double y1 = 0.0, y2 = 0.0;
        for (int k = c0 - T, kk = wrap(c0 - T, PW); k < c1; ++k, kk = wrap_next(kk, PW)) {
            y1 = line[kk] + poles.z0 * y1;
            y2 = y1 + poles.z1 * y2;
            if (k >= c0) buf[kk] = (poles.count == 2) ? y2 : y1;
        }
    }
    __syncthreads();
    if (c0 < c1) {
        double w1 = 0.0, w2 = 0.0;
        for (int k = c1 - 1 + T, kk = wrap(c1 - 1 + T, PW); k >= c0; --k, kk = wrap_prev(kk, PW)) {
            w1 = buf[kk] + poles.z0 * w1;
            w2 = w1 + poles.z1 * w2;
            if (k < c1) line[kk] = poles.gain * ((poles.count == 2) ? w2 : w1);
        }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < PW; x += blockDim.x) out[(size_t)y * PW + x] = line[x];
}

constexpr int kColChunk = 64;

__global__ void prefilter_cols_causal_kernel(const double *__restrict__ in, int PW, int PH, PoleArgs poles,
                                             double *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= PW) return;
    const int r0 = blockIdx.y * kColChunk;
    const int r1 = min(r0 + kColChunk, PH);
    double y1 = 0.0, y2 = 0.0;
    for (int k = r0 - poles.warmup, kk = wrap(r0 - poles.warmup, PH); k < r1; ++k, kk = wrap_next(kk, PH)) {
        y1 = in[(size_t)kk * PW + x] + poles.z0 * y1;
        y2 = y1 + poles.z1 * y2;
        if (k >= r0) out[(size_t)kk * PW + x] = (poles.count == 2) ? y2 : y1;
    }
}

template <typename OUT>
__global__ void prefilter_cols_anticausal_kernel(const double *__restrict__ in, int PW, int PH, PoleArgs poles,
                                                 OUT *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= PW) return;
    const int r0 = blockIdx.y * kColChunk;
    const int r1 = min(r0 + kColChunk, PH);
    double w1 = 0.0, w2 = 0.0;
    for (int k = r1 - 1 + poles.warmup, kk = wrap(r1 - 1 + poles.warmup, PH); k >= r0;
         --k, kk = wrap_prev(kk, PH)) {
        w1 = in[(size_t)kk * PW + x] + poles.z0 * w1;
        w2 = w1 + poles.z1 * w2;
        if (k < r1) out[(size_t)kk * PW + x] = (OUT)(poles.gain * ((poles.count == 2) ? w2 : w1));
    }
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

size_t prefilter_row_lds(int PW) { return sizeof(double) * 2 * (size_t)PW; }

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
    const size_t lds = prefilter_row_lds(PW);
    hipError_t e = hipSuccess;
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
    if (fp64)
        hipLaunchKernelGGL(prefilter_cols_anticausal_kernel<double>, grid, dim3(64), 0, s, tmp_b, PW, PH, poles,
                           (double *)out);
    else
        hipLaunchKernelGGL(prefilter_cols_anticausal_kernel<float>, grid, dim3(64), 0, s, tmp_b, PW, PH, poles,
                           (float *)out);
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
