// bspline_device.hpp -- B-spline basis weights for gfx950 device code.
//
// The reference evaluates a sample as a per-pixel monomial block
// qk * C * qk^T (src/interpolation/bspline.cpp:96-113, :418-424).  Because
// sum_m qk(m,c) t^m = beta_d(c - d/2 - t) exactly (Taylor expansion of a
// piecewise polynomial about a knot, left-limit convention of plus_power,
// :68-71), the same value is the tensor product
//     value(ix + tx, iy + ty) = sum_j sum_k C[iy+b-off+j][ix+b-off+k] w_j(ty) w_k(tx)
// with w_j(t) = beta_d(t + off - j), off = d/2, t in [0, 1].  The weights
// below are those pieces expanded in powers of t (exact rational
// coefficients), evaluated by Horner's rule, which keeps FP32 evaluation
// free of cancellation.
#pragma once
#include <hip/hip_runtime.h>

namespace dic {

template <int D> struct Spline;

template <> struct Spline<1> {
    static constexpr int n = 2, off = 0;
    template <typename T> __device__ __forceinline__ static void weights(T t, T *w)
    {
        w[0] = T(1) - t;
        w[1] = t;
    }
    template <typename T> __device__ __forceinline__ static void dweights(T t, T *w)
    {
        (void)t;
        w[0] = T(-1);
        w[1] = T(1);
    }
};

template <> struct Spline<3> {
    static constexpr int n = 4, off = 1;
    template <typename T> __device__ __forceinline__ static void weights(T t, T *w)
    {
        const T s = T(1) - t;
        w[0] = s * s * s * T(1.0 / 6.0);
        w[1] = T(2.0 / 3.0) + t * t * (T(-1) + t * T(0.5));
        w[2] = T(1.0 / 6.0) + t * (T(0.5) + t * (T(0.5) + t * T(-0.5)));
        w[3] = t * t * t * T(1.0 / 6.0);
    }
    template <typename T> __device__ __forceinline__ static void dweights(T t, T *w)
    {
        const T s = T(1) - t;
        w[0] = T(-0.5) * s * s;
        w[1] = t * (T(-2) + t * T(1.5));
        w[2] = T(0.5) + t * (T(1) + t * T(-1.5));
        w[3] = T(0.5) * t * t;
    }
};

template <> struct Spline<5> {
    static constexpr int n = 6, off = 2;
    template <typename T> __device__ __forceinline__ static void weights(T t, T *w)
    {
        const T s = T(1) - t;
        const T s2 = s * s, t2 = t * t;
        w[0] = s2 * s2 * s * T(1.0 / 120.0);
        w[1] = T(13.0 / 60.0) +
               t * (T(-5.0 / 12.0) + t * (T(1.0 / 6.0) + t * (T(1.0 / 6.0) + t * (T(-1.0 / 6.0) + t * T(1.0 / 24.0)))));
        w[2] = T(11.0 / 20.0) + t2 * (T(-0.5) + t2 * (T(0.25) + t * T(-1.0 / 12.0)));
        w[3] = T(13.0 / 60.0) +
               t * (T(5.0 / 12.0) + t * (T(1.0 / 6.0) + t * (T(-1.0 / 6.0) + t * (T(-1.0 / 6.0) + t * T(1.0 / 12.0)))));
        w[4] = T(1.0 / 120.0) +
               t * (T(1.0 / 24.0) + t * (T(1.0 / 12.0) + t * (T(1.0 / 12.0) + t * (T(1.0 / 24.0) + t * T(-1.0 / 24.0)))));
        w[5] = t2 * t2 * t * T(1.0 / 120.0);
    }
    template <typename T> __device__ __forceinline__ static void dweights(T t, T *w)
    {
        const T s = T(1) - t;
        const T s2 = s * s, t2 = t * t;
        w[0] = T(-1.0 / 24.0) * s2 * s2;
        w[1] = T(-5.0 / 12.0) + t * (T(1.0 / 3.0) + t * (T(0.5) + t * (T(-2.0 / 3.0) + t * T(5.0 / 24.0))));
        w[2] = t * (T(-1) + t2 * (T(1) + t * T(-5.0 / 12.0)));
        w[3] = T(5.0 / 12.0) + t * (T(1.0 / 3.0) + t * (T(-0.5) + t * (T(-2.0 / 3.0) + t * T(5.0 / 12.0))));
        w[4] = T(1.0 / 24.0) + t * (T(1.0 / 6.0) + t * (T(0.25) + t * (T(1.0 / 6.0) + t * T(-5.0 / 24.0))));
        w[5] = t2 * t2 * T(5.0 / 120.0);
    }
};

// Poles of 1 / sum_k beta_d(k) z^k and the gain prod (1 - z_i)^2
// (cubic: z = sqrt(3) - 2; quintic: z + 1/z = -13 +- sqrt(105)).
struct PrefilterPoles {
    int count;
    double z[2];
    double gain;
    int warmup; // steps after which z^steps < 1e-18 for every pole (sum over poles)
};

} // namespace dic
