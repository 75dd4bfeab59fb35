// icgn_common.hpp -- device pieces of the IC-GN kernels: the solver's LDS
// record, the Eigen-semantics LDL^T factorisation and solve operator, the
// inverse-compositional warp updates, the B-spline samplers (global memory
// and LDS tile), the wave reductions and the sample-table decoding.  Every
// definition sits in an anonymous namespace, so each translation unit that
// includes the header has its own copy.
#pragma once
#include <cfloat>
#include <cmath>

#include "bspline_device.hpp"
#include "dic_internal.hpp"

namespace dic {
namespace {

constexpr double kEps = 1e-12; // icgn.cpp:13

// One POI per wave: the lanes of a wave exchange values through LDS, and a
// wave's LDS operations execute in issue order, so the solver only has to keep
// the compiler from reordering them (no workgroup barrier).  Several POIs can
// then share a workgroup, each wave finishing on its own.
#define WAVE_SYNC() asm volatile("" ::: "memory")

// Coefficient tile in LDS shared by the POIs of a workgroup, in image cells:
// a sample whose 6 x 6 (n x n) block starts at cell (ix, iy) reads it when
// ix in [x0, x0 + w - n] and iy in [y0, y0 + h - n].
struct Tile {
    const float *t; // null: none
    int x0, y0, w, h;
};
constexpr int kGroupPitch = 72, kGroupRows = 60;

// Per-lane accumulator in the working precision.  Each lane sums at most
// ceil(Ns / 64) terms (20 for r = 20, 45 for r = 30) and the 64 lane totals
// are then added pairwise in FP64 by the butterfly reduction: pairwise
// summation with short FP32 leaves in the mixed path (error below
// (leaf + 6) * eps32 * sum |x|).  An FP64 per-lane fold would cost 2 VGPRs
// per accumulator and, at 12 + 3 accumulators, a wave per SIMD.
template <typename R> struct Acc {
    R s = R(0);
    __device__ __forceinline__ void add(R v) { s += v; }
    __device__ __forceinline__ double value() const { return (double)s; }
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
// t <= n for an integer n (NaN: false), compared in the integer domain: a
// double copy of a uniform bound kept live across the iteration loop is a
// VGPR pair, spilled to scratch memory in the FP32 second-order kernels
__device__ __forceinline__ bool le_int(double t, int n)
{
    return t <= -1e9 || (t < 1e9 && (int)ceil(t) <= n);
}

// ---- many sums over the wave at once (transposed butterfly) ----
// permlane32 swap of a pair (lo, hi): lanes 0-31 get their partner's lo in hi,
// lanes 32-63 their partner's hi in lo (gfx950 V_PERMLANE32_SWAP, both dwords)
__device__ __forceinline__ void swap_halves32(double &lo, double &hi)
{
    const auto a = __builtin_amdgcn_permlane32_swap(__double2loint(lo), __double2loint(hi), false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(__double2hiint(lo), __double2hiint(hi), false, false);
    lo = __hiloint2double((int)b[0], (int)a[0]);
    hi = __hiloint2double((int)b[1], (int)a[1]);
}
// the same between 16-lane rows 0<->1 and 2<->3 (V_PERMLANE16_SWAP)
__device__ __forceinline__ void swap_rows16(double &lo, double &hi)
{
    const auto a = __builtin_amdgcn_permlane16_swap(__double2loint(lo), __double2loint(hi), false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(__double2hiint(lo), __double2hiint(hi), false, false);
    lo = __hiloint2double((int)b[0], (int)a[0]);
    hi = __hiloint2double((int)b[1], (int)a[1]);
}
// v of the DPP partner lane (row_mirror 0x140, row_half_mirror 0x141, quad permutations)
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sums N per-lane values gen(0..N-1) over the 64 lanes and stores total i in
// dst[i] (adds it to dst[i] when `add`).  Instead of N butterflies of six
// exchanges each, every exchange level halves the list a lane is responsible
// for: lanes 0-31 keep the first half of the (padded) list and receive their
// partner's copy of it, lanes 32-63 the second half (permlane32 swap), then
// rows 0/2 against rows 1/3 (permlane16 swap), then lanes l / 15-l of a row
// (DPP row mirror); the last NP/8 values are summed over each group of 8
// lanes by full butterflies, after which the 8 lanes hold bit-identical
// totals and the first lane of each group stores them.  Called by the whole
// wave with i compile-time in gen; the caller orders LDS (WAVE_SYNC) before
// reading dst.
template <int N, typename Gen>
__device__ __forceinline__ void wave_reduce_store(Gen &&gen, int lane, double *dst, bool add)
{
    constexpr int NP = (N + 7) & ~7, H1 = NP / 2, H2 = NP / 4, H3 = NP / 8;
    double a[H1];
#pragma unroll
    for (int i = 0; i < H1; ++i) {
        double lo = (i < N) ? gen(i) : 0.0;
        double hi = (i + H1 < N) ? gen(i + H1) : 0.0;
        swap_halves32(lo, hi);
        a[i] = lo + hi;
    }
    double b[H2];
#pragma unroll
    for (int i = 0; i < H2; ++i) {
        double lo = a[i], hi = a[i + H2];
        swap_rows16(lo, hi);
        b[i] = lo + hi;
    }
    const bool up = (lane & 8) != 0;
    double c[H3];
#pragma unroll
    for (int i = 0; i < H3; ++i) {
        const double keep = up ? b[i + H3] : b[i], send = up ? b[i] : b[i + H3];
        c[i] = keep + dpp_f64<0x140>(send);
    }
#pragma unroll
    for (int i = 0; i < H3; ++i) {
        c[i] += dpp_f64<0x141>(c[i]);
        c[i] += dpp_f64<0x4E>(c[i]);
        c[i] += dpp_f64<0xB1>(c[i]);
    }
    if ((lane & 7) == 0) {
        const int base = ((lane >> 5) & 1) * H1 + ((lane >> 4) & 1) * H2 + ((lane >> 3) & 1) * H3;
#pragma unroll
        for (int i = 0; i < H3; ++i)
            if (base + i < N) dst[base + i] = add ? dst[base + i] + c[i] : c[i];
    }
}

__device__ __forceinline__ double uniform_f64(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

template <int P, typename R> __device__ __forceinline__ void steepest_descent(R gx, R gy, R lx, R ly, R *sd)
{
    sd[0] = gx;
    sd[1] = gy;
    sd[2] = gx * lx;
    sd[3] = gx * ly;
    sd[4] = gy * lx;
    sd[5] = gy * ly;
    if constexpr (P == 12) { // second_order_steepest_descent, icgn.cpp:142-165
        const R dx22 = lx * lx * R(0.5), dy22 = ly * ly * R(0.5), dxy = lx * ly;
        sd[6] = gx * dx22;
        sd[7] = gx * dxy;
        sd[8] = gx * dy22;
        sd[9] = gy * dx22;
        sd[10] = gy * dxy;
        sd[11] = gy * dy22;
    }
}

// sum over the subset of w * sd_j with sd_j = g_t(j) * m_j(dx, dy): the FP32
// kernels accumulate w gx and w gy against the monomials 1, x, y, x^2, xy, y^2
// (two products per sample instead of twelve); the halves of the second-order
// terms (icgn.cpp:142-165) are applied after the reduction, an exact scaling.
template <int P, typename R>
__device__ __forceinline__ void acc_sd(Acc<R> *acc, R wgx, R wgy, R lx, R ly, R xx, R xy, R yy)
{
    acc[0].add(wgx);
    acc[1].add(wgy);
    acc[2].add(wgx * lx);
    acc[3].add(wgx * ly);
    acc[4].add(wgy * lx);
    acc[5].add(wgy * ly);
    if constexpr (P == 12) {
        acc[6].add(wgx * xx);
        acc[7].add(wgx * xy);
        acc[8].add(wgx * yy);
        acc[9].add(wgy * xx);
        acc[10].add(wgy * xy);
        acc[11].add(wgy * yy);
    }
}
__device__ __forceinline__ constexpr double sd_half(int j) { return (j == 6 || j == 8 || j == 9 || j == 11) ? 0.5 : 1.0; }

// Parameter j of the warp as g_t(j) * dx^p(j) * dy^q(j) * sd_half(j) (the
// steepest-descent images above, t = 0 for gx, 1 for gy).  The three tables
// are packed two bits per parameter so that a run-time j is a shift and a
// mask: a constexpr array indexed at run time is materialised in per-lane
// scratch memory, once per POI.
constexpr int kParamT[12] = {0, 1, 0, 0, 1, 1, 0, 0, 0, 1, 1, 1};
constexpr int kParamP[12] = {0, 0, 1, 0, 1, 0, 2, 1, 0, 2, 1, 0};
constexpr int kParamQ[12] = {0, 0, 0, 1, 0, 1, 0, 1, 2, 0, 1, 2};
constexpr unsigned pack2(const int (&v)[12])
{
    unsigned r = 0;
    for (int j = 0; j < 12; ++j) r |= unsigned(v[j]) << (2 * j);
    return r;
}
struct ParamMonomial { int t, p, q; };
__device__ __forceinline__ ParamMonomial param_monomial(int j)
{
    constexpr unsigned kT = pack2(kParamT), kP = pack2(kParamP), kQ = pack2(kParamQ);
    return {int(kT >> (2 * j)) & 3, int(kP >> (2 * j)) & 3, int(kQ >> (2 * j)) & 3};
}

__device__ __forceinline__ float buffer_f32(__amdgpu_buffer_rsrc_t rs, int voff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0));
}

struct IcgnShared {
    double H[144];    // Hessian, then its LDL^T factor (lower triangle)
    double Hinv[144]; // H+ = LDLT solve operator, row-major
    double x[12];     // factorisation work vector, then the reciprocal pivots of D
    double p[12];     // current warp parameters (FP64 accumulator)
    double dp[12];    // increment of the current iteration
    double X[2][6];   // rows 0, 1 of G(p) G(dp)^-1 (second-order update)
    double piv[6];    // pivot magnitudes of the 6 x 6 elimination
    double red[16];   // G (P entries) and C of the current iteration
    double ssd[12], sfsd[12];
    double fmean, fnorm, sfh, sfh2;
    double mom[3][15]; // second-order Hessian moments (gx^2, gx gy, gy^2) x dx^p dy^q
    int tr[12];
    int perm[12];      // pivot order of the LDL^T: position k holds row perm[k] of H
    int ns, state;
};

// ---- Eigen 3.4 LDLT factorisation (ldlt_inplace<Lower>::unblocked), lane 0 ----
template <int P> __device__ void ldlt_factor(double *a, int *tr, double *temp, int *ok, int *is_positive)
{
    int sign = 2 /*zero*/, found_zero = 0, ret = 1;
    for (int k = 0; k < P; ++k) {
        int big = k;
        double bigv = fabs(a[k * P + k]);
        for (int i = k + 1; i < P; ++i)
            if (fabs(a[i * P + i]) > bigv) {
                bigv = fabs(a[i * P + i]);
                big = i;
            }
        tr[k] = big;
        if (k != big) {
            for (int j = 0; j < k; ++j) {
                const double t = a[k * P + j];
                a[k * P + j] = a[big * P + j];
                a[big * P + j] = t;
            }
            for (int i = big + 1; i < P; ++i) {
                const double t = a[i * P + k];
                a[i * P + k] = a[i * P + big];
                a[i * P + big] = t;
            }
            const double t = a[k * P + k];
            a[k * P + k] = a[big * P + big];
            a[big * P + big] = t;
            for (int i = k + 1; i < big; ++i) {
                const double t2 = a[i * P + k];
                a[i * P + k] = a[big * P + i];
                a[big * P + i] = t2;
            }
        }
        if (k > 0) {
            for (int j = 0; j < k; ++j) temp[j] = a[j * P + j] * a[k * P + j];
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += a[k * P + j] * temp[j];
            a[k * P + k] -= s;
            for (int i = k + 1; i < P; ++i) {
                double s2 = 0.0;
                for (int j = 0; j < k; ++j) s2 += a[i * P + j] * temp[j];
                a[i * P + k] -= s2;
            }
        }
        const double akk = a[k * P + k];
        const bool valid = fabs(akk) > 0.0;
        if (k == 0 && !valid) {
            sign = 2;
            for (int j = 0; j < P; ++j) tr[j] = j;
            ret = 0;
            break;
        }
        if (k + 1 < P && valid) {
            for (int i = k + 1; i < P; ++i) a[i * P + k] /= akk;
        } else if (k + 1 < P) {
            for (int i = k + 1; i < P; ++i) ret = ret && (a[i * P + k] == 0.0);
        }
        if (found_zero && valid) ret = 0;
        else if (!valid) found_zero = 1;
        if (sign == 0) {
            if (akk < 0.0) sign = 3;
        } else if (sign == 1) {
            if (akk > 0.0) sign = 3;
        } else if (sign == 2) {
            if (akk > 0.0) sign = 0;
            else if (akk < 0.0) sign = 1;
        }
    }
    *ok = ret;
    *is_positive = (sign == 0 || sign == 2);
}

// ---- Eigen 3.4 LDLT factorisation (ldlt_inplace<Lower>::unblocked) on the wave ----
// Lane i < P holds row i of the symmetric matrix in registers.  Step k picks
// the largest |diagonal| among rows k..P-1 (the not-yet-updated diagonal,
// first maximum), applies the symmetric transposition (exchange rows k and
// big between lanes, swap columns k and big in every row), updates column k
// left-looking (A(i,k) -= sum_j L(i,j) D(j) L(k,j)), and scales it by the
// pivot.  info / isPositive follow the same rules as the reference's Eigen
// (zero pivot handling, sign tracking).  Writes L (strict lower), D
// (diagonal) to sh.H and the transpositions to sh.tr.
template <int P> __device__ void ldlt_factor_wave(IcgnShared &sh, int lane, int &ok_out, int &pos_out)
{
    const bool own = lane < P;
    double row[P];
#pragma unroll
    for (int j = 0; j < P; ++j) row[j] = own ? sh.H[lane * P + j] : 0.0;
    double dg = own ? sh.H[lane * P + lane] : 0.0;
    int sign = 2 /*zero*/, found_zero = 0, ret = 1;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        double best = (own && lane >= k) ? fabs(dg) : -1.0;
        int key = lane;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double ob = __shfl_xor(best, m, 64);
            const int ok = __shfl_xor(key, m, 64);
            if (ob > best || (ob == best && ok < key)) {
                best = ob;
                key = ok;
            }
        }
        const int big = key;
        if (lane == 0) sh.tr[k] = big;
        if (big != k) {
            const int src = (lane == k) ? big : ((lane == big) ? k : lane);
#pragma unroll
            for (int j = 0; j < P; ++j) row[j] = __shfl(row[j], src, 64);
            dg = __shfl(dg, src, 64);
#pragma unroll
            for (int m = k + 1; m < P; ++m)
                if (big == m) {
                    const double t = row[k];
                    row[k] = row[m];
                    row[m] = t;
                }
        }
        if (k > 0) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < k; ++j) acc += row[j] * (__shfl(dg, j, 64) * __shfl(row[j], k, 64));
            row[k] -= acc;
        }
        const double akk = __shfl(row[k], k, 64);
        const bool valid = fabs(akk) > 0.0;
        if (k == 0 && !valid) {
            sign = 2;
            ret = 0;
            if (lane < P) sh.tr[lane] = lane;
            break;
        }
        if (k + 1 < P) {
            if (valid) {
                if (lane > k) row[k] /= akk;
            } else {
                ret = ret && !__any(own && lane > k && row[k] != 0.0);
            }
        }
        if (found_zero && valid) ret = 0;
        else if (!valid) found_zero = 1;
        if (sign == 0) {
            if (akk < 0.0) sign = 3;
        } else if (sign == 1) {
            if (akk > 0.0) sign = 3;
        } else if (sign == 2) {
            if (akk > 0.0) sign = 0;
            else if (akk < 0.0) sign = 1;
        }
        if (lane == k) dg = akk;
    }
    WAVE_SYNC();
    if (own) {
#pragma unroll
        for (int j = 0; j < P; ++j)
            if (j < lane) sh.H[lane * P + j] = row[j];
        sh.H[lane * P + lane] = dg;
    }
    ok_out = ret;
    pos_out = (sign == 0 || sign == 2);
    WAVE_SYNC();
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Eigen 3.4 LDLT of sh.H and its solve operator H+ (sh.Hinv), when the
// |diagonal| entries of H are distinct and none is NaN.  Eigen's pivot at
// step k is the largest |diagonal| among the rows not yet factored, and
// those rows still carry their original diagonal (step k updates only
// row k's), so the pivots sort the diagonal by magnitude, descending; with
// distinct entries the position of row i is its rank.  The permutation is
// applied once -- lane i loads row perm[i] of H with permuted columns -- and
// the left-looking factorisation then needs no exchanges: every L and D entry
// comes out of the same operations in the same order as in the stepwise
// pivoted form (ldlt_factor_wave, ldlt_factor), with row k of L broadcast by
// readlane instead of shuffles.  info / isPositive follow the same rules.
// Lane j then forms column j of H+ = P^T L^-T D^+ L^-1 P e_j in registers
// (D^+ by reciprocal pivots, zero where |d| <= DBL_MIN).  Writes L, D (in
// pivot order) to sh.H, the reciprocal pivots to sh.x, the order to sh.perm
// and, when ok and positive, H+ to sh.Hinv.  Returns false, having written
// nothing, on a tie or a NaN: the order of Eigen's exchanges then decides, and
// the caller runs the stepwise factorisation.  Called by the whole wave.
template <int P> __device__ bool ldlt_operator_fast(IcgnShared &sh, int lane, int &ok_out, int &pos_out)
{
    const bool own = lane < P;
    const double di = own ? fabs(sh.H[lane * P + lane]) : 0.0;
    int rank = 0;
    bool bad = own && !(di == di);
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const double dj = fabs(sh.H[j * P + j]);
        rank += (dj > di) ? 1 : 0;
        bad = bad || (own && j != lane && dj == di);
    }
    if (__any(bad)) return false;
    if (own) sh.perm[rank] = lane;
    WAVE_SYNC();
    double row[P];
    {
        const int pi = own ? sh.perm[lane] : 0;
#pragma unroll
        for (int j = 0; j < P; ++j) row[j] = own ? sh.H[pi * P + sh.perm[j]] : 0.0;
    }
    WAVE_SYNC();
    double D[P];
    int sign = 2 /*zero*/, found_zero = 0, ret = 1;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        if (k > 0) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < k; ++j) acc += row[j] * (D[j] * readlane_f64(row[j], k));
            row[k] -= acc;
        }
        const double akk = readlane_f64(row[k], k);
        const bool valid = fabs(akk) > 0.0;
        if (k == 0 && !valid) { // an all-zero diagonal: nothing is factored
            sign = 2;
            ret = 0;
            break;
        }
        if (k + 1 < P) {
            if (valid) {
                if (lane > k) row[k] /= akk;
            } else {
                ret = ret && !__any(own && lane > k && row[k] != 0.0);
            }
        }
        if (found_zero && valid) ret = 0;
        else if (!valid) found_zero = 1;
        if (sign == 0) {
            if (akk < 0.0) sign = 3;
        } else if (sign == 1) {
            if (akk > 0.0) sign = 3;
        } else if (sign == 2) {
            if (akk > 0.0) sign = 0;
            else if (akk < 0.0) sign = 1;
        }
        D[k] = akk;
    }
    ok_out = ret;
    pos_out = (sign == 0 || sign == 2);
    if (!(ret && pos_out)) return true;
    if (own) { // L (strict lower) and D in pivot order; lane k divides its own pivot
        double dk = 0.0;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            if (j <= lane) sh.H[lane * P + j] = row[j];
            if (j == lane) dk = D[j];
        }
        sh.x[lane] = fabs(dk) > DBL_MIN ? 1.0 / dk : 0.0;
    }
    WAVE_SYNC();
    if (own) {
        // row by row, each behind a compiler barrier, so that at most one row
        // of L is loaded ahead (hoisting all P(P-1)/2 loads would spill)
        double x[P];
#pragma unroll
        for (int i = 0; i < P; ++i) { // L^-1 P e_j: the unit vector at the pivot position of row j
            WAVE_SYNC();
            double s = (i == rank) ? 1.0 : 0.0;
#pragma unroll
            for (int l = 0; l < i; ++l) s -= sh.H[i * P + l] * x[l];
            x[i] = s;
        }
#pragma unroll
        for (int i = P - 1; i >= 0; --i) {
            WAVE_SYNC();
            double s = x[i] * sh.x[i];
#pragma unroll
            for (int l = i + 1; l < P; ++l) s -= sh.H[l * P + i] * x[l];
            x[i] = s;
        }
        WAVE_SYNC();
#pragma unroll
        for (int k = 0; k < P; ++k) sh.Hinv[sh.perm[k] * P + lane] = x[k];
    }
    WAVE_SYNC();
    return true;
}

// Column `lane` of the LDLT solve operator H+ (Eigen LDLT::_solve_impl: P b,
// L^-1, D^+ with |d| > DBL_MIN, L^-T, P^T) for lanes < P, once per POI; each
// lane works in place on its own column of sh.Hinv.  D^+ multiplies by
// reciprocal pivots (zero where |d| <= DBL_MIN), computed once per pivot.
// Called by the whole wave.
template <int P> __device__ __forceinline__ void ldlt_operator_column(IcgnShared &sh, int lane)
{
    if (lane < P) {
        const double d = sh.H[lane * P + lane];
        sh.x[lane] = fabs(d) > DBL_MIN ? 1.0 / d : 0.0;
    }
    WAVE_SYNC();
    if (lane >= P) return;
    double *x = sh.Hinv + lane; // x[i] lives at x[i * P]
#pragma unroll 1
    for (int i = 0; i < P; ++i) x[i * P] = (i == lane) ? 1.0 : 0.0;
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        const int t = sh.tr[k];
        if (t != k) {
            const double v = x[k * P];
            x[k * P] = x[t * P];
            x[t * P] = v;
        }
    }
#pragma unroll 1
    for (int i = 0; i < P; ++i) {
        double s = x[i * P];
#pragma unroll 1
        for (int j = 0; j < i; ++j) s -= sh.H[i * P + j] * x[j * P];
        x[i * P] = s;
    }
#pragma unroll 1
    for (int i = 0; i < P; ++i) x[i * P] *= sh.x[i]; // after the whole forward substitution
#pragma unroll 1
    for (int i = P - 1; i >= 0; --i) {
        double s = x[i * P];
#pragma unroll 1
        for (int j = i + 1; j < P; ++j) s -= sh.H[j * P + i] * x[j * P];
        x[i * P] = s;
    }
#pragma unroll 1
    for (int k = P - 1; k >= 0; --k) {
        const int t = sh.tr[k];
        if (t != k) {
            const double v = x[k * P];
            x[k * P] = x[t * P];
            x[t * P] = v;
        }
    }
}

// inverse_compositional_affine_update, icgn.cpp:83-120
__device__ __forceinline__ void ic_affine_update(const double *p, const double *d, double *o)
{
    const double du = p[0], dv = p[1], dudx = p[2], dudy = p[3], dvdx = p[4], dvdy = p[5];
    const double dp0 = d[0], dp1 = d[1], dp2 = d[2], dp3 = d[3], dp4 = d[4], dp5 = d[5];
    const double den = dp2 + dp5 + dp2 * dp5 - dp3 * dp4 + 1.0;
    if (fabs(den) <= kEps) {
        for (int i = 0; i < 6; ++i) o[i] = __builtin_nan("");
        return;
    }
    o[0] = du - ((dudx + 1.0) * (dp0 + dp0 * dp5 - dp1 * dp3)) / den - (dudy * (dp1 - dp0 * dp4 + dp1 * dp2)) / den;
    o[1] = dv - ((dvdy + 1.0) * (dp1 - dp0 * dp4 + dp1 * dp2)) / den - (dvdx * (dp0 + dp0 * dp5 - dp1 * dp3)) / den;
    o[2] = ((dp5 + 1.0) * (dudx + 1.0)) / den - (dp4 * dudy) / den - 1.0;
    o[3] = (dudy * (dp2 + 1.0)) / den - (dp3 * (dudx + 1.0)) / den;
    o[4] = (dvdx * (dp5 + 1.0)) / den - (dp4 * (dvdy + 1.0)) / den;
    o[5] = ((dp2 + 1.0) * (dvdy + 1.0)) / den - (dp3 * dvdx) / den - 1.0;
}

// Entry (row, col) of build_second_order_warp_matrix (icgn.cpp:196-261).
__device__ double warp_entry(const double *q, int row, int col)
{
    const double u = q[0], v = q[1], ux = q[2], uy = q[3], vx = q[4], vy = q[5];
    const double uxx = q[6], uxy = q[7], uyy = q[8], vxx = q[9], vxy = q[10], vyy = q[11];
    switch (row * 6 + col) {
    case 0: return u;
    case 1: return 1.0 + ux;
    case 2: return uy;
    case 3: return 0.5 * uxx;
    case 4: return uxy;
    case 5: return 0.5 * uyy;
    case 6: return v;
    case 7: return vx;
    case 8: return 1.0 + vy;
    case 9: return 0.5 * vxx;
    case 10: return vxy;
    case 11: return 0.5 * vyy;
    case 12: return 1.0;
    case 13: case 14: case 15: case 16: case 17: return 0.0;
    case 18: return u * u;
    case 19: return 2.0 * u * (1.0 + ux);
    case 20: return 2.0 * u * uy;
    case 21: return 1.0 + 2.0 * ux + ux * ux + u * uxx;
    case 22: return 2.0 * (uy + ux * uy + u * uxy);
    case 23: return uy * uy + u * uyy;
    case 24: return u * v;
    case 25: return v + u * vx + v * ux;
    case 26: return u + u * vy + v * uy;
    case 27: return vx + ux * vx + 0.5 * (u * vxx + v * uxx);
    case 28: return 1.0 + vy + ux + ux * vy + uy * vx + u * vxy + v * uxy;
    case 29: return uy + uy * vy + 0.5 * (u * vyy + v * uyy);
    case 30: return v * v;
    case 31: return 2.0 * v * vx;
    case 32: return 2.0 * v * (1.0 + vy);
    case 33: return vx * vx + v * vxx;
    case 34: return 2.0 * (vx + vx * vy + v * vxy);
    default: return 1.0 + 2.0 * vy + vy * vy + v * vyy;
    }
}

// inverse_compositional_second_order_update (icgn.cpp:289-357) on the whole
// wave.  G_new = G(p) G(dp)^-1 P^-1 and extract_from_warp_matrix (:265-287)
// reads only rows 0 and 1 of G_new, so it suffices to solve X G(dp) = B with
// B = rows 0, 1 of G(p): lane r*8+c holds entry (r, c) of [G(dp)^T | B^T]
// (6 x 8, lanes 0..47) and a full-pivot Gauss-Jordan elimination runs with
// FullPivLU's pivot choice (largest magnitude, first in G(dp)'s column-major
// order) and rank test (|pivot| > max|pivot| * 6 eps).  Reads sh.p, sh.dp;
// writes sh.X; returns whether G(dp) is invertible (wave-uniform).
__device__ bool second_order_solve_wave(IcgnShared &sh, int lane)
{
    const int r = lane >> 3, c = lane & 7;
    const bool active = lane < 48;
    // Fast path: with the equations taken in the order (1, 2, 0, 3, 4, 5),
    // [G(dp)^T] is the identity plus terms of the size of dp.  When every row
    // is diagonally dominant (|a_rr| >= 1/2, off-diagonal sum <= |a_rr| / 2)
    // the matrix is invertible with all full-pivoting pivots of the order of
    // one, so FullPivLU's rank test passes, and elimination without pivoting
    // gives X to rounding.  Otherwise the full-pivot elimination below runs.
    {
        const int src = (r == 0) ? 1 : ((r == 1) ? 2 : ((r == 2) ? 0 : r));
        double a = 0.0;
        if (active) a = (c < 6) ? warp_entry(sh.dp, c, src) : warp_entry(sh.p, c - 6, src);
        // off-diagonal magnitude of each row, summed over its 8 lanes
        double off = (active && c < 6 && c != r) ? fabs(a) : 0.0;
        off += __shfl_xor(off, 1, 64);
        off += __shfl_xor(off, 2, 64);
        off += __shfl_xor(off, 4, 64);
        const double dg = fabs(a);
        const bool row_ok = !(active && c == r) || (dg >= 0.5 && off <= 0.5 * dg);
        if (__all(row_ok)) {
#pragma unroll 1
            for (int k = 0; k < 6; ++k) {
                const double piv = __shfl(a, k * 8 + k, 64);
                const double fr = __shfl(a, r * 8 + k, 64);
                const double pr = __shfl(a, k * 8 + c, 64);
                const double rp = 1.0 / piv;
                if (active) a = (r == k) ? a * rp : a - fr * (rp * pr);
            }
            if (active && c >= 6) sh.X[c - 6][r] = a;
            WAVE_SYNC();
            return true;
        }
    }
    double a = 0.0;
    if (active) a = (c < 6) ? warp_entry(sh.dp, c, r) : warp_entry(sh.p, c - 6, r);
    int perm = c; // in column c < 6: which unknown it holds (tracked per lane)
    double maxpiv = 0.0;
    int nonzero = 6;
#pragma unroll 1
    for (int k = 0; k < 6; ++k) {
        // pivot: argmax |a| over rows >= k, columns k..5; ties -> smallest row-major key
        double best = (active && r >= k && c >= k && c < 6) ? fabs(a) : -1.0;
        int key = r * 6 + c;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double ob = __shfl_xor(best, m, 64);
            const int ok = __shfl_xor(key, m, 64);
            if (ob > best || (ob == best && ok < key)) {
                best = ob;
                key = ok;
            }
        }
        if (best == 0.0) {
            nonzero = k;
            break;
        }
        const int br = key / 6, bc = key % 6;
        maxpiv = fmax(maxpiv, best);
        if (lane == 0) sh.piv[k] = best;
        // swap rows k <-> br (all 8 columns), then columns k <-> bc (the 6 x 6 part)
        const int src_r = (r == k) ? br : ((r == br) ? k : r);
        a = __shfl(a, src_r * 8 + c, 64);
        const int src_c = (c < 6) ? ((c == k) ? bc : ((c == bc) ? k : c)) : c;
        a = __shfl(a, r * 8 + src_c, 64);
        perm = __shfl(perm, r * 8 + src_c, 64);
        // eliminate column k from every other row, normalise row k
        const double piv = __shfl(a, k * 8 + k, 64);
        const double fr = __shfl(a, r * 8 + k, 64);
        const double pr = __shfl(a, k * 8 + c, 64);
        if (active) a = (r == k) ? a / piv : a - (fr / piv) * pr;
    }
    WAVE_SYNC();
    int rank = 0;
#pragma unroll 1
    for (int k = 0; k < nonzero; ++k)
        if (sh.piv[k] > maxpiv * (6.0 * DBL_EPSILON)) ++rank;
    if (rank != 6) return false;
    // row k of the right-hand part holds unknown perm(column k) of X^T
    const int unknown = __shfl(perm, r * 8 + r, 64);
    if (active && c >= 6) sh.X[c - 6][unknown] = a;
    WAVE_SYNC();
    return true;
}

// Raw buffer view of the padded coefficient image: every row of the n x n
// block is loaded with the same per-sample byte offset and a wave-uniform
// row offset (j * pitch * sizeof(R)) in a scalar register, so a sample costs
// one 32-bit offset computation instead of n 64-bit addresses.  Loads past
// num_records return 0 (never taken: the warp bounds test runs first).
// Dword buffer loads of consecutive elements; the backend merges them into
// dwordx4 / dwordx2 loads that share the per-sample offset.
template <typename R> struct RowLoad;
template <> struct RowLoad<float> {
    template <int N> __device__ __forceinline__ static void row(__amdgpu_buffer_rsrc_t rs, int voff, int soff, float *c)
    {
#pragma unroll
        for (int k = 0; k < N; ++k)
            c[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff + 4 * k, soff, 0));
    }
};
template <> struct RowLoad<double> {
    template <int N> __device__ __forceinline__ static void row(__amdgpu_buffer_rsrc_t rs, int voff, int soff, double *c)
    {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int lo = (int)__builtin_amdgcn_raw_buffer_load_b32(rs, voff + 8 * k, soff, 0);
            const int hi = (int)__builtin_amdgcn_raw_buffer_load_b32(rs, voff + 8 * k + 4, soff, 0);
            c[k] = __hiloint2double(hi, lo);
        }
    }
};

// B-spline value at (ix + tx, iy + ty) from the padded coefficients (row
// pitch in elements, border b).  FP64: all n*n loads are issued before the
// weights (latency hiding at 4 waves per SIMD); FP32: row by row, which keeps
// fewer load results live and lets the mixed kernels run 5 waves per SIMD.
template <int D, typename R>
__device__ __forceinline__ R sample_at(__amdgpu_buffer_rsrc_t rs, int pitch, int b, int ix, int iy, R tx, R ty)
{
    using S = Spline<D>;
    // row x pitch with the full-rate 24-bit multiply (both below 2^23: the
    // padded image has fewer than 2^31 bytes)
    const int voff = (__mul24(iy + b - S::off, pitch) + (ix + b - S::off)) * (int)sizeof(R);
    const int rowb = pitch * (int)sizeof(R);
    R wx[S::n], wy[S::n];
    R v = R(0);
    if constexpr (sizeof(R) == 8) {
        R c[S::n][S::n];
#pragma unroll
        for (int j = 0; j < S::n; ++j) RowLoad<R>::template row<S::n>(rs, voff, j * rowb, c[j]);
        S::weights(tx, wx);
        S::weights(ty, wy);
#pragma unroll
        for (int j = 0; j < S::n; ++j) {
            R rv = R(0);
#pragma unroll
            for (int k = 0; k < S::n; ++k) rv += wx[k] * c[j][k];
            v += wy[j] * rv;
        }
    } else {
        S::weights(tx, wx);
        S::weights(ty, wy);
#pragma unroll
        for (int j = 0; j < S::n; ++j) {
            R c[S::n];
            RowLoad<R>::template row<S::n>(rs, voff, j * rowb, c);
            R rv = R(0);
#pragma unroll
            for (int k = 0; k < S::n; ++k) rv += wx[k] * c[k];
            v += wy[j] * rv;
        }
    }
    return v;
}

// Sum over the 64 lanes of a wave by DPP row operations (pairs, quads, half
// rows, rows, then rows 0+1 into 1 and 3, row 1 into 3), read back from lane
// 63 so that every lane holds the total: a pairwise tree in FP32.
__device__ __forceinline__ float wave_total_f32(float v)
{
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x140, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xA, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xC, 0xF, false));
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// d! times the centred B-spline basis on [0, 1) (integer coefficients, Horner
// forms; w_{n-1-k}(t) = w_k(1 - t)), for FP32 sampling: the sampled value
// comes out d!^2 times the B-spline value and the caller scales it once.
template <int D> struct SplineU;
template <> struct SplineU<1> {
    static constexpr int n = 2, off = 0;
    static constexpr float inv_scale2 = 1.0f;
    __device__ __forceinline__ static void weights(float t, float *w)
    {
        w[0] = 1.0f - t;
        w[1] = t;
    }
};
template <> struct SplineU<3> {
    static constexpr int n = 4, off = 1;
    static constexpr float inv_scale2 = (float)(1.0 / 36.0);
    __device__ __forceinline__ static void weights(float t, float *w)
    {
        const float s = 1.0f - t, t2 = t * t, s2 = s * s;
        w[0] = s2 * s;
        w[1] = 4.0f + t2 * (-6.0f + 3.0f * t);
        w[2] = 4.0f + s2 * (-6.0f + 3.0f * s);
        w[3] = t2 * t;
    }
};
template <> struct SplineU<5> {
    static constexpr int n = 6, off = 2;
    static constexpr float inv_scale2 = (float)(1.0 / 14400.0);
    __device__ __forceinline__ static void weights(float t, float *w)
    {
        const float s = 1.0f - t, t2 = t * t, s2 = s * s;
        w[0] = s2 * s2 * s;
        w[1] = 1.0f + s * (5.0f + s * (10.0f + s * (10.0f + s * (5.0f - 5.0f * s))));
        w[2] = 66.0f + t2 * (-60.0f + t2 * (30.0f - 10.0f * t));
        w[3] = 66.0f + s2 * (-60.0f + s2 * (30.0f - 10.0f * s));
        w[4] = 1.0f + t * (5.0f + t * (10.0f + t * (10.0f + t * (5.0f - 5.0f * t))));
        w[5] = t2 * t2 * t;
    }
};

// d!^2 times the B-spline value from the global coefficients (FP32).
template <int D>
__device__ __forceinline__ float sample_at_u(__amdgpu_buffer_rsrc_t rs, int pitch, int b, int ix, int iy, float tx,
                                             float ty)
{
    using S = SplineU<D>;
    const int voff = (__mul24(iy + b - S::off, pitch) + (ix + b - S::off)) * 4;
    const int rowb = pitch * 4;
    float wx[S::n], wy[S::n];
    S::weights(tx, wx);
    S::weights(ty, wy);
    float v = 0.0f;
#pragma unroll
    for (int j = 0; j < S::n; ++j) {
        float c[S::n];
        RowLoad<float>::template row<S::n>(rs, voff, j * rowb, c);
        float rv = 0.0f;
#pragma unroll
        for (int k = 0; k < S::n; ++k) rv += wx[k] * c[k];
        v += wy[j] * rv;
    }
    return v;
}

// d!^2 times the value from an LDS tile of pitch PITCH (same weights and
// product order as sample_at_u, so the same float).
template <int D, int PITCH> __device__ __forceinline__ float sample_tile_u(const float *blk, float tx, float ty)
{
    using S = SplineU<D>;
    float wx[S::n], wy[S::n];
    S::weights(tx, wx);
    S::weights(ty, wy);
    float v = 0.0f;
#pragma unroll
    for (int j = 0; j < S::n; ++j) {
        float rv = 0.0f;
#pragma unroll
        for (int k = 0; k < S::n; ++k) rv += wx[k] * blk[j * PITCH + k];
        v += wy[j] * rv;
    }
    return v;
}

__device__ __forceinline__ void decode(int t, int &dx, int &dy)
{
    dx = (int)(short)(t & 0xffff);
    dy = t >> 16;
}

} // namespace
} // namespace dic
