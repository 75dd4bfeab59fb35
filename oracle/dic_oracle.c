/*
 * dic_oracle.c -- CPU FP64 restatement of Traditional-DIC's subset IC-GN path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the HIP
 * product library never links or calls it.
 *
 * Reference files restated (paths inside the Traditional-DIC tree):
 *   src/interpolation/bspline.cpp            B-spline basis, qk, prefilter, blocks, sampling
 *   src/subset/solver/icgn.cpp               IC-GN 1st/2nd order x ZNSSD/SSD x plain/masked
 *   src/subset/seed/reliability_propagation.cpp  gradient cache, POI grid
 *   Eigen 3.4.0 (third-party, not vendored): LDLT (Cholesky/LDLT.h) and
 *   FullPivLU (LU/FullPivLU.h) algorithms, restated from their published source.
 *
 * One deliberate algorithmic substitution: the reference divides FFT spectra
 * (bspline.cpp:191-214, Eigen/kissfft) to deconvolve each padded periodic
 * line.  Here the same circulant system is solved by the exact periodic
 * recursive inverse (causal/anticausal first-order factors with closed-form
 * periodic initial values; Unser 1999).  Both are the exact inverse of the
 * same circulant matrix; the test suite pins this function against a numpy
 * FFT deconvolution and against a dense circulant solve.
 */
#include "dic_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* B-spline basis: bspline.cpp:51-71, :342-367                               */
/* ------------------------------------------------------------------------ */

static int factorial_i(int n)
{
    int v = 1;
    for (int i = 2; i <= n; ++i) v *= i;
    return v;
}

static int combination_i(int n, int k)
{
    if (k < 0 || k > n) return 0;
    return factorial_i(n) / (factorial_i(k) * factorial_i(n - k));
}

/* plus_power (bspline.cpp:68-71): strict x > 0, so the d-th derivative at an
 * integer knot takes the left limit. */
static double plus_power(double x, int power)
{
    return x > 0.0 ? pow(x, (double)power) : 0.0;
}

double orc_basis(double x, int m, int d)
{
    if (m < 0 || m > d) return 0.0;
    const int dfac = factorial_i(d) / factorial_i(d - m);
    double sum = 0.0;
    for (int k = 0; k <= d + 1; ++k) {
        const double sign = (k % 2 == 0) ? 1.0 : -1.0;
        const double coef = sign * (double)combination_i(d + 1, k);
        const double shift = ((double)(d + 1) / 2.0) - (double)k;
        sum += coef * (double)dfac * plus_power(x + shift, d - m);
    }
    return sum / (double)factorial_i(d);
}

/* build_qk (bspline.cpp:320-340): qk(m, c) = (-1)^m beta^(m)(c - d/2) / m! */
void orc_build_qk(int d, double *qk)
{
    const int n = d + 1, off = d / 2;
    for (int m = 0; m < n; ++m)
        for (int c = 0; c < n; ++c) {
            const double sign = (m % 2 == 0) ? 1.0 : -1.0;
            qk[m * n + c] = sign * orc_basis((double)(-off + c), m, d) / (double)factorial_i(m);
        }
}

/* symmetric_index (bspline.cpp:78-89): half-sample symmetric, -1 -> 0 */
int orc_symmetric_index(int value, int size)
{
    if (size <= 0) return 0;
    const int period = 2 * size;
    int w = value % period;
    if (w < 0) w += period;
    return w < size ? w : period - 1 - w;
}

static int clamp_i(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------ */
/* Prefilter: circular deconvolution of a padded line (bspline.cpp:177-244)  */
/* ------------------------------------------------------------------------ */

/* Poles of the sampled B-spline kernel 1/sum_k beta_d(k) z^k (|z| < 1).
 * cubic: z = sqrt(3) - 2; quintic: z + 1/z = -13 +- sqrt(105). */
static int bspline_poles(int d, double *z)
{
    if (d == 3) {
        z[0] = sqrt(3.0) - 2.0;
        return 1;
    }
    if (d == 5) {
        const double w1 = -13.0 + sqrt(105.0), w2 = -13.0 - sqrt(105.0);
        z[0] = 0.5 * (w1 + sqrt(w1 * w1 - 4.0));
        z[1] = 0.5 * (w2 + sqrt(w2 * w2 - 4.0));
        return 2;
    }
    return 0; /* degree 1: kernel is [1], identity */
}

void orc_deconvolve_periodic(double *s, int L, int d)
{
    double z[2];
    const int np = bspline_poles(d, z);
    if (np == 0 || L <= 0) return;
    double gain = 1.0;
    for (int i = 0; i < np; ++i) gain *= (1.0 - z[i]) * (1.0 - z[i]);
    for (int i = 0; i < np; ++i) {
        const double zi = z[i];
        const double zl = pow(zi, (double)L);
        /* causal: y[k] = s[k] + z y[k-1] (indices mod L) */
        double acc = 0.0, zm = 1.0;
        for (int m = 0; m < L; ++m) {
            acc += zm * s[(L - m) % L];
            zm *= zi;
        }
        s[0] = acc / (1.0 - zl);
        for (int k = 1; k < L; ++k) s[k] += zi * s[k - 1];
        /* anticausal: w[k] = y[k] + z w[k+1] (indices mod L) */
        acc = 0.0;
        zm = 1.0;
        for (int m = 0; m < L; ++m) {
            acc += zm * s[(L - 1 + m) % L];
            zm *= zi;
        }
        s[L - 1] = acc / (1.0 - zl);
        for (int k = L - 2; k >= 0; --k) s[k] += zi * s[k + 1];
    }
    for (int k = 0; k < L; ++k) s[k] *= gain;
}

/* form_coefficients (bspline.cpp:369-392) with the border clamp of the
 * BSplineImagePreprocessor constructor (:261-268). */
int orc_form_coefficients(const float *img, int W, int H, int d, int border, int exact, double *out)
{
    const int b = border < 3 ? 3 : border;
    const int PW = W + 2 * b, PH = H + 2 * b;
    for (int y = 0; y < PH; ++y) {
        const int sy = exact ? orc_symmetric_index(y - b, H) : clamp_i(y - b, 0, H - 1);
        for (int x = 0; x < PW; ++x) {
            const int sx = exact ? orc_symmetric_index(x - b, W) : clamp_i(x - b, 0, W - 1);
            out[(size_t)y * PW + x] = (double)img[(size_t)sy * W + sx];
        }
    }
    if (exact) {
        /* deconvolve_rows (:216-229) then deconvolve_columns (:231-244) */
        for (int y = 0; y < PH; ++y) orc_deconvolve_periodic(out + (size_t)y * PW, PW, d);
        double *col = (double *)malloc(sizeof(double) * (size_t)PH);
        for (int x = 0; x < PW; ++x) {
            for (int y = 0; y < PH; ++y) col[y] = out[(size_t)y * PW + x];
            orc_deconvolve_periodic(col, PH, d);
            for (int y = 0; y < PH; ++y) out[(size_t)y * PW + x] = col[y];
        }
        free(col);
    }
    return b;
}

/* ------------------------------------------------------------------------ */
/* Local polynomial blocks and sampling (bspline.cpp:96-175, :270-318, :394-500) */
/* ------------------------------------------------------------------------ */

/* extract_coefficient_block (:394-416) + build_local_polynomial_block (:418-424):
 * block = qk * C * qk^T */
static void build_block(const orc_precomputed *pre, int x, int y, double *blk)
{
    const int d = pre->degree, n = pre->n, off = d / 2, b = pre->border;
    const int PW = pre->width + 2 * b, PH = pre->height + 2 * b;
    const int top = y + b - off, left = x + b - off;
    double c[36], t[36];
    for (int r = 0; r < n; ++r) {
        const int sy = clamp_i(top + r, 0, PH - 1);
        for (int cc = 0; cc < n; ++cc) {
            const int sx = clamp_i(left + cc, 0, PW - 1);
            c[r * n + cc] = pre->coeffs[(size_t)sy * PW + sx];
        }
    }
    /* t = qk * C */
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < n; ++k) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += pre->qk[r * n + j] * c[j * n + k];
            t[r * n + k] = s;
        }
    /* blk = t * qk^T */
    for (int r = 0; r < n; ++r)
        for (int s2 = 0; s2 < n; ++s2) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += t[r * n + k] * pre->qk[s2 * n + k];
            blk[r * n + s2] = s;
        }
}

/* evaluate_polynomial (:96-113) */
static double eval_poly(const double *blk, int n, double dx, double dy)
{
    double v = 0.0, yp = 1.0;
    for (int r = 0; r < n; ++r) {
        double xp = 1.0;
        for (int c = 0; c < n; ++c) {
            v += blk[r * n + c] * yp * xp;
            xp *= dx;
        }
        yp *= dy;
    }
    return v;
}

/* evaluate_polynomial_dx (:115-132) */
static double eval_poly_dx(const double *blk, int n, double dx, double dy)
{
    double v = 0.0, yp = 1.0;
    for (int r = 0; r < n; ++r) {
        double xp = 1.0;
        for (int c = 1; c < n; ++c) {
            v += (double)c * blk[r * n + c] * yp * xp;
            xp *= dx;
        }
        yp *= dy;
    }
    return v;
}

/* evaluate_polynomial_dy (:134-151) */
static double eval_poly_dy(const double *blk, int n, double dx, double dy)
{
    double v = 0.0, yp = 1.0;
    for (int r = 1; r < n; ++r) {
        double xp = 1.0;
        for (int c = 0; c < n; ++c) {
            v += (double)r * blk[r * n + c] * yp * xp;
            xp *= dx;
        }
        yp *= dy;
    }
    return v;
}

int orc_precompute(const float *img, int W, int H, int d, int border, int exact, int with_blocks,
                   orc_precomputed *pre)
{
    memset(pre, 0, sizeof(*pre));
    if (W <= 0 || H <= 0 || img == NULL) return -1;
    if (d != 1 && d != 3 && d != 5) return -2; /* validate_degree (:42-49) */
    const int b = border < 3 ? 3 : border;
    pre->width = W;
    pre->height = H;
    pre->degree = d;
    pre->border = b;
    pre->n = d + 1;
    orc_build_qk(d, pre->qk);
    pre->coeffs = (double *)malloc(sizeof(double) * (size_t)(W + 2 * b) * (size_t)(H + 2 * b));
    pre->grad_x = (double *)calloc((size_t)W * H, sizeof(double));
    pre->grad_y = (double *)calloc((size_t)W * H, sizeof(double));
    if (!pre->coeffs || !pre->grad_x || !pre->grad_y) return -3;
    orc_form_coefficients(img, W, H, d, border, exact, pre->coeffs);
    if (with_blocks) {
        const int n = pre->n;
        pre->blocks = (double *)malloc(sizeof(double) * (size_t)W * H * n * n);
        if (!pre->blocks) return -3;
        pre->has_blocks = 1;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                double *blk = pre->blocks + ((size_t)y * W + x) * n * n;
                build_block(pre, x, y, blk);
                /* gradients at the pixel centre dx = dy = 0.5 (:289-294) */
                if (n > 1) {
                    pre->grad_x[(size_t)y * W + x] = eval_poly_dx(blk, n, 0.5, 0.5);
                    pre->grad_y[(size_t)y * W + x] = eval_poly_dy(blk, n, 0.5, 0.5);
                }
            }
    }
    return 0;
}

void orc_set_threads(int n)
{
    static int default_threads = 0; /* the calling thread's count before the first call */
    if (default_threads == 0) default_threads = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : default_threads);
}

void orc_precomputed_free(orc_precomputed *pre)
{
    free(pre->coeffs);
    free(pre->blocks);
    free(pre->grad_x);
    free(pre->grad_y);
    memset(pre, 0, sizeof(*pre));
}

/* value (:458-474): clamp floor to the image, then the local polynomial */
double orc_value(const orc_precomputed *pre, double x, double y)
{
    const int ix = clamp_i((int)floor(x), 0, pre->width - 1);
    const int iy = clamp_i((int)floor(y), 0, pre->height - 1);
    const double dx = x - (double)ix, dy = y - (double)iy;
    const int n = pre->n;
    if (pre->has_blocks) return eval_poly(pre->blocks + ((size_t)iy * pre->width + ix) * n * n, n, dx, dy);
    double blk[36];
    build_block(pre, ix, iy, blk); /* build_local_block_on_demand (:153-175) */
    return eval_poly(blk, n, dx, dy);
}

/* gradient (:476-500) */
void orc_gradient(const orc_precomputed *pre, double x, double y, double *g)
{
    const int ix = clamp_i((int)floor(x), 0, pre->width - 1);
    const int iy = clamp_i((int)floor(y), 0, pre->height - 1);
    const double dx = x - (double)ix, dy = y - (double)iy;
    const int n = pre->n;
    double tmp[36];
    const double *blk;
    if (pre->has_blocks) {
        blk = pre->blocks + ((size_t)iy * pre->width + ix) * n * n;
    } else {
        build_block(pre, ix, iy, tmp);
        blk = tmp;
    }
    g[0] = eval_poly_dx(blk, n, dx, dy);
    g[1] = eval_poly_dy(blk, n, dx, dy);
}

void orc_value_batch(const orc_precomputed *pre, const double *xy, int n, double *val, double *grad)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        if (val) val[i] = orc_value(pre, xy[2 * i], xy[2 * i + 1]);
        if (grad) orc_gradient(pre, xy[2 * i], xy[2 * i + 1], grad + 2 * i);
    }
}

/* ------------------------------------------------------------------------ */
/* Reference gradient cache: reliability_propagation.cpp:84-112              */
/* ------------------------------------------------------------------------ */
void orc_reference_gradient_cache(const float *img, int W, int H, double *gx, double *gy)
{
    for (int y = 0; y < H; ++y) {
        const int ym = y - 1 < 0 ? 0 : y - 1;
        const int yp = y + 1 > H - 1 ? H - 1 : y + 1;
        const double yden = (double)((yp - ym) > 1 ? (yp - ym) : 1);
        for (int x = 0; x < W; ++x) {
            const int xm = x - 1 < 0 ? 0 : x - 1;
            const int xp = x + 1 > W - 1 ? W - 1 : x + 1;
            const double xden = (double)((xp - xm) > 1 ? (xp - xm) : 1);
            gx[(size_t)y * W + x] = ((double)img[(size_t)y * W + xp] - (double)img[(size_t)y * W + xm]) / xden;
            gy[(size_t)y * W + x] = ((double)img[(size_t)yp * W + x] - (double)img[(size_t)ym * W + x]) / yden;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Eigen 3.4.0 LDLT (Cholesky/LDLT.h, ldlt_inplace<Lower>::unblocked and    */
/* _solve_impl_transposed), restated.  Left-looking, diagonal pivoting on   */
/* the not-yet-updated diagonal, lower triangle storage (row-major here).   */
/* ------------------------------------------------------------------------ */
enum { SIGN_POS = 0, SIGN_NEG = 1, SIGN_ZERO = 2, SIGN_INDEF = 3 };

int orc_ldlt_factor(double *a, int n, int *tr, int *is_positive)
{
    int sign = SIGN_ZERO, found_zero_pivot = 0, ret = 1;
    double temp[12];
#define A(i, j) a[(i) * n + (j)]
    for (int k = 0; k < n; ++k) {
        int big = k;
        double bigv = fabs(A(k, k));
        for (int i = k + 1; i < n; ++i)
            if (fabs(A(i, i)) > bigv) {
                bigv = fabs(A(i, i));
                big = i;
            }
        tr[k] = big;
        if (k != big) {
            for (int j = 0; j < k; ++j) {
                double t = A(k, j);
                A(k, j) = A(big, j);
                A(big, j) = t;
            }
            for (int i = big + 1; i < n; ++i) {
                double t = A(i, k);
                A(i, k) = A(i, big);
                A(i, big) = t;
            }
            double t = A(k, k);
            A(k, k) = A(big, big);
            A(big, big) = t;
            for (int i = k + 1; i < big; ++i) {
                double t2 = A(i, k);
                A(i, k) = A(big, i);
                A(big, i) = t2;
            }
        }
        const int rs = n - k - 1;
        if (k > 0) {
            for (int j = 0; j < k; ++j) temp[j] = A(j, j) * A(k, j);
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += A(k, j) * temp[j];
            A(k, k) -= s;
            for (int i = k + 1; i < n; ++i) {
                double s2 = 0.0;
                for (int j = 0; j < k; ++j) s2 += A(i, j) * temp[j];
                A(i, k) -= s2;
            }
        }
        const double akk = A(k, k);
        const int pivot_valid = fabs(akk) > 0.0;
        if (k == 0 && !pivot_valid) {
            sign = SIGN_ZERO;
            for (int j = 0; j < n; ++j) tr[j] = j;
            ret = 0;
            break;
        }
        if (rs > 0 && pivot_valid) {
            for (int i = k + 1; i < n; ++i) A(i, k) /= akk;
        } else if (rs > 0) {
            for (int i = k + 1; i < n; ++i) ret = ret && (A(i, k) == 0.0);
        }
        if (found_zero_pivot && pivot_valid) ret = 0;
        else if (!pivot_valid) found_zero_pivot = 1;
        if (sign == SIGN_POS) {
            if (akk < 0.0) sign = SIGN_INDEF;
        } else if (sign == SIGN_NEG) {
            if (akk > 0.0) sign = SIGN_INDEF;
        } else if (sign == SIGN_ZERO) {
            if (akk > 0.0) sign = SIGN_POS;
            else if (akk < 0.0) sign = SIGN_NEG;
        }
    }
#undef A
    *is_positive = (sign == SIGN_POS || sign == SIGN_ZERO);
    return ret ? 0 : 1;
}

void orc_ldlt_solve(const double *a, int n, const int *tr, const double *rhs, double *x)
{
    for (int i = 0; i < n; ++i) x[i] = rhs[i];
    for (int k = 0; k < n; ++k)
        if (tr[k] != k) {
            double t = x[k];
            x[k] = x[tr[k]];
            x[tr[k]] = t;
        }
    for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int j = 0; j < i; ++j) s -= a[i * n + j] * x[j];
        x[i] = s;
    }
    for (int i = 0; i < n; ++i) {
        const double dii = a[i * n + i];
        if (fabs(dii) > DBL_MIN) x[i] /= dii;
        else x[i] = 0.0;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int j = i + 1; j < n; ++j) s -= a[j * n + i] * x[j];
        x[i] = s;
    }
    for (int k = n - 1; k >= 0; --k)
        if (tr[k] != k) {
            double t = x[k];
            x[k] = x[tr[k]];
            x[tr[k]] = t;
        }
}

/* ------------------------------------------------------------------------ */
/* Eigen 3.4.0 FullPivLU (LU/FullPivLU.h computeInPlace, rank, _solve_impl)  */
/* restated for 6x6: complete pivoting, first maximum in column-major scan,  */
/* rank threshold = |maxpivot| * 6 * DBL_EPSILON.                            */
/* ------------------------------------------------------------------------ */
int orc_fullpivlu_inverse6(const double *m, double *inv)
{
    const int n = 6;
    double lu[36];
    int rt[6], ct[6];
    memcpy(lu, m, sizeof(lu));
    int nonzero = n;
    double maxpivot = 0.0;
    for (int k = 0; k < n; ++k) {
        int br = k, bc = k;
        double best = -1.0;
        for (int c = k; c < n; ++c)
            for (int r = k; r < n; ++r) {
                const double s = fabs(lu[r * n + c]);
                if (s > best) {
                    best = s;
                    br = r;
                    bc = c;
                }
            }
        if (best == 0.0) {
            nonzero = k;
            for (int i = k; i < n; ++i) rt[i] = ct[i] = i;
            break;
        }
        if (best > maxpivot) maxpivot = best;
        rt[k] = br;
        ct[k] = bc;
        if (k != br)
            for (int c = 0; c < n; ++c) {
                double t = lu[k * n + c];
                lu[k * n + c] = lu[br * n + c];
                lu[br * n + c] = t;
            }
        if (k != bc)
            for (int r = 0; r < n; ++r) {
                double t = lu[r * n + k];
                lu[r * n + k] = lu[r * n + bc];
                lu[r * n + bc] = t;
            }
        const double piv = lu[k * n + k];
        for (int r = k + 1; r < n; ++r) lu[r * n + k] /= piv;
        for (int r = k + 1; r < n; ++r)
            for (int c = k + 1; c < n; ++c) lu[r * n + c] -= lu[r * n + k] * lu[k * n + c];
    }
    const double thr = fabs(maxpivot) * ((double)n * DBL_EPSILON);
    int rank = 0;
    for (int i = 0; i < nonzero; ++i) rank += fabs(lu[i * n + i]) > thr;
    if (rank != n) return 0;
    for (int col = 0; col < n; ++col) {
        double c[6];
        for (int i = 0; i < n; ++i) c[i] = (i == col) ? 1.0 : 0.0;
        for (int k = 0; k < n; ++k)
            if (rt[k] != k) {
                double t = c[k];
                c[k] = c[rt[k]];
                c[rt[k]] = t;
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) c[i] -= lu[i * n + j] * c[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) c[i] -= lu[i * n + j] * c[j];
            c[i] /= lu[i * n + i];
        }
        for (int k = n - 1; k >= 0; --k)
            if (ct[k] != k) {
                double t = c[k];
                c[k] = c[ct[k]];
                c[ct[k]] = t;
            }
        for (int i = 0; i < n; ++i) inv[i * n + col] = c[i];
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Warp updates: icgn.cpp:83-120 (affine), :196-357 (second order)           */
/* ------------------------------------------------------------------------ */
void orc_ic_affine_update(const double *p, const double *d, double *out)
{
    const double du = p[0], dv = p[1], dudx = p[2], dudy = p[3], dvdx = p[4], dvdy = p[5];
    const double dp0 = d[0], dp1 = d[1], dp2 = d[2], dp3 = d[3], dp4 = d[4], dp5 = d[5];
    const double den = dp2 + dp5 + dp2 * dp5 - dp3 * dp4 + 1.0;
    if (fabs(den) <= 1e-12) {
        for (int i = 0; i < 6; ++i) out[i] = NAN;
        return;
    }
    out[0] = du - ((dudx + 1.0) * (dp0 + dp0 * dp5 - dp1 * dp3)) / den - (dudy * (dp1 - dp0 * dp4 + dp1 * dp2)) / den;
    out[1] = dv - ((dvdy + 1.0) * (dp1 - dp0 * dp4 + dp1 * dp2)) / den - (dvdx * (dp0 + dp0 * dp5 - dp1 * dp3)) / den;
    out[2] = ((dp5 + 1.0) * (dudx + 1.0)) / den - (dp4 * dudy) / den - 1.0;
    out[3] = (dudy * (dp2 + 1.0)) / den - (dp3 * (dudx + 1.0)) / den;
    out[4] = (dvdx * (dp5 + 1.0)) / den - (dp4 * (dvdy + 1.0)) / den;
    out[5] = ((dp2 + 1.0) * (dvdy + 1.0)) / den - (dp3 * dvdx) / den - 1.0;
}

/* build_second_order_warp_matrix (icgn.cpp:196-261), row-major 6x6 */
void orc_second_order_warp_matrix(const double *p, double *G)
{
    const double u = p[0], v = p[1], ux = p[2], uy = p[3], vx = p[4], vy = p[5];
    const double uxx = p[6], uxy = p[7], uyy = p[8], vxx = p[9], vxy = p[10], vyy = p[11];
    memset(G, 0, sizeof(double) * 36);
#define G_(r, c) G[(r) * 6 + (c)]
    G_(0, 0) = u; G_(0, 1) = 1.0 + ux; G_(0, 2) = uy; G_(0, 3) = 0.5 * uxx; G_(0, 4) = uxy; G_(0, 5) = 0.5 * uyy;
    G_(1, 0) = v; G_(1, 1) = vx; G_(1, 2) = 1.0 + vy; G_(1, 3) = 0.5 * vxx; G_(1, 4) = vxy; G_(1, 5) = 0.5 * vyy;
    G_(2, 0) = 1.0;
    G_(3, 0) = u * u;
    G_(3, 1) = 2.0 * u * (1.0 + ux);
    G_(3, 2) = 2.0 * u * uy;
    G_(3, 3) = 1.0 + 2.0 * ux + ux * ux + u * uxx;
    G_(3, 4) = 2.0 * (uy + ux * uy + u * uxy);
    G_(3, 5) = uy * uy + u * uyy;
    G_(4, 0) = u * v;
    G_(4, 1) = v + u * vx + v * ux;
    G_(4, 2) = u + u * vy + v * uy;
    G_(4, 3) = vx + ux * vx + 0.5 * (u * vxx + v * uxx);
    G_(4, 4) = 1.0 + vy + ux + ux * vy + uy * vx + u * vxy + v * uxy;
    G_(4, 5) = uy + uy * vy + 0.5 * (u * vyy + v * uyy);
    G_(5, 0) = v * v;
    G_(5, 1) = 2.0 * v * vx;
    G_(5, 2) = 2.0 * v * (1.0 + vy);
    G_(5, 3) = vx * vx + v * vxx;
    G_(5, 4) = 2.0 * (vx + vx * vy + v * vxy);
    G_(5, 5) = 1.0 + 2.0 * vy + vy * vy + v * vyy;
#undef G_
}

/* inverse_compositional_second_order_update (icgn.cpp:289-357):
 * G_new = G(p) * G(dp)^-1 * P^-1, FullPivLU inverse, additive fallback;
 * extract_from_warp_matrix (:265-287). */
void orc_ic_second_order_update(const double *p, const double *d, double *out)
{
    double Gp[36], Gd[36], Gi[36], T[36];
    orc_second_order_warp_matrix(p, Gp);
    orc_second_order_warp_matrix(d, Gd);
    if (!orc_fullpivlu_inverse6(Gd, Gi)) {
        for (int i = 0; i < 12; ++i) out[i] = p[i] - d[i];
        return;
    }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += Gp[r * 6 + k] * Gi[k * 6 + c];
            T[r * 6 + c] = s;
        }
    /* (T * P^-1)(r, c): column 1 <- T(r,0), column 2 <- T(r,1), column 0 <- T(r,2) */
    double Gn[2][6];
    for (int r = 0; r < 2; ++r) {
        Gn[r][0] = T[r * 6 + 2];
        Gn[r][1] = T[r * 6 + 0];
        Gn[r][2] = T[r * 6 + 1];
        Gn[r][3] = T[r * 6 + 3];
        Gn[r][4] = T[r * 6 + 4];
        Gn[r][5] = T[r * 6 + 5];
    }
    out[0] = Gn[0][0];
    out[2] = Gn[0][1] - 1.0;
    out[3] = Gn[0][2];
    out[6] = 2.0 * Gn[0][3];
    out[7] = Gn[0][4];
    out[8] = 2.0 * Gn[0][5];
    out[1] = Gn[1][0];
    out[4] = Gn[1][1];
    out[5] = Gn[1][2] - 1.0;
    out[9] = 2.0 * Gn[1][3];
    out[10] = Gn[1][4];
    out[11] = 2.0 * Gn[1][5];
}

/* ------------------------------------------------------------------------ */
/* IC-GN per POI: icgn.cpp:446-627 (1st ZNSSD), :629-818 (1st ZNSSD masked), */
/* :832-988 (1st SSD), :990-1154 (1st SSD masked), :1168-1362 (2nd ZNSSD),   */
/* :1364-1562 (2nd ZNSSD masked), :1576-1739 (2nd SSD), :1741-1912 (masked)  */
/* All eight share one body here; they differ only in P, the steepest-descent */
/* vector, the warp, the objective and the sample filter.                     */
/* ------------------------------------------------------------------------ */
static int all_finite(const double *v, int n)
{
    for (int i = 0; i < n; ++i)
        if (!isfinite(v[i])) return 0;
    return 1;
}

void orc_icgn_solve(const orc_config *cfg, const float *ref, const float *def_unused, int W, int H,
                    const double *gxp, const double *gyp, const orc_precomputed *dpre, const uint8_t *roi,
                    double px, double py, const orc_initial *init, orc_result *res)
{
    (void)def_unused;
    const int P = cfg->second_order ? 12 : 6;
    const int znssd = cfg->objective == 0;
    memset(res, 0, sizeof(*res));
    res->x = px;
    res->y = py;
    res->p[0] = init->u;
    res->p[1] = init->v;
    res->correlation = init->confidence;
    res->status = ORC_INVALID_INPUT;
    res->valid = 0;
    res->iterations = 0;

    if (W <= 0 || H <= 0 || dpre == NULL || dpre->width != W || dpre->height != H || cfg->subset_radius < 1 ||
        cfg->max_iterations < 0)
        return;

    const int cx = (int)round(px), cy = (int)round(py);
    const int r = cfg->subset_radius;
    if (roi == NULL) {
        if (cx - r < 0 || cy - r < 0 || cx + r >= W || cy + r >= H) return;
    } else {
        if (!(cx >= 0 && cy >= 0 && cx < W && cy < H && roi[(size_t)cy * W + cx])) return;
    }

    const int cap = (2 * r + 1) * (2 * r + 1);
    double *lx = (double *)malloc(sizeof(double) * cap);
    double *ly = (double *)malloc(sizeof(double) * cap);
    double *fv = (double *)malloc(sizeof(double) * cap);
    double *fn = (double *)malloc(sizeof(double) * cap);
    double *sd = (double *)malloc(sizeof(double) * cap * P);
    double *gv = (double *)malloc(sizeof(double) * cap);
    int ns = 0, full = 0;
    double fmean = 0.0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
            if (dx * dx + dy * dy > r * r) continue;
            ++full;
            const int x = cx + dx, y = cy + dy;
            if (roi != NULL) {
                if (!(x >= 0 && y >= 0 && x < W && y < H)) continue;
                if (!roi[(size_t)y * W + x]) continue;
            }
            lx[ns] = (double)dx;
            ly[ns] = (double)dy;
            fv[ns] = (double)ref[(size_t)y * W + x];
            const double gx = gxp[(size_t)y * W + x], gy = gyp[(size_t)y * W + x];
            double *s = sd + (size_t)ns * P;
            s[0] = gx;
            s[1] = gy;
            s[2] = gx * lx[ns];
            s[3] = gx * ly[ns];
            s[4] = gy * lx[ns];
            s[5] = gy * ly[ns];
            if (P == 12) { /* second_order_steepest_descent (:142-165) */
                const double dx22 = lx[ns] * lx[ns] * 0.5, dy22 = ly[ns] * ly[ns] * 0.5, dxy = lx[ns] * ly[ns];
                s[6] = gx * dx22;
                s[7] = gx * dxy;
                s[8] = gx * dy22;
                s[9] = gy * dx22;
                s[10] = gy * dxy;
                s[11] = gy * dy22;
            }
            fmean += fv[ns];
            ++ns;
        }

    int min_needed = P;
    if (roi != NULL) {
        int ms = (int)ceil(cfg->min_valid_sample_ratio * (double)full);
        if (cfg->min_valid_samples > ms) ms = cfg->min_valid_samples;
        if (ms > min_needed) min_needed = ms;
    }
    double fnorm = 1.0;
    double Hm[144];
    int tr[12], ispos = 0;
    if (ns < min_needed) goto done;

    if (znssd) {
        fmean /= (double)ns;
        double acc = 0.0;
        for (int i = 0; i < ns; ++i) {
            const double df = fv[i] - fmean;
            acc += df * df;
        }
        fnorm = sqrt(acc);
        if (fnorm <= 1e-12) {
            res->status = ORC_NUMERICAL_FAILURE;
            goto done;
        }
        for (int i = 0; i < ns; ++i) fn[i] = (fv[i] - fmean) / fnorm;
    }

    memset(Hm, 0, sizeof(Hm));
    for (int i = 0; i < ns; ++i) {
        const double *s = sd + (size_t)i * P;
        for (int a = 0; a < P; ++a)
            for (int b2 = 0; b2 < P; ++b2) Hm[a * P + b2] += s[a] * s[b2];
    }
    if (znssd) {
        const double sc = 2.0 / (fnorm * fnorm);
        for (int k = 0; k < P * P; ++k) Hm[k] *= sc;
    }
    if (orc_ldlt_factor(Hm, P, tr, &ispos) != 0 || !ispos) {
        res->status = ORC_NUMERICAL_FAILURE;
        goto done;
    }

    {
        double p[12] = {0};
        p[0] = init->u;
        p[1] = init->v;
        p[2] = init->du_dx;
        p[3] = init->du_dy;
        p[4] = init->dv_dx;
        p[5] = init->dv_dy;
        double corr = INFINITY;
        int converged = 0, it;
        const double Wm1 = (double)(W - 1), Hm1 = (double)(H - 1);
        for (it = 0; it < cfg->max_iterations; ++it) {
            res->iterations = it + 1;
            double gmean = 0.0;
            for (int i = 0; i < ns; ++i) {
                double wx, wy;
                if (P == 6) {
                    wx = (double)cx + lx[i] + p[0] + p[2] * lx[i] + p[3] * ly[i];
                    wy = (double)cy + ly[i] + p[1] + p[4] * lx[i] + p[5] * ly[i];
                } else { /* warp_second_order (:167-182) */
                    const double dx22 = lx[i] * lx[i] * 0.5, dy22 = ly[i] * ly[i] * 0.5, dxy = lx[i] * ly[i];
                    wx = (double)cx + lx[i] + p[0] + p[2] * lx[i] + p[3] * ly[i] + p[6] * dx22 + p[7] * dxy +
                         p[8] * dy22;
                    wy = (double)cy + ly[i] + p[1] + p[4] * lx[i] + p[5] * ly[i] + p[9] * dx22 + p[10] * dxy +
                         p[11] * dy22;
                }
                /* warped_point_in_bounds (:35-40) */
                if (!(wx >= 0.0 && wy >= 0.0 && wx <= Wm1 && wy <= Hm1)) {
                    res->status = ORC_INVALID_INPUT;
                    goto done;
                }
                gv[i] = orc_value(dpre, wx, wy);
                gmean += gv[i];
            }
            double G[12] = {0};
            if (znssd) {
                gmean /= (double)ns;
                double acc = 0.0; /* vector_norm (:73-81) */
                for (int i = 0; i < ns; ++i) {
                    const double dg = gv[i] - gmean;
                    acc += dg * dg;
                }
                const double gnorm = sqrt(acc);
                if (gnorm <= 1e-12) {
                    res->status = ORC_NUMERICAL_FAILURE;
                    goto done;
                }
                corr = 0.0;
                for (int i = 0; i < ns; ++i) {
                    const double e = fn[i] - (gv[i] - gmean) / gnorm;
                    const double *s = sd + (size_t)i * P;
                    for (int a = 0; a < P; ++a) G[a] += e * s[a];
                    corr += e * e;
                }
                const double sc = 2.0 / fnorm;
                for (int a = 0; a < P; ++a) G[a] *= sc;
            } else {
                corr = 0.0;
                for (int i = 0; i < ns; ++i) {
                    const double e = fv[i] - gv[i];
                    const double *s = sd + (size_t)i * P;
                    for (int a = 0; a < P; ++a) G[a] += e * s[a];
                    corr += e * e;
                }
            }
            double dp[12];
            orc_ldlt_solve(Hm, P, tr, G, dp);
            for (int a = 0; a < P; ++a) dp[a] = -dp[a];
            if (!all_finite(dp, P)) {
                res->status = ORC_NUMERICAL_FAILURE;
                goto done;
            }
            double nrm = 0.0;
            for (int a = 0; a < P; ++a) nrm += dp[a] * dp[a];
            nrm = sqrt(nrm);
            double pn[12];
            if (P == 6) orc_ic_affine_update(p, dp, pn);
            else orc_ic_second_order_update(p, dp, pn);
            memcpy(p, pn, sizeof(double) * P);
            if (!all_finite(p, P)) {
                res->status = ORC_NUMERICAL_FAILURE;
                goto done;
            }
            if (nrm < cfg->convergence_threshold) {
                converged = 1;
                break;
            }
        }
        for (int a = 0; a < P; ++a) res->p[a] = p[a];
        res->correlation = corr;
        if (!converged && isfinite(corr)) converged = 1; /* :620-622 */
        res->status = converged ? ORC_SUCCESS : ORC_NOT_CONVERGED;
        res->valid = converged;
    }
done:
    free(lx);
    free(ly);
    free(fv);
    free(fn);
    free(sd);
    free(gv);
}

void orc_icgn_solve_batch(const orc_config *cfg, const float *ref, int W, int H, const double *gx, const double *gy,
                          const orc_precomputed *dpre, const uint8_t *roi, const double *xy, const orc_initial *init,
                          int n, orc_result *out, int threads)
{
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
#else
    (void)threads;
#endif
    for (int i = 0; i < n; ++i)
        orc_icgn_solve(cfg, ref, NULL, W, H, gx, gy, dpre, roi, xy[2 * i], xy[2 * i + 1], &init[i], &out[i]);
}

/* ------------------------------------------------------------------------ */
/* Forward-compositional GN: forward_gauss_newton.cpp:188-399 (1st ZNSSD),    */
/* :401-608 (1st ZNSSD masked), :622-794 (1st SSD), :796-966 (1st SSD masked), */
/* :1069-1297 (2nd ZNSSD), :1299-1526 (2nd ZNSSD masked), :1540-1739 (2nd     */
/* SSD), :1741-1941 (2nd SSD masked).  Per iteration: warp and sample g, then  */
/* the steepest-descent vector from the deformed image's B-spline gradient at  */
/* the warped point through dW/dz of the current warp (:44-60, :986-1019), H   */
/* and G rebuilt, LDLT, dp = +H^-1 G, step cap, W(p) o W(dp).                 */
/* Variant quirks restated as they are:                                        */
/*   step cap 0.5 (1st ZNSSD, both), 0.35 (1st SSD, both; the unmasked one    */
/*   tests convergence on the norm before the cap, :763-767), 0.5 (2nd ZNSSD), */
/*   0.3 (2nd ZNSSD masked, norm before the cap, :1485-1489), 0.35 (2nd SSD,   */
/*   both);  Levenberg-Marquardt H_ii += 1e-4 max_i H_ii on 2nd ZNSSD, 2nd SSD */
/*   and 2nd SSD masked (:1242-1245, :1682-1685, :1886-1889), not on 2nd      */
/*   ZNSSD masked.                                                             */
/* ------------------------------------------------------------------------ */
void orc_fgn_compose(const double *p, const double *dp, int P, double *pn)
{
    pn[0] = dp[0] + p[0] + p[2] * dp[0] + p[3] * dp[1];
    pn[1] = dp[1] + p[1] + p[4] * dp[0] + p[5] * dp[1];
    if (P == 6) {
        pn[2] = dp[2] + p[2] * (1.0 + dp[2]) + p[3] * dp[4];
        pn[3] = dp[3] + p[2] * dp[3] + p[3] * (1.0 + dp[5]);
        pn[4] = dp[4] + p[4] * (1.0 + dp[2]) + p[5] * dp[4];
        pn[5] = dp[5] + p[4] * dp[3] + p[5] * (1.0 + dp[5]);
        return;
    }
    pn[2] = dp[2] + p[2] * (1.0 + dp[2]) + p[3] * dp[4] + p[6] * dp[0] + p[7] * dp[1];
    pn[3] = dp[3] + p[2] * dp[3] + p[3] * (1.0 + dp[5]) + p[7] * dp[0] + p[8] * dp[1];
    pn[4] = dp[4] + p[4] * (1.0 + dp[2]) + p[5] * dp[4] + p[9] * dp[0] + p[10] * dp[1];
    pn[5] = dp[5] + p[4] * dp[3] + p[5] * (1.0 + dp[5]) + p[10] * dp[0] + p[11] * dp[1];
    pn[6] = dp[6] + p[2] * dp[6] + p[3] * dp[9] + p[6] * (1.0 + 2.0 * dp[2]) + 2.0 * p[7] * dp[4];
    pn[7] = dp[7] + p[2] * dp[7] + p[3] * dp[10] + p[6] * dp[3] + p[7] * (1.0 + dp[2] + dp[5]) + p[8] * dp[4];
    pn[8] = dp[8] + p[2] * dp[8] + p[3] * dp[11] + p[7] * dp[3] + p[8] * (1.0 + 2.0 * dp[5]);
    pn[9] = dp[9] + p[4] * dp[6] + p[5] * dp[9] + p[9] * (1.0 + 2.0 * dp[2]) + 2.0 * p[10] * dp[4];
    pn[10] = dp[10] + p[4] * dp[7] + p[5] * dp[10] + p[9] * dp[3] + p[10] * (1.0 + dp[2] + dp[5]) + p[11] * dp[4];
    pn[11] = dp[11] + p[4] * dp[8] + p[5] * dp[11] + p[10] * dp[3] + p[11] * (1.0 + 2.0 * dp[5]);
}

void orc_fgn_solve(const orc_config *cfg, const float *ref, int W, int H, const orc_precomputed *dpre,
                   const uint8_t *roi, double px, double py, const orc_initial *init, orc_result *res)
{
    const int P = cfg->second_order ? 12 : 6;
    const int znssd = cfg->objective == 0;
    const int masked = roi != NULL;
    memset(res, 0, sizeof(*res));
    res->x = px;
    res->y = py;
    res->p[0] = init->u;
    res->p[1] = init->v;
    res->correlation = init->confidence;
    res->status = ORC_INVALID_INPUT;
    if (W <= 0 || H <= 0 || dpre == NULL || dpre->width != W || dpre->height != H || cfg->subset_radius < 1 ||
        cfg->max_iterations < 0)
        return;
    const int cx = (int)round(px), cy = (int)round(py);
    const int r = cfg->subset_radius;
    if (!masked) {
        if (cx - r < 0 || cy - r < 0 || cx + r >= W || cy + r >= H) return;
    } else {
        if (!(cx >= 0 && cy >= 0 && cx < W && cy < H && roi[(size_t)cy * W + cx])) return;
    }
    /* step cap and its convergence norm, Levenberg-Marquardt */
    double cap = 0.5;
    int precap_norm = 0, lm = 0;
    if (P == 6) {
        cap = znssd ? 0.5 : 0.35;
        precap_norm = !znssd && !masked;
    } else if (znssd) {
        cap = masked ? 0.3 : 0.5;
        precap_norm = masked;
        lm = !masked;
    } else {
        cap = 0.35;
        lm = 1;
    }
    const int cnt = (2 * r + 1) * (2 * r + 1);
    double *lx = (double *)malloc(sizeof(double) * cnt);
    double *ly = (double *)malloc(sizeof(double) * cnt);
    double *fv = (double *)malloc(sizeof(double) * cnt);
    double *fn = (double *)malloc(sizeof(double) * cnt);
    double *gv = (double *)malloc(sizeof(double) * cnt);
    int ns = 0, full = 0;
    double fmean = 0.0, fnorm = 1.0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
            if (dx * dx + dy * dy > r * r) continue;
            ++full;
            const int x = cx + dx, y = cy + dy;
            if (masked && !(x >= 0 && y >= 0 && x < W && y < H && roi[(size_t)y * W + x])) continue;
            lx[ns] = (double)dx;
            ly[ns] = (double)dy;
            fv[ns] = (double)ref[(size_t)y * W + x];
            fmean += fv[ns];
            ++ns;
        }
    int min_needed = P;
    if (masked) {
        int ms = (int)ceil(cfg->min_valid_sample_ratio * (double)full);
        if (cfg->min_valid_samples > ms) ms = cfg->min_valid_samples;
        if (ms > min_needed) min_needed = ms;
    }
    if (ns < min_needed) goto done;
    if (znssd) {
        fmean /= (double)ns;
        double acc = 0.0;
        for (int i = 0; i < ns; ++i) acc += (fv[i] - fmean) * (fv[i] - fmean);
        fnorm = sqrt(acc);
        if (fnorm <= 1e-12) {
            res->status = ORC_NUMERICAL_FAILURE;
            goto done;
        }
        for (int i = 0; i < ns; ++i) fn[i] = (fv[i] - fmean) / fnorm;
    }
    {
        double p[12] = {0};
        p[0] = init->u;
        p[1] = init->v;
        p[2] = init->du_dx;
        p[3] = init->du_dy;
        p[4] = init->dv_dx;
        p[5] = init->dv_dy;
        double corr = INFINITY;
        int converged = 0;
        const double Wm1 = (double)(W - 1), Hm1 = (double)(H - 1);
        for (int it = 0; it < cfg->max_iterations; ++it) {
            res->iterations = it + 1;
            double gmean = 0.0;
            for (int i = 0; i < ns; ++i) {
                double wx = (double)cx + lx[i] + p[0] + p[2] * lx[i] + p[3] * ly[i];
                double wy = (double)cy + ly[i] + p[1] + p[4] * lx[i] + p[5] * ly[i];
                if (P == 12) {
                    const double dx22 = lx[i] * lx[i] * 0.5, dy22 = ly[i] * ly[i] * 0.5, dxy = lx[i] * ly[i];
                    wx += p[6] * dx22 + p[7] * dxy + p[8] * dy22;
                    wy += p[9] * dx22 + p[10] * dxy + p[11] * dy22;
                }
                if (!(wx >= 0.0 && wy >= 0.0 && wx <= Wm1 && wy <= Hm1)) {
                    res->status = ORC_INVALID_INPUT;
                    goto done;
                }
                gv[i] = orc_value(dpre, wx, wy);
                gmean += gv[i];
            }
            double gnorm = 1.0;
            if (znssd) {
                gmean /= (double)ns;
                double acc = 0.0;
                for (int i = 0; i < ns; ++i) acc += (gv[i] - gmean) * (gv[i] - gmean);
                gnorm = sqrt(acc);
                if (gnorm <= 1e-12) {
                    res->status = ORC_NUMERICAL_FAILURE;
                    goto done;
                }
            }
            double Hm[144], G[12];
            memset(Hm, 0, sizeof(Hm));
            memset(G, 0, sizeof(G));
            corr = 0.0;
            for (int i = 0; i < ns; ++i) {
                const double dx22 = lx[i] * lx[i] * 0.5, dy22 = ly[i] * ly[i] * 0.5, dxy = lx[i] * ly[i];
                double wx = (double)cx + lx[i] + p[0] + p[2] * lx[i] + p[3] * ly[i];
                double wy = (double)cy + ly[i] + p[1] + p[4] * lx[i] + p[5] * ly[i];
                if (P == 12) {
                    wx += p[6] * dx22 + p[7] * dxy + p[8] * dy22;
                    wy += p[9] * dx22 + p[10] * dxy + p[11] * dy22;
                }
                double g2[2];
                orc_gradient(dpre, wx, wy, g2);
                /* dW/dz of the current warp at (lx, ly) */
                double a00 = 1.0 + p[2], a01 = p[3], a10 = p[4], a11 = 1.0 + p[5];
                if (P == 12) {
                    a00 += p[6] * lx[i] + p[7] * ly[i];
                    a01 += p[7] * lx[i] + p[8] * ly[i];
                    a10 += p[9] * lx[i] + p[10] * ly[i];
                    a11 += p[10] * lx[i] + p[11] * ly[i];
                }
                const double ge = g2[0] * a00 + g2[1] * a10, he = g2[0] * a01 + g2[1] * a11;
                double sd[12] = {ge, he, ge * lx[i], ge * ly[i], he * lx[i], he * ly[i],
                                 ge * dx22, ge * dxy, ge * dy22, he * dx22, he * dxy, he * dy22};
                const double e = znssd ? fn[i] - (gv[i] - gmean) / gnorm : fv[i] - gv[i];
                for (int a = 0; a < P; ++a) {
                    for (int b2 = 0; b2 < P; ++b2) Hm[a * P + b2] += sd[a] * sd[b2];
                    G[a] += e * sd[a];
                }
                corr += e * e;
            }
            if (znssd) {
                const double sh = 2.0 / (fnorm * fnorm), sg = 2.0 / fnorm;
                for (int k = 0; k < P * P; ++k) Hm[k] *= sh;
                for (int a = 0; a < P; ++a) G[a] *= sg;
            }
            if (lm) {
                double mx = Hm[0];
                for (int a = 1; a < P; ++a)
                    if (Hm[a * P + a] > mx) mx = Hm[a * P + a];
                const double lambda = 1e-4 * mx;
                for (int a = 0; a < P; ++a) Hm[a * P + a] += lambda;
            }
            int tr[12], ispos = 0;
            if (orc_ldlt_factor(Hm, P, tr, &ispos) != 0 || !ispos) {
                res->status = ORC_NUMERICAL_FAILURE;
                goto done;
            }
            double dp[12] = {0};
            orc_ldlt_solve(Hm, P, tr, G, dp);
            if (!all_finite(dp, P)) {
                res->status = ORC_NUMERICAL_FAILURE;
                goto done;
            }
            double nrm = 0.0;
            for (int a = 0; a < P; ++a) nrm += dp[a] * dp[a];
            nrm = sqrt(nrm);
            double test = nrm;
            if (nrm > cap) {
                for (int a = 0; a < P; ++a) dp[a] *= cap / nrm;
                if (!precap_norm) test = cap;
            }
            double pn[12] = {0};
            orc_fgn_compose(p, dp, P, pn);
            memcpy(p, pn, sizeof(double) * P);
            if (!all_finite(p, P)) {
                res->status = ORC_NUMERICAL_FAILURE;
                goto done;
            }
            if (test < cfg->convergence_threshold) {
                converged = 1;
                break;
            }
        }
        for (int a = 0; a < P; ++a) res->p[a] = p[a];
        res->correlation = corr;
        if (!converged && isfinite(corr)) converged = 1;
        res->status = converged ? ORC_SUCCESS : ORC_NOT_CONVERGED;
        res->valid = converged;
    }
done:
    free(lx);
    free(ly);
    free(fv);
    free(fn);
    free(gv);
}

void orc_fgn_solve_batch(const orc_config *cfg, const float *ref, int W, int H, const orc_precomputed *dpre,
                         const uint8_t *roi, const double *xy, const orc_initial *init, int n, orc_result *out,
                         int threads)
{
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
#else
    (void)threads;
#endif
    for (int i = 0; i < n; ++i) orc_fgn_solve(cfg, ref, W, H, dpre, roi, xy[2 * i], xy[2 * i + 1], &init[i], &out[i]);
}

/* ------------------------------------------------------------------------ */
/* Reliability-guided propagation, reliability_propagation.cpp:130-394: each */
/* seed (aligned to the nearest grid point, :206-262) starts a best-first    */
/* walk over the 4-neighbours (up, right, down, left) of the popped point     */
/* (min-heap on correlation, :44-53); a neighbour is visited once (:298);     */
/* outside the ROI or the subset margin it is only marked (:300-311); its     */
/* guess is the parent's affine extrapolation (:314-327); it is accepted if   */
/* Success, valid, correlation < max_znssd, |du|, |dv| < step from the guess  */
/* and finite (:359-368), and only accepted points are stored and pushed.     */
/* ------------------------------------------------------------------------ */
typedef struct {
    double corr;
    int gx, gy;
    double u, v, ux, uy, vx, vy;
} prop_item;

static void heap_push(prop_item *h, int *n, prop_item it)
{
    int i = (*n)++;
    h[i] = it;
    while (i > 0) {
        const int pa = (i - 1) / 2;
        if (!(h[i].corr < h[pa].corr)) break;
        prop_item t = h[i];
        h[i] = h[pa];
        h[pa] = t;
        i = pa;
    }
}

static prop_item heap_pop(prop_item *h, int *n)
{
    prop_item top = h[0];
    h[0] = h[--(*n)];
    int i = 0;
    for (;;) {
        const int l = 2 * i + 1, r = l + 1;
        int m = i;
        if (l < *n && h[l].corr < h[m].corr) m = l;
        if (r < *n && h[r].corr < h[m].corr) m = r;
        if (m == i) break;
        prop_item t = h[i];
        h[i] = h[m];
        h[m] = t;
        i = m;
    }
    return top;
}

static void prop_solve(const orc_config *cfg, int optimizer, const float *ref, int W, int H, const double *gx,
                       const double *gy, const orc_precomputed *dpre, const uint8_t *roi, double x, double y,
                       const orc_initial *ini, orc_result *out)
{
    if (optimizer == 1) orc_fgn_solve(cfg, ref, W, H, dpre, roi, x, y, ini, out);
    else orc_icgn_solve(cfg, ref, NULL, W, H, gx, gy, dpre, roi, x, y, ini, out);
}

int orc_propagate(const orc_config *cfg, int optimizer, int spacing, double max_znssd, int truncate_roi,
                  const float *ref, int W, int H, const double *gxp, const double *gyp, const orc_precomputed *dpre,
                  const uint8_t *roi, const double *seeds, int n_seeds, orc_result *grid)
{
    if (spacing < 0) spacing = 0;
    const int step = spacing + 1, r = cfg->subset_radius;
    const int gw = (W + step - 1) / step, gh = (H + step - 1) / step, total = gw * gh;
    for (int i = 0; i < total; ++i) {
        memset(&grid[i], 0, sizeof(orc_result));
        grid[i].x = (double)((i % gw) * step);
        grid[i].y = (double)((i / gw) * step);
        grid[i].status = ORC_INVALID_INPUT;
    }
    if (W <= 0 || H <= 0 || roi == NULL || r < 1 || n_seeds <= 0) return 0;
    const uint8_t *sroi = truncate_roi ? roi : NULL;
    uint8_t *calc = (uint8_t *)calloc((size_t)total, 1);
    prop_item *heap = (prop_item *)malloc(sizeof(prop_item) * ((size_t)total + 1));
    int computed = 0;
    for (int k = 0; k < n_seeds; ++k) {
        const double *sd = seeds + 9 * k;
        const int sxr = (int)round(sd[0]), syr = (int)round(sd[1]);
        if (!(sxr >= 0 && syr >= 0 && sxr < W && syr < H) || !roi[(size_t)syr * W + sxr]) continue;
        int gxs = (int)llround((double)sxr / step), gys = (int)llround((double)syr / step);
        gxs = gxs < 0 ? 0 : (gxs > gw - 1 ? gw - 1 : gxs);
        gys = gys < 0 ? 0 : (gys > gh - 1 ? gh - 1 : gys);
        const int sx = gxs * step, sy = gys * step;
        if (!(sx >= 0 && sy >= 0 && sx < W && sy < H) || !roi[(size_t)sy * W + sx]) continue;
        const int sidx = gys * gw + gxs;
        if (calc[sidx]) continue;
        orc_result al;
        memset(&al, 0, sizeof(al));
        al.p[0] = sd[2], al.p[1] = sd[3], al.p[2] = sd[4], al.p[3] = sd[5], al.p[4] = sd[6], al.p[5] = sd[7];
        al.correlation = sd[8];
        if (sx != sxr || sy != syr) {
            orc_initial ini = {sd[2], sd[3], sd[4], sd[5], sd[6], sd[7], sd[8]};
            prop_solve(cfg, optimizer, ref, W, H, gxp, gyp, dpre, sroi, (double)sx, (double)sy, &ini, &al);
            if (!al.valid || al.status != ORC_SUCCESS || !all_finite(al.p, 6) || !isfinite(al.correlation)) continue;
        }
        calc[sidx] = 1;
        grid[sidx] = al;
        grid[sidx].x = (double)sx;
        grid[sidx].y = (double)sy;
        grid[sidx].valid = 1;
        grid[sidx].status = ORC_SUCCESS;
        ++computed;
        int hn = 0;
        prop_item it0 = {al.correlation > 0.0 ? al.correlation : 0.0, gxs, gys, al.p[0], al.p[1], al.p[2],
                         al.p[3], al.p[4], al.p[5]};
        heap_push(heap, &hn, it0);
        while (hn > 0) {
            const prop_item it = heap_pop(heap, &hn);
            const int px = it.gx * step, py = it.gy * step;
            const int nb[4][2] = {{px, py - step}, {px + step, py}, {px, py + step}, {px - step, py}};
            for (int q = 0; q < 4; ++q) {
                const int nx = nb[q][0], ny = nb[q][1];
                if (!(nx >= 0 && ny >= 0 && nx < W && ny < H)) continue;
                const int gnx = nx / step, gny = ny / step;
                if (gnx < 0 || gny < 0 || gnx >= gw || gny >= gh) continue;
                const int nidx = gny * gw + gnx;
                if (calc[nidx]) continue;
                if (!roi[(size_t)ny * W + nx]) {
                    calc[nidx] = 1;
                    continue;
                }
                if (nx - r < 0 || ny - r < 0 || nx + r >= W || ny + r >= H) {
                    calc[nidx] = 1;
                    continue;
                }
                const double ddx = (double)(nx - px), ddy = (double)(ny - py);
                const double iu = it.u + it.ux * ddx + it.uy * ddy;
                const double iv = it.v + it.vx * ddx + it.vy * ddy;
                orc_initial ini = {iu, iv, it.ux, it.uy, it.vx, it.vy, 0.0};
                orc_result res;
                prop_solve(cfg, optimizer, ref, W, H, gxp, gyp, dpre, sroi, (double)nx, (double)ny, &ini, &res);
                calc[nidx] = 1;
                const int ok = res.valid && res.status == ORC_SUCCESS && res.correlation < max_znssd &&
                               fabs(iu - res.p[0]) < (double)step && fabs(iv - res.p[1]) < (double)step &&
                               all_finite(res.p, 6) && isfinite(res.correlation);
                if (ok) {
                    grid[nidx] = res;
                    grid[nidx].x = (double)nx;
                    grid[nidx].y = (double)ny;
                    ++computed;
                    prop_item nxt = {res.correlation, gnx, gny, res.p[0], res.p[1], res.p[2], res.p[3], res.p[4],
                                     res.p[5]};
                    heap_push(heap, &hn, nxt);
                }
            }
        }
    }
    free(calc);
    free(heap);
    return computed;
}

int orc_poi_grid(int W, int H, int spacing, double *xy)
{
    if (spacing < 0) spacing = 0;
    const int step = spacing + 1;
    const int gw = (W + step - 1) / step, gh = (H + step - 1) / step;
    if (xy)
        for (int i = 0; i < gw * gh; ++i) {
            xy[2 * i] = (double)((i % gw) * step);
            xy[2 * i + 1] = (double)((i / gw) * step);
        }
    return gw * gh;
}

/* ======================================================================
 * Seeding: integer-pixel ZNCC search and seed selection
 * ====================================================================== */

#define ORC_SEARCH_EPS 1e-12    /* kEpsilon, integer_search.cpp:10 */
#define ORC_LOCAL_MARGIN 0.02   /* kLocalPreferenceZnccMargin, :11 */

typedef struct {
    int dx, dy;
    double val;
} orc_sample;

typedef struct {
    int u, v;
    double z;
} orc_cand;

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* full_circle_sample_count, :35-47 */
static int circle_count(int r, int step)
{
    step = imax(1, step);
    int c = 0;
    for (int dy = -r; dy <= r; dy += step)
        for (int dx = -r; dx <= r; dx += step)
            if (dx * dx + dy * dy <= r * r) ++c;
    return c;
}

/* min_required_samples, :49-52 */
static int min_required(int r, int step)
{
    return imax(12, (int)ceil(0.5 * (double)circle_count(r, step)));
}

/* collect_circular_reference_samples (:62-79) when roi is NULL, else
 * collect_masked_reference_samples (:81-105).  The unmasked form returns -1
 * where a sample leaves the image (Image::at throws there, image.cpp:154-160). */
static int collect_samples(const float *ref, int W, int H, const uint8_t *roi, int cx, int cy, int r, int step,
                           orc_sample *s)
{
    step = imax(1, step);
    int n = 0;
    for (int dy = -r; dy <= r; dy += step)
        for (int dx = -r; dx <= r; dx += step) {
            if (dx * dx + dy * dy > r * r) continue;
            const int x = cx + dx, y = cy + dy;
            const int inside = x >= 0 && y >= 0 && x < W && y < H;
            if (roi) {
                if (!inside || !roi[(size_t)y * W + x]) continue;
            } else if (!inside) {
                return -1;
            }
            s[n].dx = dx;
            s[n].dy = dy;
            s[n].val = (double)ref[(size_t)y * W + x];
            ++n;
        }
    return n;
}

/* sample_mean / sample_norm, :107-125 */
static void sample_stats(const orc_sample *s, int n, double *mean, double *norm)
{
    double m = 0.0;
    for (int i = 0; i < n; ++i) m += s[i].val;
    m = n > 0 ? m / (double)n : 0.0;
    double q = 0.0;
    for (int i = 0; i < n; ++i) {
        const double d = s[i].val - m;
        q += d * d;
    }
    *mean = m;
    *norm = q;
}

/* circular_subset_zncc, :127-165 */
static double shift_zncc(const float *def, int W, int H, const orc_sample *s, int n, double rm, double rn, int cx,
                         int cy, int u, int v)
{
    if (n == 0 || rn <= ORC_SEARCH_EPS) return -INFINITY;
    double dm = 0.0;
    for (int i = 0; i < n; ++i) {
        const int x = cx + u + s[i].dx, y = cy + v + s[i].dy;
        if (!(x >= 0 && y >= 0 && x < W && y < H)) return -INFINITY;
        dm += (double)def[(size_t)y * W + x];
    }
    dm /= (double)n;
    double num = 0.0, dn = 0.0;
    for (int i = 0; i < n; ++i) {
        const double rd = s[i].val - rm;
        const double dd = (double)def[(size_t)(cy + v + s[i].dy) * W + (cx + u + s[i].dx)] - dm;
        num += rd * dd;
        dn += dd * dd;
    }
    if (dn <= ORC_SEARCH_EPS) return -INFINITY;
    return num / sqrt(rn * dn);
}

/* search_displacements, :167-206: row-major scan, the first strict maximum wins */
static orc_cand search_window(const float *def, int W, int H, const orc_sample *s, int n, double rm, double rn,
                              int cx, int cy, int r, int u0, int u1, int v0, int v1, int step)
{
    orc_cand best = {0, 0, -INFINITY};
    step = imax(1, step);
    for (int v = v0; v <= v1; v += step)
        for (int u = u0; u <= u1; u += step) {
            const int x = cx + u, y = cy + v;
            if (!(x - r >= 0 && y - r >= 0 && x + r < W && y + r < H)) continue; /* subset_in_bounds :54-60 */
            const double z = shift_zncc(def, W, H, s, n, rm, rn, cx, cy, u, v);
            if (z > best.z) {
                best.u = u;
                best.v = v;
                best.z = z;
            }
        }
    return best;
}

/* search_ncorr_multigrid, :208-265 */
static orc_cand search_multigrid(const float *def, int W, int H, const orc_sample *s, int n, double rm, double rn,
                                 const orc_sample *cs_, int cn, double crm, double crn, int cx, int cy, int r,
                                 int coarse_step, int refine_radius)
{
    orc_cand none = {0, 0, -INFINITY};
    const int min_u = r - cx, max_u = W - 1 - r - cx;
    const int min_v = r - cy, max_v = H - 1 - r - cy;
    if (min_u > max_u || min_v > max_v) return none;
    coarse_step = imax(1, coarse_step);
    const orc_cand coarse =
        search_window(def, W, H, cs_, cn, crm, crn, cx, cy, r, min_u, max_u, min_v, max_v, coarse_step);
    if (!isfinite(coarse.z)) return coarse;
    const int rr = imax(refine_radius, coarse_step);
    return search_window(def, W, H, s, n, rm, rn, cx, cy, r, imax(min_u, coarse.u - rr), imin(max_u, coarse.u + rr),
                         imax(min_v, coarse.v - rr), imin(max_v, coarse.v + rr), 1);
}

/* prefer_local_when_ambiguous, :267-280 */
static orc_cand prefer_local(orc_cand global, orc_cand local)
{
    if (!isfinite(local.z)) return global;
    if (!isfinite(global.z)) return local;
    if (local.z + ORC_LOCAL_MARGIN >= global.z) return local;
    return global;
}

/* znssd_from_zncc, :22-28 */
static double znssd_from_zncc(double z)
{
    if (!isfinite(z)) return INFINITY;
    const double c = z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
    const double q = 2.0 * (1.0 - c);
    return q > 0.0 ? q : 0.0;
}

static void estimate_from(orc_cand best, orc_estimate *out)
{
    memset(out, 0, sizeof(*out));
    if (!isfinite(best.z)) return;
    out->u = (double)best.u;
    out->v = (double)best.v;
    out->confidence = znssd_from_zncc(best.z);
    out->valid = 1;
    out->zncc = best.z;
    out->znssd = out->confidence;
}

/* One point of IntegerSearchInitializer (constructor clamps :293-301). */
static void integer_search_one(const orc_search_params *p, const float *ref, const float *def, int W, int H,
                               const uint8_t *roi, double px, double py, const double *prior, orc_sample *s,
                               orc_sample *cs_, orc_estimate *out)
{
    memset(out, 0, sizeof(*out));
    const int r = imax(1, p->subset_radius);
    const int sr = imax(0, p->search_radius);
    const int scale = imax(1, p->pyramid_scale);
    const int refine_param = imax(1, p->pyramid_refinement_radius);
    const uint8_t *mroi = p->masked ? roi : NULL;
    /* a point or prior that is not a finite image coordinate is invalid (the
     * reference's integer conversion of it is undefined) */
    if (!(fabs(px) < 1e9 && fabs(py) < 1e9)) return;
    if (p->around_prior) {
        /* estimate_around_displacement :519-579 / _with_mask_ :581-645 */
        if (!(fabs(prior[0]) < 1e9 && fabs(prior[1]) < 1e9)) return;
        const int cx = (int)llround(px), cy = (int)llround(py);
        const int n = collect_samples(ref, W, H, mroi, cx, cy, r, 1, s);
        if (n < min_required(r, 1)) return;
        double rm, rn;
        sample_stats(s, n, &rm, &rn);
        if (rn <= ORC_SEARCH_EPS) return;
        const int pu = (int)llround(prior[0]), pv = (int)llround(prior[1]);
        const int min_u = r - cx, max_u = W - 1 - r - cx, min_v = r - cy, max_v = H - 1 - r - cy;
        const orc_cand best = search_window(def, W, H, s, n, rm, rn, cx, cy, r, imax(min_u, pu - sr),
                                            imin(max_u, pu + sr), imax(min_v, pv - sr), imin(max_v, pv + sr), 1);
        estimate_from(best, out);
        return;
    }
    const int cx = (int)round(px), cy = (int)round(py);
    int n, cn;
    double rm, rn, crm, crn;
    const int coarse_step = p->pyramid_enabled ? scale : 1;
    const int refine_radius = imax(refine_param, (int)ceil(1.5 * (double)r));
    orc_cand best = {0, 0, -INFINITY};
    if (!mroi) {
        /* estimate_with_interpolators :318-403 */
        if (!(cx - r >= 0 && cy - r >= 0 && cx + r < W && cy + r < H)) return;
        n = collect_samples(ref, W, H, NULL, cx, cy, r, 1, s);
        sample_stats(s, n, &rm, &rn);
        if (n <= 0 || rn <= ORC_SEARCH_EPS) return;
        cn = collect_samples(ref, W, H, NULL, cx, cy, r, coarse_step, cs_);
        sample_stats(cs_, cn, &crm, &crn);
        if (cn > 0 && crn > ORC_SEARCH_EPS)
            best = search_multigrid(def, W, H, s, n, rm, rn, cs_, cn, crm, crn, cx, cy, r, coarse_step, refine_radius);
    } else {
        /* estimate_with_mask_interpolators :425-517 */
        if (!(cx >= 0 && cy >= 0 && cx < W && cy < H) || !mroi[(size_t)cy * W + cx]) return;
        n = collect_samples(ref, W, H, mroi, cx, cy, r, 1, s);
        if (n < min_required(r, 1)) return;
        sample_stats(s, n, &rm, &rn);
        if (rn <= ORC_SEARCH_EPS) return;
        cn = collect_samples(ref, W, H, mroi, cx, cy, r, coarse_step, cs_);
        sample_stats(cs_, cn, &crm, &crn);
        if (cn >= min_required(r, coarse_step) && crn > ORC_SEARCH_EPS)
            best = search_multigrid(def, W, H, s, n, rm, rn, cs_, cn, crm, crn, cx, cy, r, coarse_step, refine_radius);
    }
    if (sr > 0) {
        const orc_cand local = search_window(def, W, H, s, n, rm, rn, cx, cy, r, -sr, sr, -sr, sr, 1);
        best = prefer_local(best, local);
    }
    estimate_from(best, out);
}

void orc_integer_search(const orc_search_params *p, const float *ref, const float *def, int W, int H,
                        const uint8_t *roi, const double *xy, const double *prior_uv, int n, orc_estimate *out,
                        int threads)
{
    const int r = imax(1, p->subset_radius);
    const size_t cap = (size_t)(2 * r + 1) * (size_t)(2 * r + 1);
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
    {
        orc_sample *s = (orc_sample *)malloc(cap * sizeof(orc_sample));
        orc_sample *c = (orc_sample *)malloc(cap * sizeof(orc_sample));
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < n; ++i)
            integer_search_one(p, ref, def, W, H, roi, xy[2 * i], xy[2 * i + 1],
                               prior_uv ? prior_uv + 2 * i : NULL, s, c, out + i);
        free(s);
        free(c);
    }
}

double orc_subset_zncc(const float *ref, const float *def, int W, int H, int r, int cx, int cy, int u, int v)
{
    orc_sample *s = (orc_sample *)malloc((size_t)(2 * r + 1) * (2 * r + 1) * sizeof(orc_sample));
    const int n = collect_samples(ref, W, H, NULL, cx, cy, r, 1, s);
    double z = -INFINITY;
    if (n > 0) {
        double rm, rn;
        sample_stats(s, n, &rm, &rn);
        z = shift_zncc(def, W, H, s, n, rm, rn, cx, cy, u, v);
    }
    free(s);
    return z;
}

/* ---- seed selection (seed_selector.cpp) ---- */

double orc_texture_std(const float *ref, int W, int H, int x, int y, int r)
{
    if (x - r < 0 || y - r < 0 || x + r >= W || y + r >= H) return 0.0; /* :425-430 */
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

/* deterministic_sample :34-48.  With limit == 1 the reference's step is
 * (n-1)/0 = inf, 0 * inf = NaN and llround(NaN) converts to an index past the
 * end, which min() clamps to the last point (x86-64). */
int orc_deterministic_sample(int n, int limit, int *idx)
{
    if (limit <= 0 || n <= limit) {
        for (int i = 0; i < n; ++i) idx[i] = i;
        return n;
    }
    if (limit == 1) {
        idx[0] = n - 1;
        return 1;
    }
    const double step = (double)(n - 1) / (double)(limit - 1);
    for (int i = 0; i < limit; ++i) {
        long long k = llround((double)i * step);
        idx[i] = k > n - 1 ? n - 1 : (int)k;
    }
    return limit;
}

static double dist2(double x0, double y0, double x1, double y1)
{
    const double dx = x0 - x1, dy = y0 - y1;
    return dx * dx + dy * dy;
}

/* initialize_centers (:50-68) + run_kmeans (:70-110) */
int orc_kmeans(const int *pts, int n, int requested, int iterations, double *c)
{
    if (n <= 0 || requested <= 0) return 0;
    const int k = imin(requested, n);
    const double step = k == 1 ? 0.0 : (double)(n - 1) / (double)(k - 1);
    for (int i = 0; i < k; ++i) {
        long long j = llround((double)i * step);
        if (j > n - 1) j = n - 1;
        c[2 * i] = (double)pts[2 * j];
        c[2 * i + 1] = (double)pts[2 * j + 1];
    }
    if (iterations <= 0) return k;
    double *sums = (double *)malloc(sizeof(double) * 2 * k);
    int *counts = (int *)malloc(sizeof(int) * k);
    for (int it = 0; it < iterations; ++it) {
        memset(sums, 0, sizeof(double) * 2 * k);
        memset(counts, 0, sizeof(int) * k);
        for (int p = 0; p < n; ++p) {
            const double px = (double)pts[2 * p], py = (double)pts[2 * p + 1];
            int best = 0;
            double bd = INFINITY;
            for (int i = 0; i < k; ++i) {
                const double d = dist2(px, py, c[2 * i], c[2 * i + 1]);
                if (d < bd) {
                    bd = d;
                    best = i;
                }
            }
            sums[2 * best] += px;
            sums[2 * best + 1] += py;
            ++counts[best];
        }
        for (int i = 0; i < k; ++i)
            if (counts[i] > 0) {
                c[2 * i] = sums[2 * i] / (double)counts[i];
                c[2 * i + 1] = sums[2 * i + 1] / (double)counts[i];
            }
    }
    free(sums);
    free(counts);
    return k;
}

static int seed_radius(const orc_seed_params *p)
{
    return imax(p->search.subset_radius, p->subpixel_radius);
}

/* is_candidate_margin_valid, :385-419 */
static int margin_valid(const orc_seed_params *p, const uint8_t *roi, int W, int H, int x, int y)
{
    const int radius = seed_radius(p);
    const int margin = radius + (p->truncate_roi_subsets ? p->search.search_radius : 0);
    if (x - margin < 0 || y - margin < 0 || x + margin >= W || y + margin >= H) return 0;
    if (!p->truncate_roi_subsets) return 1;
    int valid = 0, full = 0;
    for (int yy = y - radius; yy <= y + radius; ++yy)
        for (int xx = x - radius; xx <= x + radius; ++xx) {
            const int dx = xx - x, dy = yy - y;
            if (dx * dx + dy * dy <= radius * radius) {
                ++full;
                if (xx >= 0 && yy >= 0 && xx < W && yy < H && roi[(size_t)yy * W + xx]) ++valid;
            }
        }
    const int need = imax(p->min_valid_samples, (int)ceil(p->min_valid_sample_ratio * (double)full));
    return valid >= imax(6, need);
}

static int roi_valid(const uint8_t *roi, int W, int H, int x, int y)
{
    return x >= 0 && y >= 0 && x < W && y < H && roi[(size_t)y * W + x];
}

typedef struct {
    double d;
    int i;
} orc_dist_index;

static int cmp_dist_index(const void *a, const void *b)
{
    const orc_dist_index *x = (const orc_dist_index *)a, *y = (const orc_dist_index *)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

/* generate_candidates, :208-316.  The fallback's nearest-point order uses
 * std::sort (:293-298), which leaves the order of equally distant points
 * unspecified; here equal distances keep the sampled order. */
int orc_seed_candidates(const orc_seed_params *p, const float *ref, int W, int H, const uint8_t *roi,
                        double *xy_out, int cap)
{
    const int radius = seed_radius(p);
    const int margin = radius + (p->truncate_roi_subsets ? p->search.search_radius : 0);
    int *pts = (int *)malloc(sizeof(int) * 2 * (size_t)W * H);
    int n = 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            if (roi[(size_t)y * W + x] && x - margin >= 0 && y - margin >= 0 && x + margin < W && y + margin < H) {
                pts[2 * n] = x;
                pts[2 * n + 1] = y;
                ++n;
            }
    if (n == 0)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                if (roi[(size_t)y * W + x]) {
                    pts[2 * n] = x;
                    pts[2 * n + 1] = y;
                    ++n;
                }
    int count = 0;
    if (n == 0) {
        free(pts);
        return 0;
    }
    const int requested = imax(1, p->seed_count);
    const int limit = imax(1, p->kmeans_sample_limit);
    int *idx = (int *)malloc(sizeof(int) * (size_t)(n < limit ? n : limit));
    const int ns = orc_deterministic_sample(n, limit, idx);
    int *sp = (int *)malloc(sizeof(int) * 2 * (size_t)ns);
    for (int i = 0; i < ns; ++i) {
        sp[2 * i] = pts[2 * idx[i]];
        sp[2 * i + 1] = pts[2 * idx[i] + 1];
    }
    free(idx);
    free(pts);
    const int tr = p->search.subset_radius;
    if (requested == 1) {
        double best = -1.0;
        int found = 0, bx = 0, by = 0;
        for (int i = 0; i < ns; ++i) {
            if (!margin_valid(p, roi, W, H, sp[2 * i], sp[2 * i + 1])) continue;
            const double t = orc_texture_std(ref, W, H, sp[2 * i], sp[2 * i + 1], tr);
            if (!found || t > best) {
                best = t;
                bx = sp[2 * i];
                by = sp[2 * i + 1];
                found = 1;
            }
        }
        if (found && cap > 0) {
            xy_out[0] = (double)bx;
            xy_out[1] = (double)by;
            count = 1;
        }
        free(sp);
        return count;
    }
    double *centers = (double *)malloc(sizeof(double) * 2 * (size_t)imin(requested, ns));
    const int k = orc_kmeans(sp, ns, requested, imax(0, p->kmeans_iterations), centers);
    orc_dist_index *order = (orc_dist_index *)malloc(sizeof(orc_dist_index) * (size_t)ns);
    for (int c = 0; c < k && count < cap; ++c) {
        const double cx = centers[2 * c], cy = centers[2 * c + 1];
        const int rx = (int)llround(cx), ry = (int)llround(cy);
        if (roi_valid(roi, W, H, rx, ry) && margin_valid(p, roi, W, H, rx, ry) &&
            orc_texture_std(ref, W, H, rx, ry, tr) >= p->min_texture_std) {
            xy_out[2 * count] = (double)rx;
            xy_out[2 * count + 1] = (double)ry;
            ++count;
            continue;
        }
        for (int i = 0; i < ns; ++i) {
            order[i].d = dist2((double)sp[2 * i], (double)sp[2 * i + 1], cx, cy);
            order[i].i = i;
        }
        qsort(order, (size_t)ns, sizeof(orc_dist_index), cmp_dist_index);
        const int lim = imin(50, ns);
        double best = -1.0;
        int bi = order[0].i;
        for (int i = 0; i < lim; ++i) {
            const int j = order[i].i;
            const double t = orc_texture_std(ref, W, H, sp[2 * j], sp[2 * j + 1], tr);
            if (t > best) {
                best = t;
                bi = j;
            }
        }
        xy_out[2 * count] = (double)sp[2 * bi];
        xy_out[2 * count + 1] = (double)sp[2 * bi + 1];
        ++count;
    }
    free(order);
    free(centers);
    free(sp);
    return count;
}

/* quality_passes, :448-460 */
static int quality_passes(const orc_seed_params *p, double q)
{
    if (p->quality_metric == 0) return q >= p->min_zncc;
    if (p->quality_metric == 2) return q <= p->max_ssd;
    return q <= p->max_znssd;
}

/* seed_is_better, :112-136 */
static int seed_is_better(const orc_seed_eval *a, const orc_seed_eval *cur, int metric)
{
    const double eps = 1.0e-9;
    if (!cur->valid) return 1;
    if (metric == 0) {
        if (a->quality > cur->quality + eps) return 1;
        if (a->quality < cur->quality - eps) return 0;
    } else {
        if (a->quality < cur->quality - eps) return 1;
        if (a->quality > cur->quality + eps) return 0;
    }
    return a->displacement_norm > cur->displacement_norm;
}

/* zncc_from_znssd, subset_initializer.cpp:13-19 */
static double zncc_from_znssd(double q)
{
    if (!isfinite(q)) return -1.0;
    const double z = 1.0 - 0.5 * q;
    return z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
}

/* select_best_seed (:146-206) with evaluate_candidate (:466-539) on the
 * integer-search path (no SIFT matches) and SubsetInitializer's refinement
 * (subset_initializer.cpp:128-264): IC-GN on the non-prefiltered lazy spline
 * of the deformed image with central-difference reference gradients, or
 * forward Gauss-Newton on the configured spline. */
int orc_select_seed(const orc_seed_params *p, const float *ref, const float *def, int W, int H, const uint8_t *roi,
                    orc_seed_eval *out, int cap, int *n_out)
{
    *n_out = 0;
    const int kmax = imax(1, p->seed_count);
    double *xy = (double *)malloc(sizeof(double) * 2 * (size_t)kmax);
    const int k = orc_seed_candidates(p, ref, W, H, roi, xy, imin(cap, kmax));
    orc_search_params sp = p->search;
    sp.around_prior = 0;
    sp.masked = p->truncate_roi_subsets ? 1 : 0;
    orc_estimate *est = (orc_estimate *)calloc((size_t)(k > 0 ? k : 1), sizeof(orc_estimate));
    orc_integer_search(&sp, ref, def, W, H, roi, xy, NULL, k, est, 0);
    if (p->subpixel_enabled && k > 0) {
        orc_config cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.subset_radius = imax(1, p->subpixel_radius);
        cfg.max_iterations = imax(0, p->subpixel_max_iterations);
        cfg.convergence_threshold = p->subpixel_convergence_threshold;
        cfg.second_order = p->subpixel_shape_order == 2;
        cfg.objective = p->subpixel_objective;
        cfg.min_valid_samples = p->min_valid_samples;
        cfg.min_valid_sample_ratio = p->min_valid_sample_ratio;
        orc_precomputed pre;
        const int fgn = p->subpixel_optimizer == 1;
        /* IC-GN: compute_lazy with use_exact_prefilter = false (seed_selector.cpp:163-173);
         * forward Gauss-Newton: its own interpolator on the configured spline (forward_gauss_newton.cpp:98) */
        orc_precompute(def, W, H, p->degree, p->border, fgn ? 1 : 0, 0, &pre);
        double *gx = NULL, *gy = NULL;
        if (!fgn) {
            gx = (double *)malloc(sizeof(double) * (size_t)W * H);
            gy = (double *)malloc(sizeof(double) * (size_t)W * H);
            orc_reference_gradient_cache(ref, W, H, gx, gy); /* central_difference_gradient, icgn.cpp:42-53 */
        }
        for (int i = 0; i < k; ++i) {
            if (!est[i].valid) continue;
            const orc_initial ini = {est[i].u, est[i].v, est[i].du_dx, est[i].du_dy, est[i].dv_dx, est[i].dv_dy,
                                     est[i].confidence};
            orc_result res;
            if (fgn)
                orc_fgn_solve(&cfg, ref, W, H, &pre, NULL, xy[2 * i], xy[2 * i + 1], &ini, &res);
            else
                orc_icgn_solve(&cfg, ref, NULL, W, H, gx, gy, &pre, p->truncate_roi_subsets ? roi : NULL,
                               xy[2 * i], xy[2 * i + 1], &ini, &res);
            if (res.valid && res.status == ORC_SUCCESS) {
                est[i].u = res.p[0];
                est[i].v = res.p[1];
                est[i].du_dx = res.p[2];
                est[i].du_dy = res.p[3];
                est[i].dv_dx = res.p[4];
                est[i].dv_dy = res.p[5];
                est[i].confidence = res.correlation;
                est[i].zncc = zncc_from_znssd(res.correlation);
                est[i].znssd = res.correlation;
            }
        }
        free(gx);
        free(gy);
        orc_precomputed_free(&pre);
    }
    int best = -1;
    for (int i = 0; i < k; ++i) {
        orc_seed_eval *e = out + i;
        memset(e, 0, sizeof(*e));
        e->x = xy[2 * i];
        e->y = xy[2 * i + 1];
        e->displacement.status = ORC_INVALID_INPUT; /* Displacement2D defaults, result.hpp:32-51 */
        const orc_estimate *q = est + i;
        if (q->valid && isfinite(q->u) && isfinite(q->v) && isfinite(q->confidence)) {
            e->displacement.x = e->x;
            e->displacement.y = e->y;
            e->displacement.p[0] = q->u;
            e->displacement.p[1] = q->v;
            e->displacement.p[2] = q->du_dx;
            e->displacement.p[3] = q->du_dy;
            e->displacement.p[4] = q->dv_dx;
            e->displacement.p[5] = q->dv_dy;
            e->displacement.correlation = q->confidence;
            e->displacement.status = ORC_SUCCESS;
            e->displacement.valid = 1;
            e->quality = p->quality_metric == 0 ? q->zncc : q->znssd;
            e->displacement_norm = hypot(q->u, q->v);
            e->quality_passed = quality_passes(p, e->quality);
            e->valid = 1;
        }
        if (e->valid && e->quality_passed && e->displacement_norm >= p->min_displacement_norm)
            if (best < 0 || seed_is_better(e, out + best, p->quality_metric)) best = i;
    }
    *n_out = k;
    free(est);
    free(xy);
    return best;
}
