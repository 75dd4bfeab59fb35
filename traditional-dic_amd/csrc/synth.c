/*
 * synth.c -- seeded synthetic speckle images for the correlation workloads.
 *
 * I(X) = background + sum_k A_k exp(-|X - c_k|^2 / s_k^2), blobs drawn from a
 * PCG32 stream (O'Neill 2014): c_k ~ U(-8, W+8) x U(-8, H+8), s_k ~ U(smin, smax)
 * (s_k is the 1/e radius), A_k ~ U(amin, amax).  A blob is evaluated out to
 * 5 s_k (the dropped tail is below 2e-9 grey levels).
 *
 * The deformed image is the analytic render I_def(x) = I(X(x)) with X solving
 * x = X + u(X): closed form for an affine field, 20 fixed-point steps for
 * the sinusoidal field, and a closed form for the pinhole plane pair.
 * Summation per pixel runs in blob-index order, so output is bit-reproducible
 * for any thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "synth.h"

typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg32_next(pcg32 *r)
{
    const uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

static void pcg32_seed(pcg32 *r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0U;
    r->inc = (initseq << 1u) | 1u;
    pcg32_next(r);
    r->state += initstate;
    pcg32_next(r);
}

static double pcg32_double(pcg32 *r)
{
    const uint32_t a = pcg32_next(r) >> 5, b = pcg32_next(r) >> 6;
    return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
}

static double pcg32_normal(pcg32 *r)
{
    /* Box-Muller, one value per call */
    double u1 = pcg32_double(r);
    const double u2 = pcg32_double(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct { double cx, cy, s, a; } blob;

static blob *make_blobs(const synth_speckle *sp, int W, int H)
{
    blob *bl = (blob *)malloc(sizeof(blob) * (size_t)(sp->nblobs > 0 ? sp->nblobs : 1));
    pcg32 r;
    pcg32_seed(&r, sp->seed, 0x5eedULL);
    for (int k = 0; k < sp->nblobs; ++k) {
        bl[k].cx = -8.0 + (W + 16.0) * pcg32_double(&r);
        bl[k].cy = -8.0 + (H + 16.0) * pcg32_double(&r);
        bl[k].s = sp->smin + (sp->smax - sp->smin) * pcg32_double(&r);
        bl[k].a = sp->amin + (sp->amax - sp->amin) * pcg32_double(&r);
    }
    return bl;
}

/* reference-frame point X for a deformed-frame pixel x */
static void inverse_map(const synth_field *f, double x, double y, double *X, double *Y)
{
    if (f == NULL || f->kind == SYNTH_IDENTITY) {
        *X = x;
        *Y = y;
        return;
    }
    if (f->kind == SYNTH_AFFINE) {
        /* x = X + u0 + ux (X-xc) + uy (Y-yc); y = Y + v0 + vx (X-xc) + vy (Y-yc) */
        const double a = 1.0 + f->ux, b = f->uy, c = f->vx, d = 1.0 + f->vy;
        const double rx = x - f->u0 - f->xc, ry = y - f->v0 - f->yc;
        const double det = a * d - b * c;
        *X = f->xc + (d * rx - b * ry) / det;
        *Y = f->yc + (-c * rx + a * ry) / det;
        return;
    }
    if (f->kind == SYNTH_SINUSOID) {
        /* u = au sin(2 pi X / lu), v = av cos(2 pi Y / lv) */
        double Xc = x, Yc = y;
        for (int i = 0; i < 20; ++i) {
            Xc = x - f->u0 * sin(6.283185307179586 * Xc / f->ux);
            Yc = y - f->v0 * cos(6.283185307179586 * Yc / f->vy);
        }
        *X = Xc;
        *Y = Yc;
        return;
    }
    if (f->kind == SYNTH_AFFINE_SINUSOID) {
        /* affine part (u0,v0,ux,uy,vx,vy about xc,yc) plus a sinusoid of
         * amplitude (au, av) and wavelengths (lu, lv); fixed point */
        double Xc = x, Yc = y;
        for (int i = 0; i < 30; ++i) {
            const double u = f->u0 + f->ux * (Xc - f->xc) + f->uy * (Yc - f->yc) +
                             f->au * sin(6.283185307179586 * Xc / f->lu);
            const double v = f->v0 + f->vx * (Xc - f->xc) + f->vy * (Yc - f->yc) +
                             f->av * cos(6.283185307179586 * Yc / f->lv);
            Xc = x - u;
            Yc = y - v;
        }
        *X = Xc;
        *Y = Yc;
        return;
    }
    *X = x;
    *Y = y;
}

void synth_displacement(const synth_field *f, double X, double Y, double *u, double *v, double *grad4)
{
    double uu = 0, vv = 0, g[4] = {0, 0, 0, 0};
    if (f != NULL && f->kind == SYNTH_AFFINE) {
        uu = f->u0 + f->ux * (X - f->xc) + f->uy * (Y - f->yc);
        vv = f->v0 + f->vx * (X - f->xc) + f->vy * (Y - f->yc);
        g[0] = f->ux;
        g[1] = f->uy;
        g[2] = f->vx;
        g[3] = f->vy;
    } else if (f != NULL && f->kind == SYNTH_SINUSOID) {
        const double k1 = 6.283185307179586 / f->ux, k2 = 6.283185307179586 / f->vy;
        uu = f->u0 * sin(k1 * X);
        vv = f->v0 * cos(k2 * Y);
        g[0] = f->u0 * k1 * cos(k1 * X);
        g[3] = -f->v0 * k2 * sin(k2 * Y);
    } else if (f != NULL && f->kind == SYNTH_AFFINE_SINUSOID) {
        const double k1 = 6.283185307179586 / f->lu, k2 = 6.283185307179586 / f->lv;
        uu = f->u0 + f->ux * (X - f->xc) + f->uy * (Y - f->yc) + f->au * sin(k1 * X);
        vv = f->v0 + f->vx * (X - f->xc) + f->vy * (Y - f->yc) + f->av * cos(k2 * Y);
        g[0] = f->ux + f->au * k1 * cos(k1 * X);
        g[1] = f->uy;
        g[2] = f->vx;
        g[3] = f->vy - f->av * k2 * sin(k2 * Y);
    }
    if (u) *u = uu;
    if (v) *v = vv;
    if (grad4) memcpy(grad4, g, sizeof(g));
}

int synth_render(int W, int H, const synth_speckle *sp, const synth_field *field, float *out)
{
    if (W <= 0 || H <= 0 || sp == NULL || out == NULL) return -1;
    blob *bl = make_blobs(sp, W, H);
    const size_t npx = (size_t)W * H;
    double *acc = (double *)malloc(sizeof(double) * npx);
    double *mx = (double *)malloc(sizeof(double) * npx);
    double *my = (double *)malloc(sizeof(double) * npx);
    if (!acc || !mx || !my) {
        free(bl);
        free(acc);
        free(mx);
        free(my);
        return -2;
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            inverse_map(field, (double)x, (double)y, &mx[(size_t)y * W + x], &my[(size_t)y * W + x]);
            acc[(size_t)y * W + x] = sp->background;
        }
    /* the forward displacement bounds the window where a blob lands */
    double maxdisp = 0.0;
    for (int y = 0; y < H; y += 8)
        for (int x = 0; x < W; x += 8) {
            const double dx = x - mx[(size_t)y * W + x], dy = y - my[(size_t)y * W + x];
            const double d = sqrt(dx * dx + dy * dy);
            if (d > maxdisp) maxdisp = d;
        }
    const double margin = 2.0 + 0.02 * maxdisp;
    const int band = 16;
    const int nb = (H + band - 1) / band;
#pragma omp parallel for schedule(dynamic, 1)
    for (int bi = 0; bi < nb; ++bi) {
        const int y0 = bi * band, y1 = (y0 + band < H) ? y0 + band : H;
        for (int k = 0; k < sp->nblobs; ++k) {
            const blob b = bl[k];
            double fx, fy; /* approximate landing point of the blob centre */
            double uu = 0.0, vv = 0.0;
            if (field != NULL && field->kind != SYNTH_IDENTITY) synth_displacement(field, b.cx, b.cy, &uu, &vv, NULL);
            fx = b.cx + uu;
            fy = b.cy + vv;
            const double rad = 5.0 * b.s * 1.05 + margin;
            int ya = (int)floor(fy - rad), yb = (int)ceil(fy + rad);
            if (yb < y0 || ya >= y1) continue;
            if (ya < y0) ya = y0;
            if (yb > y1 - 1) yb = y1 - 1;
            int xa = (int)floor(fx - rad), xb = (int)ceil(fx + rad);
            if (xa < 0) xa = 0;
            if (xb > W - 1) xb = W - 1;
            const double inv_s2 = 1.0 / (b.s * b.s), cut = 25.0;
            for (int y = ya; y <= yb; ++y)
                for (int x = xa; x <= xb; ++x) {
                    const size_t i = (size_t)y * W + x;
                    const double ddx = mx[i] - b.cx, ddy = my[i] - b.cy;
                    const double q = (ddx * ddx + ddy * ddy) * inv_s2;
                    if (q <= cut) acc[i] += b.a * exp(-q);
                }
        }
    }
    for (size_t i = 0; i < npx; ++i) out[i] = (float)acc[i];
    free(bl);
    free(acc);
    free(mx);
    free(my);
    return 0;
}

void synth_add_noise(float *img, size_t n, double sigma, uint64_t seed, uint64_t stream)
{
    if (sigma <= 0.0) return;
    pcg32 r;
    pcg32_seed(&r, seed, stream);
    for (size_t i = 0; i < n; ++i) img[i] = (float)((double)img[i] + sigma * pcg32_normal(&r));
}

void synth_quantize_unit(float *img, size_t n)
{
    /* clip to [0,255], round to 8 bit, then divide by the maximum in float32
     * (the read_gray convention of workflows/subset_2d.py:21-27) */
    float mx = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        double v = img[i];
        if (v < 0.0) v = 0.0;
        if (v > 255.0) v = 255.0;
        img[i] = (float)floor(v + 0.5);
        if (img[i] > mx) mx = img[i];
    }
    if (mx > 0.0f)
        for (size_t i = 0; i < n; ++i) img[i] = img[i] / mx;
}

void synth_scale(float *img, size_t n, float s)
{
    for (size_t i = 0; i < n; ++i) img[i] *= s;
}

uint32_t synth_pcg32_first(uint64_t seed, uint64_t stream)
{
    pcg32 r;
    pcg32_seed(&r, seed, stream);
    return pcg32_next(&r);
}

/* ---- stereo pair of a speckled surface (pinhole cameras) ---- */

static void stereo_pose(const synth_stereo *st, int cam, double C[3], double R[3][3])
{
    const double th = 0.5 * st->converge_deg * 3.141592653589793 / 180.0;
    const double phi = cam == 0 ? -th : th; /* R = R_y(phi) maps the optical axis to +z */
    C[0] = cam == 0 ? -0.5 * st->baseline : 0.5 * st->baseline;
    C[1] = 0.0;
    C[2] = 0.0;
    const double c = cos(phi), s = sin(phi);
    R[0][0] = c, R[0][1] = 0.0, R[0][2] = s;
    R[1][0] = 0.0, R[1][1] = 1.0, R[1][2] = 0.0;
    R[2][0] = -s, R[2][1] = 0.0, R[2][2] = c;
}

static double stereo_height(const synth_stereo *st, double x, double y)
{
    const double s2 = st->bulge_sigma * st->bulge_sigma;
    return st->bulge_amp * exp(-(x * x + y * y) / (2.0 * s2));
}

void synth_stereo_pixel_to_world(const synth_stereo *st, int cam, double u, double v, int with_bulge, double *X)
{
    double C[3], R[3][3];
    stereo_pose(st, cam, C, R);
    const double dc[3] = {(u - st->cx) / st->f, (v - st->cy) / st->f, 1.0};
    double d[3]; /* R^T dc */
    for (int i = 0; i < 3; ++i) d[i] = R[0][i] * dc[0] + R[1][i] * dc[1] + R[2][i] * dc[2];
    double h = 0.0, t = 0.0;
    for (int it = 0; it < (with_bulge ? 8 : 1); ++it) {
        t = (st->distance - h - C[2]) / d[2];
        if (with_bulge) h = stereo_height(st, C[0] + t * d[0], C[1] + t * d[1]);
    }
    t = (st->distance - h - C[2]) / d[2];
    for (int i = 0; i < 3; ++i) X[i] = C[i] + t * d[i];
}

void synth_stereo_world_to_pixel(const synth_stereo *st, int cam, const double *X, double *u, double *v)
{
    double C[3], R[3][3];
    stereo_pose(st, cam, C, R);
    double q[3];
    for (int i = 0; i < 3; ++i) q[i] = R[i][0] * (X[0] - C[0]) + R[i][1] * (X[1] - C[1]) + R[i][2] * (X[2] - C[2]);
    *u = st->f * q[0] / q[2] + st->cx;
    *v = st->f * q[1] / q[2] + st->cy;
}

int synth_render_stereo(int W, int H, const synth_stereo *st, int cam, float *out)
{
    if (W <= 0 || H <= 0 || st == NULL || out == NULL || st->nblobs < 0) return -1;
    const size_t npx = (size_t)W * H;
    double *sx = (double *)malloc(sizeof(double) * npx), *sy = (double *)malloc(sizeof(double) * npx);
    blob *bl = (blob *)malloc(sizeof(blob) * (size_t)(st->nblobs > 0 ? st->nblobs : 1));
    if (!sx || !sy || !bl) {
        free(sx);
        free(sy);
        free(bl);
        return -2;
    }
    pcg32 r;
    pcg32_seed(&r, st->seed, 0x57e0ULL);
    for (int k = 0; k < st->nblobs; ++k) {
        bl[k].cx = st->x0 + (st->x1 - st->x0) * pcg32_double(&r);
        bl[k].cy = st->y0 + (st->y1 - st->y0) * pcg32_double(&r);
        bl[k].s = st->smin + (st->smax - st->smin) * pcg32_double(&r);
        bl[k].a = st->amin + (st->amax - st->amin) * pcg32_double(&r);
    }
    /* surface coordinates seen by every pixel */
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double X[3];
            synth_stereo_pixel_to_world(st, cam, (double)x, (double)y, 1, X);
            sx[(size_t)y * W + x] = X[0];
            sy[(size_t)y * W + x] = X[1];
        }
    /* blobs binned on the surface grid; each pixel sums the blobs of the
     * 3 x 3 cells around its surface point, in blob-index order */
    const double cell = 5.0 * st->smax;
    const int gw = (int)ceil((st->x1 - st->x0) / cell) + 1, gh = (int)ceil((st->y1 - st->y0) / cell) + 1;
    int *count = (int *)calloc((size_t)gw * gh + 1, sizeof(int));
    int *idx = (int *)malloc(sizeof(int) * (size_t)(st->nblobs > 0 ? st->nblobs : 1));
    if (!count || !idx) {
        free(sx);
        free(sy);
        free(bl);
        free(count);
        free(idx);
        return -2;
    }
    for (int k = 0; k < st->nblobs; ++k) {
        const int gx = (int)((bl[k].cx - st->x0) / cell), gy = (int)((bl[k].cy - st->y0) / cell);
        count[gy * gw + gx + 1]++;
    }
    for (int i = 0; i < gw * gh; ++i) count[i + 1] += count[i];
    int *fill = (int *)malloc(sizeof(int) * (size_t)gw * gh);
    memcpy(fill, count, sizeof(int) * (size_t)gw * gh);
    for (int k = 0; k < st->nblobs; ++k) {
        const int gx = (int)((bl[k].cx - st->x0) / cell), gy = (int)((bl[k].cy - st->y0) / cell);
        idx[fill[gy * gw + gx]++] = k;
    }
    free(fill);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const double px = sx[i], py = sy[i];
            const int gx = (int)floor((px - st->x0) / cell), gy = (int)floor((py - st->y0) / cell);
            double acc = st->background;
            for (int cy = gy - 1; cy <= gy + 1; ++cy) {
                if (cy < 0 || cy >= gh) continue;
                for (int cx = gx - 1; cx <= gx + 1; ++cx) {
                    if (cx < 0 || cx >= gw) continue;
                    const int c = cy * gw + cx;
                    for (int j = count[c]; j < count[c + 1]; ++j) {
                        const blob b = bl[idx[j]];
                        const double ddx = px - b.cx, ddy = py - b.cy;
                        const double q = (ddx * ddx + ddy * ddy) / (b.s * b.s);
                        if (q <= 25.0) acc += b.a * exp(-q);
                    }
                }
            }
            out[i] = (float)acc;
        }
    free(count);
    free(idx);
    free(sx);
    free(sy);
    free(bl);
    return 0;
}
