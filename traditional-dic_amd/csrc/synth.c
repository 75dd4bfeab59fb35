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
