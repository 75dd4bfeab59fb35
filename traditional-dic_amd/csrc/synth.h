/* synth.h -- seeded synthetic speckle images (workload generator). */
#ifndef DIC_SYNTH_H
#define DIC_SYNTH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SYNTH_IDENTITY = 0, SYNTH_AFFINE = 1, SYNTH_SINUSOID = 2, SYNTH_AFFINE_SINUSOID = 3 };

typedef struct synth_speckle {
    int32_t nblobs;
    int32_t reserved;
    uint64_t seed;
    double background, smin, smax, amin, amax;
} synth_speckle;

/* Displacement field u(X), v(X) of the deformed frame.
 * AFFINE:   u = u0 + ux (X-xc) + uy (Y-yc),  v = v0 + vx (X-xc) + vy (Y-yc)
 * SINUSOID: u = u0 sin(2 pi X / ux),          v = v0 cos(2 pi Y / vy)
 * AFFINE_SINUSOID: the affine field plus au sin(2 pi X / lu), av cos(2 pi Y / lv) */
typedef struct synth_field {
    int32_t kind;
    int32_t reserved;
    double u0, v0, ux, uy, vx, vy, xc, yc;
    double au, lu, av, lv;
} synth_field;

int synth_render(int W, int H, const synth_speckle *sp, const synth_field *field, float *out);
void synth_displacement(const synth_field *f, double X, double Y, double *u, double *v, double *grad4);
void synth_add_noise(float *img, size_t n, double sigma, uint64_t seed, uint64_t stream);
void synth_quantize_unit(float *img, size_t n);
void synth_scale(float *img, size_t n, float s);
uint32_t synth_pcg32_first(uint64_t seed, uint64_t stream);

/* Stereo pair of a speckled surface seen by two pinhole cameras.  World frame
 * in mm: cameras at (-baseline/2, 0, 0) (camera 0, left) and (+baseline/2, 0, 0)
 * (camera 1, right), each turned by converge_deg/2 about the y axis towards
 * the other's side; the surface is z = distance - bulge_amp exp(-(x^2+y^2) /
 * (2 bulge_sigma^2)) (the bulge comes towards the cameras).  Its texture is a
 * speckle pattern in surface (x, y) coordinates: blobs with 1/e radius
 * U(smin, smax) mm and peak U(amin, amax) over `background`, centres uniform in
 * [x0, x1] x [y0, y1] mm.  Pixel (u, v) of a camera sees the surface point hit
 * by its ray (pinhole: f, cx, cy in px; camera y = world y). */
typedef struct synth_stereo {
    double f, cx, cy;
    double baseline, converge_deg, distance;
    double bulge_amp, bulge_sigma;
    double x0, x1, y0, y1;
    int32_t nblobs, reserved;
    uint64_t seed;
    double background, smin, smax, amin, amax;
} synth_stereo;

int synth_render_stereo(int W, int H, const synth_stereo *st, int camera, float *out);
/* Surface point (with or without the bulge) seen by pixel (u, v) of a camera. */
void synth_stereo_pixel_to_world(const synth_stereo *st, int camera, double u, double v, int with_bulge, double *X);
void synth_stereo_world_to_pixel(const synth_stereo *st, int camera, const double *X, double *u, double *v);

#ifdef __cplusplus
}
#endif
#endif
