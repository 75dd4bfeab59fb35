// dic_abi.hip -- extern "C" entry points of libdic_hip.so (declared in
// include/dic_hip.h).  Host code only: validation, handle ownership, HBM
// accounting and kernel dispatch.  No exception crosses the boundary.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "dic_internal.hpp"

namespace dic {

int set_error(dic_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->error = msg;
    return code;
}

int check_hip(dic_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess) return DIC_OK;
    // a failed runtime call also leaves its code as the thread's last error;
    // consume it so that the next launch's hipGetLastError does not report it
    (void)hipGetLastError();
    return set_error(ctx, DIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t solve_dynamic_lds(bool fgn, int precision, int objective, int nfull)
{
    if (fgn) return 0; // FGN keeps no per-sample values
    const bool twopass = precision == DIC_PRECISION_FP64 && objective == DIC_OBJECTIVE_ZNSSD;
    return twopass ? sizeof(double) * (size_t)nfull : 0;
}

int check_prefilter_size(dic_ctx *ctx, int W, int H, int border, int precision, const char *who)
{
    const int b = border < 3 ? 3 : border;
    const size_t PW = (size_t)W + 2 * (size_t)b, PH = (size_t)H + 2 * (size_t)b;
    if (PW * PH * sizeof(double) >= ((size_t)1 << 31))
        return set_error(ctx, DIC_ERR_UNSUPPORTED,
                         std::string(who) + ": padded coefficient image of 2 GiB or more (" + std::to_string(PW) +
                             " x " + std::to_string(PH) + ")");
    const size_t need = prefilter_row_lds((int)PW, precision);
    if (ctx->max_lds != 0 && need > ctx->max_lds)
        return set_error(ctx, DIC_ERR_UNSUPPORTED,
                         std::string(who) + ": a padded row of " + std::to_string(PW) + " samples needs " +
                             std::to_string(need) + " bytes of LDS per workgroup; the device has " +
                             std::to_string(ctx->max_lds));
    return DIC_OK;
}

int check_solve_lds(dic_ctx *ctx, bool fgn, int precision, int objective, int radius, int nfull)
{
    const size_t need = solve_static_lds() + solve_dynamic_lds(fgn, precision, objective, nfull);
    if (ctx->max_lds == 0 || need <= ctx->max_lds) return DIC_OK;
    return set_error(ctx, DIC_ERR_UNSUPPORTED,
                     "subset radius " + std::to_string(radius) + " (" + std::to_string(nfull) + " samples) needs " +
                         std::to_string(need) + " bytes of LDS per POI with " + (fgn ? "FGN" : "IC-GN") + " in " +
                         (precision == DIC_PRECISION_FP64 ? "FP64" : "mixed precision") + "; the device allows " +
                         std::to_string(ctx->max_lds));
}

hipError_t ctx_malloc(dic_ctx *ctx, void **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    if (ctx->budget != 0 && ctx->in_use.load() + bytes > ctx->budget) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipSuccess) ctx->in_use += bytes;
    return e;
}

void ctx_free(dic_ctx *ctx, void *ptr, size_t bytes)
{
    if (!ptr) return;
    if (bytes == 0) bytes = 1;
    (void)hipFree(ptr);
    ctx->in_use -= bytes;
}

// Circular subset in the reference's order: dy outer, dx inner,
// dx^2 + dy^2 <= r^2 (icgn.cpp:485-489); entries packed (dy << 16) | (dx & 0xffff).
int get_subset_table(dic_ctx *ctx, int radius, const int **table, int *count)
{
    std::lock_guard<std::mutex> lock(ctx->table_mutex);
    auto it = ctx->subset_tables.find(radius);
    if (it != ctx->subset_tables.end()) {
        *table = it->second.first;
        *count = it->second.second;
        return DIC_OK;
    }
    std::vector<int> host;
    if (radius >= 1) {
        if (radius > 2000) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "subset_radius too large (max 2000)");
        for (int dy = -radius; dy <= radius; ++dy)
            for (int dx = -radius; dx <= radius; ++dx)
                if (dx * dx + dy * dy <= radius * radius) host.push_back((int)(((unsigned)dy << 16) | (dx & 0xffff)));
    }
    int *dev = nullptr;
    if (!host.empty()) {
        hipError_t e = ctx_malloc(ctx, (void **)&dev, sizeof(int) * host.size());
        if (e != hipSuccess) return check_hip(ctx, e, "subset table allocation");
        e = hipMemcpy(dev, host.data(), sizeof(int) * host.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return check_hip(ctx, e, "subset table upload");
    }
    ctx->subset_tables[radius] = {dev, (int)host.size()};
    *table = dev;
    *count = (int)host.size();
    return DIC_OK;
}

static int ensure_staging(dic_ctx *ctx, size_t bytes)
{
    if (ctx->staging_bytes >= bytes) return DIC_OK;
    if (ctx->staging) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return check_hip(ctx, e, "stream synchronize");
        ctx_free(ctx, ctx->staging, ctx->staging_bytes);
        ctx->staging = nullptr;
        ctx->staging_bytes = 0;
    }
    hipError_t e = ctx_malloc(ctx, &ctx->staging, bytes);
    if (e == hipErrorOutOfMemory && ctx->budget != 0)
        return set_error(ctx, DIC_ERR_BUDGET, "HBM budget exceeded by staging buffer");
    if (e != hipSuccess) return check_hip(ctx, e, "staging allocation");
    ctx->staging_bytes = bytes;
    return DIC_OK;
}

static int alloc_checked(dic_ctx *ctx, void **p, size_t bytes, const char *what)
{
    hipError_t e = ctx_malloc(ctx, p, bytes);
    if (e == hipErrorOutOfMemory && ctx->budget != 0)
        return set_error(ctx, DIC_ERR_BUDGET, std::string("HBM budget exceeded: ") + what);
    return check_hip(ctx, e, what);
}

static bool valid_degree(int d) { return d == 1 || d == 3 || d == 5; }

} // namespace dic

using namespace dic;

extern "C" {

const char *dic_version(void) { return "dic_hip 0.1.0 gfx950"; }

int dic_device_count(int *count)
{
    if (!count) return DIC_ERR_INVALID_ARGUMENT;
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return DIC_ERR_NO_DEVICE;
    }
    return DIC_OK;
}

int dic_ctx_create(int device, uint64_t hbm_budget_bytes, dic_ctx **out)
{
    if (!out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DIC_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return DIC_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return DIC_ERR_HIP;
    dic_ctx *ctx = new (std::nothrow) dic_ctx();
    if (!ctx) return DIC_ERR_HIP;
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && max_lds > 0)
        ctx->max_lds = (size_t)max_lds;
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return DIC_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    if (noalloc_setup(ctx) != DIC_OK) { // the device address of the no-allocation entry's table
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return DIC_ERR_HIP;
    }
    *out = ctx;
    return DIC_OK;
}

void dic_ctx_destroy(dic_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->subset_tables) ctx_free(ctx, kv.second.first, sizeof(int) * (size_t)kv.second.second);
    for (auto &kv : ctx->buffers) ctx_free(ctx, kv.first, kv.second);
    if (ctx->staging) ctx_free(ctx, ctx->staging, ctx->staging_bytes);
    destroy_stager(ctx);
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char *dic_last_error(const dic_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->error.c_str();
}

int dic_ctx_set_stream(dic_ctx *ctx, void *hip_stream)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return DIC_OK;
}

void *dic_ctx_stream(dic_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int dic_ctx_synchronize(dic_ctx *ctx)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "stream synchronize");
}

uint64_t dic_ctx_bytes_in_use(const dic_ctx *ctx) { return ctx ? ctx->in_use.load() : 0; }

int dic_buffer_alloc(dic_ctx *ctx, size_t bytes, void **device_ptr)
{
    if (!ctx || !device_ptr) return DIC_ERR_INVALID_ARGUMENT;
    *device_ptr = nullptr;
    int rc = alloc_checked(ctx, device_ptr, bytes, "buffer allocation");
    if (rc != DIC_OK) return rc;
    std::lock_guard<std::mutex> lock(ctx->table_mutex);
    ctx->buffers[*device_ptr] = bytes;
    return DIC_OK;
}

void dic_buffer_free(dic_ctx *ctx, void *device_ptr)
{
    if (!ctx || !device_ptr) return;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lock(ctx->table_mutex);
        auto it = ctx->buffers.find(device_ptr);
        if (it == ctx->buffers.end()) return;
        bytes = it->second;
        ctx->buffers.erase(it);
    }
    (void)hipStreamSynchronize(ctx->stream);
    ctx_free(ctx, device_ptr, bytes);
}

int dic_buffer_upload(dic_ctx *ctx, void *device_dst, const void *host_src, size_t bytes)
{
    if (!ctx || (bytes && (!device_dst || !host_src))) return DIC_ERR_INVALID_ARGUMENT;
    int rc = check_hip(ctx, upload_host(ctx, device_dst, host_src, bytes, ctx->stream), "buffer upload");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "buffer upload sync");
    return rc;
}

int dic_buffer_download(dic_ctx *ctx, void *host_dst, const void *device_src, size_t bytes)
{
    if (!ctx || (bytes && (!device_src || !host_dst))) return DIC_ERR_INVALID_ARGUMENT;
    int rc = check_hip(ctx, download_host(ctx, host_dst, device_src, bytes, ctx->stream), "buffer download");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "buffer download sync");
    return rc;
}

int dic_buffer_copy(dic_ctx *ctx, void *device_dst, const void *device_src, size_t bytes)
{
    if (!ctx || (bytes && (!device_src || !device_dst))) return DIC_ERR_INVALID_ARGUMENT;
    return check_hip(ctx, hipMemcpyAsync(device_dst, device_src, bytes, hipMemcpyDeviceToDevice, ctx->stream),
                     "buffer copy");
}

int dic_event_record(dic_ctx *ctx, void **event)
{
    if (!ctx || !event) return DIC_ERR_INVALID_ARGUMENT;
    hipEvent_t e = nullptr;
    int rc = check_hip(ctx, hipEventCreate(&e), "event create");
    if (rc != DIC_OK) return rc;
    rc = check_hip(ctx, hipEventRecord(e, ctx->stream), "event record");
    if (rc != DIC_OK) {
        (void)hipEventDestroy(e);
        return rc;
    }
    *event = (void *)e;
    return DIC_OK;
}

int dic_event_elapsed_ms(void *start, void *end, float *ms)
{
    if (!start || !end || !ms) return DIC_ERR_INVALID_ARGUMENT;
    if (hipEventSynchronize((hipEvent_t)end) != hipSuccess) return DIC_ERR_HIP;
    return hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end) == hipSuccess ? DIC_OK : DIC_ERR_HIP;
}

void dic_event_destroy(void *event)
{
    if (event) (void)hipEventDestroy((hipEvent_t)event);
}

int dic_image_upload(dic_ctx *ctx, const float *host, int width, int height, dic_image **out)
{
    if (!ctx || !out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!host || width <= 0 || height <= 0)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "image must be non-empty with a host pointer");
    dic_image *img = new (std::nothrow) dic_image();
    if (!img) return set_error(ctx, DIC_ERR_HIP, "host allocation failed");
    img->ctx = ctx;
    img->width = width;
    img->height = height;
    img->bytes = sizeof(float) * (size_t)width * height;
    img->owned = true;
    int rc = alloc_checked(ctx, (void **)&img->data, img->bytes, "image allocation");
    if (rc != DIC_OK) {
        delete img;
        return rc;
    }
    rc = check_hip(ctx, upload_host(ctx, img->data, host, img->bytes, ctx->stream), "image upload");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "image upload sync");
    if (rc != DIC_OK) {
        ctx_free(ctx, img->data, img->bytes);
        delete img;
        return rc;
    }
    *out = img;
    return DIC_OK;
}

int dic_image_wrap_device(dic_ctx *ctx, const float *device_ptr, int width, int height, dic_image **out)
{
    if (!ctx || !out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!device_ptr || width <= 0 || height <= 0)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "image must be non-empty with a device pointer");
    dic_image *img = new (std::nothrow) dic_image();
    if (!img) return set_error(ctx, DIC_ERR_HIP, "host allocation failed");
    img->ctx = ctx;
    img->width = width;
    img->height = height;
    img->data = const_cast<float *>(device_ptr);
    img->owned = false;
    *out = img;
    return DIC_OK;
}

int dic_image_download(dic_ctx *ctx, const dic_image *img, float *host)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    if (!img || !img->data || !host) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "image and host buffer required");
    const size_t bytes = sizeof(float) * (size_t)img->width * img->height;
    return check_hip(ctx, download_host(ctx, host, img->data, bytes, ctx->stream), "image download");
}

void dic_image_free(dic_image *img)
{
    if (!img) return;
    if (img->owned) ctx_free(img->ctx, img->data, img->bytes);
    delete img;
}

static int run_prefilter(dic_ctx *ctx, const dic_image *img, dic_coeffs *c)
{
    const size_t tmp = sizeof(double) * (size_t)c->pitch * c->rows;
    int rc = ensure_staging(ctx, 2 * tmp);
    if (rc != DIC_OK) return rc;
    double *ta = (double *)ctx->staging;
    double *tb = ta + (size_t)c->pitch * c->rows;
    return check_hip(ctx,
                     launch_prefilter(img->data, img->width, img->height, c->degree, c->border, c->exact,
                                      c->precision, c->data, ta, tb, ctx->stream),
                     "prefilter launch");
}

int dic_bspline_prefilter_into(dic_ctx *ctx, const dic_image *img, dic_coeffs *out)
{
    if (!ctx || !img || !out) return DIC_ERR_INVALID_ARGUMENT;
    if (img->width != out->width || img->height != out->height)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "image and coefficient handle sizes differ");
    return run_prefilter(ctx, img, out);
}

int dic_bspline_prefilter(dic_ctx *ctx, const dic_image *img, int degree, int border, int exact, int precision,
                          dic_coeffs **out)
{
    if (!ctx || !out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!img || !img->data) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null image");
    if (!valid_degree(degree))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT,
                         "Only B-spline degree 1, 3, and 5 are supported; got " + std::to_string(degree) + ".");
    if (precision != DIC_PRECISION_FP64 && precision != DIC_PRECISION_MIXED)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown precision");
    const int b = border < 3 ? 3 : border;
    const int PW = img->width + 2 * b, PH = img->height + 2 * b;
    const int rc0 = check_prefilter_size(ctx, img->width, img->height, b, precision, "B-spline prefilter");
    if (rc0 != DIC_OK) return rc0;
    dic_coeffs *c = new (std::nothrow) dic_coeffs();
    if (!c) return set_error(ctx, DIC_ERR_HIP, "host allocation failed");
    c->ctx = ctx;
    c->width = img->width;
    c->height = img->height;
    c->degree = degree;
    c->border = b;
    c->precision = precision;
    c->pitch = PW;
    c->rows = PH;
    c->bytes = (precision == DIC_PRECISION_FP64 ? sizeof(double) : sizeof(float)) * (size_t)PW * PH;
    int rc = alloc_checked(ctx, &c->data, c->bytes, "coefficient allocation");
    if (rc != DIC_OK) {
        delete c;
        return rc;
    }
    c->exact = exact ? 1 : 0;
    rc = run_prefilter(ctx, img, c);
    if (rc != DIC_OK) {
        ctx_free(ctx, c->data, c->bytes);
        delete c;
        return rc;
    }
    *out = c;
    return DIC_OK;
}

int dic_coeffs_info(const dic_coeffs *c, int *width, int *height, int *degree, int *border, int *precision)
{
    if (!c) return DIC_ERR_INVALID_ARGUMENT;
    if (width) *width = c->width;
    if (height) *height = c->height;
    if (degree) *degree = c->degree;
    if (border) *border = c->border;
    if (precision) *precision = c->precision;
    return DIC_OK;
}

int dic_coeffs_download(dic_ctx *ctx, const dic_coeffs *c, double *host)
{
    if (!ctx || !c || !host) return DIC_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)c->pitch * c->rows;
    if (c->precision == DIC_PRECISION_FP64) {
        return check_hip(ctx, download_host(ctx, host, c->data, sizeof(double) * n, ctx->stream),
                         "coefficient download");
    }
    std::vector<float> tmp(n);
    int rc = check_hip(ctx, download_host(ctx, tmp.data(), c->data, sizeof(float) * n, ctx->stream),
                       "coefficient download");
    if (rc != DIC_OK) return rc;
    for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
    return DIC_OK;
}

void dic_coeffs_free(dic_coeffs *c)
{
    if (!c) return;
    ctx_free(c->ctx, c->data, c->bytes);
    delete c;
}

int dic_ref_gradient(dic_ctx *ctx, const dic_image *ref, int mode, int degree, int border, int precision,
                     dic_grad **out)
{
    if (!ctx || !out) return DIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!ref || !ref->data) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null image");
    if (precision != DIC_PRECISION_FP64 && precision != DIC_PRECISION_MIXED)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown precision");
    if (mode != DIC_GRADIENT_CENTRAL && mode != DIC_GRADIENT_BSPLINE_HALF)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown gradient mode");
    dic_grad *g = new (std::nothrow) dic_grad();
    if (!g) return set_error(ctx, DIC_ERR_HIP, "host allocation failed");
    g->ctx = ctx;
    g->width = ref->width;
    g->height = ref->height;
    g->precision = precision;
    const size_t plane = (precision == DIC_PRECISION_FP64 ? sizeof(double) : sizeof(float)) * (size_t)ref->width *
                         ref->height;
    g->bytes = 2 * plane;
    void *mem = nullptr;
    int rc = alloc_checked(ctx, &mem, g->bytes, "gradient allocation");
    if (rc != DIC_OK) {
        delete g;
        return rc;
    }
    g->gx = mem;
    g->gy = (char *)mem + plane;
    if (mode == DIC_GRADIENT_CENTRAL) {
        rc = check_hip(ctx, launch_gradient_central(ref->data, ref->width, ref->height, precision, g->gx, g->gy,
                                                    ctx->stream),
                       "gradient launch");
    } else {
        dic_coeffs *c = nullptr;
        rc = dic_bspline_prefilter(ctx, ref, degree, border, 1, DIC_PRECISION_FP64, &c);
        if (rc == DIC_OK) {
            rc = check_hip(ctx, launch_gradient_bspline(c, precision, g->gx, g->gy, ctx->stream), "gradient launch");
            if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "gradient sync");
            dic_coeffs_free(c);
        }
    }
    if (rc != DIC_OK) {
        ctx_free(ctx, mem, g->bytes);
        delete g;
        return rc;
    }
    *out = g;
    return DIC_OK;
}

int dic_ref_gradient_into(dic_ctx *ctx, const dic_image *ref, dic_grad *out)
{
    if (!ctx || !ref || !out) return DIC_ERR_INVALID_ARGUMENT;
    if (ref->width != out->width || ref->height != out->height)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "image and gradient handle sizes differ");
    return check_hip(ctx, launch_gradient_central(ref->data, ref->width, ref->height, out->precision, out->gx, out->gy,
                                                  ctx->stream),
                     "gradient launch");
}

int dic_grad_download(dic_ctx *ctx, const dic_grad *g, double *gx_host, double *gy_host)
{
    if (!ctx || !g || !gx_host || !gy_host) return DIC_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)g->width * g->height;
    if (g->precision == DIC_PRECISION_FP64) {
        int rc = check_hip(ctx, download_host(ctx, gx_host, g->gx, sizeof(double) * n, ctx->stream),
                           "gradient download");
        if (rc == DIC_OK)
            rc = check_hip(ctx, download_host(ctx, gy_host, g->gy, sizeof(double) * n, ctx->stream),
                           "gradient download");
        return rc;
    }
    std::vector<float> tx(n), ty(n);
    int rc = check_hip(ctx, download_host(ctx, tx.data(), g->gx, sizeof(float) * n, ctx->stream),
                       "gradient download");
    if (rc == DIC_OK)
        rc = check_hip(ctx, download_host(ctx, ty.data(), g->gy, sizeof(float) * n, ctx->stream),
                       "gradient download");
    if (rc != DIC_OK) return rc;
    for (size_t i = 0; i < n; ++i) {
        gx_host[i] = tx[i];
        gy_host[i] = ty[i];
    }
    return DIC_OK;
}

void dic_grad_free(dic_grad *g)
{
    if (!g) return;
    ctx_free(g->ctx, g->gx, g->bytes);
    delete g;
}

int dic_bspline_sample(dic_ctx *ctx, const dic_coeffs *c, const double *xy, int n, double *val, double *grad)
{
    if (!ctx || !c || n < 0 || (n > 0 && !xy)) return DIC_ERR_INVALID_ARGUMENT;
    if (n == 0) return DIC_OK;
    const size_t bxy = sizeof(double) * 2 * (size_t)n, bv = sizeof(double) * (size_t)n;
    int rc = ensure_staging(ctx, bxy + bv + bxy);
    if (rc != DIC_OK) return rc;
    char *st = (char *)ctx->staging;
    double *dxy = (double *)st, *dv = (double *)(st + bxy), *dg = (double *)(st + bxy + bv);
    rc = check_hip(ctx, hipMemcpyAsync(dxy, xy, bxy, hipMemcpyHostToDevice, ctx->stream), "sample upload");
    if (rc == DIC_OK)
        rc = check_hip(ctx, launch_sample(c, dxy, n, val ? dv : nullptr, grad ? dg : nullptr, ctx->stream),
                       "sample launch");
    if (rc == DIC_OK && val)
        rc = check_hip(ctx, hipMemcpyAsync(val, dv, bv, hipMemcpyDeviceToHost, ctx->stream), "sample download");
    if (rc == DIC_OK && grad)
        rc = check_hip(ctx, hipMemcpyAsync(grad, dg, bxy, hipMemcpyDeviceToHost, ctx->stream), "sample download");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sample sync");
    return rc;
}

} // extern "C"

namespace dic {

int validate_solve(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                          const dic_coeffs *def, int n)
{
    if (!params || !ref || !def || n < 0)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null argument or negative POI count");
    if (params->optimizer != DIC_OPTIMIZER_ICGN && params->optimizer != DIC_OPTIMIZER_FGN)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown optimizer");
    // forward Gauss-Newton differentiates the deformed image; IC-GN needs the reference gradients
    const bool need_grad = params->optimizer == DIC_OPTIMIZER_ICGN;
    if (need_grad && !grad) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "IC-GN needs reference gradients");
    if (params->shape_order != 1 && params->shape_order != 2)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "shape_order must be 1 or 2");
    if (params->objective != DIC_OBJECTIVE_ZNSSD && params->objective != DIC_OBJECTIVE_SSD)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "unknown objective");
    if (params->precision != def->precision || (grad && params->precision != grad->precision))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "precision of params, coefficients and gradients differ");
    if (grad && (grad->width != ref->width || grad->height != ref->height))
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "gradient planes do not match the reference image");
    return DIC_OK;
}

int host_tile_hint(const dic_icgn_params *params, const dic_coeffs *def, const double *xy, const dic_initial *init,
                   int n)
{
    const bool candidate = params->shape_order == 1 && params->precision != DIC_PRECISION_FP64 && !params->masked &&
                           params->optimizer == DIC_OPTIMIZER_ICGN && grouped_radius(params->subset_radius);
    if (!candidate) return 0;
    return tile_fit_share(xy, init, n, params->subset_radius, def->degree) >= 0.5 ? 1 : 0;
}

int run_solve(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                     const dic_coeffs *def, const uint8_t *roi_dev, const double *xy, const dic_initial *init, int n,
                     dic_result *out, hipStream_t s, int tile_hint)
{
    const int *table = nullptr;
    int count = 0;
    if (params->subset_radius >= 1) {
        int rc = get_subset_table(ctx, params->subset_radius, &table, &count);
        if (rc != DIC_OK) return rc;
        rc = check_solve_lds(ctx, params->optimizer == DIC_OPTIMIZER_FGN, params->precision, params->objective,
                             params->subset_radius, count);
        if (rc != DIC_OK) return rc;
    }
    return launch_solve(ctx, params, ref, grad, def, roi_dev, xy, init, n, out, s, tile_hint, table, count);
}

int launch_solve(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                 const dic_coeffs *def, const uint8_t *roi_dev, const double *xy, const dic_initial *init, int n,
                 dic_result *out, hipStream_t s, int tile_hint, const int *table, int count)
{
    IcgnLaunch L;
    L.params = params;
    // size mismatch between reference and deformed: every POI reports
    // InvalidInput, as icgn.cpp:464-470 does (width 0 takes that exit)
    const bool same = ref->width == def->width && ref->height == def->height;
    L.width = same ? ref->width : 0;
    L.height = same ? ref->height : 0;
    L.ref = ref->data;
    L.gx = grad ? grad->gx : nullptr;
    L.gy = grad ? grad->gy : nullptr;
    L.coeffs = def->data;
    L.coeff_bytes = def->bytes;
    L.pitch = def->pitch;
    L.border = def->border;
    L.degree = def->degree;
    L.precision = def->precision;
    L.roi = roi_dev;
    L.table = table;
    L.table_count = count;
    L.poi_xy = xy;
    L.init = init;
    L.out = out;
    L.n = n;
    L.tile_hint = tile_hint;
    return check_hip(ctx, launch_icgn(L, s), "icgn launch");
}

} // namespace dic

extern "C" {

int dic_icgn_solve_batch_device(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref,
                                const dic_grad *grad, const dic_coeffs *def, const uint8_t *roi_device,
                                const double *poi_xy_device, const dic_initial *init_device, int n,
                                dic_result *out_device, void *stream)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    int rc = validate_solve(ctx, params, ref, grad, def, n);
    if (rc != DIC_OK) return rc;
    if (params->masked && !roi_device) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "masked solve needs a roi");
    if (n == 0) return DIC_OK;
    if (!poi_xy_device || !init_device || !out_device)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null POI, initial-guess or result array");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return run_solve(ctx, params, ref, grad, def, params->masked ? roi_device : nullptr, poi_xy_device, init_device, n,
                     out_device, s);
}

int dic_icgn_solve_batch(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                         const dic_coeffs *def, const uint8_t *roi_host, const double *poi_xy,
                         const dic_initial *init, int n, dic_result *out)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    int rc = validate_solve(ctx, params, ref, grad, def, n);
    if (rc != DIC_OK) return rc;
    if (params->masked && !roi_host) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "masked solve needs a roi");
    if (n == 0) return DIC_OK;
    if (!poi_xy || !init || !out)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null POI, initial-guess or result array");
    const size_t bxy = sizeof(double) * 2 * (size_t)n;
    const size_t bin = sizeof(dic_initial) * (size_t)n;
    const size_t bout = sizeof(dic_result) * (size_t)n;
    const size_t broi = params->masked ? (size_t)ref->width * ref->height : 0;
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    rc = ensure_staging(ctx, align(bxy) + align(bin) + align(bout) + align(broi));
    if (rc != DIC_OK) return rc;
    char *st = (char *)ctx->staging;
    double *dxy = (double *)st;
    dic_initial *din = (dic_initial *)(st + align(bxy));
    dic_result *dout = (dic_result *)(st + align(bxy) + align(bin));
    uint8_t *droi = broi ? (uint8_t *)(st + align(bxy) + align(bin) + align(bout)) : nullptr;
    hipStream_t s = ctx->stream;
    rc = check_hip(ctx, upload_host(ctx, dxy, poi_xy, bxy, s), "POI upload");
    if (rc == DIC_OK) rc = check_hip(ctx, upload_host(ctx, din, init, bin, s), "init upload");
    if (rc == DIC_OK && droi) rc = check_hip(ctx, upload_host(ctx, droi, roi_host, broi, s), "roi upload");
    if (rc == DIC_OK) rc = run_solve(ctx, params, ref, grad, def, droi, dxy, din, n, dout, s,
                                     host_tile_hint(params, def, poi_xy, init, n));
    if (rc == DIC_OK) rc = check_hip(ctx, download_host(ctx, out, dout, bout, s), "result download");
    if (rc == DIC_OK) rc = check_hip(ctx, hipStreamSynchronize(s), "solve sync");
    return rc;
}

} // extern "C"
