// noalloc.hip -- dic_icgn_solve_batch_device_noalloc: the device-pointer
// IC-GN / FGN entry without any device allocation.  The only device memory
// the library allocates for that entry is the circular-subset offset table
// (dic_abi.hip allocates one per radius on first use); here the table lives in
// the file-scope array `scratch` below and a one-workgroup kernel rebuilds it
// on the caller's stream before every solve.  The solve kernels are the other
// entry's, so the results are the same bits.
#include <map>
#include <mutex>

#include "dic_internal.hpp"

namespace dic {
namespace {

static __device__ __align__(256) int scratch[DIC_NOALLOC_MAX_SAMPLES];

// Device address of `scratch` per device, taken once per device with that
// device current.  Host code never passes `scratch` itself to a kernel or a
// copy: in host code the name is a stand-in, not a device address.
std::mutex g_scratch_mutex;
std::map<int, int *> g_scratch_ptr;

// Samples of the circle dx^2 + dy^2 <= r^2 (the table's length).
int circle_count(int r)
{
    if (r < 1) return 0;
    long long n = 0;
    for (int dy = -r; dy <= r; ++dy) {
        int hw = 0;
        while ((hw + 1) * (hw + 1) + dy * dy <= r * r) ++hw;
        n += 2 * hw + 1;
        if (n > DIC_NOALLOC_MAX_SAMPLES) return (int)n;
    }
    return (int)n;
}

// Largest hw with hw^2 + dy^2 <= r^2 (integers only).
__device__ int half_width(int r, int dy)
{
    const int rem = r * r - dy * dy;
    int hw = (int)sqrtf((float)rem);
    while (hw * hw > rem) --hw;
    while ((hw + 1) * (hw + 1) <= rem) ++hw;
    return hw;
}

// The table of dic_abi.hip's get_subset_table: (dy << 16) | (dx & 0xffff),
// dy outer, dx inner, ascending.  Thread t writes the rows dy = -r + t + k*256;
// a row's start is the sum of the lengths of the rows above it.
__global__ __launch_bounds__(256) void subset_table_kernel(int *table, int r)
{
    for (int row = threadIdx.x; row <= 2 * r; row += blockDim.x) {
        const int dy = row - r;
        int start = 0;
        for (int k = -r; k < dy; ++k) start += 2 * half_width(r, k) + 1;
        const int hw = half_width(r, dy);
        for (int dx = -hw; dx <= hw; ++dx) table[start + dx + hw] = (int)(((unsigned)dy << 16) | (unsigned)(dx & 0xffff));
    }
}

} // namespace

int noalloc_setup(dic_ctx *ctx)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_ptr.find(ctx->device);
    if (it == g_scratch_ptr.end()) {
        int *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return check_hip(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
        it = g_scratch_ptr.emplace(ctx->device, ptr).first;
    }
    ctx->noalloc_table = it->second;
    return DIC_OK;
}

} // namespace dic

using namespace dic;

extern "C" int dic_icgn_solve_batch_device_noalloc(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref,
                                                   const dic_grad *grad, const dic_coeffs *def,
                                                   const uint8_t *roi_device, const double *poi_xy_device,
                                                   const dic_initial *init_device, int n, dic_result *out_device,
                                                   void *stream)
{
    if (!ctx) return DIC_ERR_INVALID_ARGUMENT;
    int rc = validate_solve(ctx, params, ref, grad, def, n);
    if (rc != DIC_OK) return rc;
    if (params->masked && !roi_device) return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "masked solve needs a roi");
    if (n == 0) return DIC_OK;
    if (!poi_xy_device || !init_device || !out_device)
        return set_error(ctx, DIC_ERR_INVALID_ARGUMENT, "null POI, initial-guess or result array");
    if (!ctx->noalloc_table) return set_error(ctx, DIC_ERR_HIP, "no scratch table address on this device");
    const int r = params->subset_radius;
    const int count = circle_count(r);
    if (count > DIC_NOALLOC_MAX_SAMPLES)
        return set_error(ctx, DIC_ERR_UNSUPPORTED,
                         "subset_radius " + std::to_string(r) + " needs more than " +
                             std::to_string(DIC_NOALLOC_MAX_SAMPLES) +
                             " table entries; the no-allocation entry accepts radii up to " +
                             std::to_string(DIC_NOALLOC_MAX_RADIUS));
    if (r >= 1) {
        rc = check_solve_lds(ctx, params->optimizer == DIC_OPTIMIZER_FGN, params->precision, params->objective, r,
                             count);
        if (rc != DIC_OK) return rc;
    }
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    int *table = ctx->noalloc_table;
    if (r >= 1) {
        hipLaunchKernelGGL(subset_table_kernel, dim3(1), dim3(256), 0, s, table, r);
        rc = check_hip(ctx, hipGetLastError(), "subset table kernel");
        if (rc != DIC_OK) return rc;
    }
    return launch_solve(ctx, params, ref, grad, def, params->masked ? roi_device : nullptr, poi_xy_device,
                        init_device, n, out_device, s, -1, r >= 1 ? table : nullptr, count);
}
