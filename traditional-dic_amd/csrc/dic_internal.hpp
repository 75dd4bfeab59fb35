// dic_internal.hpp -- handle layouts and kernel launchers shared by the
// translation units of libdic_hip.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>

#include "dic_hip.h"

namespace dic {
struct HostStager;
}

struct dic_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::atomic<uint64_t> in_use{0};
    std::string error;
    std::mutex table_mutex;
    // circular-subset offset tables, one per radius: packed (dx, dy) int16 pairs
    std::map<int, std::pair<int *, int>> subset_tables;
    // sizes of dic_buffer_alloc allocations, for budget accounting
    std::map<void *, size_t> buffers;
    // scratch used by the host-pointer entry points
    void *staging = nullptr;
    size_t staging_bytes = 0;
    // LDS one workgroup may use on this device (hipDeviceAttributeMaxSharedMemoryPerBlock)
    size_t max_lds = 0;
    // device address of the no-allocation entry's table array on this device
    int *noalloc_table = nullptr;
    // pinned double buffer and copy threads of large host copies (hostcopy.hip)
    dic::HostStager *stager = nullptr;
    bool stager_failed = false;
};

struct dic_image {
    dic_ctx *ctx = nullptr;
    int width = 0, height = 0;
    float *data = nullptr;
    bool owned = false;
    size_t bytes = 0;
};

struct dic_grad {
    dic_ctx *ctx = nullptr;
    int width = 0, height = 0, precision = 0;
    void *gx = nullptr, *gy = nullptr;
    size_t bytes = 0;
};

struct dic_coeffs {
    dic_ctx *ctx = nullptr;
    int width = 0, height = 0, degree = 5, border = 3, precision = 0, exact = 1;
    int pitch = 0, rows = 0; // padded width and height
    void *data = nullptr;
    size_t bytes = 0;
};

namespace dic {

// device memory accounting against the context budget
hipError_t ctx_malloc(dic_ctx *ctx, void **ptr, size_t bytes);
void ctx_free(dic_ctx *ctx, void *ptr, size_t bytes);
int set_error(dic_ctx *ctx, int code, const std::string &msg);
int check_hip(dic_ctx *ctx, hipError_t e, const char *what);

// hostcopy.hip: copies between pageable host memory and the device.  An
// upload is ordered on s (the host buffer may be reused on return); a
// download returns with the data on the host.
hipError_t upload_host(dic_ctx *ctx, void *device, const void *host, size_t bytes, hipStream_t s);
hipError_t download_host(dic_ctx *ctx, void *host, const void *device, size_t bytes, hipStream_t s);
void destroy_stager(dic_ctx *ctx);

// dic_abi.hip: argument checks and one solve launch over device arrays
int validate_solve(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                   const dic_coeffs *def, int n);
int run_solve(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
              const dic_coeffs *def, const uint8_t *roi_dev, const double *xy, const dic_initial *init, int n,
              dic_result *out, hipStream_t s, int tile_hint = -1);
// run_solve's launch with the circular-subset table given (count entries,
// dy-outer / dx-inner order, packed (dy << 16) | (dx & 0xffff)).
int launch_solve(dic_ctx *ctx, const dic_icgn_params *params, const dic_image *ref, const dic_grad *grad,
                 const dic_coeffs *def, const uint8_t *roi_dev, const double *xy, const dic_initial *init, int n,
                 dic_result *out, hipStream_t s, int tile_hint, const int *table, int count);
// Takes the device address of the no-allocation entry's table array on the
// context's device (current at the call); dic_ctx_create calls it.
int noalloc_setup(dic_ctx *ctx);

// LDS of one solve workgroup: the static per-POI block (icgn.hip) plus the
// per-sample buffer of FP64 two-pass ZNSSD IC-GN (one value; FGN keeps none);
// check_solve_lds turns a subset too large for the device into
// DIC_ERR_UNSUPPORTED before any launch
size_t solve_static_lds();
// the same for the cooperative kernel (one POI per workgroup of several
// waves, with their per-wave slots)
size_t coop_static_lds();
size_t solve_dynamic_lds(bool fgn, int precision, int objective, int nfull);
int check_solve_lds(dic_ctx *ctx, bool fgn, int precision, int objective, int radius, int nfull);

// one-line sample table of a circular subset (dy outer, dx inner, dx^2+dy^2 <= r^2)
int get_subset_table(dic_ctx *ctx, int radius, const int **table, int *count);

// bspline.hip
// Bytes of LDS one workgroup of the exact prefilter's row pass needs for a
// padded row of PW samples in the given storage precision; check_prefilter_size refuses (DIC_ERR_UNSUPPORTED,
// with the bytes needed) a frame whose padded rows exceed the device's LDS or
// whose padded FP64 image reaches 2 GiB.
size_t prefilter_row_lds(int PW, int precision);
int check_prefilter_size(dic_ctx *ctx, int W, int H, int border, int precision, const char *who);
hipError_t launch_prefilter(const float *img, int W, int H, int degree, int border, int exact, int precision,
                            void *out, double *tmp_a, double *tmp_b, hipStream_t s);
hipError_t launch_gradient_central(const float *img, int W, int H, int precision, void *gx, void *gy, hipStream_t s);
hipError_t launch_gradient_bspline(const dic_coeffs *c, int precision, void *gx, void *gy, hipStream_t s);
hipError_t launch_sample(const dic_coeffs *c, const double *xy, int n, double *val, double *grad, hipStream_t s);

// icgn.hip
// Kernel arguments of one (pair, parameter set): images, planes, table, POIs.
struct IcgnArgs {
    int W, H, pitch, b;
    const float *ref;
    const void *gx, *gy, *coeffs;
    const uint8_t *roi;
    const int *table;
    int nfull;
    int coeff_bytes; // size of the padded coefficient image (< 2^31)
    int radius, max_it;
    double thr;
    int min_needed;
    int fgn; // forward-compositional Gauss-Newton instead of IC-GN
    const double *xy;
    const dic_initial *init;
    dic_result *out;
    int n;
};

// First-order FP32 solves of the per-pair and serving launches run this many
// POIs per workgroup (one per wave) over a shared LDS tile of the
// coefficients (72 x 60 cells); with a subset radius above 24 four POIs never
// fit one tile, and the per-pair launch keeps one POI per workgroup.
constexpr int kGroupPois = 4;
inline bool grouped_radius(int radius) { return radius <= 24; }

// One descriptor of the persistent batch kernel.
struct BatchPair {
    IcgnArgs a;
    int shape_order;
    int reserved;
};

// Descriptors launched together: same optimizer and shape order (and, where
// the solve buffers per-sample values in LDS, the same LDS size class); their
// items are items[first, first + count).
struct BatchGroup {
    int shape_order;
    int max_nfull;
    int first, count;
    int fgn;  // forward Gauss-Newton descriptors (one POI per workgroup)
    int tile; // first-order FP32 descriptors of r <= 24 under the TILE mapping: four POIs per workgroup
};
constexpr int kMaxBatchGroups = 64; // queue counters in the batch's 256-byte counter buffer

struct BatchLaunch {
    int objective, degree, precision;
    const BatchPair *pairs;
    const int2 *items; // (descriptor, POI), group after group
    const BatchGroup *groups;
    int ngroups;
    bool persistent;   // work-queue kernels on a fixed grid, else one workgroup per item
    int grid;          // persistent workgroups
    int chunk;         // items per queue grab
    unsigned *counter; // one queue counter per group (persistent)
};
hipError_t launch_batch(const BatchLaunch &L, hipStream_t s);

// One launch over descriptors of one shape order with contiguous POI ranges
// [offsets[k], offsets[k+1]) (serving ticks).
struct RangeLaunch {
    int objective, degree, precision, shape_order;
    const BatchPair *pairs; // device
    const int *offsets;     // device, npairs + 1 prefix sums
    int npairs;
    int total;              // offsets[npairs]
    int max_nfull;
    int fgn = 0;            // forward Gauss-Newton descriptors
};
hipError_t launch_range(const RangeLaunch &L, hipStream_t s);
hipError_t launch_init_from_results(const dic_result *prev, dic_initial *init, int n, hipStream_t s);
// guesses 2 a - b (u, v, gradients) where both a and b are Success, a otherwise
hipError_t launch_init_extrapolated(const dic_result *a, const dic_result *b, dic_initial *init, int n,
                                    hipStream_t s);
hipError_t add_batch_nodes(const BatchLaunch &L, hipGraph_t g);

struct IcgnLaunch {
    const dic_icgn_params *params;
    int width, height;
    const float *ref;
    const void *gx, *gy;
    const void *coeffs;
    size_t coeff_bytes;
    int pitch, border, degree, precision;
    const uint8_t *roi;
    const int *table;
    int table_count;
    const double *poi_xy;
    const dic_initial *init;
    dic_result *out;
    int n;
    // launch shape: kShapeTile four POIs per workgroup over an LDS tile
    // (first-order FP32 IC-GN), kShapeWave one POI per 64-lane workgroup,
    // kShapeCoop one POI per workgroup of several waves (IC-GN; a few POIs at
    // a time), -1 decide by the radius (grouped_radius)
    int tile_hint = -1;
};
enum { kShapeWave = 0, kShapeTile = 1, kShapeCoop = 2 };
hipError_t launch_icgn(const IcgnLaunch &L, hipStream_t s);


// Share of the quads of consecutive POIs (POI positions and guesses on the
// host) whose starting subsets fit one LDS tile; the host-pointer entry points
// pass tile_hint = share >= 1/2.
double tile_fit_share(const double *xy, const dic_initial *init, int n, int radius, int degree);
// tile_hint of a solve whose POIs and guesses are on the host (0 where the
// tile kernel cannot run: second order, FP64, masked, FGN, radius > 24)
int host_tile_hint(const dic_icgn_params *params, const dic_coeffs *def, const double *xy, const dic_initial *init,
                   int n);

} // namespace dic
