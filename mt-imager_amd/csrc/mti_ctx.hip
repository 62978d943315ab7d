// mti_ctx.hip -- context, validation and the device-wide exclusive scan.
#include "mti_internal.hpp"

#include <cmath>

extern "C" const char *mti_version(void) { return "mtimager_amd 0.1 (gfx950)"; }

extern "C" mti_status mti_create(int device, mti_ctx **out)
{
    if (!out) return MTI_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return MTI_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return MTI_ERR_HIP;
    mti_ctx *c = new mti_ctx();
    c->device = device;
    mti::static_scratch_setup(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (c->num_cus <= 0) c->num_cus = 256;
    *out = c;
    return MTI_OK;
}

extern "C" void mti_destroy(mti_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (auto &e : ctx->stage_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &s : ctx->scratch)
        if (s.ptr) (void)hipFree(s.ptr);
    for (auto &t : ctx->twiddles)
        if (t.second) (void)hipFree(t.second);
    for (auto &t : ctx->gen_tables)
        if (t.ptr) (void)hipFree(t.ptr);
    delete ctx;
}

extern "C" mti_status mti_set_workspace(mti_ctx *ctx, void *ptr, size_t bytes)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, ctx->call_depth == 0, "mti_set_workspace during a call");
    MTI_CHECK(ctx, (ptr == nullptr) == (bytes == 0), "workspace pointer and size must be set (or cleared) together");
    ctx->ws_ptr = (char *)ptr;
    ctx->ws_bytes = bytes;
    return MTI_OK;
}

extern "C" mti_status mti_set_tuning(mti_ctx *ctx, const mti_tuning *t)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, t, "null tuning");
    const int u = t->unit_pairs ? t->unit_pairs : 2048;
    MTI_CHECK(ctx, u >= 256 && u <= 16384 && (u & (u - 1)) == 0, "unit_pairs must be a power of two in [256, 16384]");
    const int w = t->whole_columns_ny;
    MTI_CHECK(ctx, w == 0 || (w >= 16 && w <= 256 && (w & (w - 1)) == 0),
              "whole_columns_ny must be 0 or a power of two in [16, 256]");
    const int x = t->xcd_chunk;
    MTI_CHECK(ctx, x == 0 || (x >= 1 && x <= 4096 && (x & (x - 1)) == 0),
              "xcd_chunk must be 0 or a power of two in [1, 4096]");
    const int r = t->region_width ? t->region_width : 8;
    MTI_CHECK(ctx, r == 4 || r == 8 || r == 16, "region_width must be 0, 4, 8 or 16");
    ctx->unit_pairs = u;
    ctx->whole_columns_ny = w;
    ctx->xcd_chunk = x;
    ctx->region_width = r;
    return MTI_OK;
}

extern "C" mti_status mti_get_tuning(const mti_ctx *ctx, mti_tuning *t)
{
    if (!ctx || !t) return MTI_ERR_INVALID;
    t->unit_pairs = ctx->unit_pairs;
    t->whole_columns_ny = ctx->whole_columns_ny;
    t->xcd_chunk = ctx->xcd_chunk;
    t->region_width = ctx->region_width;
    return MTI_OK;
}

extern "C" mti_status mti_set_transform_storage(mti_ctx *ctx, int32_t grid_precision)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, ctx->call_depth == 0, "mti_set_transform_storage during a call");
    MTI_CHECK(ctx, grid_precision == MTI_GRID_C64 || grid_precision == MTI_GRID_C128,
              "transform storage must be MTI_GRID_C64 or MTI_GRID_C128");
    ctx->xform_c64 = grid_precision == MTI_GRID_C64 ? 1 : 0;
    return MTI_OK;
}

extern "C" mti_status mti_get_transform_storage(const mti_ctx *ctx, int32_t *grid_precision)
{
    if (!ctx || !grid_precision) return MTI_ERR_INVALID;
    *grid_precision = ctx->xform_c64 ? MTI_GRID_C64 : MTI_GRID_C128;
    return MTI_OK;
}

extern "C" mti_status mti_context_device_bytes(const mti_ctx *ctx, size_t *bytes)
{
    if (!ctx || !bytes) return MTI_ERR_INVALID;
    size_t t = 0;
    for (const auto &s : ctx->scratch) t += s.bytes;
    for (const auto &tw : ctx->twiddles) t += (size_t)tw.first * sizeof(double2);
    for (const auto &gt : ctx->gen_tables) t += gt.bytes;
    *bytes = t;
    return MTI_OK;
}

extern "C" mti_status mti_grid_out_shape(const mti_image_params *p, int32_t npol, int32_t nchan, int32_t grid_layout,
                                         int64_t grid_dims[4], int64_t norm_dims[2])
{
    if (!p || !grid_dims || !norm_dims || mti::validate_params(nullptr, p) != MTI_OK ||
        !(npol == 1 || npol == 2 || npol == 4) || nchan < 1 ||
        !(grid_layout == MTI_GRID_VUP || grid_layout == MTI_GRID_PVU))
        return MTI_ERR_INVALID;
    const int64_t vup[4] = {nchan, p->ny, p->nx, npol}, pvu[4] = {nchan, npol, p->ny, p->nx};
    for (int i = 0; i < 4; i++) grid_dims[i] = grid_layout == MTI_GRID_VUP ? vup[i] : pvu[i];
    norm_dims[0] = nchan;
    norm_dims[1] = npol;
    return MTI_OK;
}

extern "C" mti_status mti_preprocess_out_shape(int64_t nrecords, int64_t bins_dims[2])
{
    if (!bins_dims || nrecords < 0 || nrecords >= (int64_t(1) << 31)) return MTI_ERR_INVALID;
    bins_dims[0] = nrecords;
    bins_dims[1] = 8;
    return MTI_OK;
}

extern "C" const char *mti_last_error(const mti_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" mti_status mti_set_profiling(mti_ctx *ctx, int32_t enable)
{
    if (!ctx) return MTI_ERR_INVALID;
    ctx->profiling = enable != 0;
    return MTI_OK;
}

extern "C" mti_status mti_stage_times(mti_ctx *ctx, double *ms, int64_t *launches)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    for (auto &e : ctx->stage_events) {
        MTI_HIP(ctx, hipEventSynchronize(e.b));
        float t = 0.0f;
        MTI_HIP(ctx, hipEventElapsedTime(&t, e.a, e.b));
        if (e.stage >= 0 && e.stage < MTI_STAGE_COUNT) {
            if (ms) ms[e.stage] += t;
            if (launches) launches[e.stage] += 1;
        }
        ctx->event_pool.push_back(e.a);
        ctx->event_pool.push_back(e.b);
    }
    ctx->stage_events.clear();
    return MTI_OK;
}

namespace mti {

static hipEvent_t pool_event(mti_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

StageTimer::StageTimer(mti_ctx *c, int s, hipStream_t stream) : ctx(c), st(stream), stage(s)
{
    if (!ctx || !ctx->profiling || ctx->capturing) return;
    a = pool_event(ctx);
    b = pool_event(ctx);
    if (a) (void)hipEventRecord(a, st);
}

StageTimer::~StageTimer()
{
    if (!a || !b) return;
    (void)hipEventRecord(b, st);
    ctx->stage_events.push_back({stage, a, b});
}

// Validation mirrors the reference's validate() checks: odd taper support
// (op/CasaGridSF.cpp:65-66), known mode (comp/ConfigurationManager.cpp:106-107),
// positive image size (comp/WImager.cpp:262-263), pol in {1,2,4}
// (comp/WImager.cpp:249-257).
mti_status validate_params(mti_ctx *ctx, const mti_image_params *p)
{
    MTI_CHECK(ctx, p, "null image parameters");
    MTI_CHECK(ctx, p->nx >= 2 && p->ny >= 2, "v or u pixels do not make sense (image.nx/image.ny)");
    MTI_CHECK(ctx, p->nx <= 32768 && p->ny <= 32768, "image.nx/image.ny above 32768 unsupported");
    MTI_CHECK(ctx, p->cell_l > 0.0 && p->cell_m > 0.0 && std::isfinite(p->cell_l) &&
                       std::isfinite(p->cell_m),
              "image.lintervalsize/mintervalsize must be positive");
    MTI_CHECK(ctx, p->mode == MTI_MODE_NORMAL || p->mode == MTI_MODE_WPROJ,
              "mode must be normal or wproj");
    MTI_CHECK(ctx, p->sampling >= 1 && p->sampling <= 64, "sampling must be in [1,64]");
    if (p->mode == MTI_MODE_WPROJ) {
        MTI_CHECK(ctx, p->wplanes >= 1 && p->wplanes <= 4096, "wproj.wplanes must be in [1,4096]");
        int mn = p->nx < p->ny ? p->nx : p->ny;
        int st = mn / p->sampling;
        if (st % 2 == 0) st++;
        MTI_CHECK(ctx, st >= 3, "image too small for the requested sampling (trial support < 3)");
    } else {
        MTI_CHECK(ctx, p->taper_support >= 1, "support requested is less then 1");
        MTI_CHECK(ctx, p->taper_support % 2 == 1, "Support is expected to be odd");
    }
    return MTI_OK;
}

mti_status validate_vis(mti_ctx *ctx, const mti_vis *v)
{
    MTI_CHECK(ctx, v, "null visibility descriptor");
    MTI_CHECK(ctx, v->npol == 1 || v->npol == 2 || v->npol == 4,
              "No of polarisations requested is not supported, it must be 1,2 or 4");
    MTI_CHECK(ctx, v->nrows >= 0 && v->nchan >= 1, "bad row/channel counts");
    MTI_CHECK(ctx, v->nrows < (int64_t(1) << 31), "more than 2^31-1 rows per call");
    if (v->nrows > 0) {
        MTI_CHECK(ctx, v->uvw && v->inv_lambda && v->flags && v->weight,
                  "null uvw/frequency/flags/weight pointer");
    }
    if (v->row_index) MTI_CHECK(ctx, v->nindex >= 0 && v->nindex <= v->nrows, "bad nindex");
    return MTI_OK;
}

// ---- exclusive scan of uint32 ---------------------------------------------
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_PER_THREAD = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Block-exclusive scan of SCAN_TILE elements; writes the tile total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(uint32_t *data, int64_t n,
                                                            uint32_t *tile_sums)
{
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; i++) {
        v[i] = (base + i < n) ? data[base + i] : 0u;
        s += v[i];
    }
    uint32_t incl = wave_incl_scan(s);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = (lane < SCAN_THREADS / 64) ? wsum[lane] : 0u;
        uint32_t wi = wave_incl_scan(w);
        if (lane < SCAN_THREADS / 64) wsum[lane] = wi - w;
        if (lane == SCAN_THREADS / 64 - 1 && tile_sums) tile_sums[blockIdx.x] = wi;
    }
    __syncthreads();
    uint32_t run = wsum[wave] + incl - s;
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; i++) {
        if (base + i < n) data[base + i] = run;
        run += v[i];
    }
}

__global__ void k_scan_add(uint32_t *data, int64_t n, const uint32_t *tile_offsets)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] += tile_offsets[i / SCAN_TILE];
}

__global__ void k_scan_total(const uint32_t *scanned, const uint32_t *orig_last, int64_t n,
                             uint32_t *total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) total[0] = n > 0 ? scanned[n - 1] + orig_last[0] : 0u;
}

__global__ void k_copy_last(const uint32_t *data, int64_t n, uint32_t *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = n > 0 ? data[n - 1] : 0u;
}

static mti_status scan_rec(mti_ctx *ctx, uint32_t *d, int64_t n, uint32_t *scratch,
                           hipStream_t st)
{
    int64_t tiles = ceil_div(n, SCAN_TILE);
    if (tiles <= 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, st, d, n, nullptr);
        MTI_HIP(ctx, hipGetLastError());
        return MTI_OK;
    }
    uint32_t *sums = scratch;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, d, n, sums);
    MTI_HIP(ctx, hipGetLastError());
    mti_status r = scan_rec(ctx, sums, tiles, scratch + tiles, st);
    if (r != MTI_OK) return r;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, d, n, sums);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

// scratch words of a scan of n: one for the original last element + tile
// sums per level
static int64_t scan_words(int64_t n)
{
    int64_t need = 2;
    for (int64_t m = ceil_div(n, SCAN_TILE); m > 1; m = ceil_div(m, SCAN_TILE)) need += m + 1;
    return need;
}

void plan_scan(int64_t n, ScratchPlan &pl) { pl.need(SLOT_MISC, (size_t)scan_words(n) * sizeof(uint32_t)); }

mti_status exclusive_scan_u32(mti_ctx *ctx, uint32_t *d_data, int64_t n, uint32_t *d_total,
                              hipStream_t st)
{
    void *p = nullptr;
    mti_status r = scratch_get(ctx, SLOT_MISC, (size_t)scan_words(n) * sizeof(uint32_t), &p);
    if (r != MTI_OK) return r;
    uint32_t *w = (uint32_t *)p;
    hipLaunchKernelGGL(k_copy_last, dim3(1), dim3(64), 0, st, d_data, n, w);
    MTI_HIP(ctx, hipGetLastError());
    if (n > 0) {
        r = scan_rec(ctx, d_data, n, w + 2, st);
        if (r != MTI_OK) return r;
    }
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, st, d_data, w, n, d_total);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

}  // namespace mti
