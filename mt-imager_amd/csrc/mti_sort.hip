// mti_sort.hip -- stable LSD radix sort of (uint32 key, int32 value) pairs and
// the baseline-major row order.
//
// vis/SortIndex.cpp:46-87 buckets rows by baseline key Nant^2 - a1*Nant - a2 - 1
// on one host thread and keeps time order inside a baseline.  Here the same
// order comes from a stable 8-bit-digit radix sort on the device; rows with
// a1 == a2 get the out-of-range key Nant^2 and sort behind every kept row.
// The same sort orders gridding records by (channel, uv tile) while keeping
// their baseline-major order inside each tile.
#include "mti_internal.hpp"
#include "mti_sort.hpp"

#include <mutex>

// Device scratch of the allocation-free row sort (mti_sort_rows_static):
// 52 MiB per GPU, enough for the plan of MTI_SORT_STATIC_MAX_ROWS rows (keys
// and values 12 B per row, radix digit counts and scan words).  In host code
// `scratch` names a host-side stand-in; kernels get the device address that
// hipGetSymbolAddress returns for the current device (g_scratch_ptr).
constexpr size_t kStaticSortBytes = size_t(52) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticSortBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
}  // namespace

namespace mti {

// Looks the static scratch up once per device, with that device current
// (called by mti_create after hipSetDevice).  A failure leaves the device's
// pointer null: mti_sort_rows_static then fails with a message.
void static_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    if (e != hipSuccess) (void)hipGetLastError();
    g_scratch_done[device] = true;
}

constexpr int RX_THREADS = 256;
constexpr int RX_ROUNDS = 16;
constexpr int RX_ITEMS = RX_THREADS * RX_ROUNDS;
constexpr int RX_WAVES = RX_THREADS / 64;

__global__ __launch_bounds__(RX_THREADS) void k_radix_hist(const uint32_t *keys, int64_t n,
                                                          int shift, uint32_t *counts,
                                                          int nblocks)
{
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RX_ITEMS;
    for (int r = 0; r < RX_ROUNDS; r++) {
        int64_t i = base + r * RX_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(RX_THREADS) void k_radix_scatter(const uint32_t *keys,
                                                             const int32_t *vals, int64_t n,
                                                             int shift, const uint32_t *offsets,
                                                             int nblocks, uint32_t *keys_out,
                                                             int32_t *vals_out)
{
    __shared__ uint32_t cursor[256];
    __shared__ uint32_t wcount[RX_WAVES][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    cursor[tid] = offsets[(int64_t)tid * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RX_WAVES; w++) wcount[w][tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RX_ITEMS;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < RX_ROUNDS; r++) {
        const int64_t i = base + r * RX_THREADS + tid;
        const bool valid = i < n;
        uint32_t key = valid ? keys[i] : 0u;
        int32_t val = valid ? vals[i] : 0;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcount[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = cursor[d] + rank;
            for (int w = 0; w < wave; w++) pos += wcount[w][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < RX_WAVES; w++) {
            tot += wcount[w][tid];
            wcount[w][tid] = 0;
        }
        cursor[tid] += tot;
        __syncthreads();
    }
}

static int64_t radix_counts(int64_t n) { return ceil_div(n, RX_ITEMS) * 256; }

void plan_radix_sort(int64_t n, ScratchPlan &pl)
{
    if (n <= 0) return;
    pl.need(SLOT_RADIX, (size_t)(radix_counts(n) + 64) * sizeof(uint32_t));
    plan_scan(radix_counts(n), pl);
}

void plan_sort_rows_device(int64_t nrows, ScratchPlan &pl) { plan_radix_sort(nrows, pl); }

mti_status radix_sort_pairs(mti_ctx *ctx, uint32_t *keys, int32_t *vals, uint32_t *keys_tmp,
                            int32_t *vals_tmp, int64_t n, int key_bits, hipStream_t st,
                            bool *result_in_tmp)
{
    *result_in_tmp = false;
    if (n <= 0) return MTI_OK;
    MTI_CHECK(ctx, n < (int64_t(1) << 32), "radix sort above 2^32 elements");
    const int nblocks = (int)ceil_div(n, RX_ITEMS);
    const int64_t ncounts = radix_counts(n);
    void *p = nullptr;
    mti_status r = scratch_get(ctx, SLOT_RADIX, (size_t)(ncounts + 64) * sizeof(uint32_t), &p);
    if (r != MTI_OK) return r;
    uint32_t *counts = (uint32_t *)p;
    uint32_t *total = counts + ncounts;
    int passes = (key_bits + 7) / 8;
    if (passes < 1) passes = 1;
    uint32_t *ka = keys, *kb = keys_tmp;
    int32_t *va = vals, *vb = vals_tmp;
    for (int ps = 0; ps < passes; ps++) {
        int shift = ps * 8;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(RX_THREADS), 0, st, ka, n, shift, counts,
                           nblocks);
        MTI_HIP(ctx, hipGetLastError());
        r = exclusive_scan_u32(ctx, counts, ncounts, total, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(RX_THREADS), 0, st, ka, va, n, shift,
                           counts, nblocks, kb, vb);
        MTI_HIP(ctx, hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    *result_in_tmp = (ka == keys_tmp);
    return MTI_OK;
}

__global__ void k_baseline_keys(const int32_t *a1, const int32_t *a2, int64_t n, int nant,
                                uint32_t *keys, int32_t *vals, unsigned long long *kept)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    if (i < n) {
        int x = a1[i], y = a2[i];
        const uint32_t nk = (uint32_t)nant * (uint32_t)nant;
        keep = (x != y) && x >= 0 && y >= 0 && x < nant && y < nant;
        keys[i] = keep ? nk - (uint32_t)x * (uint32_t)nant - (uint32_t)y - 1u : nk;
        vals[i] = (int32_t)i;
    }
    uint64_t b = __ballot(keep);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(kept, (unsigned long long)__popcll(b));
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

mti_status sort_rows_device(mti_ctx *ctx, const int32_t *ant1, const int32_t *ant2, int64_t nrows, int nant,
                            int32_t *index_out, uint32_t *tmp, unsigned long long *kept, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, kept, 1);
    if (nrows == 0) return MTI_OK;
    uint32_t *k0 = tmp, *k1 = tmp + nrows;
    int32_t *v1 = (int32_t *)(k1 + nrows);
    hipLaunchKernelGGL(k_baseline_keys, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, st, ant1, ant2, nrows,
                       nant, k0, index_out, kept);
    MTI_HIP(ctx, hipGetLastError());
    const uint32_t nk = (uint32_t)nant * (uint32_t)nant;
    int bits = 1;
    while (bits < 32 && (nk >> bits) != 0) bits++;
    bool in_tmp = false;
    mti_status r = radix_sort_pairs(ctx, k0, index_out, k1, v1, nrows, bits, st, &in_tmp);
    if (r != MTI_OK) return r;
    if (in_tmp)
        MTI_HIP(ctx, hipMemcpyAsync(index_out, v1, (size_t)nrows * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    return MTI_OK;
}

}  // namespace mti

using namespace mti;

namespace mti {
static void plan_sort_rows(int64_t nrows, ScratchPlan &pl)
{
    if (nrows <= 0) return;
    pl.need(SLOT_KEYS, (size_t)nrows * 12 + 64);
    plan_sort_rows_device(nrows, pl);
}
}  // namespace mti

extern "C" mti_status mti_sort_rows_workspace_size(int64_t nrows, size_t *bytes)
{
    if (!bytes || nrows < 0 || nrows >= (int64_t(1) << 31)) return MTI_ERR_INVALID;
    ScratchPlan pl;
    plan_sort_rows(nrows, pl);
    *bytes = pl.total();
    return MTI_OK;
}

extern "C" mti_status mti_sort_rows(mti_ctx *ctx, const int32_t *ant1, const int32_t *ant2,
                                    int64_t nrows, int32_t nant, int32_t *index_out,
                                    int64_t *kept_out, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, nant >= 1 && nant <= 46340, "antenna count out of range");
    MTI_CHECK(ctx, nrows >= 0 && nrows < (int64_t(1) << 31), "row count out of range");
    MTI_CHECK(ctx, kept_out && (nrows == 0 || (ant1 && ant2 && index_out)), "null pointer");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    CallScope cs(ctx);
    if (cs.planning()) {
        ScratchPlan pl;
        plan_sort_rows(nrows, pl);
        mti_status rp = cs.begin(pl);
        if (rp != MTI_OK) return rp;
    }
    StageTimer tm(ctx, MTI_STAGE_SORT_ROWS, st);
    if (nrows == 0) {
        *kept_out = 0;
        return MTI_OK;
    }
    void *p = nullptr;
    size_t need = (size_t)nrows * 12 + 64;
    mti_status r = scratch_get(ctx, SLOT_KEYS, need, &p);
    if (r != MTI_OK) return r;
    uint32_t *tmp = (uint32_t *)p;
    unsigned long long *kept = (unsigned long long *)(((uintptr_t)(tmp + 3 * nrows) + 15) & ~uintptr_t(15));
    // the value array in slot 0 is the caller's output buffer
    r = sort_rows_device(ctx, ant1, ant2, nrows, nant, index_out, tmp, kept, st);
    if (r != MTI_OK) return r;
    unsigned long long h = 0;
    MTI_HIP(ctx, hipMemcpyAsync(&h, kept, sizeof(h), hipMemcpyDeviceToHost, st));
    MTI_HIP(ctx, hipStreamSynchronize(st));
    *kept_out = (int64_t)h;
    return MTI_OK;
}

extern "C" mti_status mti_sort_rows_static(mti_ctx *ctx, const int32_t *ant1, const int32_t *ant2, int64_t nrows,
                                           int32_t nant, int32_t *index_out, int64_t *kept_out, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, nrows >= 0 && nrows <= MTI_SORT_STATIC_MAX_ROWS,
              "mti_sort_rows_static: nrows outside [0, MTI_SORT_STATIC_MAX_ROWS = 4194304]");
    MTI_CHECK(ctx, ctx->call_depth == 0, "mti_sort_rows_static inside another call");
    void *ptr = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
    MTI_CHECK(ctx, ptr, "mti_sort_rows_static: the static scratch of this device is unavailable");
    ScratchPlan pl;
    plan_sort_rows(nrows, pl);
    MTI_CHECK(ctx, pl.total() <= kStaticSortBytes, "mti_sort_rows_static: plan above the static scratch");
    // run mti_sort_rows inside the static array as a caller workspace
    char *const saved_ptr = ctx->ws_ptr;
    const size_t saved_bytes = ctx->ws_bytes;
    ctx->ws_ptr = (char *)ptr;
    ctx->ws_bytes = kStaticSortBytes;
    const mti_status r = mti_sort_rows(ctx, ant1, ant2, nrows, nant, index_out, kept_out, stream);
    ctx->ws_ptr = saved_ptr;
    ctx->ws_bytes = saved_bytes;
    return r;
}
