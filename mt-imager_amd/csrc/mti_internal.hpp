// mti_internal.hpp -- shared definitions of the gfx950 imaging library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mtimager_amd.h"

namespace mti {

// Scratch buffer owned by a context; grows on demand, never shrinks.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Status {
    mti_status code = MTI_OK;
    std::string msg;
};

// Scratch slots.
enum ScratchSlot {
    SLOT_KEYS = 0,
    SLOT_DESC = 1,
    SLOT_SORTED = 2,
    SLOT_BUCKETS = 3,
    SLOT_ITEMS = 4,
    SLOT_DENSITY = 5,
    SLOT_MISC = 6,
    SLOT_FFT = 7,
    SLOT_RADIX = 8,
    SLOT_NORM = 9,
    SLOT_MISC2 = 10,
    SLOT_RECVIS = 11,   // row order and kept counts of jobs sorted inside mti_grid_jobs
    SLOT_CRECVIS = 12,  // per-record visibilities + weights
    SLOT_CHANDESC = 13, // descriptors of the finalize (packed planes)
    SLOT_CORR = 14,     // per-job correction vectors of the batched finalize
    SLOT_FCORR = 15,    // correction vectors of mti_finalize
    SLOT_TWIDDLE = 16,  // FFT twiddles of a call that runs in a caller workspace
    SLOT_GTAB = 17,     // general-size transform tables of a call in a caller workspace
    SLOT_COUNT = 18
};

// Device scratch one call takes: the largest request of each slot.
struct ScratchPlan {
    size_t bytes[SLOT_COUNT] = {};
    void need(int slot, size_t b)
    {
        if (b > bytes[slot]) bytes[slot] = b;
    }
    void merge(const ScratchPlan &o)
    {
        for (int i = 0; i < SLOT_COUNT; i++) need(i, o.bytes[i]);
    }
    // bytes of a caller workspace holding every slot 256-byte aligned
    size_t total() const
    {
        size_t t = 256;  // room to align the workspace base
        for (size_t b : bytes)
            if (b) t += (b + 255) / 256 * 256;
        return t;
    }
};

}  // namespace mti

struct mti_ctx {
    int device = 0;
    int num_cus = 0;
    std::string last_error;
    // profiling: (stage, start, end) events recorded on the caller's stream
    bool profiling = false;
    struct StageEv {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<StageEv> stage_events;
    std::vector<hipEvent_t> event_pool;
    mti::Scratch scratch[mti::SLOT_COUNT];
    // FFT twiddle tables keyed by length (FP64, exp(+2 pi i k / n)).
    std::vector<std::pair<int, double2 *>> twiddles;
    // general-size transform tables (twiddles, Bluestein chirps) keyed by (nx, ny)
    struct GenTable {
        int nx, ny;
        double2 *ptr;
        size_t bytes;
    };
    std::vector<GenTable> gen_tables;
    // caller workspace (mti_set_workspace): while a planned call runs
    // (ws_active), every slot is a fixed part of it and nothing is allocated
    char *ws_ptr = nullptr;
    size_t ws_bytes = 0;
    bool ws_active = false;
    int call_depth = 0;
    size_t ws_off[mti::SLOT_COUNT] = {};
    size_t ws_len[mti::SLOT_COUNT] = {};
    // launch-shape tuning (mti_set_tuning)
    int unit_pairs = 2048;       // (record, tile) pairs per gridder work unit
    int whole_columns_ny = 256;  // largest ny transformed by whole columns when the split also applies
    int xcd_chunk = 0;           // gridder units dealt to each XCD in runs of this many (0: round-robin)
    int region_width = 8;        // gridder wave regions of region_width x (64 / region_width) cells
    // packed planes between the transform passes stored as complex64 (1) or
    // complex128 (0) (mti_set_transform_storage)
    int xform_c64 = 0;
    // HIP-graph capture (mti_graph_create): calls enqueue device work only --
    // no host synchronisation, no statistics read-back, host tables staged
    // in pinned memory that lives as long as the graph
    bool capturing = false;
    char *stage_ptr = nullptr;
    size_t stage_bytes = 0, stage_used = 0, stage_need = 0;
};

struct mti_kernels {
    int device = 0;
    mti_kernel_info info{};
    std::vector<int32_t> support;    // host copy
    std::vector<int64_t> base;       // host copy
    void *d_mem = nullptr;           // one allocation holding the three arrays below
    int2 *d_convinfo = nullptr;      // {support, base} per plane (base < 2^31)
    void *d_table = nullptr;         // complex samples, precision-dependent
    float *d_sums = nullptr;         // [nplanes][T*T]
};

namespace mti {


// Device-pointer lookup of the allocation-free sort's static scratch, once
// per device (mti_sort.hip).
void static_scratch_setup(int device);

inline mti_status set_error(mti_ctx *ctx, mti_status code, const std::string &m)
{
    if (ctx) ctx->last_error = m;
    return code;
}

#define MTI_HIP(ctx, call)                                                     \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return ::mti::set_error((ctx), MTI_ERR_HIP,                        \
                                    std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

#define MTI_CHECK(ctx, cond, msg)                                              \
    do {                                                                       \
        if (!(cond)) return ::mti::set_error((ctx), MTI_ERR_INVALID, (msg));   \
    } while (0)

inline mti_status scratch_get(mti_ctx *ctx, int slot, size_t bytes, void **out)
{
    if (slot < 0 || slot >= SLOT_COUNT)
        return set_error(ctx, MTI_ERR_INVALID, "scratch slot out of range");
    if (ctx->ws_active) {
        // a slot of the caller workspace, sized by the call's plan
        if (bytes > ctx->ws_len[slot])
            return set_error(ctx, MTI_ERR_INVALID,
                             "scratch request of " + std::to_string(bytes) + " bytes in slot " +
                                 std::to_string(slot) + " above the call's workspace plan (" +
                                 std::to_string(ctx->ws_len[slot]) + ")");
        *out = ctx->ws_ptr + ctx->ws_off[slot];
        return MTI_OK;
    }
    Scratch &s = ctx->scratch[slot];
    if (s.bytes < bytes) {
        if (s.ptr) {
            hipError_t e = hipFree(s.ptr);
            (void)e;
            s.ptr = nullptr;
            s.bytes = 0;
        }
        size_t want = bytes + bytes / 8 + 256;
        hipError_t e = hipMalloc(&s.ptr, want);
        if (e != hipSuccess)
            return set_error(ctx, MTI_ERR_NOMEM,
                             std::string("scratch hipMalloc failed: ") + hipGetErrorString(e));
        s.bytes = want;
    }
    *out = s.ptr;
    return MTI_OK;
}

// One entry-point call.  The outermost call on a context with a caller
// workspace lays the workspace out by the call's plan (begin); nested
// entry-point calls run inside that layout.  Without a workspace it does
// nothing: scratch comes from the context.
struct CallScope {
    mti_ctx *ctx;
    bool outer;
    explicit CallScope(mti_ctx *c) : ctx(c), outer(c->call_depth == 0) { c->call_depth++; }
    ~CallScope()
    {
        ctx->call_depth--;
        if (outer) ctx->ws_active = false;
    }
    bool planning() const { return outer && ctx->ws_ptr != nullptr; }
    mti_status begin(const ScratchPlan &pl)
    {
        const uintptr_t base = ((uintptr_t)ctx->ws_ptr + 255) & ~uintptr_t(255);
        size_t off = base - (uintptr_t)ctx->ws_ptr;
        for (int i = 0; i < SLOT_COUNT; i++) {
            ctx->ws_off[i] = off;
            ctx->ws_len[i] = pl.bytes[i];
            off += (pl.bytes[i] + 255) / 256 * 256;
        }
        if (off > ctx->ws_bytes)
            return set_error(ctx, MTI_ERR_NOMEM,
                             "caller workspace of " + std::to_string(ctx->ws_bytes) + " bytes is below the " +
                                 std::to_string(pl.total()) + " bytes this call needs");
        ctx->ws_active = true;
        return MTI_OK;
    }
};

// Host-to-device copy of a call's small tables (job tables, descriptors,
// twiddles).  Outside graph capture the call synchronises before the host
// copy goes away; while a graph is captured the bytes go through the
// graph's pinned staging, which the captured copy then reads at every launch.
inline mti_status upload(mti_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t st)
{
    if (bytes == 0) return MTI_OK;
    if (ctx->capturing) {
        const size_t off = (ctx->stage_used + 255) / 256 * 256;
        ctx->stage_used = ctx->stage_need = off + bytes;
        // past the end of the staging: keep counting and record nothing; the
        // capture is then discarded and repeated with stage_need bytes
        if (off + bytes > ctx->stage_bytes) return MTI_OK;
        std::memcpy(ctx->stage_ptr + off, src, bytes);
        src = ctx->stage_ptr + off;
    }
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return set_error(ctx, MTI_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    return MTI_OK;
}

// Waits for the stream unless a graph is being captured.
inline mti_status host_sync(mti_ctx *ctx, hipStream_t st)
{
    if (ctx->capturing) return MTI_OK;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(ctx, MTI_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return MTI_OK;
}

inline hipStream_t to_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// dw of cfg/WProjectionConvFunction.cpp:105-112 in FP64.
inline double w_increment(const mti_image_params *p)
{
    if (p->mode != MTI_MODE_WPROJ || p->wplanes <= 1) return 1.0;
    double wmax = 0.25 / p->cell_l;
    double wm1 = (double)(p->wplanes - 1);
    return wmax / (wm1 * wm1);
}

// RAII bracket of one stage with HIP events (no-op unless profiling).
struct StageTimer {
    mti_ctx *ctx;
    hipStream_t st;
    int stage;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(mti_ctx *c, int s, hipStream_t stream);
    ~StageTimer();
};

mti_status validate_params(mti_ctx *ctx, const mti_image_params *p);
mti_status validate_vis(mti_ctx *ctx, const mti_vis *v);

// Exclusive scan of n uint32 counts (in place) on the stream; returns total
// in *d_total (device).  Uses scratch SLOT_MISC.
mti_status exclusive_scan_u32(mti_ctx *ctx, uint32_t *d_data, int64_t n,
                              uint32_t *d_total, hipStream_t st);
void plan_scan(int64_t n, ScratchPlan &pl);

}  // namespace mti

// Device helpers --------------------------------------------------------------
namespace mti_dev {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// psi(nu) of op/CasaGridSF.cu:41-85 (Schwab M=6, alpha=1), FP64.
__device__ __forceinline__ double gridsf(double nu)
{
    const double P0[5] = {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1};
    const double P1[5] = {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2};
    const double Q0[3] = {1.0, 8.212018e-1, 2.078043e-1};
    const double Q1[3] = {1.0, 9.599102e-1, 2.918724e-1};
    bool hi = !(nu < 0.75);
    const double *P = hi ? P1 : P0;
    const double *Q = hi ? Q1 : Q0;
    double d = nu * nu - (hi ? 1.0 : 0.5625);
    double top = P[0], bot = Q[0], dk = 1.0;
#pragma unroll
    for (int k = 1; k < 5; k++) {
        dk *= d;
        top += P[k] * dk;
    }
    dk = 1.0;
#pragma unroll
    for (int k = 1; k < 3; k++) {
        dk *= d;
        bot += Q[k] * dk;
    }
    return bot != 0.0 ? top / bot : 0.0;
}

}  // namespace mti_dev
