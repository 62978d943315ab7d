// mti_grid.hip -- convolutional gridding of many channels (rows a5-a7, a15, a21).
//
// Reference: ConvolutionGridder (op/ConvolutionGridder.cu:183-550) scatters
// grid[corner_y+sy][corner_x+sx][p] += K[conv_idx + sy*S + sx] * V_p, with
// V_p = vis_p * w_p * (1-flag_p) and the imaginary part negated for w<0
// (op/common.hcu:80-104), in FP32 with float atomics; the normaliser is the
// sum of sums[k][o] * w_p * (1-flag_p) (op/WProjPreProcessUVW.cu:84-108,
// comp/WImager.cpp:196-197).  Its work plan (AccumulatedSum,
// CreateIndexAndReorder, ValueChangeDetect, CreateBlockDataIndex,
// NotSameSupportManipulate, CreateCompressionPlan, CompressVisibilities)
// exists to feed that gridder runs of records of equal support.
//
// MI355X design:
//   1. binning pass (mti_bin.hip): one record per (row, channel), FP64 bins,
//      key = (channel, uv tile of the support-box corner);
//   2. stable radix sort of records by key (mti_sort.hip) -- keeps the
//      baseline-major order inside every tile;
//   3. tile gridder: a work item is (tile, <= R records).  A 512-thread
//      workgroup zeroes an LDS sub-grid covering the tile plus a halo of
//      S_max-1 cells (planar [re/im x pol][cell] layout, bank-skewed), each
//      wave walks its records with the 64 lanes spread over the S^2 taps
//      (kernel index conv_idx + t needs no division), accumulates with LDS
//      float atomics, then the sub-grid is added to HBM once with coalesced
//      float atomics on non-zero cells only.  Persistent grid of workgroups,
//      every wave exits when the item counter is exhausted.
//   4. optional uniform/Briggs weights from a per-channel density grid.
#include "mti_grid.hpp"
#include "mti_sort.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

namespace mti {

constexpr int GRID_THREADS = 512;
constexpr int GRID_WAVES = GRID_THREADS / 64;
constexpr int RECS_PER_ITEM = 2048;

struct TileArgs {
    VisView v;
    int chan0;
    const uint32_t *sorted_keys;  // unused by the gridder (kept for debugging)
    const int32_t *sorted_vals;
    const int4 *desc;
    const uint32_t *bucket_start;   // [nbuckets]
    const uint32_t *bucket_count;   // [nbuckets]
    const uint32_t *item_prefix;    // [nbuckets] exclusive scan of items per bucket
    const uint32_t *total_items;    // [1]
    uint32_t nbuckets;
    const void *table;
    const float *sums;
    const int2 *convinfo;
    int T2;
    int TS, LW, PS, ntx, ntiles;
    int nx, ny;
    float *grids;                   // [nb][ny][nx][P][2]
    double *norm;                   // [nb][P]
    int psf;
    int weighting;
    const float *density;           // [nb][P][ny][nx]
    const double *f2;               // [nb][P]
};

template <int PREC>
__device__ __forceinline__ float2 ksample(const void *table, int idx)
{
    if (PREC == MTI_TABLE_F32) return ((const float2 *)table)[idx];
    if (PREC == MTI_TABLE_F16) return __half22float2(((const __half2 *)table)[idx]);
    __hip_bfloat162 h = ((const __hip_bfloat162 *)table)[idx];
    return make_float2(__bfloat162float(h.x), __bfloat162float(h.y));
}

// Effective weight of correlation p of a record (flag, uniform/Briggs).
__device__ __forceinline__ float eff_weight(const TileArgs &a, int64_t row, int chan, int c_local,
                                            int p, int cell_x, int cell_y)
{
    if (vis_flag(a.v, row, chan, p)) return 0.0f;
    float w = a.v.weight[row * a.v.w_rs + p];
    if (a.weighting == MTI_WEIGHT_NATURAL) return w;
    const int npol = a.v.npol;
    const float d = a.density[(((int64_t)c_local * npol + p) * a.ny + cell_y) * a.nx + cell_x];
    if (a.weighting == MTI_WEIGHT_UNIFORM) return d > 0.0f ? w / d : 0.0f;
    return (float)((double)w / (1.0 + a.f2[c_local * npol + p] * (double)d));
}

template <int P, int PREC>
__global__ __launch_bounds__(GRID_THREADS) void k_grid_tiles(TileArgs a)
{
    extern __shared__ float lds[];
    __shared__ int s_item[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t total = *a.total_items;
    const int PS = a.PS, LW = a.LW;
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        if (tid == 0) {
            // bucket b: last index with item_prefix[b] <= item
            uint32_t lo = 0, hi = a.nbuckets - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (a.item_prefix[mid] <= item) lo = mid;
                else hi = mid - 1;
            }
            const uint32_t b = lo;
            const uint32_t chunk = item - a.item_prefix[b];
            const uint32_t st = a.bucket_start[b] + chunk * RECS_PER_ITEM;
            const uint32_t en = min(a.bucket_start[b] + a.bucket_count[b], st + RECS_PER_ITEM);
            s_item[0] = (int)b;
            s_item[1] = (int)st;
            s_item[2] = (int)en;
        }
        for (int i = tid; i < 2 * P * PS; i += GRID_THREADS) lds[i] = 0.0f;
        __syncthreads();
        const int b = s_item[0], st = s_item[1], en = s_item[2];
        const int c_local = b / a.ntiles;
        const int tile = b - c_local * a.ntiles;
        const int ty = tile / a.ntx, tx = tile - ty * a.ntx;
        const int x0 = tx * a.TS, y0 = ty * a.TS;
        const int chan = a.chan0 + c_local;
        double nacc[P];
#pragma unroll
        for (int p = 0; p < P; p++) nacc[p] = 0.0;

        for (int j = st + wave; j < en; j += GRID_WAVES) {
            const int e = a.sorted_vals[j];
            const int4 d = a.desc[e];
            const int64_t row = d.x;
            const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16);
            const int conv_idx = d.z;
            const int plane = (int)((unsigned)d.w >> 16);
            const bool conj = (d.w >> 15) & 1;
            const int S = d.w & 0x7fff;
            float vr[P], vi[P];
#pragma unroll
            for (int p = 0; p < P; p++) {
                const float w = eff_weight(a, row, chan, c_local, p, cx + S / 2, cy + S / 2);
                float re = 1.0f, im = 0.0f;
                if (!a.psf) {
                    const float2 vv =
                        *(const float2 *)(a.v.vis + 2 * (row * a.v.vis_rs + (int64_t)chan * a.v.vis_cs + p));
                    re = vv.x;
                    im = vv.y;
                }
                vr[p] = re * w;
                vi[p] = conj ? -(im * w) : im * w;
                if (lane == 0) {
                    const int S2 = S * S;
                    const int o = (conv_idx - a.convinfo[plane].y) / S2;
                    nacc[p] += (double)a.sums[plane * a.T2 + o] * (double)w;
                }
            }
            const int cell_base = (cy - y0) * LW + (cx - x0);
            const int S2 = S * S;
            const float invS = 1.0f / (float)S;
            for (int t = lane; t < S2; t += 64) {
                const int sy = (int)(((float)t + 0.5f) * invS);
                const int sx = t - sy * S;
                const float2 k = ksample<PREC>(a.table, conv_idx + t);
                const int cell = cell_base + sy * LW + sx;
#pragma unroll
                for (int p = 0; p < P; p++) {
                    atomicAdd(&lds[(2 * p) * PS + cell], vr[p] * k.x - vi[p] * k.y);
                    atomicAdd(&lds[(2 * p + 1) * PS + cell], vi[p] * k.x + vr[p] * k.y);
                }
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int p = 0; p < P; p++)
                if (nacc[p] != 0.0) unsafeAtomicAdd(&a.norm[c_local * P + p], nacc[p]);
        }
        __syncthreads();
        // flush the sub-grid: coalesced along grid rows, [cell][pol][re,im]
        const int rows = min(LW, a.ny - y0), cols = min(LW, a.nx - x0);
        float *g = a.grids + (size_t)c_local * a.nx * a.ny * (2 * P);
        constexpr int C = 2 * P;
        for (int yy = wave; yy < rows; yy += GRID_WAVES) {
            float *grow = g + ((size_t)(y0 + yy) * a.nx + x0) * C;
            for (int x = lane; x < cols * C; x += 64) {
                const int comp = x & (C - 1);
                const int xx = x / C;
                const float val = lds[comp * PS + yy * LW + xx];
                if (val != 0.0f) unsafeAtomicAdd(&grow[x], val);
            }
        }
        __syncthreads();
    }
}

// Fallback for supports too large for an LDS sub-grid: one wave per record,
// taps straight to HBM with float atomics (the reference's scatter pattern).
template <int P, int PREC>
__global__ __launch_bounds__(256) void k_grid_global(TileArgs a, int64_t nrec, const uint32_t *keys)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave_id = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t e = wave_id; e < nrec; e += nwaves) {
        const uint32_t key = keys[e];
        if (key >= a.nbuckets) continue;
        const int c_local = (int)(key / (uint32_t)a.ntiles);
        const int chan = a.chan0 + c_local;
        const int4 d = a.desc[e];
        const int64_t row = d.x;
        const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16);
        const int conv_idx = d.z;
        const int plane = (int)((unsigned)d.w >> 16);
        const bool conj = (d.w >> 15) & 1;
        const int S = d.w & 0x7fff;
        float vr[P], vi[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            const float w = eff_weight(a, row, chan, c_local, p, cx + S / 2, cy + S / 2);
            float re = 1.0f, im = 0.0f;
            if (!a.psf) {
                const float2 vv =
                    *(const float2 *)(a.v.vis + 2 * (row * a.v.vis_rs + (int64_t)chan * a.v.vis_cs + p));
                re = vv.x;
                im = vv.y;
            }
            vr[p] = re * w;
            vi[p] = conj ? -(im * w) : im * w;
            if (lane == 0) {
                const int o = (conv_idx - a.convinfo[plane].y) / (S * S);
                unsafeAtomicAdd(&a.norm[c_local * P + p], (double)a.sums[plane * a.T2 + o] * (double)w);
            }
        }
        float *g = a.grids + (size_t)c_local * a.nx * a.ny * (2 * P);
        const int S2 = S * S;
        const float invS = 1.0f / (float)S;
        for (int t = lane; t < S2; t += 64) {
            const int sy = (int)(((float)t + 0.5f) * invS);
            const int sx = t - sy * S;
            const float2 k = ksample<PREC>(a.table, conv_idx + t);
            float *c = g + ((size_t)(cy + sy) * a.nx + (cx + sx)) * (2 * P);
#pragma unroll
            for (int p = 0; p < P; p++) {
                unsafeAtomicAdd(&c[2 * p], vr[p] * k.x - vi[p] * k.y);
                unsafeAtomicAdd(&c[2 * p + 1], vi[p] * k.x + vr[p] * k.y);
            }
        }
    }
}

__global__ void k_fill_f32(float *p, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_items(const uint32_t *counts, uint32_t n, uint32_t *items)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) items[i] = (counts[i] + RECS_PER_ITEM - 1) / RECS_PER_ITEM;
}

// Per (channel, pol) plane: sum W and sum W^2 of the density grid.
__global__ __launch_bounds__(256) void k_density_stats(const float *dens, int64_t plane_sz,
                                                      double *acc /* [planes][2] */)
{
    const int plane = blockIdx.y;
    const float *d = dens + (int64_t)plane * plane_sz;
    double s1 = 0, s2 = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < plane_sz;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = d[i];
        s1 += v;
        s2 += v * v;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    __shared__ double r1[4], r2[4];
    if ((threadIdx.x & 63) == 0) {
        r1[threadIdx.x >> 6] = s1;
        r2[threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = r1[0] + r1[1] + r1[2] + r1[3], t2 = r2[0] + r2[1] + r2[2] + r2[3];
        unsafeAtomicAdd(&acc[2 * plane], t1);
        unsafeAtomicAdd(&acc[2 * plane + 1], t2);
    }
}

// Briggs f^2 = (5*10^-R)^2 / (sum W^2 / sum w)
__global__ void k_briggs_f2(const double *acc, int n, double robust, double *f2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sw = acc[2 * i], sw2 = acc[2 * i + 1];
    double k = 5.0 * pow(10.0, -robust);
    f2[i] = sw2 > 0.0 ? (k * k) / (sw2 / sw) : 0.0;
}

static int lds_plane_stride(int LW, int P)
{
    int lw2 = LW * LW;
    int want = 64 / (2 * P);  // bank skew between component planes
    int ps = (lw2 + 63) / 64 * 64 + want;
    return ps;
}

template <int P, int PREC>
static mti_status launch_tiles(mti_ctx *ctx, const TileArgs &a, size_t lds_bytes, hipStream_t st)
{
    auto kern = k_grid_tiles<P, PREC>;
    MTI_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_bytes));
    int per_cu = 0;
    MTI_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, GRID_THREADS, lds_bytes));
    if (per_cu < 1) per_cu = 1;
    int blocks = ctx->num_cus * per_cu;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(GRID_THREADS), lds_bytes, st, a);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int P>
static mti_status dispatch_tiles(mti_ctx *ctx, int prec, const TileArgs &a, size_t lds, hipStream_t st)
{
    if (prec == MTI_TABLE_F32) return launch_tiles<P, MTI_TABLE_F32>(ctx, a, lds, st);
    if (prec == MTI_TABLE_F16) return launch_tiles<P, MTI_TABLE_F16>(ctx, a, lds, st);
    return launch_tiles<P, MTI_TABLE_BF16>(ctx, a, lds, st);
}

template <int P>
static mti_status dispatch_global(mti_ctx *ctx, int prec, const TileArgs &a, int64_t nrec,
                                  const uint32_t *keys, hipStream_t st)
{
    dim3 gr((unsigned)std::min<int64_t>(ceil_div(nrec, 4), (int64_t)ctx->num_cus * 16)), bl(256);
    if (prec == MTI_TABLE_F32)
        hipLaunchKernelGGL((k_grid_global<P, MTI_TABLE_F32>), gr, bl, 0, st, a, nrec, keys);
    else if (prec == MTI_TABLE_F16)
        hipLaunchKernelGGL((k_grid_global<P, MTI_TABLE_F16>), gr, bl, 0, st, a, nrec, keys);
    else
        hipLaunchKernelGGL((k_grid_global<P, MTI_TABLE_BF16>), gr, bl, 0, st, a, nrec, keys);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

static inline size_t align_up(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_grid(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                               const mti_vis *v, const mti_grid_opts *o, float *grids, double *norm,
                               mti_grid_stats *stats, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    r = validate_vis(ctx, v);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, k && o && grids && norm, "null kernels/options/outputs");
    MTI_CHECK(ctx, k->info.sampling == p->sampling, "kernels built for another sampling");
    MTI_CHECK(ctx, o->nchan >= 1, "opts.nchan must be >= 1");
    if (v->inv_lambda_per_row)
        MTI_CHECK(ctx, o->nchan == 1, "per-row 1/lambda implies a single channel per call");
    else
        MTI_CHECK(ctx, o->chan0 >= 0 && o->chan0 + o->nchan <= v->nchan, "channel range out of bounds");
    MTI_CHECK(ctx, o->weighting >= MTI_WEIGHT_NATURAL && o->weighting <= MTI_WEIGHT_BRIGGS, "bad weighting");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);

    const int P = v->npol, nb = o->nchan;
    const int64_t gsz = (int64_t)p->nx * p->ny;
    VisView vw = make_view(v);
    const int64_t nrec = vw.nidx * nb;
    MTI_CHECK(ctx, nrec < (int64_t(1) << 31), "more than 2^31-1 records (rows x channels) per call");

    if (!o->accumulate) {
        hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st, grids, gsz * nb * 2 * P, 0.0f);
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, norm, (int64_t)nb * P, 0.0);
        MTI_HIP(ctx, hipGetLastError());
    }
    if (stats) *stats = mti_grid_stats{};
    if (nrec == 0) return MTI_OK;

    // tile size: largest multiple of 8 whose LDS sub-grid fits the budget
    const int smax = std::max(1, k->info.max_support);
    int TS = 0, LW = 0, PS = 0;
    size_t lds = 0;
    for (size_t budget : {size_t(80) << 10, size_t(160) << 10}) {
        for (int ts = 256; ts >= 8; ts -= 8) {
            int lw = ts + smax - 1;
            int ps = lds_plane_stride(lw, P);
            size_t need = (size_t)2 * P * ps * sizeof(float);
            if (need <= budget) {
                TS = ts;
                LW = lw;
                PS = ps;
                lds = need;
                break;
            }
        }
        if (TS) break;
    }
    const bool global_path = (TS == 0);
    if (global_path) TS = 64;
    const int ntx = (int)ceil_div(p->nx, TS), nty = (int)ceil_div(p->ny, TS);
    const int64_t ntiles = (int64_t)ntx * nty;
    const int64_t nbuckets64 = ntiles * nb;
    MTI_CHECK(ctx, nbuckets64 < (int64_t(1) << 31), "too many (channel, tile) buckets");
    const uint32_t nbuckets = (uint32_t)nbuckets64;
    int key_bits = 1;
    while (key_bits < 32 && (((uint64_t)nbuckets) >> key_bits) != 0) key_bits++;

    // scratch layout
    void *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *w4 = nullptr;
    r = scratch_get(ctx, SLOT_KEYS, align_up(nrec * 4) * 4, &w1);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_DESC, (size_t)nrec * sizeof(int4), &w2);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_BUCKETS, align_up(nbuckets64 * 4) * 3 + 1024, &w3);
    if (r != MTI_OK) return r;
    uint32_t *keys = (uint32_t *)w1;
    int32_t *vals = (int32_t *)((char *)w1 + align_up(nrec * 4));
    uint32_t *keys_t = (uint32_t *)((char *)w1 + 2 * align_up(nrec * 4));
    int32_t *vals_t = (int32_t *)((char *)w1 + 3 * align_up(nrec * 4));
    int4 *desc = (int4 *)w2;
    uint32_t *counts = (uint32_t *)w3;
    uint32_t *starts = (uint32_t *)((char *)w3 + align_up(nbuckets64 * 4));
    uint32_t *items = (uint32_t *)((char *)w3 + 2 * align_up(nbuckets64 * 4));
    char *tail = (char *)w3 + 3 * align_up(nbuckets64 * 4);
    uint32_t *tot_rec = (uint32_t *)tail;
    uint32_t *tot_items = (uint32_t *)(tail + 256);
    unsigned long long *dstats = (unsigned long long *)(tail + 512);

    float *density = nullptr;
    double *dacc = nullptr, *f2 = nullptr;
    if (o->weighting != MTI_WEIGHT_NATURAL) {
        r = scratch_get(ctx, SLOT_DENSITY, (size_t)gsz * nb * P * sizeof(float), &w4);
        if (r != MTI_OK) return r;
        density = (float *)w4;
        void *w5 = nullptr;
        r = scratch_get(ctx, SLOT_NORM, (size_t)nb * P * 3 * sizeof(double) + 256, &w5);
        if (r != MTI_OK) return r;
        dacc = (double *)w5;
        f2 = dacc + 2 * nb * P;
        hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st, density, gsz * nb * P, 0.0f);
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, dacc, (int64_t)nb * P * 2, 0.0);
    }
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<int64_t>(ceil_div(nbuckets64, 256), 4096)),
                       dim3(256), 0, st, counts, nbuckets64, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (uint32_t *)dstats, 8, 0u);
    MTI_HIP(ctx, hipGetLastError());

    BinPassArgs ba;
    ba.v = vw;
    ba.g = make_geom(k, p);
    ba.chan0 = o->chan0;
    ba.nb = nb;
    ba.TS = TS;
    ba.ntx = ntx;
    ba.ntiles = (int)ntiles;
    ba.nbuckets = nbuckets;
    ba.keys = keys;
    ba.vals = vals;
    ba.desc = desc;
    ba.counts = counts;
    ba.density = density;
    ba.stats = dstats;
    {
        StageTimer tm(ctx, MTI_STAGE_BIN, st);
        r = launch_bin_pass(ctx, ba, nrec, st);
        if (r != MTI_OK) return r;
        if (density) {
            dim3 gr(64, (unsigned)(nb * P));
            hipLaunchKernelGGL(k_density_stats, gr, dim3(256), 0, st, density, gsz, dacc);
            hipLaunchKernelGGL(k_briggs_f2, dim3((unsigned)ceil_div(nb * P, 64)), dim3(64), 0, st, dacc,
                               nb * P, o->robust, f2);
            MTI_HIP(ctx, hipGetLastError());
        }
    }

    TileArgs ta;
    ta.v = vw;
    ta.chan0 = o->chan0;
    ta.desc = desc;
    ta.nbuckets = nbuckets;
    ta.table = k->d_table;
    ta.sums = k->d_sums;
    ta.convinfo = k->d_convinfo;
    ta.T2 = k->info.T * k->info.T;
    ta.TS = TS;
    ta.LW = LW;
    ta.PS = PS;
    ta.ntx = ntx;
    ta.ntiles = (int)ntiles;
    ta.nx = p->nx;
    ta.ny = p->ny;
    ta.grids = grids;
    ta.norm = norm;
    ta.psf = o->psf;
    ta.weighting = o->weighting;
    ta.density = density;
    ta.f2 = f2;

    const int prec = k->info.precision;
    if (global_path) {
        StageTimer tm(ctx, MTI_STAGE_GRID, st);
        ta.sorted_keys = keys;
        ta.sorted_vals = vals;
        switch (P) {
        case 1: r = dispatch_global<1>(ctx, prec, ta, nrec, keys, st); break;
        case 2: r = dispatch_global<2>(ctx, prec, ta, nrec, keys, st); break;
        default: r = dispatch_global<4>(ctx, prec, ta, nrec, keys, st); break;
        }
        if (r != MTI_OK) return r;
    } else {
        bool in_tmp = false;
        {
            StageTimer tm(ctx, MTI_STAGE_SORT_RECORDS, st);
            r = radix_sort_pairs(ctx, keys, vals, keys_t, vals_t, nrec, key_bits, st, &in_tmp);
            if (r != MTI_OK) return r;
        }
        ta.sorted_keys = in_tmp ? keys_t : keys;
        ta.sorted_vals = in_tmp ? vals_t : vals;
        {
            StageTimer tm(ctx, MTI_STAGE_PLAN, st);
            MTI_HIP(ctx, hipMemcpyAsync(starts, counts, nbuckets64 * 4, hipMemcpyDeviceToDevice, st));
            r = exclusive_scan_u32(ctx, starts, nbuckets64, tot_rec, st);
            if (r != MTI_OK) return r;
            hipLaunchKernelGGL(k_items, dim3((unsigned)ceil_div(nbuckets64, 256)), dim3(256), 0, st, counts,
                               nbuckets, items);
            MTI_HIP(ctx, hipGetLastError());
            r = exclusive_scan_u32(ctx, items, nbuckets64, tot_items, st);
            if (r != MTI_OK) return r;
        }
        ta.bucket_start = starts;
        ta.bucket_count = counts;
        ta.item_prefix = items;
        ta.total_items = tot_items;
        StageTimer tm(ctx, MTI_STAGE_GRID, st);
        switch (P) {
        case 1: r = dispatch_tiles<1>(ctx, prec, ta, lds, st); break;
        case 2: r = dispatch_tiles<2>(ctx, prec, ta, lds, st); break;
        default: r = dispatch_tiles<4>(ctx, prec, ta, lds, st); break;
        }
        if (r != MTI_OK) return r;
    }
    if (stats) {
        unsigned long long h[2] = {0, 0};
        uint32_t ni = 0;
        MTI_HIP(ctx, hipMemcpyAsync(h, dstats, sizeof(h), hipMemcpyDeviceToHost, st));
        if (!global_path) MTI_HIP(ctx, hipMemcpyAsync(&ni, tot_items, 4, hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipStreamSynchronize(st));
        stats->gridded_records = (int64_t)h[0];
        stats->gridded_points = (int64_t)h[1];
        stats->work_items = ni;
        stats->tile = global_path ? 0 : TS;
        stats->path = global_path ? 1 : 0;
    }
    return MTI_OK;
}
