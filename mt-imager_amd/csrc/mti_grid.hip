// mti_grid.hip -- convolutional gridding of many channels (rows a5-a7, a15, a21).
//
// Reference: ConvolutionGridder (op/ConvolutionGridder.cu:183-550) scatters
// grid[corner_y+sy][corner_x+sx][p] += K[conv_idx + sy*S + sx] * V_p, with
// V_p = vis_p * w_p * (1-flag_p) and the imaginary part negated for w<0
// (op/common.hcu:80-104); the normaliser is the sum of sums[k][o] * w_p *
// (1-flag_p) (op/WProjPreProcessUVW.cu:84-108, comp/WImager.cpp:196-197).  Its
// thread (bu,bv) owns the grid points congruent to (bu,bv) mod S and keeps a
// register accumulator while consecutive records leave its point unchanged
// (:102-129, :394-441), fed by a work plan of <= 440-record runs of equal
// support (AccumulatedSum, CreateIndexAndReorder, ValueChangeDetect,
// CreateBlockDataIndex, NotSameSupportManipulate).
//
// MI355X design:
//   1. binning pass (mti_bin.hip): one record per (row, channel) in
//      channel-major, baseline-major order; FP64 bins; weight density for
//      uniform/Briggs with wave-segmented atomics;
//   2. stable compaction of gridded records (exclusive scan + scatter) and
//      the work list: items are runs of <= 512 records of one channel with
//      equal support (scan of boundary flags);
//   3. gridder: persistent 256-thread workgroups; an item's records (weighted
//      visibilities, corner, kernel offset) are staged in LDS once; every
//      lane owns tap classes (bu,bv) mod S -- several sub-streams of the run
//      in parallel when S^2 is small, up to 5 classes per lane when it is
//      large -- and accumulates in registers while its grid point is
//      unchanged, flushing to HBM with float atomics only when it moves.
//      The planar grid layout makes a wave's flush one contiguous segment.
#include "mti_grid.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

namespace mti {

constexpr int RG_THREADS = 256;
constexpr int RG_RECS = 512;     // records per item
constexpr int RG_JMAX = 5;       // tap classes per lane
constexpr int RG_CLASSES = RG_THREADS * RG_JMAX;

struct RomeinArgs {
    VisView v;
    int chan0;
    const int4 *cdesc;              // compacted record descriptors [M]
    const int32_t *cchan;           // channel (local) of each compacted record
    const uint32_t *item_start;     // [NI]
    const uint32_t *n_items;        // device scalar NI
    const uint32_t *n_rec;          // device scalar M
    const void *table;
    const float *sums;
    const int2 *convinfo;
    int T2;
    int nx, ny;
    int64_t cs, ps;                 // complex strides of cell and correlation
    int64_t chan_stride;            // complex elements per channel grid
    double *grid;                   // FP64 accumulation grid
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
__device__ __forceinline__ float eff_weight(const RomeinArgs &a, int64_t row, int chan, int c_local,
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

template <int P>
__device__ __forceinline__ void flush_point(double *g, int cell, int64_t cs, int64_t ps, const double *acc)
{
    if (cell < 0) return;
#pragma unroll
    for (int p = 0; p < P; p++) {
        double *c = g + 2 * ((int64_t)cell * cs + (int64_t)p * ps);
        if (acc[2 * p] != 0.0) unsafeAtomicAdd(c, acc[2 * p]);
        if (acc[2 * p + 1] != 0.0) unsafeAtomicAdd(c + 1, acc[2 * p + 1]);
    }
}

// One work unit = (item, class chunk).  JMAX = tap classes per lane; UNR =
// records whose kernel samples are loaded ahead of the accumulation.
template <int P, int PREC, int JMAX, int UNR>
__global__ __launch_bounds__(RG_THREADS) void k_grid_romein(RomeinArgs a, const uint32_t *unit_prefix,
                                                           const uint32_t *n_units)
{
    __shared__ int4 s_rec[RG_RECS];            // corner cell, kernel offset, corner mod S
    __shared__ double s_vis[2 * P][RG_RECS];   // weighted visibility
    __shared__ double s_norm[RG_THREADS / 64][P];
    __shared__ int s_hdr[6];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t NU = *n_units, NI = *a.n_items, M = *a.n_rec;

    for (uint32_t unit = blockIdx.x; unit < NU; unit += gridDim.x) {
        if (tid == 0) {
            uint32_t lo = 0, hi = NI - 1;  // last item with unit_prefix <= unit
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (unit_prefix[mid] <= unit) lo = mid;
                else hi = mid - 1;
            }
            const uint32_t st = a.item_start[lo];
            const uint32_t en = (lo + 1 < NI) ? a.item_start[lo + 1] : M;
            s_hdr[0] = (int)st;
            s_hdr[1] = (int)(en - st);
            s_hdr[2] = (int)(unit - unit_prefix[lo]);  // class chunk
            s_hdr[3] = a.cdesc[st].w & 0x7fff;          // support
            s_hdr[4] = a.cchan[st];                     // channel (local)
        }
        __syncthreads();
        const int st = s_hdr[0], n = s_hdr[1], chunk = s_hdr[2], S = s_hdr[3], c_local = s_hdr[4];
        const int chan = a.chan0 + c_local;
        const int S2 = S * S;
        if (S == 0) {  // plane without kernel samples: contributes nothing
            __syncthreads();
            continue;
        }

        // ---- stage the item's records -------------------------------------
        double np[P];
#pragma unroll
        for (int p = 0; p < P; p++) np[p] = 0.0;
        for (int j = tid; j < n; j += RG_THREADS) {
            const int4 d = a.cdesc[st + j];
            const int64_t row = d.x;
            const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16);
            const int plane = (int)((unsigned)d.w >> 16);
            const bool conj = (d.w >> 15) & 1;
            s_rec[j] = make_int4(cy * a.nx + cx, d.z, (cx % S) | ((cy % S) << 16), 0);
            const int o = (d.z - a.convinfo[plane].y) / S2;
            const double csum = (double)a.sums[plane * a.T2 + o];
#pragma unroll
            for (int p = 0; p < P; p++) {
                const float w = eff_weight(a, row, chan, c_local, p, cx + S / 2, cy + S / 2);
                double re = 1.0, im = 0.0;
                if (!a.psf) {
                    const float2 vv =
                        *(const float2 *)(a.v.vis + 2 * (row * a.v.vis_rs + (int64_t)chan * a.v.vis_cs + p));
                    re = vv.x;
                    im = vv.y;
                }
                s_vis[2 * p][j] = re * (double)w;
                s_vis[2 * p + 1][j] = conj ? -(im * (double)w) : im * (double)w;
                np[p] += csum * (double)w;
            }
        }
        if (chunk == 0) {
#pragma unroll
            for (int p = 0; p < P; p++) {
                double v = np[p];
                for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) s_norm[wave][p] = v;
            }
        }
        __syncthreads();
        if (chunk == 0 && tid < P) {
            double t = 0;
            for (int w = 0; w < RG_THREADS / 64; w++) t += s_norm[w][tid];
            if (t != 0.0) unsafeAtomicAdd(&a.norm[c_local * P + tid], t);
        }

        // ---- Romein walk ---------------------------------------------------
        int r0, r1, cls0, J;
        if (JMAX == 1) {
            const int K = RG_THREADS / S2;  // parallel sub-streams of the run
            const int sub = tid / S2;
            cls0 = tid - sub * S2;
            if (sub < K) {
                J = 1;
                r0 = (int)(((int64_t)n * sub) / K);
                r1 = (int)(((int64_t)n * (sub + 1)) / K);
            } else {
                J = 0;
                r0 = r1 = 0;
            }
        } else {
            cls0 = chunk * RG_THREADS * JMAX + tid;
            J = JMAX;
            r0 = 0;
            r1 = n;
        }
        double *g = a.grid + 2 * (int64_t)c_local * a.chan_stride;
        int bx[JMAX], by[JMAX], cur[JMAX];
        double acc[JMAX][2 * P];
#pragma unroll
        for (int j = 0; j < JMAX; j++) {
            const int c = cls0 + j * RG_THREADS;
            const bool on = (j < J) && (c < S2);
            by[j] = on ? c / S : -1;
            bx[j] = on ? c - (c / S) * S : 0;
            cur[j] = -1;
#pragma unroll
            for (int q = 0; q < 2 * P; q++) acc[j][q] = 0.0;
        }
        for (int r = r0; r < r1; r += UNR) {
            int cellv[UNR][JMAX];
            float2 kv[UNR][JMAX];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const int rr = r + u;
                const int4 rec = s_rec[rr < r1 ? rr : r];
                const int rx = rec.z & 0xffff, ry = rec.z >> 16;
#pragma unroll
                for (int j = 0; j < JMAX; j++) {
                    int sx = bx[j] - rx, sy = by[j] - ry;
                    sx += (sx < 0) ? S : 0;
                    sy += (sy < 0) ? S : 0;
                    cellv[u][j] = rec.x + sy * a.nx + sx;
                    kv[u][j] = (by[j] >= 0 && rr < r1) ? ksample<PREC>(a.table, rec.y + sy * S + sx)
                                                        : make_float2(0.0f, 0.0f);
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const int rr = r + u;
                if (rr >= r1) break;
                double vr[P], vi[P];
#pragma unroll
                for (int p = 0; p < P; p++) {
                    vr[p] = s_vis[2 * p][rr];
                    vi[p] = s_vis[2 * p + 1][rr];
                }
#pragma unroll
                for (int j = 0; j < JMAX; j++) {
                    if (by[j] < 0) continue;
                    const int cell = cellv[u][j];
                    if (cell != cur[j]) {
                        flush_point<P>(g, cur[j], a.cs, a.ps, acc[j]);
                        cur[j] = cell;
#pragma unroll
                        for (int q = 0; q < 2 * P; q++) acc[j][q] = 0.0;
                    }
                    const double kx = kv[u][j].x, ky = kv[u][j].y;
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        acc[j][2 * p] += vr[p] * kx - vi[p] * ky;
                        acc[j][2 * p + 1] += vi[p] * kx + vr[p] * ky;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < JMAX; j++) flush_point<P>(g, cur[j], a.cs, a.ps, acc[j]);
        __syncthreads();
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

// Stable compaction.  pos is the exclusive scan of the valid flags, so
// record e is kept exactly when the next position differs from its own.
__global__ void k_compact(const uint32_t *pos, const uint32_t *n_rec, const int4 *desc, int64_t nrec,
                          int64_t nidx, int4 *cdesc, int32_t *cchan)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrec) return;
    const uint32_t j = pos[e];
    const uint32_t next = (e + 1 < nrec) ? pos[e + 1] : *n_rec;
    if (next == j) return;
    cdesc[j] = desc[e];
    cchan[j] = (int32_t)(e / nidx);
}

// Item boundaries: channel change, support change, or every RG_RECS records.
__global__ void k_item_flags(const int4 *cdesc, const int32_t *cchan, const uint32_t *n_rec, int64_t cap,
                             uint32_t *flag)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap) return;
    const int64_t M = *n_rec;
    uint32_t f = 0;
    if (j < M) {
        f = (j == 0) || (j % RG_RECS == 0) || (cchan[j] != cchan[j - 1]) ||
            ((cdesc[j].w & 0x7fff) != (cdesc[j - 1].w & 0x7fff));
    }
    flag[j] = f;
}

// ids = exclusive scan of the boundary flags: record j starts item ids[j]
// when the next id differs.  Writes item_start and the work units of every
// item for the two gridder variants: S^2 <= 256 (one tap class per lane,
// sub-streams) or larger (chunks of RG_CLASSES classes).
__global__ void k_item_starts(const uint32_t *ids, const uint32_t *n_items, const int4 *cdesc,
                              const uint32_t *n_rec, int64_t cap, uint32_t *item_start, uint32_t *units_small,
                              uint32_t *units_large)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap || j >= (int64_t)*n_rec) return;
    const uint32_t id = ids[j];
    const uint32_t next = (j + 1 < cap) ? ids[j + 1] : *n_items;
    if (next == id) return;
    item_start[id] = (uint32_t)j;
    const int S = cdesc[j].w & 0x7fff;
    if (S * S <= RG_THREADS) {
        units_small[id] = 1;
    } else {
        units_large[id] = (uint32_t)((S * S + RG_CLASSES - 1) / RG_CLASSES);
    }
}

// out = (accumulate ? out : 0) + work, FP64 work grid -> complex64 output.
__global__ void k_round_grid(const double *work, float *out, int64_t n, int accumulate)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = accumulate ? (float)((double)out[i] + work[i]) : (float)work[i];
}

__global__ void k_fill_f64x(double *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
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

template <int P, int PREC, int JMAX, int UNR>
static mti_status launch_romein(mti_ctx *ctx, const RomeinArgs &a, const uint32_t *prefix,
                                const uint32_t *nunits, hipStream_t st)
{
    auto kern = k_grid_romein<P, PREC, JMAX, UNR>;
    int per_cu = 0;
    MTI_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, RG_THREADS, 0));
    if (per_cu < 1) per_cu = 1;
    const int blocks = ctx->num_cus * per_cu;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(RG_THREADS), 0, st, a, prefix, nunits);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int P, int PREC>
static mti_status romein_pair(mti_ctx *ctx, const RomeinArgs &a, const uint32_t *ps, const uint32_t *ns,
                              const uint32_t *pl, const uint32_t *nl, hipStream_t st)
{
    mti_status r = launch_romein<P, PREC, 1, 4>(ctx, a, ps, ns, st);
    if (r != MTI_OK) return r;
    return launch_romein<P, PREC, RG_JMAX, 2>(ctx, a, pl, nl, st);
}

template <int P>
static mti_status dispatch_romein(mti_ctx *ctx, int prec, const RomeinArgs &a, const uint32_t *ps,
                                  const uint32_t *ns, const uint32_t *pl, const uint32_t *nl, hipStream_t st)
{
    if (prec == MTI_TABLE_F32) return romein_pair<P, MTI_TABLE_F32>(ctx, a, ps, ns, pl, nl, st);
    if (prec == MTI_TABLE_F16) return romein_pair<P, MTI_TABLE_F16>(ctx, a, ps, ns, pl, nl, st);
    return romein_pair<P, MTI_TABLE_BF16>(ctx, a, ps, ns, pl, nl, st);
}

static inline size_t align_up(size_t x) { return (x + 255) / 256 * 256; }

// Grids channels [chan0, chan0+nb) into the FP64 grid G (already zeroed or
// holding previous chunks); norm is accumulated.
static mti_status grid_batch(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                             const VisView &vw, const mti_grid_opts *o, int chan0, int nb, double *G,
                             double *norm, unsigned long long *dstats, uint32_t *ni_total, hipStream_t st)
{
    const int P = vw.npol;
    const int64_t gsz = (int64_t)p->nx * p->ny;
    const int64_t nrec = vw.nidx * nb;
    mti_status r;
    // scratch: valid, flags, item_start, units_small, units_large, cchan (u32 x nrec) + scalars;
    // desc + cdesc (int4 x nrec)
    const size_t u32n = align_up((size_t)nrec * 4);
    void *w1 = nullptr, *w2 = nullptr;
    r = scratch_get(ctx, SLOT_KEYS, u32n * 6 + 4096, &w1);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_DESC, align_up((size_t)nrec * 16) * 2, &w2);
    if (r != MTI_OK) return r;
    uint32_t *valid = (uint32_t *)w1;
    uint32_t *flag = (uint32_t *)((char *)w1 + u32n);
    uint32_t *istart = (uint32_t *)((char *)w1 + 2 * u32n);
    uint32_t *us = (uint32_t *)((char *)w1 + 3 * u32n);
    uint32_t *ul = (uint32_t *)((char *)w1 + 4 * u32n);
    int32_t *cchan = (int32_t *)((char *)w1 + 5 * u32n);
    uint32_t *scal = (uint32_t *)((char *)w1 + 6 * u32n);
    uint32_t *M_d = scal, *NI_d = scal + 64, *NS_d = scal + 128, *NL_d = scal + 192;
    int4 *desc = (int4 *)w2;
    int4 *cdesc = (int4 *)((char *)w2 + align_up((size_t)nrec * 16));

    float *density = nullptr;
    double *dacc = nullptr, *f2 = nullptr;
    if (o->weighting != MTI_WEIGHT_NATURAL) {
        void *w4 = nullptr, *w5 = nullptr;
        r = scratch_get(ctx, SLOT_DENSITY, (size_t)gsz * nb * P * sizeof(float), &w4);
        if (r != MTI_OK) return r;
        density = (float *)w4;
        r = scratch_get(ctx, SLOT_NORM, (size_t)nb * P * 3 * sizeof(double) + 256, &w5);
        if (r != MTI_OK) return r;
        dacc = (double *)w5;
        f2 = dacc + 2 * nb * P;
        hipLaunchKernelGGL(k_fill_f32, dim3(4096), dim3(256), 0, st, density, gsz * nb * P, 0.0f);
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, dacc, (int64_t)nb * P * 2, 0.0);
        MTI_HIP(ctx, hipGetLastError());
    }
    BinPassArgs ba;
    ba.v = vw;
    ba.g = make_geom(k, p);
    ba.chan0 = chan0;
    ba.nb = nb;
    ba.valid = valid;
    ba.desc = desc;
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
    const unsigned gblocks = (unsigned)ceil_div(nrec, 256);
    {
        StageTimer tm(ctx, MTI_STAGE_SORT_RECORDS, st);
        // stable compaction of gridded records (keeps baseline-major order)
        r = exclusive_scan_u32(ctx, valid, nrec, M_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_compact, dim3(gblocks), dim3(256), 0, st, valid, M_d, desc, nrec, vw.nidx, cdesc,
                           cchan);
        MTI_HIP(ctx, hipGetLastError());
    }
    {
        StageTimer tm(ctx, MTI_STAGE_PLAN, st);
        hipLaunchKernelGGL(k_item_flags, dim3(gblocks), dim3(256), 0, st, cdesc, cchan, M_d, nrec, flag);
        MTI_HIP(ctx, hipGetLastError());
        r = exclusive_scan_u32(ctx, flag, nrec, NI_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, st, us, nrec, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, st, ul, nrec, 0u);
        hipLaunchKernelGGL(k_item_starts, dim3(gblocks), dim3(256), 0, st, flag, NI_d, cdesc, M_d, nrec, istart,
                           us, ul);
        MTI_HIP(ctx, hipGetLastError());
        r = exclusive_scan_u32(ctx, us, nrec, NS_d, st);
        if (r != MTI_OK) return r;
        r = exclusive_scan_u32(ctx, ul, nrec, NL_d, st);
        if (r != MTI_OK) return r;
    }
    RomeinArgs ra;
    ra.v = vw;
    ra.chan0 = chan0;
    ra.cdesc = cdesc;
    ra.cchan = cchan;
    ra.item_start = istart;
    ra.n_items = NI_d;
    ra.n_rec = M_d;
    ra.table = k->d_table;
    ra.sums = k->d_sums;
    ra.convinfo = k->d_convinfo;
    ra.T2 = k->info.T * k->info.T;
    ra.nx = p->nx;
    ra.ny = p->ny;
    if (o->grid_layout == MTI_GRID_VUP) {
        ra.cs = P;
        ra.ps = 1;
    } else {
        ra.cs = 1;
        ra.ps = gsz;
    }
    ra.chan_stride = gsz * P;
    ra.grid = G;
    ra.norm = norm;
    ra.psf = o->psf;
    ra.weighting = o->weighting;
    ra.density = density;
    ra.f2 = f2;
    {
        StageTimer tm(ctx, MTI_STAGE_GRID, st);
        const int prec = k->info.precision;
        switch (P) {
        case 1: r = dispatch_romein<1>(ctx, prec, ra, us, NS_d, ul, NL_d, st); break;
        case 2: r = dispatch_romein<2>(ctx, prec, ra, us, NS_d, ul, NL_d, st); break;
        default: r = dispatch_romein<4>(ctx, prec, ra, us, NS_d, ul, NL_d, st); break;
        }
        if (r != MTI_OK) return r;
    }
    if (ni_total) {
        uint32_t ni = 0;
        MTI_HIP(ctx, hipMemcpyAsync(&ni, NI_d, 4, hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipStreamSynchronize(st));
        *ni_total += ni;
    }
    return MTI_OK;
}

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_grid(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                               const mti_vis *v, const mti_grid_opts *o, void *grids, double *norm,
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
    MTI_CHECK(ctx, o->grid_layout == MTI_GRID_VUP || o->grid_layout == MTI_GRID_PVU, "bad grid_layout");
    MTI_CHECK(ctx, o->grid_precision == MTI_GRID_C64 || o->grid_precision == MTI_GRID_C128,
              "bad grid_precision");
    MTI_CHECK(ctx, (int64_t)p->nx * p->ny < (int64_t(1) << 30), "grid above 2^30 cells");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);

    const int P = v->npol, nb = o->nchan;
    const int64_t gsz = (int64_t)p->nx * p->ny;
    const int64_t chan_floats = gsz * P * 2;  // reals per channel grid
    VisView vw = make_view(v);
    MTI_CHECK(ctx, vw.nidx * nb < (int64_t(1) << 31), "more than 2^31-1 records (rows x channels) per call");

    if (!o->accumulate) {
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, norm, (int64_t)nb * P, 0.0);
        if (o->grid_precision == MTI_GRID_C128)
            hipLaunchKernelGGL(k_fill_f64x, dim3(8192), dim3(256), 0, st, (double *)grids, chan_floats * nb);
        MTI_HIP(ctx, hipGetLastError());
    }
    void *sw = nullptr;
    r = scratch_get(ctx, SLOT_MISC2, 256, &sw);
    if (r != MTI_OK) return r;
    unsigned long long *dstats = (unsigned long long *)sw;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t *)dstats, 8, 0u);
    MTI_HIP(ctx, hipGetLastError());
    uint32_t ni = 0;

    if (o->grid_precision == MTI_GRID_C128) {
        if (vw.nidx > 0) {
            r = grid_batch(ctx, k, p, vw, o, o->chan0, nb, (double *)grids, norm, dstats, stats ? &ni : nullptr, st);
            if (r != MTI_OK) return r;
        }
    } else {
        // FP64 work grid for a batch of channels, rounded once into the output
        const size_t chan_bytes = (size_t)chan_floats * sizeof(double);
        const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nb, (size_t(8) << 30) / chan_bytes));
        void *gw = nullptr;
        r = scratch_get(ctx, SLOT_GRID64, chan_bytes * batch, &gw);
        if (r != MTI_OK) return r;
        for (int c0 = 0; c0 < nb; c0 += batch) {
            const int n = std::min(batch, nb - c0);
            hipLaunchKernelGGL(k_fill_f64x, dim3(8192), dim3(256), 0, st, (double *)gw, chan_floats * n);
            MTI_HIP(ctx, hipGetLastError());
            if (vw.nidx > 0) {
                r = grid_batch(ctx, k, p, vw, o, o->chan0 + c0, n, (double *)gw, norm + (size_t)c0 * P, dstats,
                               stats ? &ni : nullptr, st);
                if (r != MTI_OK) return r;
            }
            hipLaunchKernelGGL(k_round_grid, dim3(8192), dim3(256), 0, st, (const double *)gw,
                               (float *)grids + (size_t)c0 * chan_floats, chan_floats * n, o->accumulate);
            MTI_HIP(ctx, hipGetLastError());
        }
    }
    if (stats) {
        unsigned long long h[2] = {0, 0};
        MTI_HIP(ctx, hipMemcpyAsync(h, dstats, sizeof(h), hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipStreamSynchronize(st));
        *stats = mti_grid_stats{};
        stats->gridded_records = (int64_t)h[0];
        stats->gridded_points = (int64_t)h[1];
        stats->work_items = ni;
    }
    return MTI_OK;
}
