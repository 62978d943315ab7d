// mti_grid.hip -- convolutional gridding of many channels (rows a5-a7, a15, a21).
//
// Reference: ConvolutionGridder (op/ConvolutionGridder.cu:183-550) scatters
// grid[corner_y+sy][corner_x+sx][p] += K[conv_idx + sy*S + sx] * V_p, with
// V_p = vis_p * w_p * (1-flag_p) and the imaginary part negated for w<0
// (op/common.hcu:80-104); the normaliser is the sum of sums[k][o] * w_p *
// (1-flag_p) (op/WProjPreProcessUVW.cu:84-108, comp/WImager.cpp:196-197).  The
// reference keeps per-thread register accumulators fed by a work plan of runs
// of equal support (AccumulatedSum, CreateIndexAndReorder, ValueChangeDetect,
// CreateBlockDataIndex, NotSameSupportManipulate) and flushes with float
// atomics to the grid in HBM.
//
// MI355X design (tile ownership):
//   1. binning pass (mti_bin.hip): one record per (row, channel), FP64 bins;
//   2. stable compaction of gridded records (exclusive scan + scatter);
//   3. counting sort of (record, tile) pairs: every record is listed under each
//      32x32-cell tile its support box touches (wave-aggregated counters, one
//      scan, scatter); a tile's list is cut into work units of <= TG_CHUNK
//      records;
//   4. uniform/Briggs: the weight density is built the same way, per tile in
//      LDS, from the records whose nearest cell lies in the tile;
//   5. gridder: one 512-thread workgroup per unit holds the tile of every
//      correlation in LDS as FP64 (ds_add_f64), stages records in batches
//      (weighted visibility, clipped box, kernel offset) and lets each wave
//      take one record at a time with its lanes over the clipped taps; the tile
//      is then added to the FP64 grid once -- plain read-modify-write when the
//      unit is the tile's only one, FP64 atomics otherwise.
// HBM sees each touched grid cell once per unit instead of once per tap.
#include "mti_grid.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

namespace mti {

constexpr int TG_T = 32;             // tile side in cells
constexpr int TG_TT = TG_T * TG_T;
constexpr int TG_THREADS = 1024;     // 16 waves, one 8x8 region each
constexpr int TG_REGIONS = 16;
constexpr int TG_BATCH = 256;        // records staged per round
static_assert(TG_T == 32 && TG_THREADS == 64 * TG_REGIONS && TG_BATCH <= 256, "region layout");
constexpr int TG_CHUNK = 2048;       // records per work unit

struct TileArgs {
    VisView v;
    int chan0;
    const int4 *cdesc;          // compacted record descriptors
    const int32_t *cchan;       // local channel of each compacted record
    const uint32_t *n_rec;      // device scalar: compacted records
    float2 *wvis;               // [M][P] visibilities, weighted in place by k_record_prep
    float *wt;                  // [M][P] natural weights, 0 where flagged
    const uint32_t *pairs;      // record index per (tile, record) pair, grouped by tile
    const int4 *units;          // {key, first pair, pairs, 1 if the tile's only unit}
    const uint32_t *n_units;    // device scalar
    const void *table;
    const float *sums;
    const int2 *convinfo;
    int T2;
    int nx, ny, ntx, nty;
    int64_t cs, ps;             // complex strides of cell and correlation
    int64_t chan_stride;        // complex elements per channel grid
    double *grid;               // FP64 accumulation grid
    double *norm;               // [nb][P]
    int psf;
    int weighting;
    float *density;             // [nb][ny][nx][P] (written by the density pass)
    const double *f2;           // [nb][P]
};

template <int PREC>
__device__ __forceinline__ float2 ksample(const void *table, int idx)
{
    if (PREC == MTI_TABLE_F32) return ((const float2 *)table)[idx];
    if (PREC == MTI_TABLE_F16) return __half22float2(((const __half2 *)table)[idx]);
    __hip_bfloat162 h = ((const __hip_bfloat162 *)table)[idx];
    return make_float2(__bfloat162float(h.x), __bfloat162float(h.y));
}

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---- tile lists ---------------------------------------------------------------

struct TileRange {
    int tx0, ty0, ntx, nty;  // tiles touched by a support box
};

__device__ __forceinline__ TileRange tile_range(int4 d)
{
    const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16), S = d.w & 0x7fff;
    TileRange t;
    t.tx0 = cx / TG_T;
    t.ty0 = cy / TG_T;
    t.ntx = S > 0 ? (cx + S - 1) / TG_T - t.tx0 + 1 : 0;  // a plane without samples grids nothing
    t.nty = S > 0 ? (cy + S - 1) / TG_T - t.ty0 + 1 : 0;
    return t;
}

// Runs of equal keys in consecutive lanes: head lanes and their run lengths.
struct WaveRuns {
    bool head;
    int len;   // valid on head lanes
    int hl;    // head lane of this lane's run
};

__device__ __forceinline__ WaveRuns wave_runs(int key)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    WaveRuns r;
    r.head = (lane == 0) || (prev != key);
    const uint64_t heads = __ballot(r.head);
    const uint64_t above = (lane == 63) ? 0ull : (heads & ~((2ull << lane) - 1ull));
    const int end = above ? (__ffsll((unsigned long long)above) - 1) : 64;
    r.len = end - lane;
    const uint64_t upto = (lane == 63) ? heads : (heads & ((2ull << lane) - 1ull));
    r.hl = 63 - __clzll((long long)upto);
    return r;
}

// Key of tile (tx, ty) of local channel c.
__device__ __forceinline__ int tile_key(int c, int ty, int tx, int nty, int ntx)
{
    return (c * nty + ty) * ntx + tx;
}

// Pass 0: count the pairs of every tile.  Pass 1: scatter record indices.
template <int PASS>
__global__ __launch_bounds__(256) void k_tile_pairs(const int4 *cdesc, const int32_t *cchan, const uint32_t *n_rec,
                                                    int64_t cap, int nty, int ntx, uint32_t *cnt,
                                                    const uint32_t *start, uint32_t *pairs)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)blockIdx.x * blockDim.x >= (int64_t)*n_rec) return;  // whole block past the end
    const bool in = j < cap && j < (int64_t)*n_rec;
    TileRange t{0, 0, 0, 0};
    int c = 0;
    if (in) {
        t = tile_range(cdesc[j]);
        c = cchan[j];
    }
    const int nq = t.ntx * t.nty;
    int qmax = nq;
    for (int o = 32; o >= 1; o >>= 1) qmax = max(qmax, __shfl_xor(qmax, o, 64));
    for (int q = 0; q < qmax; q++) {
        int key = -1;
        if (q < nq) {
            const int dy = q / t.ntx, dx = q - dy * t.ntx;
            key = tile_key(c, t.ty0 + dy, t.tx0 + dx, nty, ntx);
        }
        const WaveRuns r = wave_runs(key);
        if (PASS == 0) {
            if (r.head && key >= 0) atomicAdd(&cnt[key], (uint32_t)r.len);
        } else {
            uint32_t base = 0;
            if (r.head && key >= 0) base = atomicAdd(&cnt[key], (uint32_t)r.len);
            base = __shfl(base, r.hl, 64);
            if (key >= 0) pairs[start[key] + base + (uint32_t)((threadIdx.x & 63) - r.hl)] = (uint32_t)j;
        }
    }
}

// Units per tile (ceil(pairs / TG_CHUNK)); start is the exclusive scan of
// the pair counts, total the sum.
__global__ void k_unit_counts(const uint32_t *start, const uint32_t *total, int64_t nkeys, uint32_t *nunits)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const uint32_t e = (k + 1 < nkeys) ? start[k + 1] : *total;
    nunits[k] = (e - start[k] + TG_CHUNK - 1) / TG_CHUNK;
}

__global__ void k_unit_list(const uint32_t *start, const uint32_t *total, const uint32_t *uprefix,
                            const uint32_t *n_units, int64_t nkeys, int4 *units)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const uint32_t s = start[k];
    const uint32_t e = (k + 1 < nkeys) ? start[k + 1] : *total;
    const uint32_t u0 = uprefix[k];
    const uint32_t u1 = (k + 1 < nkeys) ? uprefix[k + 1] : *n_units;
    const int excl = (u1 - u0) == 1;
    for (uint32_t u = u0; u < u1; u++) {
        const uint32_t f = s + (u - u0) * TG_CHUNK;
        units[u] = make_int4((int)k, (int)f, (int)min((uint32_t)TG_CHUNK, e - f), excl);
    }
}

// ---- weight density (uniform / Briggs) ----------------------------------------

// Per unit: the weights of the records whose nearest cell (box centre) lies in
// the tile, summed per correlation in LDS and added to the density grid
// [channel][v][u][P].
template <int P>
__global__ __launch_bounds__(256) void k_density_tiles(TileArgs a)
{
    __shared__ float s_d[TG_TT * P];
    const int tid = threadIdx.x;
    const uint32_t NU = *a.n_units;
    for (uint32_t u = blockIdx.x; u < NU; u += gridDim.x) {
        const int4 un = a.units[u];
        const int key = un.x;
        const int tx = key % a.ntx, t2 = key / a.ntx, ty = t2 % a.nty, c_local = t2 / a.nty;
        const int x0 = tx * TG_T, y0 = ty * TG_T;
        for (int i = tid; i < P * TG_TT; i += 256) s_d[i] = 0.0f;
        __syncthreads();
        for (int i = tid; i < un.z; i += 256) {
            const uint32_t j = a.pairs[un.y + i];
            const int4 d = a.cdesc[j];
            const int S = d.w & 0x7fff;
            const int mx = (d.y & 0xffff) + S / 2 - x0, my = (int)((unsigned)d.y >> 16) + S / 2 - y0;
            if (mx < 0 || mx >= TG_T || my < 0 || my >= TG_T) continue;
#pragma unroll
            for (int p = 0; p < P; p++) {
                const float w = a.wt[(int64_t)j * P + p];
                if (w != 0.0f) atomicAdd(&s_d[(my * TG_T + mx) * P + p], w);
            }
        }
        __syncthreads();
        float *dch = a.density + (int64_t)c_local * a.nx * a.ny * P;
        for (int i = tid; i < TG_TT * P; i += 256) {
            const int cell = i / P, p = i - cell * P;
            const int gx = x0 + (cell % TG_T), gy = y0 + cell / TG_T;
            const float v = s_d[i];
            if (v == 0.0f || gx >= a.nx || gy >= a.ny) continue;
            float *o = dch + ((int64_t)gy * a.nx + gx) * P + p;
            if (un.w) *o += v;
            else unsafeAtomicAdd(o, v);
        }
        __syncthreads();
    }
}

// ---- per-record weights -------------------------------------------------------

// Per compacted record: effective weights (natural, uniform 1/W or Briggs
// 1/(1 + f^2 W) at the record's nearest cell), visibilities weighted in place
// and the normaliser sums[k][o] * w_p, reduced per block and channel.
template <int P>
__global__ __launch_bounds__(256) void k_record_prep(TileArgs a)
{
    const uint32_t M = *a.n_rec;
    const int lane = threadIdx.x & 63;
    double np[P];
#pragma unroll
    for (int p = 0; p < P; p++) np[p] = 0.0;
    int c_prev = -1;
    for (uint32_t j0 = blockIdx.x * 256u; j0 < M; j0 += gridDim.x * 256u) {
        const uint32_t j = j0 + threadIdx.x;
        // records are channel-major: partials of the previous channel are
        // flushed when a block starts on a new one
        const int c_first = a.cchan[j0];
        if (c_first != c_prev && c_prev >= 0) {
#pragma unroll
            for (int p = 0; p < P; p++) {
                double v = np[p];
                for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0 && v != 0.0) unsafeAtomicAdd(&a.norm[c_prev * P + p], v);
                np[p] = 0.0;
            }
        }
        c_prev = c_first;
        if (j < M) {
            const int4 d = a.cdesc[j];
            const int c_local = a.cchan[j];
            const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16);
            const int plane = (int)((unsigned)d.w >> 16);
            const int S = d.w & 0x7fff;
            const double csum = S > 0 ? (double)a.sums[plane * a.T2 + (d.z - a.convinfo[plane].y) / (S * S)] : 0.0;
            float dens[P];
            if (a.weighting != MTI_WEIGHT_NATURAL) {
                const float *dp = a.density + (((int64_t)c_local * a.ny + cy + S / 2) * a.nx + cx + S / 2) * P;
#pragma unroll
                for (int p = 0; p < P; p++) dens[p] = dp[p];
            }
#pragma unroll
            for (int p = 0; p < P; p++) {
                float w = a.wt[(int64_t)j * P + p];
                if (a.weighting == MTI_WEIGHT_UNIFORM)
                    w = dens[p] > 0.0f ? w / dens[p] : 0.0f;
                else if (a.weighting == MTI_WEIGHT_BRIGGS)
                    w = (float)((double)w / (1.0 + a.f2[c_local * P + p] * (double)dens[p]));
                const float2 vv = a.wvis[(int64_t)j * P + p];
                a.wvis[(int64_t)j * P + p] = make_float2(vv.x * w, vv.y * w);
                if (c_local == c_first) np[p] += csum * (double)w;
                else unsafeAtomicAdd(&a.norm[c_local * P + p], csum * (double)w);
            }
        }
    }
    if (c_prev >= 0) {
#pragma unroll
        for (int p = 0; p < P; p++) {
            double v = np[p];
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0 && v != 0.0) unsafeAtomicAdd(&a.norm[c_prev * P + p], v);
        }
    }
}

// ---- gridder ------------------------------------------------------------------

// One in-flight visit of a wave to a record: staged record, lane active,
// kernel sample.
struct Visit {
    int r;
    bool on;
    float2 k;
};

// Visit i of the wave's region list: the lane's cell (gx, gy) against the
// record's support box; kernel index 0 (a valid dummy) when outside.
template <int PREC>
__device__ __forceinline__ Visit visit_issue(int i, int n, const uint8_t *list, const int4 *s_rec, int gx, int gy,
                                             const void *table)
{
    Visit v;
    v.r = -1;
    v.on = false;
    int idx = 0;
    if (i < n) {
        v.r = list[i];
        const int4 rec = s_rec[v.r];  // corner x | y << 16, kernel offset, S
        const int dx = gx - (rec.x & 0xffff), dy = gy - (int)((unsigned)rec.x >> 16);
        v.on = (unsigned)dx < (unsigned)rec.z && (unsigned)dy < (unsigned)rec.z;
        idx = v.on ? rec.y + dy * rec.z + dx : 0;
    }
    v.k = ksample<PREC>(table, idx);
    return v;
}

// Gridder: one 1024-thread workgroup per unit; wave w owns the 8x8 region w
// of the 32x32 tile and each lane one cell, accumulated in FP64 registers.
// Records are staged in LDS per batch together with, for every region, the
// list of records whose support box reaches it; a wave walks only its list,
// with kernel samples loaded four visits ahead.  No LDS traffic per tap.
template <int P, int PREC>
__global__ __launch_bounds__(TG_THREADS) void k_grid_tiles(TileArgs a)
{
    __shared__ int4 s_rec[TG_BATCH];                 // corner, kernel offset, S
    __shared__ double s_vd[TG_BATCH][2 * P];         // weighted visibility
    __shared__ uint8_t s_list[TG_REGIONS][TG_BATCH]; // records per region
    __shared__ int s_cnt[TG_REGIONS];
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_uniform(tid >> 6);
    const int rx = (wave & 3) * 8 + (lane & 7), ry = (wave >> 2) * 8 + (lane >> 3);  // cell in the tile
    const uint32_t NU = *a.n_units;

    for (uint32_t u = blockIdx.x; u < NU; u += gridDim.x) {
        const int4 un = a.units[u];
        const int key = un.x, first = un.y, npairs = un.z, excl = un.w;
        const int tx = key % a.ntx, t2 = key / a.ntx, ty = t2 % a.nty, c_local = t2 / a.nty;
        const int x0 = tx * TG_T, y0 = ty * TG_T;
        const int gx = x0 + rx, gy = y0 + ry;
        double acc[2 * P];
#pragma unroll
        for (int q = 0; q < 2 * P; q++) acc[q] = 0.0;

        for (int b0 = 0; b0 < npairs; b0 += TG_BATCH) {
            const int nbat = min(TG_BATCH, npairs - b0);
            __syncthreads();  // previous batch consumed
            if (tid < TG_REGIONS) s_cnt[tid] = 0;
            __syncthreads();
            // staging: four threads per record (one correlation each; the
            // first also stages the geometry and the region lists)
            {
                const int rr = tid >> 2, part = tid & 3;
                if (rr < nbat) {
                    const uint32_t j = a.pairs[first + b0 + rr];
                    if (part < P) {
                        const float2 v = a.wvis[(int64_t)j * P + part];
                        s_vd[rr][2 * part] = v.x;
                        s_vd[rr][2 * part + 1] = v.y;
                    }
                    if (part == 0) {
                        const int4 d = a.cdesc[j];
                        const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16);
                        const int S = d.w & 0x7fff;
                        s_rec[rr] = make_int4(d.y, d.z, S, 0);
                        const int ix0 = max(cx, x0) - x0, ix1 = min(cx + S, x0 + TG_T) - x0;
                        const int iy0 = max(cy, y0) - y0, iy1 = min(cy + S, y0 + TG_T) - y0;
                        for (int gyr = iy0 >> 3; gyr <= (iy1 - 1) >> 3; gyr++)
                            for (int gxr = ix0 >> 3; gxr <= (ix1 - 1) >> 3; gxr++) {
                                const int reg = gyr * 4 + gxr;
                                const int pos = atomicAdd(&s_cnt[reg], 1);
                                s_list[reg][pos] = (uint8_t)rr;
                            }
                    }
                }
            }
            __syncthreads();
            const int n = wave_uniform(s_cnt[wave]);
            const uint8_t *list = s_list[wave];
            Visit va = visit_issue<PREC>(0, n, list, s_rec, gx, gy, a.table);
            Visit vb = visit_issue<PREC>(1, n, list, s_rec, gx, gy, a.table);
            Visit vc = visit_issue<PREC>(2, n, list, s_rec, gx, gy, a.table);
            Visit vd = visit_issue<PREC>(3, n, list, s_rec, gx, gy, a.table);
            int next = 4;
            auto visit = [&](Visit &v) {
                if (v.on) {
                    const double kx = v.k.x, ky = v.k.y;
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        const double vr = s_vd[v.r][2 * p], vi = s_vd[v.r][2 * p + 1];
                        acc[2 * p] = fma(-vi, ky, fma(vr, kx, acc[2 * p]));
                        acc[2 * p + 1] = fma(vr, ky, fma(vi, kx, acc[2 * p + 1]));
                    }
                }
                v = visit_issue<PREC>(next++, n, list, s_rec, gx, gy, a.table);
            };
            for (int i = 0; i < n; i += 4) {
                visit(va);
                visit(vb);
                visit(vc);
                visit(vd);
            }
        }
        // add the cell to the grid
        if (gx < a.nx && gy < a.ny) {
            double *g = a.grid + 2 * ((int64_t)c_local * a.chan_stride + ((int64_t)gy * a.nx + gx) * a.cs);
            if (excl) {
                double2 old[P];
#pragma unroll
                for (int p = 0; p < P; p++)
                    old[p] = (acc[2 * p] != 0.0 || acc[2 * p + 1] != 0.0) ? *(const double2 *)(g + 2 * p * a.ps)
                                                                         : make_double2(0.0, 0.0);
#pragma unroll
                for (int p = 0; p < P; p++)
                    if (acc[2 * p] != 0.0 || acc[2 * p + 1] != 0.0)
                        *(double2 *)(g + 2 * p * a.ps) = make_double2(old[p].x + acc[2 * p], old[p].y + acc[2 * p + 1]);
            } else {
#pragma unroll
                for (int p = 0; p < P; p++) {
                    if (acc[2 * p] != 0.0) unsafeAtomicAdd(g + 2 * p * a.ps, acc[2 * p]);
                    if (acc[2 * p + 1] != 0.0) unsafeAtomicAdd(g + 2 * p * a.ps + 1, acc[2 * p + 1]);
                }
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

// Stable compaction.  pos is the exclusive scan of the valid flags, so
// record e is kept exactly when the next position differs from its own.
__global__ void k_compact(const uint32_t *pos, const uint32_t *n_rec, const int4 *desc, const float2 *wv,
                          const float *wt, int64_t nrec, int64_t nidx, int P, int4 *cdesc, int32_t *cchan,
                          float2 *cwv, float *cwt)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrec) return;
    const uint32_t j = pos[e];
    const uint32_t next = (e + 1 < nrec) ? pos[e + 1] : *n_rec;
    if (next == j) return;
    cdesc[j] = desc[e];
    cchan[j] = (int32_t)(e / nidx);
    for (int p = 0; p < P; p++) {
        cwv[(int64_t)j * P + p] = wv[e * P + p];
        cwt[(int64_t)j * P + p] = wt[e * P + p];
    }
}

// out = (accumulate ? out : 0) + work, FP64 work grid -> complex64 output.
__global__ void k_round_grid(const double *work, float *out, int64_t n, int accumulate)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = accumulate ? (float)((double)out[i] + work[i]) : (float)work[i];
}

// Per (channel, correlation): sum W and sum W^2 of the density grid
// [channel][cell][P]; one block row per channel.
__global__ __launch_bounds__(256) void k_density_stats(const float *dens, int64_t cells, int P,
                                                      double *acc /* [nb][P][2] */)
{
    const int c = blockIdx.y;
    const float *d = dens + (int64_t)c * cells * P;
    double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int p = 0; p < P; p++) {
            const double v = d[i * P + p];
            s1[p] += v;
            s2[p] += v * v;
        }
    }
    __shared__ double r[4][4][2];
    for (int p = 0; p < P; p++) {
        double a1 = s1[p], a2 = s2[p];
        for (int o = 32; o >= 1; o >>= 1) {
            a1 += __shfl_xor(a1, o, 64);
            a2 += __shfl_xor(a2, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            r[threadIdx.x >> 6][p][0] = a1;
            r[threadIdx.x >> 6][p][1] = a2;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * P) {
        const int p = threadIdx.x >> 1, k = threadIdx.x & 1;
        const double t = r[0][p][k] + r[1][p][k] + r[2][p][k] + r[3][p][k];
        unsafeAtomicAdd(&acc[2 * (c * P + p) + k], t);
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

template <typename K>
static int persistent_blocks(mti_ctx *ctx, K kern, int threads)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return ctx->num_cus * per_cu;
}

template <int P, int PREC>
static mti_status launch_tiles(mti_ctx *ctx, const TileArgs &a, hipStream_t st)
{
    auto kern = k_grid_tiles<P, PREC>;
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx, kern, TG_THREADS)), dim3(TG_THREADS), 0, st, a);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int P>
static mti_status dispatch_tiles(mti_ctx *ctx, int prec, const TileArgs &a, hipStream_t st)
{
    if (prec == MTI_TABLE_F32) return launch_tiles<P, MTI_TABLE_F32>(ctx, a, st);
    if (prec == MTI_TABLE_F16) return launch_tiles<P, MTI_TABLE_F16>(ctx, a, st);
    return launch_tiles<P, MTI_TABLE_BF16>(ctx, a, st);
}

template <int P>
static mti_status launch_prep(mti_ctx *ctx, const TileArgs &a, hipStream_t st)
{
    auto kern = k_record_prep<P>;
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx, kern, 256)), dim3(256), 0, st, a);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int P>
static mti_status launch_density(mti_ctx *ctx, const TileArgs &a, hipStream_t st)
{
    auto kern = k_density_tiles<P>;
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx, kern, 256)), dim3(256), 0, st, a);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

static inline size_t align_up(size_t x) { return (x + 255) / 256 * 256; }

// Tiles a support box of side S can touch along one axis.
static inline int tiles_per_axis(int S) { return 1 + (std::max(S, 1) - 1 + TG_T - 1) / TG_T; }

// Grids channels [chan0, chan0+nb) into the FP64 grid G (already zeroed or
// holding previous chunks); norm is accumulated.
static mti_status grid_batch(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                             const VisView &vw, const mti_grid_opts *o, int chan0, int nb, double *G,
                             double *norm, unsigned long long *dstats, uint32_t *nu_total, hipStream_t st)
{
    const int P = vw.npol;
    const int64_t gsz = (int64_t)p->nx * p->ny;
    const int64_t nrec = vw.nidx * nb;
    const int ntx = (p->nx + TG_T - 1) / TG_T, nty = (p->ny + TG_T - 1) / TG_T;
    const int64_t nkeys = (int64_t)nb * ntx * nty;
    const int tpa = tiles_per_axis(k->info.max_support);
    const int64_t pair_cap = nrec * tpa * tpa;
    MTI_CHECK(ctx, nkeys < (int64_t(1) << 31), "too many grid tiles per call");
    MTI_CHECK(ctx, pair_cap < (int64_t(1) << 32), "too many (record, tile) pairs per call");
    const int64_t unit_cap = nkeys + pair_cap / TG_CHUNK + 1;
    mti_status r;
    // scratch: valid, cchan (u32 x nrec); desc, cdesc (int4 x nrec);
    // start, fill, nunits (u32 x nkeys); pairs (u32 x pair_cap); units (int4)
    const size_t u32n = align_up((size_t)nrec * 4);
    const size_t u32k = align_up((size_t)nkeys * 4);
    void *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *w4 = nullptr;
    r = scratch_get(ctx, SLOT_KEYS, u32n * 2 + 4096, &w1);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_DESC, align_up((size_t)nrec * 16) * 2, &w2);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_BUCKETS, u32k * 3, &w3);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_ITEMS, align_up((size_t)pair_cap * 4) + align_up((size_t)unit_cap * 16), &w4);
    if (r != MTI_OK) return r;
    uint32_t *valid = (uint32_t *)w1;
    int32_t *cchan = (int32_t *)((char *)w1 + u32n);
    uint32_t *scal = (uint32_t *)((char *)w1 + 2 * u32n);
    uint32_t *M_d = scal, *NP_d = scal + 64, *NU_d = scal + 128;
    int4 *desc = (int4 *)w2;
    int4 *cdesc = (int4 *)((char *)w2 + align_up((size_t)nrec * 16));
    uint32_t *start = (uint32_t *)w3;
    uint32_t *fill = (uint32_t *)((char *)w3 + u32k);
    uint32_t *nunits = (uint32_t *)((char *)w3 + 2 * u32k);
    uint32_t *pairs = (uint32_t *)w4;
    int4 *units = (int4 *)((char *)w4 + align_up((size_t)pair_cap * 4));

    BinPassArgs ba;
    ba.v = vw;
    ba.g = make_geom(k, p);
    ba.chan0 = chan0;
    ba.nb = nb;
    ba.psf = o->psf;
    ba.valid = valid;
    ba.desc = desc;
    ba.stats = dstats;
    // per-record visibilities and weights, binning order and compacted
    void *w7 = nullptr, *w8 = nullptr;
    const size_t wv_bytes = align_up((size_t)nrec * P * sizeof(float2));
    const size_t wt_bytes = align_up((size_t)nrec * P * sizeof(float));
    r = scratch_get(ctx, SLOT_RECVIS, wv_bytes + wt_bytes, &w7);
    if (r != MTI_OK) return r;
    r = scratch_get(ctx, SLOT_CRECVIS, wv_bytes + wt_bytes, &w8);
    if (r != MTI_OK) return r;
    ba.wv = (float2 *)w7;
    ba.wt = (float *)((char *)w7 + wv_bytes);
    float2 *cwv = (float2 *)w8;
    float *cwt = (float *)((char *)w8 + wv_bytes);
    {
        StageTimer tm(ctx, MTI_STAGE_BIN, st);
        r = launch_bin_pass(ctx, ba, nrec, st);
        if (r != MTI_OK) return r;
    }
    const unsigned gblocks = (unsigned)ceil_div(nrec, 256);
    const unsigned kblocks = (unsigned)ceil_div(nkeys, 256);
    {
        StageTimer tm(ctx, MTI_STAGE_SORT_RECORDS, st);
        // stable compaction of gridded records (keeps baseline-major order)
        r = exclusive_scan_u32(ctx, valid, nrec, M_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_compact, dim3(gblocks), dim3(256), 0, st, valid, M_d, desc, ba.wv, ba.wt, nrec,
                           vw.nidx, P, cdesc, cchan, cwv, cwt);
        // counting sort of (record, tile) pairs by tile
        hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, start, nkeys, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, fill, nkeys, 0u);
        hipLaunchKernelGGL(k_tile_pairs<0>, dim3(gblocks), dim3(256), 0, st, cdesc, cchan, M_d, nrec, nty, ntx,
                           start, (const uint32_t *)nullptr, (uint32_t *)nullptr);
        MTI_HIP(ctx, hipGetLastError());
        r = exclusive_scan_u32(ctx, start, nkeys, NP_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_tile_pairs<1>, dim3(gblocks), dim3(256), 0, st, cdesc, cchan, M_d, nrec, nty, ntx,
                           fill, (const uint32_t *)start, pairs);
        MTI_HIP(ctx, hipGetLastError());
    }
    {
        StageTimer tm(ctx, MTI_STAGE_PLAN, st);
        hipLaunchKernelGGL(k_unit_counts, dim3(kblocks), dim3(256), 0, st, start, NP_d, nkeys, nunits);
        MTI_HIP(ctx, hipGetLastError());
        r = exclusive_scan_u32(ctx, nunits, nkeys, NU_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_unit_list, dim3(kblocks), dim3(256), 0, st, start, NP_d, nunits, NU_d, nkeys, units);
        MTI_HIP(ctx, hipGetLastError());
    }
    TileArgs ta;
    ta.v = vw;
    ta.chan0 = chan0;
    ta.cdesc = cdesc;
    ta.cchan = cchan;
    ta.n_rec = M_d;
    ta.wvis = cwv;
    ta.wt = cwt;
    ta.pairs = pairs;
    ta.units = units;
    ta.n_units = NU_d;
    ta.table = k->d_table;
    ta.sums = k->d_sums;
    ta.convinfo = k->d_convinfo;
    ta.T2 = k->info.T * k->info.T;
    ta.nx = p->nx;
    ta.ny = p->ny;
    ta.ntx = ntx;
    ta.nty = nty;
    if (o->grid_layout == MTI_GRID_VUP) {
        ta.cs = P;
        ta.ps = 1;
    } else {
        ta.cs = 1;
        ta.ps = gsz;
    }
    ta.chan_stride = gsz * P;
    ta.grid = G;
    ta.norm = norm;
    ta.psf = o->psf;
    ta.weighting = o->weighting;
    ta.density = nullptr;
    ta.f2 = nullptr;
    if (o->weighting != MTI_WEIGHT_NATURAL) {
        StageTimer tm(ctx, MTI_STAGE_PLAN, st);
        void *w5 = nullptr, *w6 = nullptr;
        r = scratch_get(ctx, SLOT_DENSITY, (size_t)gsz * nb * P * sizeof(float), &w5);
        if (r != MTI_OK) return r;
        r = scratch_get(ctx, SLOT_NORM, (size_t)nb * P * 3 * sizeof(double) + 256, &w6);
        if (r != MTI_OK) return r;
        ta.density = (float *)w5;
        double *dacc = (double *)w6;
        double *f2 = dacc + 2 * nb * P;
        ta.f2 = f2;
        hipLaunchKernelGGL(k_fill_f32, dim3(8192), dim3(256), 0, st, ta.density, gsz * nb * P, 0.0f);
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, dacc, (int64_t)nb * P * 2, 0.0);
        MTI_HIP(ctx, hipGetLastError());
        switch (P) {
        case 1: r = launch_density<1>(ctx, ta, st); break;
        case 2: r = launch_density<2>(ctx, ta, st); break;
        default: r = launch_density<4>(ctx, ta, st); break;
        }
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_density_stats, dim3(256, (unsigned)nb), dim3(256), 0, st, ta.density, gsz, P, dacc);
        hipLaunchKernelGGL(k_briggs_f2, dim3((unsigned)ceil_div(nb * P, 64)), dim3(64), 0, st, dacc, nb * P,
                           o->robust, f2);
        MTI_HIP(ctx, hipGetLastError());
    }
    {
        StageTimer tm(ctx, MTI_STAGE_PLAN, st);
        switch (P) {
        case 1: r = launch_prep<1>(ctx, ta, st); break;
        case 2: r = launch_prep<2>(ctx, ta, st); break;
        default: r = launch_prep<4>(ctx, ta, st); break;
        }
        if (r != MTI_OK) return r;
    }
    {
        StageTimer tm(ctx, MTI_STAGE_GRID, st);
        const int prec = k->info.precision;
        switch (P) {
        case 1: r = dispatch_tiles<1>(ctx, prec, ta, st); break;
        case 2: r = dispatch_tiles<2>(ctx, prec, ta, st); break;
        default: r = dispatch_tiles<4>(ctx, prec, ta, st); break;
        }
        if (r != MTI_OK) return r;
    }
    if (nu_total) {
        uint32_t h[2] = {0, 0};
        MTI_HIP(ctx, hipMemcpyAsync(&h[0], NU_d, 4, hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipMemcpyAsync(&h[1], NP_d, 4, hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipStreamSynchronize(st));
        nu_total[0] += h[0];
        nu_total[1] += h[1];
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
    MTI_CHECK(ctx, p->nx <= 65535 && p->ny <= 65535, "grid side above 65535");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);

    const int P = v->npol, nb = o->nchan;
    const int64_t gsz = (int64_t)p->nx * p->ny;
    const int64_t chan_floats = gsz * P * 2;  // reals per channel grid
    VisView vw = make_view(v);
    MTI_CHECK(ctx, vw.nidx < (int64_t(1) << 31), "more than 2^31-1 rows per call");

    if (!o->accumulate) {
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, norm, (int64_t)nb * P, 0.0);
        if (o->grid_precision == MTI_GRID_C128)
            hipLaunchKernelGGL(k_fill_f64, dim3(8192), dim3(256), 0, st, (double *)grids, chan_floats * nb, 0.0);
        MTI_HIP(ctx, hipGetLastError());
    }
    void *sw = nullptr;
    r = scratch_get(ctx, SLOT_MISC2, 256, &sw);
    if (r != MTI_OK) return r;
    unsigned long long *dstats = (unsigned long long *)sw;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t *)dstats, 8, 0u);
    MTI_HIP(ctx, hipGetLastError());
    uint32_t nu[2] = {0, 0};

    // Channel batches keep records below 2^31 and (record, tile) pairs below 2^32.
    const int tpa = tiles_per_axis(k->info.max_support);
    const int64_t per_chan = std::max<int64_t>(1, vw.nidx * tpa * tpa);
    const int64_t cap_chans = std::min<int64_t>(((int64_t(1) << 32) - 1) / per_chan,
                                                ((int64_t(1) << 31) - 1) / std::max<int64_t>(1, vw.nidx));
    MTI_CHECK(ctx, cap_chans >= 1, "one channel has too many (record, tile) pairs");
    const int pair_batch = (int)std::min<int64_t>(nb, cap_chans);

    if (o->grid_precision == MTI_GRID_C128) {
        for (int c0 = 0; c0 < nb && vw.nidx > 0; c0 += pair_batch) {
            const int n = std::min(pair_batch, nb - c0);
            r = grid_batch(ctx, k, p, vw, o, o->chan0 + c0, n, (double *)grids + (size_t)c0 * chan_floats,
                           norm + (size_t)c0 * P, dstats, stats ? nu : nullptr, st);
            if (r != MTI_OK) return r;
        }
    } else {
        // FP64 work grid for a batch of channels, rounded once into the output
        const size_t chan_bytes = (size_t)chan_floats * sizeof(double);
        int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nb, (size_t(8) << 30) / chan_bytes));
        batch = std::min(batch, pair_batch);
        void *gw = nullptr;
        r = scratch_get(ctx, SLOT_GRID64, chan_bytes * batch, &gw);
        if (r != MTI_OK) return r;
        for (int c0 = 0; c0 < nb; c0 += batch) {
            const int n = std::min(batch, nb - c0);
            hipLaunchKernelGGL(k_fill_f64, dim3(8192), dim3(256), 0, st, (double *)gw, chan_floats * n, 0.0);
            MTI_HIP(ctx, hipGetLastError());
            if (vw.nidx > 0) {
                r = grid_batch(ctx, k, p, vw, o, o->chan0 + c0, n, (double *)gw, norm + (size_t)c0 * P, dstats,
                               stats ? nu : nullptr, st);
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
        stats->work_items = nu[0];
        stats->flushes = nu[1];
    }
    return MTI_OK;
}
