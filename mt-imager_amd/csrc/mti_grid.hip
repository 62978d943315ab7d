// mti_grid.hip -- convolutional gridding of many channels and many jobs.
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
// MI355X design (tile ownership), for one or many imaging jobs per pass:
//   1. binning pass (mti_bin.hip): one record per (row, channel), FP64 bins,
//      written in place (job, channel, baseline order); a record that is not
//      gridded has support 0, so it yields no tile pairs and no weight;
//   2. (no compaction: ungridded records are a few percent, and every later
//      pass reaches records through tile pairs or skips support 0);
//   3. counting sort of (record, tile) pairs: every record is listed under each
//      32x32-cell tile its support box touches (wave-aggregated counters in the
//      binning pass, one scan, scatter); a tile's list is cut into work units of <= unit_pairs
//      records.  Tile keys are numbered job-major, so the units of all jobs of
//      a pass form one list and one persistent launch grids them all;
//   4. uniform/Briggs: one block per tile key gathers the weights of the
//      records whose nearest cell lies in the tile into LDS, adds the tile's
//      sum W and sum W^2 to its channel and hands each such record its W (no
//      density grid is stored);
//   5. per-record effective weights and the normaliser;
//   6. gridder: one 1024-thread workgroup per unit; wave w owns the 8x8 region
//      w of the tile and each lane one cell, accumulated in FP64 registers
//      (see k_grid_tiles); the tile is added to the FP64 grid once -- stored
//      when the unit is the tile's only one and the grid was not pre-filled
//      (k_zero_tiles zeroes every other tile first), read-modify-write when
//      the unit is the tile's only one of an accumulating call, FP64 atomics
//      otherwise.
#include "mti_grid.hpp"
#include "mti_sort.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace mti {

constexpr int TG_TT = TG_T * TG_T;
constexpr int TG_THREADS = 1024;     // 16 waves, one 8x8 region each
constexpr int TG_REGIONS = 16;
// A wave's region: RW x (64 / RW) cells (64 lanes), RW = 4, 8 or 16 (the
// context's tuning, mti_tuning.region_width).  Wide regions read each
// support-box row of kernel samples as a longer contiguous run, but a box
// touches more of them: 16 x 4 measured slower than 8 x 8 on the 4096^2
// full-polarisation cube's gridder (147.5 ms against 141), which is why 8 x 8
// is the default and the others are left to the tuner per workload.
template <int RW>
struct Region {
    static constexpr int W = RW, H = 64 / RW, X = TG_T / RW;  // cells; regions along a tile row
    static_assert((RW == 4 || RW == 8 || RW == 16) && X * (TG_T / H) == TG_REGIONS, "region shape");
};
constexpr int TG_BATCH = 256;        // records staged per round
static_assert(TG_T == 32 && TG_THREADS == 64 * TG_REGIONS && TG_BATCH <= 256, "region layout");
constexpr int TG_CHUNK = 2048;       // default (record, tile) pairs per work unit
constexpr int TG_CHUNK_MIN = 256;    // smallest unit a context may be tuned to

// Arrays of a gridding pass shared by the kernels after the binning pass.
struct GridArgs {
    JobSet js;
    const int4 *cdesc;          // record descriptors: x sub-kernel (plane T^2 + oy T + ox, the sums
                                // index), y corner (cy << 16 | cx), z kernel sample offset, w plane << 16 |
                                // conj << 15 | support (support 0: not gridded)
    const int32_t *cjc;         // job << 16 | local channel of each record
    const uint32_t *n_rec;      // device scalar: records
    float2 *wvis;               // [M][P] visibilities, weighted in place by k_record_prep
    float *wt;                  // [M][P] natural weights, 0 where flagged
    float *rdens;               // [M][P] uniform/Briggs: weight density at the record's nearest cell
    const uint32_t *pairs;      // record index (| PAIR_HOME) per (tile, record) pair, grouped by tile
    const int4 *units;          // {key, first pair, pairs, 1 if the tile's only unit}
    const uint32_t *n_units;    // device scalar
    int fresh;                  // grids not pre-filled: a tile's only unit stores its cells,
                                // every other tile was zeroed by k_zero_tiles
    int xcd_log2;               // gridder units dealt to each XCD in runs of 2^xcd_log2 (-1: round-robin)
    uint32_t *unit_queue;       // gridder: MTI_XCDS counters, UQ_STRIDE words apart (zero before the launch)
};

constexpr uint32_t MTI_XCDS = 8;  // XCDs of an MI355X; workgroup b runs on XCD b % 8
constexpr uint32_t UQ_STRIDE = 16;  // words between the gridder's unit-queue counters

// Global-memory views of pointers read from the device-resident job table.
// Such pointers are generic to the compiler, and loads through a generic
// pointer become FLAT instructions, which also count against the LDS wait
// counter: every LDS wait would then drain the prefetched kernel samples.
#define MTI_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ const MTI_GLOBAL T *gconst(const void *p)
{
    return (const MTI_GLOBAL T *)p;
}

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// A pointer known to be the same in every lane (kept in scalar registers).
__device__ __forceinline__ const void *uniform_ptr(const void *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const void *)(((uint64_t)hi << 32) | lo);
}

// Tile coordinates of a global (job, channel, tile) key.
struct TileAt {
    int job, c_local, tx, ty;
};

__device__ __forceinline__ TileAt tile_at(const JobSet &js, int key)
{
    TileAt t;
    t.job = job_of(js.key0, js.njobs, key);
    const JobDev &jd = js.jobs[t.job];
    const int k = key - (int)js.key0[t.job];
    t.tx = k % jd.ntx;
    const int r = k / jd.ntx;
    t.ty = r % jd.nty;
    t.c_local = r / jd.nty;
    return t;
}

// ---- tile lists ---------------------------------------------------------------

// A pair is the record index; bit 31 marks the record's home tile, the one
// holding its nearest cell (box centre), which the weight density uses.
constexpr uint32_t PAIR_HOME = 0x80000000u, PAIR_INDEX = 0x7fffffffu;

// Scatters the record indices into the tiles' pair lists (the binning pass
// counted the pairs of every tile, and start is the exclusive scan of the
// counts; cnt counts again from zero to place each pair).
__global__ __launch_bounds__(256) void k_tile_pairs(GridArgs a, int64_t cap, uint32_t *cnt, const uint32_t *start,
                                                    uint32_t *pairs)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)blockIdx.x * blockDim.x >= (int64_t)*a.n_rec) return;  // whole block past the end
    const bool in = j < cap && j < (int64_t)*a.n_rec;
    TileRange t{0, 0, 0, 0};
    int base = 0, ntx = 1, home = -1;
    if (in) {
        const int4 d = a.cdesc[j];
        t = tile_range(d);
        const int S = d.w & 0x7fff;
        home = ((((unsigned)d.y >> 16) + S / 2) / TG_T - t.ty0) * t.ntx + ((d.y & 0xffff) + S / 2) / TG_T - t.tx0;
        const int jc = a.cjc[j], job = jc >> 16, c = jc & 0xffff;
        const JobDev &jd = a.js.jobs[job];
        ntx = jd.ntx;
        base = (int)a.js.key0[job] + c * jd.nty * jd.ntx;
    }
    const int nq = t.ntx * t.nty;
    int qmax = nq;
    for (int o = 32; o >= 1; o >>= 1) qmax = max(qmax, __shfl_xor(qmax, o, 64));
    for (int q = 0; q < qmax; q++) {
        int key = -1;
        if (q < nq) {
            const int dy = q / t.ntx, dx = q - dy * t.ntx;
            key = base + (t.ty0 + dy) * ntx + t.tx0 + dx;
        }
        const WaveRuns r = wave_runs(key);
        uint32_t b = 0;
        if (r.head && key >= 0) b = atomicAdd(&cnt[key], (uint32_t)r.len);
        b = __shfl(b, r.hl, 64);
        if (key >= 0)
            pairs[start[key] + b + (uint32_t)((threadIdx.x & 63) - r.hl)] = (uint32_t)j | (q == home ? PAIR_HOME : 0u);
    }
}

// Units per tile (ceil(pairs / chunk)); start is the exclusive scan of
// the pair counts, total the sum.
__global__ void k_unit_counts(const uint32_t *start, const uint32_t *total, int64_t nkeys, uint32_t chunk,
                              uint32_t *nunits)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const uint32_t e = (k + 1 < nkeys) ? start[k + 1] : *total;
    nunits[k] = (e - start[k] + chunk - 1) / chunk;
}

__global__ void k_unit_list(const uint32_t *start, const uint32_t *total, const uint32_t *uprefix,
                            const uint32_t *n_units, int64_t nkeys, uint32_t chunk, int4 *units)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const uint32_t s = start[k];
    const uint32_t e = (k + 1 < nkeys) ? start[k + 1] : *total;
    const uint32_t u0 = uprefix[k];
    const uint32_t u1 = (k + 1 < nkeys) ? uprefix[k + 1] : *n_units;
    const int excl = (u1 - u0) == 1;
    for (uint32_t u = u0; u < u1; u++) {
        const uint32_t f = s + (u - u0) * chunk;
        units[u] = make_int4((int)k, (int)f, (int)min(chunk, e - f), excl);
    }
}

// ---- weight density (uniform / Briggs) ----------------------------------------

// The density W of a cell is the sum of the weights of the records whose
// nearest cell (box centre) it is, per channel and correlation.  A block
// takes whole tiles (all pairs of a tile key), so the tile's densities are
// complete in LDS: it adds sum W and sum W^2 of the tile's cells to the
// channel's totals and gives every record whose nearest cell lies in the
// tile its W (each record has exactly one such tile).  No density grid is
// stored.  acc: [global channel][P][2] FP64 sums; jc0: channel prefix per job.
template <int P>
__global__ __launch_bounds__(256) void k_density_keys(GridArgs a, const uint32_t *start, const uint32_t *np_total,
                                                     int64_t nkeys, const int64_t *jc0, double *acc,
                                                     uint32_t *queue)
{
    __shared__ float s_d[TG_TT * P];
    __shared__ double s_r[4][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t NP = *np_total;
    // per-thread partial sums of the channel being visited, flushed on change;
    // thread tid only ever sees correlation p = tid % P (256 is a multiple of P)
    const int pc = tid % P;
    double s1 = 0.0, s2 = 0.0;
    int ch_cur = -1;
    auto flush = [&]() {
        // block reduction of (s1, s2) per correlation, then 2P atomics
        for (int q = 0; q < P; q++) {
            double v1 = pc == q ? s1 : 0.0, v2 = pc == q ? s2 : 0.0;
            for (int o = 32; o >= 1; o >>= 1) {
                v1 += __shfl_xor(v1, o, 64);
                v2 += __shfl_xor(v2, o, 64);
            }
            if (lane == 0) {
                s_r[wave][0] = v1;
                s_r[wave][1] = v2;
            }
            __syncthreads();
            if (tid < 2) {
                const double t = s_r[0][tid] + s_r[1][tid] + s_r[2][tid] + s_r[3][tid];
                if (t != 0.0) unsafeAtomicAdd(&acc[2 * ((int64_t)ch_cur * P + q) + tid], t);
            }
            __syncthreads();
        }
        s1 = s2 = 0.0;
    };
    // Blocks take DQ consecutive keys at a time from one atomic counter (zero
    // before the launch): the crowded inner-uv tiles of every channel hold
    // most of the pairs, and a fixed stride over the keys left the blocks
    // whose keys fall on them working long after the others.  16 keys a grab
    // measured faster than 4 (more atomics) and 64 (coarser balance).
    constexpr uint32_t DQ = 16;
    __shared__ uint32_t s_next;
    for (;;) {
        __syncthreads();  // every thread has read the previous s_next
        if (tid == 0) s_next = atomicAdd(queue, DQ);
        __syncthreads();
        const int64_t kq = s_next;
        if (kq >= nkeys) break;
        const int64_t kend = min(kq + (int64_t)DQ, nkeys);
        for (int64_t k = kq; k < kend; k++) {
            const uint32_t p0 = start[k], p1 = (k + 1 < nkeys) ? start[k + 1] : NP;
            if (p0 == p1) continue;  // empty tile (uniform in the block)
            const TileAt ta = tile_at(a.js, (int)k);
            const int ch = (int)jc0[ta.job] + ta.c_local;
            if (ch != ch_cur) {
                if (ch_cur >= 0) flush();
                ch_cur = ch;
            }
            const int x0 = ta.tx * TG_T, y0 = ta.ty * TG_T;
            for (int i = tid; i < P * TG_TT; i += 256) s_d[i] = 0.0f;
            __syncthreads();
            // DK pairs a thread at a time (pairs i0 + tid + 256 k): their records
            // and the records' nearest cells in the tile (-1: not the record's home
            // tile, which alone gathers it), read in phases -- the pairs, then the
            // home records' descriptors -- so each phase's DK reads are in flight
            // together.  The pass is bound by the latency of these phases; one
            // pair a thread keeps the kernel at 8 waves a SIMD, and measured
            // faster than 2 (7 waves), 3 or 4 a thread (full-polarisation cube:
            // plan stage 21.7-22.6 ms against 22.7-23.1, 23.6-23.7 and 24.3-24.6).
            constexpr int DK = 1;
            auto walk = [&](uint32_t i0, uint32_t (&kj)[DK], int (&kc)[DK]) {
                int4 kd[DK];
#pragma unroll
                for (int k = 0; k < DK; k++) {
                    const uint32_t i = i0 + tid + 256u * k;
                    const uint32_t pj = i < p1 ? a.pairs[i] : 0u;  // past the tile: no home bit
                    kj[k] = pj & PAIR_INDEX;
                    kc[k] = (pj & PAIR_HOME) ? 0 : -1;
                }
#pragma unroll
                for (int k = 0; k < DK; k++) kd[k] = kc[k] >= 0 ? a.cdesc[kj[k]] : make_int4(0, 0, 0, 0);
#pragma unroll
                for (int k = 0; k < DK; k++)
                    if (kc[k] >= 0) {
                        const int S = kd[k].w & 0x7fff;
                        const int mx = (kd[k].y & 0xffff) + S / 2 - x0;
                        const int my = (int)((unsigned)kd[k].y >> 16) + S / 2 - y0;
                        kc[k] = my * TG_T + mx;
                    }
            };
            // consecutive pairs are consecutive time samples of a baseline and
            // often share a cell (crowded inner-uv cells): a wave first sums each
            // run of equal cells in its lanes (segmented shuffle sum down to the
            // run's head), then issues one LDS atomic per run -- atomics to one
            // address serialise
            auto gather = [&](const uint32_t (&kj)[DK], const int (&kc)[DK]) {
                float kw[DK][P];
#pragma unroll
                for (int k = 0; k < DK; k++)
#pragma unroll
                    for (int p = 0; p < P; p++) kw[k][p] = kc[k] >= 0 ? a.wt[(int64_t)kj[k] * P + p] : 0.0f;
                const int lane = tid & 63;
#pragma unroll
                for (int k = 0; k < DK; k++) {
                    const WaveRuns r = wave_runs(kc[k]);
                    if (__ballot(!r.head)) {  // some run is longer than one lane
                        const int end = lane + r.len;
                        for (int d = 1; d < 64; d <<= 1)
#pragma unroll
                            for (int p = 0; p < P; p++) {
                                const float o = __shfl_down(kw[k][p], d, 64);
                                if (lane + d < end) kw[k][p] += o;
                            }
                    }
                    if (r.head && kc[k] >= 0) {
#pragma unroll
                        for (int p = 0; p < P; p++)
                            if (kw[k][p] != 0.0f) atomicAdd(&s_d[kc[k] * P + p], kw[k][p]);
                    }
                }
            };
            auto hand_back = [&](const uint32_t (&kj)[DK], const int (&kc)[DK]) {
#pragma unroll
                for (int k = 0; k < DK; k++)
                    if (kc[k] >= 0) {
#pragma unroll
                        for (int p = 0; p < P; p++) a.rdens[(int64_t)kj[k] * P + p] = s_d[kc[k] * P + p];
                    }
            };
            // the first chunk's (record, cell) stay in registers for the hand-back;
            // later chunks are walked again
            uint32_t cj[DK];
            int cc[DK];
            walk(p0, cj, cc);
            gather(cj, cc);
            for (uint32_t i0 = p0 + 256u * DK; i0 < p1; i0 += 256u * DK) {
                uint32_t kj[DK];
                int kc[DK];
                walk(i0, kj, kc);
                gather(kj, kc);
            }
            __syncthreads();
            for (int i = tid; i < TG_TT * P; i += 256) {
                const double v = s_d[i];
                s1 += v;
                s2 += v * v;
            }
            hand_back(cj, cc);
            for (uint32_t i0 = p0 + 256u * DK; i0 < p1; i0 += 256u * DK) {
                uint32_t kj[DK];
                int kc[DK];
                walk(i0, kj, kc);
                hand_back(kj, kc);
            }
            __syncthreads();  // s_d is cleared for the next tile
        }
    }
    if (ch_cur >= 0) flush();
}

// ---- per-record weights -------------------------------------------------------

// Per gridded record: effective weights (natural, uniform 1/W or Briggs
// 1/(1 + f^2 W) at the record's nearest cell), visibilities weighted in place
// and the normaliser sums[k][o] * w_p, reduced per block and (job, channel).
template <int P>
__global__ __launch_bounds__(256) void k_record_prep(GridArgs a)
{
    // the block's weighted visibilities, stored from LDS in 16-byte runs: a
    // record's P samples stored lane by lane (8 bytes at P * 8-byte strides)
    // left partly written lines that went out to memory once per store
    __shared__ __align__(16) float2 s_vis[256 * P];
    const uint32_t M = *a.n_rec;
    const int lane = threadIdx.x & 63;
    double np[P];
#pragma unroll
    for (int p = 0; p < P; p++) np[p] = 0.0;
    int jc_prev = -1;
    auto flush = [&](int jc) {
        const JobDev &jd = a.js.jobs[jc >> 16];
#pragma unroll
        for (int p = 0; p < P; p++) {
            double v = np[p];
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0 && v != 0.0) unsafeAtomicAdd(&jd.norm[(jc & 0xffff) * P + p], v);
            np[p] = 0.0;
        }
    };
    // The next block of records' descriptor, tag, weights, visibilities and
    // density are read while this block's are weighted (two reads in flight a
    // thread; the sums and Briggs factors depend on them and are read after).
    // Reads of records that are not gridded (or past M) are in bounds and
    // unused.
    struct In {
        int4 d;
        int jc;
        float w[P], dens[P];
        float2 v[P];
    };
    const bool dense_w = a.rdens != nullptr;
    auto fetch = [&](uint32_t j, In &x) {
        x.d = j < M ? a.cdesc[j] : make_int4(0, 0, 0, 0);
        x.jc = j < M ? a.cjc[j] : 0;
#pragma unroll
        for (int p = 0; p < P; p++) {
            x.w[p] = j < M ? a.wt[(int64_t)j * P + p] : 0.0f;
            x.v[p] = j < M ? a.wvis[(int64_t)j * P + p] : make_float2(0.0f, 0.0f);
            x.dens[p] = j < M && dense_w ? a.rdens[(int64_t)j * P + p] : 0.0f;
        }
    };
    const uint32_t stride = gridDim.x * 256u;
    In cur;
    fetch(blockIdx.x * 256u + threadIdx.x, cur);
    for (uint32_t j0 = blockIdx.x * 256u; j0 < M; j0 += stride) {
        const uint32_t j = j0 + threadIdx.x;
        In nxt;
        fetch(j + stride, nxt);
        // records are ordered by (job, channel): partials are flushed when a
        // block starts on a new one
        const int jc_first = a.cjc[j0];
        if (jc_first != jc_prev && jc_prev >= 0) flush(jc_prev);
        jc_prev = jc_first;
        const int4 d = cur.d;
        const int S = d.w & 0x7fff;
        // records that are not gridded (support 0) hold no visibilities or weights
        if (S > 0) {
            const int jc = cur.jc;
            const JobDev &jd = a.js.jobs[jc >> 16];
            const int c_local = jc & 0xffff;
            const double csum = (double)jd.sums[d.x];
#pragma unroll
            for (int p = 0; p < P; p++) {
                float w = cur.w[p];
                const float dens = cur.dens[p];
                if (jd.weighting == MTI_WEIGHT_UNIFORM)
                    w = dens > 0.0f ? w / dens : 0.0f;
                else if (jd.weighting == MTI_WEIGHT_BRIGGS)
                    w = (float)((double)w / (1.0 + jd.f2[c_local * P + p] * (double)dens));
                // a zero weight (flagged) gives an exact zero, whatever the
                // visibility holds (a flagged sample may be NaN)
                const float2 vv = cur.v[p];
                s_vis[threadIdx.x * P + p] = w != 0.0f ? make_float2(vv.x * w, vv.y * w) : make_float2(0.0f, 0.0f);
                if (jc == jc_first) np[p] += csum * (double)w;
                else unsafeAtomicAdd(&jd.norm[c_local * P + p], csum * (double)w);
            }
        } else {
            // not gridded: no kernel reads its visibilities
#pragma unroll
            for (int p = 0; p < P; p++) s_vis[threadIdx.x * P + p] = make_float2(0.0f, 0.0f);
        }
        __syncthreads();
        const int64_t e0 = (int64_t)j0 * P, ne = (int64_t)min(256u, M - j0) * P;  // the block's samples
        for (int q = threadIdx.x; 2 * q < ne; q += 256) {
            if (2 * q + 1 < ne) *(float4 *)&a.wvis[e0 + 2 * q] = ((const float4 *)s_vis)[q];
            else a.wvis[e0 + 2 * q] = s_vis[2 * q];
        }
        __syncthreads();
        cur = nxt;
    }
    if (jc_prev >= 0) flush(jc_prev);
}

// ---- gridder ------------------------------------------------------------------

// One in-flight visit of a wave to a record: staged record, visit inside the
// list (the same in every lane), the lane's kernel sample.
struct Visit {
    int vo;  // byte offset of the record's visibilities in LDS
    bool live;
    float2 k;
};

// Kernel samples are read through a buffer resource over the job's table:
// the lane's byte offset is one 32-bit register (no 64-bit address per
// visit), and the buffer's range check makes an offset at or past the
// table's end read as zero -- a lane outside the record's support box reads
// a zero sample and adds exactly nothing (staged visibilities are finite:
// zero-weight samples are staged as zeros), so a visit needs no branch on
// the lane.  Tables are at most 2^32 - 16 bytes, so KS_OUTSIDE is past every
// table.
constexpr uint32_t KS_OUTSIDE = 0xfffffff0u;
constexpr int MTI_BUFFER_DWORD3 = 0x00020000;  // raw buffer, gfx9-family dword 3

typedef unsigned int mti_u32x2 __attribute__((ext_vector_type(2)));

template <int PREC>
__device__ __forceinline__ float2 ksample_buf(__amdgpu_buffer_rsrc_t tab, uint32_t off)
{
    if (PREC == MTI_TABLE_F32) {
        const mti_u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(tab, off, 0, 0);
        return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
    }
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(tab, off, 0, 0);
    if (PREC == MTI_TABLE_F16)
        return make_float2(__half2float(__ushort_as_half((unsigned short)(w & 0xffffu))),
                           __half2float(__ushort_as_half((unsigned short)(w >> 16))));
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

// Region-list entry (12 bytes): the record's corner x | y << 16, its kernel
// offset, and S | (visibility byte offset) << 16.
struct ListEntry {
    int xy, koff, s_vo;
};

// Visit i of the wave's region list: the lane's cell (gx, gy) against the
// record's support box; a zero sample when outside (or past the list).  The
// entry carries everything a visit needs, so it is one LDS read with no
// dependent read; dy * S (|dy| < 2^23) uses the full-rate 24-bit multiply.
template <int PREC>
__device__ __forceinline__ Visit visit_issue(int i, int n, const ListEntry *list, int gx, int gy,
                                             __amdgpu_buffer_rsrc_t tab)
{
    constexpr uint32_t esz = PREC == MTI_TABLE_F32 ? 8u : 4u;
    Visit v;
    v.vo = 0;
    v.live = i < n;
    uint32_t off = KS_OUTSIDE;
    if (v.live) {
        const ListEntry e = list[i];
        const int S = e.s_vo & 0xffff;
        const int dx = gx - (e.xy & 0xffff), dy = gy - (int)((unsigned)e.xy >> 16);
        const uint32_t t = (uint32_t)(e.koff + __mul24(dy, S) + dx);
        v.vo = (int)((unsigned)e.s_vo >> 16);
        off = max((unsigned)dx, (unsigned)dy) < (unsigned)S ? t * esz : KS_OUTSIDE;
    }
    v.k = ksample_buf<PREC>(tab, off);
    return v;
}

// Adds a cell's FP64 sums acc (re, im per correlation) to the grid at g
// (correlations ps complex elements apart): stored when the unit is the
// tile's only one and the grid was not pre-filled, read-modify-write when it
// is the only one of an accumulating call, atomics otherwise.  complex64
// grids round each sum once (the read-modify-write adds in FP64 first).
template <int P, typename R>
__device__ __forceinline__ void flush_cell(R *g, int64_t ps, const double (&acc)[2 * P], int excl, int fresh)
{
    if (excl && fresh) {
#pragma unroll
        for (int p = 0; p < P; p++) {
            g[2 * p * ps] = (R)acc[2 * p];
            g[2 * p * ps + 1] = (R)acc[2 * p + 1];
        }
    } else if (excl) {
        R old[2 * P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            const bool nz = acc[2 * p] != 0.0 || acc[2 * p + 1] != 0.0;
            old[2 * p] = nz ? g[2 * p * ps] : (R)0;
            old[2 * p + 1] = nz ? g[2 * p * ps + 1] : (R)0;
        }
#pragma unroll
        for (int p = 0; p < P; p++)
            if (acc[2 * p] != 0.0 || acc[2 * p + 1] != 0.0) {
                g[2 * p * ps] = (R)((double)old[2 * p] + acc[2 * p]);
                g[2 * p * ps + 1] = (R)((double)old[2 * p + 1] + acc[2 * p + 1]);
            }
    } else {
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (acc[2 * p] != 0.0) unsafeAtomicAdd(g + 2 * p * ps, (R)acc[2 * p]);
            if (acc[2 * p + 1] != 0.0) unsafeAtomicAdd(g + 2 * p * ps + 1, (R)acc[2 * p + 1]);
        }
    }
}

// Gridder: one 1024-thread workgroup per unit; wave w owns the 8x8 region w
// of the 32x32 tile and each lane one cell, accumulated in FP64 registers.
// Records are staged in LDS per batch: their weighted visibilities, and for
// every region the list of records whose support box reaches it, each entry
// a copy of the record's geometry.  A wave walks only its list, with kernel
// samples loaded four visits ahead.  No LDS traffic per tap.
// f32 tables: two 1024-lane workgroups per CU (8 waves per SIMD, at most 64
// VGPRs) is a requirement, not a hope -- one more register halves the
// occupancy.  The f16/bf16 forms convert samples and need more registers.
//
// ACC selects the accumulation: MTI_ACC_F64 (FP64 registers), or
// MTI_ACC_F32_KAHAN -- every four visits of a wave summed in FP32 registers,
// and those partial sums added with Kahan's compensated summation into an
// FP32 (sum, compensation) pair per cell component; the cell's value
// sum - compensation is then added to the grid.  Visibilities are staged as
// FP32 and samples stay FP32 (no conversions).
template <int P, int PREC, int ACC, int RW>
__global__ __launch_bounds__(TG_THREADS, PREC == MTI_TABLE_F32 || ACC == MTI_ACC_F32_KAHAN ? 8 : 1) void
k_grid_tiles(GridArgs a)
{
    using RG = Region<RW>;
    constexpr bool F32 = ACC == MTI_ACC_F32_KAHAN;
    constexpr int VB = F32 ? (int)sizeof(float) : (int)sizeof(double);  // bytes per staged component
    __shared__ __align__(16) double s_vd[TG_BATCH][2 * P];  // weighted visibility (FP64, or FP32 packed)
    __shared__ ListEntry s_list[TG_REGIONS][TG_BATCH];  // records per region
    __shared__ int s_cnt[TG_REGIONS];
    static_assert(TG_BATCH * 2 * P * sizeof(double) < 65536, "visibility offsets fit 16 bits");
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_uniform(tid >> 6);
    const int rx = (wave % RG::X) * RG::W + (lane % RG::W), ry = (wave / RG::X) * RG::H + (lane / RG::W);  // cell in the tile
    const uint32_t NU = *a.n_units;
    // XCD-aware order: runs of 2^lg consecutive units (neighbouring tiles of
    // one channel, whose records share w-planes) are dealt to the XCDs in turn
    // and an XCD's workgroups walk its runs in order, so that XCD's L2 serves
    // the kernel samples the run has in common.  (xcd, i) -> u is a bijection:
    // every unit is gridded once.  lg < 0 (or a grid not a multiple of 8):
    // round-robin.  All of it is workgroup-uniform (scalar registers).  A
    // workgroup takes its next i from a counter (its XCD's, or one for all),
    // so one that drew light units takes more of them instead of idling.
    const int lg = (gridDim.x % MTI_XCDS) ? -1 : a.xcd_log2;
    const uint32_t xcd = blockIdx.x % MTI_XCDS;
    uint32_t *queue = a.unit_queue + (lg >= 0 ? xcd : 0u) * UQ_STRIDE;
    __shared__ uint32_t s_i;
    for (;;) {
        // thread 0 writes s_i only after every thread has passed the previous
        // unit's batch barriers, so after every thread has read it
        if (tid == 0) s_i = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t i = s_i;
        const uint32_t u = lg >= 0 ? ((((i >> lg) * MTI_XCDS + xcd) << lg) | (i & ((1u << lg) - 1u))) : i;
        if (u >= NU) break;
        const int4 un = a.units[u];
        const int first = un.y, npairs = un.z, excl = un.w;
        const TileAt ta = tile_at(a.js, un.x);
        const JobDev &jd = a.js.jobs[ta.job];
        const __amdgpu_buffer_rsrc_t table = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<void *>(uniform_ptr(jd.table)), (short)0, (int)wave_uniform((int)jd.table_bytes),
            MTI_BUFFER_DWORD3);
        const int x0 = ta.tx * TG_T, y0 = ta.ty * TG_T;
        const int gx = x0 + rx, gy = y0 + ry;
        double acc[2 * P];
        float accf[2 * P], ksum[2 * P], kcmp[2 * P];
#pragma unroll
        for (int q = 0; q < 2 * P; q++) {
            acc[q] = 0.0;
            accf[q] = ksum[q] = kcmp[q] = 0.0f;
        }

        for (int b0 = 0; b0 < npairs; b0 += TG_BATCH) {
            const int nbat = min(TG_BATCH, npairs - b0);
            __syncthreads();  // previous batch consumed
            if (tid < TG_REGIONS) s_cnt[tid] = 0;
            __syncthreads();
            // staging: four threads per record (one correlation each; the
            // first also appends the record to the lists of its regions)
            {
                const int rr = tid >> 2, part = tid & 3;
                if (rr < nbat) {
                    const uint32_t j = a.pairs[first + b0 + rr] & PAIR_INDEX;
                    if (part < P) {
                        const float2 v = a.wvis[(int64_t)j * P + part];
                        if (F32) {
                            ((float2 *)&s_vd[0][0])[rr * P + part] = v;
                        } else {
                            s_vd[rr][2 * part] = v.x;
                            s_vd[rr][2 * part + 1] = v.y;
                        }
                    }
                    if (part == 0) {
                        const int4 d = a.cdesc[j];
                        const int cx = d.y & 0xffff, cy = (int)((unsigned)d.y >> 16);
                        const int S = d.w & 0x7fff;
                        ListEntry e;
                        e.xy = d.y;
                        e.koff = d.z;
                        e.s_vo = S | ((rr * 2 * P * VB) << 16);
                        const int ix0 = max(cx, x0) - x0, ix1 = min(cx + S, x0 + TG_T) - x0;
                        const int iy0 = max(cy, y0) - y0, iy1 = min(cy + S, y0 + TG_T) - y0;
                        for (int gyr = iy0 / RG::H; gyr <= (iy1 - 1) / RG::H; gyr++)
                            for (int gxr = ix0 / RG::W; gxr <= (ix1 - 1) / RG::W; gxr++) {
                                const int reg = gyr * RG::X + gxr;
                                const int pos = atomicAdd(&s_cnt[reg], 1);
                                s_list[reg][pos] = e;
                            }
                    }
                }
            }
            __syncthreads();
            const int n = wave_uniform(s_cnt[wave]);
            const ListEntry *list = s_list[wave];
            Visit va = visit_issue<PREC>(0, n, list, gx, gy, table);
            Visit vb = visit_issue<PREC>(1, n, list, gx, gy, table);
            Visit vc = visit_issue<PREC>(2, n, list, gx, gy, table);
            Visit vd = visit_issue<PREC>(3, n, list, gx, gy, table);
            int next = 4;
            auto visit = [&](Visit &v) {
                if (v.live) {
                    if (F32) {
                        const float kx = v.k.x, ky = v.k.y;
                        const float2 *vis = (const float2 *)((const char *)&s_vd[0][0] + v.vo);
#pragma unroll
                        for (int p = 0; p < P; p++) {
                            const float2 x = vis[p];
                            accf[2 * p] = fmaf(-x.y, ky, fmaf(x.x, kx, accf[2 * p]));
                            accf[2 * p + 1] = fmaf(x.x, ky, fmaf(x.y, kx, accf[2 * p + 1]));
                        }
                    } else {
                        const double kx = v.k.x, ky = v.k.y;
                        // 16-byte aligned pairs: one ds_read_b128 per correlation
                        const double2 *vis = (const double2 *)((const char *)&s_vd[0][0] + v.vo);
#pragma unroll
                        for (int p = 0; p < P; p++) {
                            const double2 x = vis[p];
                            const double vr = x.x, vi = x.y;
                            acc[2 * p] = fma(-vi, ky, fma(vr, kx, acc[2 * p]));
                            acc[2 * p + 1] = fma(vr, ky, fma(vi, kx, acc[2 * p + 1]));
                        }
                    }
                }
                v = visit_issue<PREC>(next++, n, list, gx, gy, table);
            };
            for (int i = 0; i < n; i += 4) {
                visit(va);
                visit(vb);
                visit(vc);
                visit(vd);
                if (F32) {
                    // Kahan: the four visits' FP32 sum into (ksum, kcmp)
#pragma unroll
                    for (int q = 0; q < 2 * P; q++) {
                        const float y = accf[q] - kcmp[q];
                        const float t = ksum[q] + y;
                        kcmp[q] = (t - ksum[q]) - y;
                        ksum[q] = t;
                        accf[q] = 0.0f;
                    }
                }
            }
        }
        if (F32) {
#pragma unroll
            for (int q = 0; q < 2 * P; q++) acc[q] = (double)ksum[q] - (double)kcmp[q];
        }
        // add the cell to the grid
        if (gx < jd.g.nx && gy < jd.g.ny) {
            const int64_t e = (int64_t)ta.c_local * jd.chan_stride + ((int64_t)gy * jd.g.nx + gx) * jd.cs;
            if (jd.grid_c64)
                flush_cell<P>((float *)jd.grid + 2 * e, jd.ps, acc, excl, a.fresh);
            else
                flush_cell<P>((double *)jd.grid + 2 * e, jd.ps, acc, excl, a.fresh);
        }
    }
}

// ---- small kernels ------------------------------------------------------------

// Zeroes the tiles that no single unit owns (no unit at all, or several units
// adding with atomics), so that a tile's only unit can store its cells
// without a prior fill of the grid.  uprefix is the exclusive scan of the
// units per tile key, total the unit count.  A job with a tile mask leaves
// the tiles no unit reaches unwritten and marks every tile (1: written);
// only its tiles of several units are zeroed.  A lane per key finds what its
// tile needs; the wave then zeroes the tiles its lanes found, all 64 lanes on
// one tile at a time.
__global__ __launch_bounds__(256) void k_zero_tiles(JobSet js, const uint32_t *uprefix, const uint32_t *total,
                                                   int64_t nkeys)
{
    const int P = js.P, lane = threadIdx.x & 63;
    const uint32_t NU = *total;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    // k0: the wave's first key (the same in every lane of the wave)
    for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); k0 < nkeys; k0 += step) {
        const int64_t k = k0 + lane;
        bool zero = false;
        if (k < nkeys) {
            const uint32_t nu = ((k + 1 < nkeys) ? uprefix[k + 1] : NU) - uprefix[k];
            if (js.masks || nu != 1) {
                const int job = job_of(js.key0, js.njobs, k);
                const JobDev &jd = js.jobs[job];
                if (jd.tile_mask) {
                    jd.tile_mask[k - js.key0[job]] = nu > 0 ? 1 : 0;
                    zero = nu > 1;
                } else {
                    zero = nu != 1;
                }
            }
        }
        for (uint64_t todo = __ballot(zero); todo; todo &= todo - 1) {
            const int l = __ffsll((unsigned long long)todo) - 1;
            const TileAt ta = tile_at(js, (int)(k0 + l));
            const JobDev &jd = js.jobs[ta.job];
            const int nx = jd.g.nx, ny = jd.g.ny, x0 = ta.tx * TG_T, y0 = ta.ty * TG_T;
            const int w = min(TG_T, nx - x0), h = min(TG_T, ny - y0), cells = w * h;
            const int64_t g0 = (int64_t)ta.c_local * jd.chan_stride;
            for (int i = lane; i < cells * P; i += 64) {
                int p, cell;
                if (jd.ps == 1) {  // [v][u][P]: correlations of a cell adjacent
                    p = i % P;
                    cell = i / P;
                } else {           // [P][v][u]: cells of a plane adjacent
                    p = i / cells;
                    cell = i - p * cells;
                }
                const int cy = cell / w, cx = cell - cy * w;
                const int64_t e = g0 + ((int64_t)(y0 + cy) * nx + x0 + cx) * jd.cs + p * jd.ps;
                if (jd.grid_c64)
                    ((float2 *)jd.grid)[e] = make_float2(0.0f, 0.0f);
                else
                    ((double2 *)jd.grid)[e] = make_double2(0.0, 0.0);
            }
        }
    }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Briggs f^2 = (5*10^-R)^2 / (sum W^2 / sum w) per (job, channel, correlation).
__global__ void k_briggs_f2(JobSet js, const int64_t *jc0, int nch, const double *acc, double robust)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = js.P;
    if (i >= nch * P) return;
    const int y = i / P, p = i - y * P;
    const int job = job_of(jc0, js.njobs, y);
    const int c = y - (int)jc0[job];
    const double sw = acc[2 * i], sw2 = acc[2 * i + 1];
    const double k = 5.0 * pow(10.0, -robust);
    js.jobs[job].f2[c * P + p] = sw2 > 0.0 ? (k * k) / (sw2 / sw) : 0.0;
}

template <typename K>
static int persistent_blocks(mti_ctx *ctx, K kern, int threads)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return ctx->num_cus * per_cu;
}

template <int P, int PREC, int ACC, int RW>
static mti_status launch_tiles_rw(mti_ctx *ctx, const GridArgs &a, hipStream_t st)
{
    auto kern = k_grid_tiles<P, PREC, ACC, RW>;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, a.unit_queue, (int64_t)(MTI_XCDS * UQ_STRIDE), 0u);
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx, kern, TG_THREADS)), dim3(TG_THREADS), 0, st, a);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int P, int PREC, int ACC>
static mti_status launch_tiles(mti_ctx *ctx, const GridArgs &a, hipStream_t st)
{
    if (ctx->region_width == 4) return launch_tiles_rw<P, PREC, ACC, 4>(ctx, a, st);
    if (ctx->region_width == 16) return launch_tiles_rw<P, PREC, ACC, 16>(ctx, a, st);
    return launch_tiles_rw<P, PREC, ACC, 8>(ctx, a, st);
}

template <int P, int ACC>
static mti_status dispatch_tiles_acc(mti_ctx *ctx, int prec, const GridArgs &a, hipStream_t st)
{
    if (prec == MTI_TABLE_F32) return launch_tiles<P, MTI_TABLE_F32, ACC>(ctx, a, st);
    if (prec == MTI_TABLE_F16) return launch_tiles<P, MTI_TABLE_F16, ACC>(ctx, a, st);
    return launch_tiles<P, MTI_TABLE_BF16, ACC>(ctx, a, st);
}

template <int P>
static mti_status dispatch_tiles(mti_ctx *ctx, int prec, int acc, const GridArgs &a, hipStream_t st)
{
    if (acc == MTI_ACC_F32_KAHAN) return dispatch_tiles_acc<P, MTI_ACC_F32_KAHAN>(ctx, prec, a, st);
    return dispatch_tiles_acc<P, MTI_ACC_F64>(ctx, prec, a, st);
}

template <int P>
static mti_status launch_prep(mti_ctx *ctx, const GridArgs &a, hipStream_t st)
{
    auto kern = k_record_prep<P>;
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx, kern, 256)), dim3(256), 0, st, a);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int P>
static mti_status launch_density(mti_ctx *ctx, const GridArgs &a, const uint32_t *start, const uint32_t *np_total,
                                 int64_t nkeys, const int64_t *jc0, double *acc, uint32_t *queue, hipStream_t st)
{
    auto kern = k_density_keys<P>;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, queue, (int64_t)1, 0u);
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx, kern, 256)), dim3(256), 0, st, a, start, np_total, nkeys,
                       jc0, acc, queue);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

static inline size_t align_up(size_t x) { return (x + 255) / 256 * 256; }

// Tiles a support box of side S can touch along one axis.
static inline int tiles_per_axis(int S) { return 1 + (std::max(S, 1) - 1 + TG_T - 1) / TG_T; }

// One job (or channel range of a job) of a pass, host side.
struct PassJob {
    const mti_kernels *k;
    mti_image_params p;
    VisView vw;
    int chan0, nb;
    void *grid;        // grid of chan0 in the job's layout (complex128, or complex64 if c64)
    int layout;
    double *norm;      // [nb][P]
    uint8_t *mask;     // [nb][nty][ntx] tile mask (null: complete grids)
    int c64 = 0;       // complex64 grids
    int mfs = 0;       // multi-frequency synthesis: the nb channels share one grid
};

// Grids (and normalisers, tile masks, density channels) a job piece writes.
static int grids_of(const PassJob &j) { return j.mfs ? 1 : j.nb; }

// Limits of one pass: records < 2^31, (record, tile) pairs < 2^32, keys < 2^31.
static int64_t job_records(const PassJob &j) { return j.vw.nidx * j.nb; }
static int64_t job_pairs(const PassJob &j)
{
    const int t = tiles_per_axis(j.k->info.max_support);
    return job_records(j) * t * t;
}
static int64_t job_keys(const PassJob &j)
{
    return (int64_t)grids_of(j) * ((j.p.nx + TG_T - 1) / TG_T) * ((j.p.ny + TG_T - 1) / TG_T);
}

static PassJob one_channel(PassJob j)
{
    j.nb = 1;
    return j;
}

// Scratch of one gridding pass over a set of job pieces; grid_pass lays its
// buffers out from these sizes and plan_pass asks for the same.
struct PassSizes {
    int64_t nrec = 0, nkeys = 0, nch = 0, pair_cap = 0, unit_cap = 0;
    size_t u32n = 0, u32k = 0, pre = 0, tab = 0, wv_bytes = 0, wt_bytes = 0;
};

static PassSizes pass_sizes(const std::vector<PassJob> &jobs, int P, int chunk)
{
    PassSizes z;
    const int nj = (int)jobs.size();
    for (const PassJob &j : jobs) {
        z.nrec += job_records(j);
        z.nkeys += job_keys(j);
        z.nch += grids_of(j);
        z.pair_cap += job_pairs(j);
    }
    z.unit_cap = z.nkeys + z.pair_cap / chunk + 1;
    z.u32n = align_up((size_t)z.nrec * 4);
    z.u32k = align_up((size_t)z.nkeys * 4);
    z.pre = align_up((size_t)(nj + 1) * 8);
    z.tab = align_up((size_t)nj * sizeof(JobDev)) + 4 * z.pre + align_up((size_t)z.nch * P * 3 * sizeof(double));
    z.wv_bytes = align_up((size_t)z.nrec * P * sizeof(float2));
    z.wt_bytes = align_up((size_t)z.nrec * P * sizeof(float));
    return z;
}

static void plan_pass(const std::vector<PassJob> &jobs, int P, bool weighted, ScratchPlan &pl)
{
    // the smallest unit a context can be tuned to: the most units
    const PassSizes z = pass_sizes(jobs, P, TG_CHUNK_MIN);
    if (jobs.empty() || z.nrec == 0) return;
    pl.need(SLOT_KEYS, z.u32n + 4096);
    pl.need(SLOT_DESC, align_up((size_t)z.nrec * 16));
    pl.need(SLOT_BUCKETS, z.u32k * 3);
    pl.need(SLOT_ITEMS, align_up((size_t)z.pair_cap * 4) + align_up((size_t)z.unit_cap * 16));
    pl.need(SLOT_CRECVIS, z.wv_bytes + z.wt_bytes);
    pl.need(SLOT_NORM, z.tab);
    if (weighted) pl.need(SLOT_DENSITY, z.wt_bytes);
    plan_scan(z.nkeys, pl);
}

// Grids a set of jobs (the same npol and table precision) into their FP64
// grids (zeroed or holding earlier chunks) and accumulates their normalisers.
// nu_total[0] += work units, [1] += (record, tile) pairs.
static mti_status grid_pass(mti_ctx *ctx, const std::vector<PassJob> &jobs, const mti_grid_opts *o, int P, int prec,
                            bool fresh, unsigned long long *dstats, uint64_t *nu_total, hipStream_t st)
{
    const int nj = (int)jobs.size();
    if (nj == 0) return MTI_OK;
    std::vector<JobDev> jd(nj);
    std::vector<int64_t> rec0(nj + 1, 0), key0(nj + 1, 0), blk0(nj + 1, 0), jc0(nj + 1, 0);
    int64_t pair_cap = 0;
    const bool weighted = o->weighting != MTI_WEIGHT_NATURAL;
    for (int i = 0; i < nj; i++) {
        const PassJob &j = jobs[i];
        MTI_CHECK(ctx, j.k->info.precision == prec, "jobs of one call must share the table precision");
        JobDev &d = jd[i];
        d = JobDev{};
        d.v = j.vw;
        d.g = make_geom(j.k, &j.p);
        d.chan0 = j.chan0;
        d.nb = j.nb;
        d.ntx = (j.p.nx + TG_T - 1) / TG_T;
        d.nty = (j.p.ny + TG_T - 1) / TG_T;
        d.table = j.k->d_table;
        d.table_bytes = (uint32_t)((uint64_t)j.k->info.total * (prec == MTI_TABLE_F32 ? 8u : 4u));
        d.sums = j.k->d_sums;
        d.T2 = j.k->info.T * j.k->info.T;
        d.grid = j.grid;
        d.grid_c64 = j.c64;
        const int64_t gsz = (int64_t)j.p.nx * j.p.ny;
        d.cs = j.layout == MTI_GRID_VUP ? P : 1;
        d.ps = j.layout == MTI_GRID_VUP ? 1 : gsz;
        d.chan_stride = gsz * P;
        d.norm = j.norm;
        d.psf = o->psf;
        d.weighting = o->weighting;
        d.tile_mask = j.mask;
        d.mfs = j.mfs;
        rec0[i + 1] = rec0[i] + job_records(j);
        key0[i + 1] = key0[i] + job_keys(j);
        blk0[i + 1] = blk0[i] + bin_blocks(j.vw.nidx, j.nb);
        jc0[i + 1] = jc0[i] + grids_of(j);
        pair_cap += job_pairs(j);
        d.rec0 = rec0[i];
        d.key0 = key0[i];
    }
    const int chunk = ctx->unit_pairs;
    const PassSizes z = pass_sizes(jobs, P, chunk);
    const int64_t nrec = z.nrec, nkeys = z.nkeys, nch = z.nch;
    if (nrec == 0) {
        // nothing to grid: a grid that was not pre-filled is all zeros (or,
        // with a tile mask, all of its tiles are marked empty)
        if (fresh)
            for (const PassJob &j : jobs) {
                if (j.mask)
                    hipLaunchKernelGGL(k_fill_u8, dim3(64), dim3(256), 0, st, j.mask, job_keys(j), (uint8_t)0);
                else if (j.c64)  // 0.0f is all zero bits
                    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, (uint32_t *)j.grid,
                                       (int64_t)j.p.nx * j.p.ny * P * 2 * grids_of(j), 0u);
                else
                    hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, st, (double *)j.grid,
                                       (int64_t)j.p.nx * j.p.ny * P * 2 * grids_of(j), 0.0);
            }
        MTI_HIP(ctx, hipGetLastError());
        return MTI_OK;
    }
    MTI_CHECK(ctx, nrec < (int64_t(1) << 31) && nkeys < (int64_t(1) << 31) && pair_cap < (int64_t(1) << 32),
              "gridding pass above the record/tile limits");
    const int64_t unit_cap = z.unit_cap;
    mti_status r;
    // scratch (the sizes plan_pass asks for)
    const size_t u32n = z.u32n, u32k = z.u32k, pre = z.pre, tab = z.tab;
    const size_t wv_bytes = z.wv_bytes, wt_bytes = z.wt_bytes;
    void *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *w4 = nullptr, *w6 = nullptr, *w7 = nullptr;
    if ((r = scratch_get(ctx, SLOT_KEYS, u32n + 4096, &w1)) != MTI_OK) return r;
    if ((r = scratch_get(ctx, SLOT_DESC, align_up((size_t)nrec * 16), &w2)) != MTI_OK) return r;
    if ((r = scratch_get(ctx, SLOT_BUCKETS, u32k * 3, &w3)) != MTI_OK) return r;
    if ((r = scratch_get(ctx, SLOT_ITEMS, align_up((size_t)pair_cap * 4) + align_up((size_t)unit_cap * 16), &w4)) !=
        MTI_OK)
        return r;
    if ((r = scratch_get(ctx, SLOT_CRECVIS, wv_bytes + wt_bytes, &w6)) != MTI_OK) return r;
    if ((r = scratch_get(ctx, SLOT_NORM, tab, &w7)) != MTI_OK) return r;
    int32_t *cjc = (int32_t *)w1;
    uint32_t *scal = (uint32_t *)((char *)w1 + u32n);
    uint32_t *M_d = scal, *NP_d = scal + 64, *NU_d = scal + 128, *DQ_d = scal + 192;  // DQ_d: density queue
    uint32_t *UQ_d = scal + 256;  // gridder unit queues [MTI_XCDS][UQ_STRIDE]
    int4 *cdesc = (int4 *)w2;
    uint32_t *start = (uint32_t *)w3;
    uint32_t *fill = (uint32_t *)((char *)w3 + u32k);
    uint32_t *nunits = (uint32_t *)((char *)w3 + 2 * u32k);
    uint32_t *pairs = (uint32_t *)w4;
    int4 *units = (int4 *)((char *)w4 + align_up((size_t)pair_cap * 4));
    float2 *cwv = (float2 *)w6;
    float *cwt = (float *)((char *)w6 + wv_bytes);
    char *tp = (char *)w7;
    JobDev *d_jobs = (JobDev *)tp;
    tp += align_up((size_t)nj * sizeof(JobDev));
    int64_t *d_rec0 = (int64_t *)tp;
    int64_t *d_key0 = (int64_t *)(tp + pre);
    int64_t *d_blk0 = (int64_t *)(tp + 2 * pre);
    int64_t *d_jc0 = (int64_t *)(tp + 3 * pre);
    double *dacc = (double *)(tp + 4 * pre);  // [nch][P][2] density sums, then f2 [nch][P]
    double *f2 = dacc + 2 * nch * P;
    float *rdens = nullptr;
    if (weighted) {
        void *wd = nullptr;
        if ((r = scratch_get(ctx, SLOT_DENSITY, wt_bytes, &wd)) != MTI_OK) return r;
        rdens = (float *)wd;
        for (int i = 0; i < nj; i++) jd[i].f2 = f2 + jc0[i] * P;
    }
    // the job table (the host vectors live until the synchronisation below)
    if ((r = upload(ctx, d_jobs, jd.data(), (size_t)nj * sizeof(JobDev), st)) != MTI_OK) return r;
    if ((r = upload(ctx, d_rec0, rec0.data(), (size_t)(nj + 1) * 8, st)) != MTI_OK) return r;
    if ((r = upload(ctx, d_key0, key0.data(), (size_t)(nj + 1) * 8, st)) != MTI_OK) return r;
    if ((r = upload(ctx, d_blk0, blk0.data(), (size_t)(nj + 1) * 8, st)) != MTI_OK) return r;
    if ((r = upload(ctx, d_jc0, jc0.data(), (size_t)(nj + 1) * 8, st)) != MTI_OK) return r;
    JobSet js;
    js.jobs = d_jobs;
    js.rec0 = d_rec0;
    js.key0 = d_key0;
    js.blk0 = d_blk0;
    js.njobs = nj;
    js.P = P;
    js.masks = 0;
    for (const PassJob &j : jobs)
        if (j.mask) js.masks = 1;

    BinPassArgs ba;
    ba.js = js;
    ba.cjc = cjc;
    ba.desc = cdesc;
    ba.wv = cwv;
    ba.wt = cwt;
    ba.stats = dstats;
    ba.tile_cnt = start;
    {
        StageTimer tm(ctx, MTI_STAGE_BIN, st);
        hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, start, nkeys, 0u);
        r = launch_bin_pass(ctx, ba, blk0[nj], st);
        if (r != MTI_OK) return r;
    }
    const unsigned gblocks = (unsigned)ceil_div(nrec, 256);
    const unsigned kblocks = (unsigned)ceil_div(nkeys, 256);
    GridArgs ga;
    ga.js = js;
    ga.cdesc = cdesc;
    ga.cjc = cjc;
    ga.n_rec = M_d;
    ga.wvis = cwv;
    ga.wt = cwt;
    ga.rdens = rdens;
    ga.pairs = pairs;
    ga.units = units;
    ga.unit_queue = UQ_d;
    ga.n_units = NU_d;
    ga.fresh = fresh ? 1 : 0;
    ga.xcd_log2 = -1;
    for (int k = 0; k <= 12; k++)
        if (ctx->xcd_chunk == (1 << k)) ga.xcd_log2 = k;
    {
        StageTimer tm(ctx, MTI_STAGE_SORT_RECORDS, st);
        // every record stays in place (job, channel, baseline order); records
        // that are not gridded have support 0 and produce no pairs
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, M_d, (int64_t)1, (uint32_t)nrec);
        // counting sort of (record, tile) pairs by tile (the binning pass
        // counted them into start)
        hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, fill, nkeys, 0u);
        r = exclusive_scan_u32(ctx, start, nkeys, NP_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_tile_pairs, dim3(gblocks), dim3(256), 0, st, ga, nrec, fill, (const uint32_t *)start,
                           pairs);
        MTI_HIP(ctx, hipGetLastError());
    }
    {
        StageTimer tm(ctx, MTI_STAGE_PLAN, st);
        hipLaunchKernelGGL(k_unit_counts, dim3(kblocks), dim3(256), 0, st, start, NP_d, nkeys, (uint32_t)chunk,
                           nunits);
        MTI_HIP(ctx, hipGetLastError());
        r = exclusive_scan_u32(ctx, nunits, nkeys, NU_d, st);
        if (r != MTI_OK) return r;
        hipLaunchKernelGGL(k_unit_list, dim3(kblocks), dim3(256), 0, st, start, NP_d, nunits, NU_d, nkeys,
                           (uint32_t)chunk, units);
        MTI_HIP(ctx, hipGetLastError());
        if (fresh) {
            const unsigned zblocks = (unsigned)std::min<int64_t>(ceil_div(nkeys, 256), 8192);
            hipLaunchKernelGGL(k_zero_tiles, dim3(zblocks), dim3(256), 0, st, js, (const uint32_t *)nunits,
                               (const uint32_t *)NU_d, nkeys);
            MTI_HIP(ctx, hipGetLastError());
        }
        if (weighted) {
            hipLaunchKernelGGL(k_fill_f64, dim3(64), dim3(256), 0, st, dacc, (int64_t)nch * P * 2, 0.0);
            MTI_HIP(ctx, hipGetLastError());
            switch (P) {
            case 1: r = launch_density<1>(ctx, ga, start, NP_d, nkeys, d_jc0, dacc, DQ_d, st); break;
            case 2: r = launch_density<2>(ctx, ga, start, NP_d, nkeys, d_jc0, dacc, DQ_d, st); break;
            default: r = launch_density<4>(ctx, ga, start, NP_d, nkeys, d_jc0, dacc, DQ_d, st); break;
            }
            if (r != MTI_OK) return r;
            hipLaunchKernelGGL(k_briggs_f2, dim3((unsigned)ceil_div(nch * P, 64)), dim3(64), 0, st, js,
                               (const int64_t *)d_jc0, (int)nch, (const double *)dacc, o->robust);
            MTI_HIP(ctx, hipGetLastError());
        }
        switch (P) {
        case 1: r = launch_prep<1>(ctx, ga, st); break;
        case 2: r = launch_prep<2>(ctx, ga, st); break;
        default: r = launch_prep<4>(ctx, ga, st); break;
        }
        if (r != MTI_OK) return r;
    }
    {
        StageTimer tm(ctx, MTI_STAGE_GRID, st);
        switch (P) {
        case 1: r = dispatch_tiles<1>(ctx, prec, o->accumulation, ga, st); break;
        case 2: r = dispatch_tiles<2>(ctx, prec, o->accumulation, ga, st); break;
        default: r = dispatch_tiles<4>(ctx, prec, o->accumulation, ga, st); break;
        }
        if (r != MTI_OK) return r;
    }
    if (ctx->capturing) return MTI_OK;  // the tables are staged; no counts read back
    uint32_t h[2] = {0, 0};
    MTI_HIP(ctx, hipMemcpyAsync(&h[0], NU_d, 4, hipMemcpyDeviceToHost, st));
    MTI_HIP(ctx, hipMemcpyAsync(&h[1], NP_d, 4, hipMemcpyDeviceToHost, st));
    // also keeps the job table's host copies alive until the copies are done
    MTI_HIP(ctx, hipStreamSynchronize(st));
    if (nu_total) {
        nu_total[0] += h[0];
        nu_total[1] += h[1];
    }
    return MTI_OK;
}

static mti_status check_opts(mti_ctx *ctx, const mti_grid_opts *o)
{
    MTI_CHECK(ctx, o, "null options");
    MTI_CHECK(ctx, o->weighting >= MTI_WEIGHT_NATURAL && o->weighting <= MTI_WEIGHT_BRIGGS, "bad weighting");
    MTI_CHECK(ctx, o->grid_layout == MTI_GRID_VUP || o->grid_layout == MTI_GRID_PVU, "bad grid_layout");
    MTI_CHECK(ctx, o->grid_precision == MTI_GRID_C64 || o->grid_precision == MTI_GRID_C128, "bad grid_precision");
    MTI_CHECK(ctx, o->mfs == 0 || o->mfs == 1, "mfs must be 0 or 1");
    MTI_CHECK(ctx, o->accumulation == MTI_ACC_F64 || o->accumulation == MTI_ACC_F32_KAHAN,
              "accumulation must be MTI_ACC_F64 or MTI_ACC_F32_KAHAN");
    return MTI_OK;
}

static mti_status check_job_geometry(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p, const mti_vis *v)
{
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    r = validate_vis(ctx, v);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, k, "null kernels");
    MTI_CHECK(ctx, k->info.sampling == p->sampling, "kernels built for another sampling");
    MTI_CHECK(ctx, (int64_t)p->nx * p->ny < (int64_t(1) << 30), "grid above 2^30 cells");
    MTI_CHECK(ctx, p->nx <= 65535 && p->ny <= 65535, "grid side above 65535");
    return MTI_OK;
}

static mti_status stats_out(mti_ctx *ctx, unsigned long long *dstats, const uint64_t *nu, mti_grid_stats *stats,
                            hipStream_t st)
{
    if (!stats) return MTI_OK;
    unsigned long long h[2 * BIN_STAT_STRIPES];
    MTI_HIP(ctx, hipMemcpyAsync(h, dstats, sizeof(h), hipMemcpyDeviceToHost, st));
    MTI_HIP(ctx, hipStreamSynchronize(st));
    unsigned long long s[2] = {0, 0};
    for (int k = 0; k < BIN_STAT_STRIPES; k++) {
        s[0] += h[k];
        s[1] += h[BIN_STAT_STRIPES + k];
    }
    *stats = mti_grid_stats{};
    stats->gridded_records = (int64_t)s[0];
    stats->gridded_points = (int64_t)s[1];
    stats->work_items = (int64_t)nu[0];
    stats->flushes = (int64_t)nu[1];
    return MTI_OK;
}

// Largest channel count of one job piece that keeps a pass within its limits.
static int64_t channels_per_pass(const PassJob &j)
{
    PassJob one = j;
    one.nb = 1;
    const int64_t pc = std::max<int64_t>(1, job_pairs(one));
    return std::min<int64_t>(((int64_t(1) << 32) - 1) / pc, ((int64_t(1) << 31) - 1) / std::max<int64_t>(1, j.vw.nidx));
}

// Passes of a batched call: consecutive job pieces while the pass stays
// within its limits (a job too large for one pass is split by channel).
template <class F>
static mti_status for_each_pass(const std::vector<PassJob> &all, int P, F fn)
{
    std::vector<PassJob> pass;
    int64_t in_r = 0, in_p = 0, in_k = 0;
    for (const PassJob &j : all) {
        // an MFS job's channels share one density and grid: never split (the
        // entry points check that it fits one pass)
        const int64_t cap = j.mfs ? j.nb : channels_per_pass(j);
        const int64_t gch = (int64_t)j.p.nx * j.p.ny * P;  // complex elements per channel grid
        for (int c0 = 0; c0 < j.nb;) {
            PassJob piece = j;
            piece.chan0 = j.chan0 + c0;
            piece.nb = (int)std::min<int64_t>(j.nb - c0, cap);
            piece.grid = j.grid ? (char *)j.grid + (size_t)gch * c0 * (j.c64 ? 8 : 16) : nullptr;
            piece.norm = j.norm ? j.norm + (size_t)c0 * P : nullptr;
            piece.mask = j.mask ? j.mask + (size_t)c0 * job_keys(one_channel(j)) : nullptr;
            if (!pass.empty() &&
                (in_r + job_records(piece) >= (int64_t(1) << 31) || in_p + job_pairs(piece) >= (int64_t(1) << 32) ||
                 in_k + job_keys(piece) >= (int64_t(1) << 31) || pass.size() >= 32767)) {
                mti_status r = fn(pass);
                if (r != MTI_OK) return r;
                pass.clear();
                in_r = in_p = in_k = 0;
            }
            pass.push_back(piece);
            in_r += job_records(piece);
            in_p += job_pairs(piece);
            in_k += job_keys(piece);
            c0 += piece.nb;
        }
    }
    return pass.empty() ? MTI_OK : fn(pass);
}

// The scratch plan of one mti_grid_jobs call: statistics, the in-call row
// sorts and every pass.
static void plan_grid_jobs(const std::vector<PassJob> &all, int32_t njobs, const mti_job *jobs, int P, bool weighted,
                           ScratchPlan &pl)
{
    pl.need(SLOT_MISC2, 2 * BIN_STAT_STRIPES * sizeof(unsigned long long));
    int64_t sort_rows_total = 0;
    size_t maxrows = 0;
    for (int i = 0; i < njobs; i++)
        if (jobs[i].ant1) {
            sort_rows_total += jobs[i].vis.nrows;
            maxrows = std::max<size_t>(maxrows, (size_t)jobs[i].vis.nrows);
            plan_sort_rows_device(jobs[i].vis.nrows, pl);
        }
    if (sort_rows_total > 0) {
        pl.need(SLOT_RECVIS, align_up((size_t)sort_rows_total * 4) + align_up((size_t)njobs * 8));
        pl.need(SLOT_KEYS, maxrows * 12 + 4096);
    }
    for_each_pass(all, P, [&](const std::vector<PassJob> &pass) {
        plan_pass(pass, P, weighted, pl);
        return MTI_OK;
    });
}

// Channels per pass of one mti_grid call (record / pair limits).
static int grid_batch(const PassJob &base, int nb)
{
    if (base.mfs) return nb;
    return (int)std::min<int64_t>(nb, std::max<int64_t>(1, channels_per_pass(base)));
}

// The scratch plan of one mti_grid call (the same batches and passes).
static void plan_grid(const PassJob &base, int nb, int P, const mti_grid_opts *o, ScratchPlan &pl)
{
    pl.need(SLOT_MISC2, 2 * BIN_STAT_STRIPES * sizeof(unsigned long long));
    const bool weighted = o->weighting != MTI_WEIGHT_NATURAL;
    const int step = grid_batch(base, nb);
    for (int c0 = 0; c0 < nb && base.vw.nidx > 0; c0 += step) {
        PassJob j = base;
        j.nb = std::min(step, nb - c0);
        plan_pass({j}, P, weighted, pl);
    }
}

// Sizes-only checks of a workspace query (pointers are not needed).
static mti_status check_sizes(const mti_image_params *p, const mti_vis *v, int32_t chan0, int32_t nchan)
{
    if (!p || !v || validate_params(nullptr, p) != MTI_OK) return MTI_ERR_INVALID;
    if (!(v->npol == 1 || v->npol == 2 || v->npol == 4) || v->nrows < 0 || v->nrows >= (int64_t(1) << 31) ||
        v->nchan < 1 || (v->row_index && (v->nindex < 0 || v->nindex > v->nrows)))
        return MTI_ERR_INVALID;
    if (nchan < 1 || nchan > 65535 || (!v->inv_lambda_per_row && (chan0 < 0 || chan0 + nchan > v->nchan)))
        return MTI_ERR_INVALID;
    if ((int64_t)p->nx * p->ny >= (int64_t(1) << 30)) return MTI_ERR_INVALID;
    return MTI_OK;
}

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_grid_workspace_size(const mti_kernel_info *ki, const mti_image_params *p, const mti_vis *v,
                                              const mti_grid_opts *o, size_t *bytes)
{
    if (!ki || !bytes || !o || check_opts(nullptr, o) != MTI_OK || check_sizes(p, v, o->chan0, o->nchan) != MTI_OK)
        return MTI_ERR_INVALID;
    mti_kernels kk;
    kk.info = *ki;
    PassJob base{&kk, *p, make_view(v), o->chan0, 1, nullptr, o->grid_layout, nullptr};
    base.mfs = o->mfs;
    ScratchPlan pl;
    plan_grid(base, o->nchan, v->npol, o, pl);
    *bytes = pl.total();
    return MTI_OK;
}

extern "C" mti_status mti_grid(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                               const mti_vis *v, const mti_grid_opts *o, void *grids, double *norm,
                               mti_grid_stats *stats, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = check_job_geometry(ctx, k, p, v);
    if (r != MTI_OK) return r;
    r = check_opts(ctx, o);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, grids && norm, "null outputs");
    MTI_CHECK(ctx, o->nchan >= 1 && o->nchan <= 65535, "opts.nchan must be in [1, 65535]");
    if (o->tile_mask) MTI_CHECK(ctx, !o->accumulate, "a tile mask (sparse grids) needs accumulate = 0");
    if (v->inv_lambda_per_row)
        MTI_CHECK(ctx, o->nchan == 1, "per-row 1/lambda implies a single channel per call");
    else
        MTI_CHECK(ctx, o->chan0 >= 0 && o->chan0 + o->nchan <= v->nchan, "channel range out of bounds");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);

    const int P = v->npol, nb = o->nchan;
    const int c64 = o->grid_precision == MTI_GRID_C64 ? 1 : 0;
    const int64_t gsz = (int64_t)p->nx * p->ny;
    const int64_t chan_reals = gsz * P * 2;  // reals per channel grid
    VisView vw = make_view(v);
    MTI_CHECK(ctx, vw.nidx < (int64_t(1) << 31), "more than 2^31-1 rows per call");
    // channel batches keep a pass within the record / pair limits
    PassJob base{k, *p, vw, o->chan0, 1, nullptr, o->grid_layout, nullptr};
    base.c64 = c64;
    base.mfs = o->mfs;
    // multi-frequency synthesis: every channel into one grid and normaliser,
    // one density (the whole channel range in one pass)
    const int ngrids = o->mfs ? 1 : nb;
    if (o->mfs)
        MTI_CHECK(ctx, channels_per_pass(base) >= nb, "MFS channel range above one pass's record/tile limits");
    CallScope cs(ctx);
    if (cs.planning()) {
        ScratchPlan pl;
        plan_grid(base, nb, P, o, pl);
        if ((r = cs.begin(pl)) != MTI_OK) return r;
    }

    if (!o->accumulate) {
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, norm, (int64_t)ngrids * P, 0.0);
        // with rows, the gridding pass zeroes exactly the tiles it does not store whole
        if (vw.nidx == 0) {
            if (o->tile_mask)
                hipLaunchKernelGGL(k_fill_u8, dim3(64), dim3(256), 0, st, o->tile_mask,
                                   (int64_t)ngrids * ((p->nx + TG_T - 1) / TG_T) * ((p->ny + TG_T - 1) / TG_T),
                                   (uint8_t)0);
            else if (c64)  // 0.0f is all zero bits
                hipLaunchKernelGGL(k_fill_u32, dim3(8192), dim3(256), 0, st, (uint32_t *)grids, chan_reals * ngrids,
                                   0u);
            else
                hipLaunchKernelGGL(k_fill_f64, dim3(8192), dim3(256), 0, st, (double *)grids, chan_reals * ngrids,
                                   0.0);
        }
        MTI_HIP(ctx, hipGetLastError());
    }
    void *sw = nullptr;
    r = scratch_get(ctx, SLOT_MISC2, 2 * BIN_STAT_STRIPES * sizeof(unsigned long long), &sw);
    if (r != MTI_OK) return r;
    unsigned long long *dstats = (unsigned long long *)sw;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (uint32_t *)dstats, 4 * BIN_STAT_STRIPES, 0u);
    MTI_HIP(ctx, hipGetLastError());
    uint64_t nu[2] = {0, 0};

    MTI_CHECK(ctx, channels_per_pass(base) >= 1, "one channel has too many (record, tile) pairs");
    const int pair_batch = grid_batch(base, nb);
    const int prec = k->info.precision;
    for (int c0 = 0; c0 < nb && vw.nidx > 0; c0 += pair_batch) {
        PassJob j = base;
        j.chan0 = o->chan0 + c0;
        j.nb = std::min(pair_batch, nb - c0);
        j.grid = (char *)grids + (size_t)c0 * chan_reals * (c64 ? 4 : 8);
        j.norm = norm + (size_t)c0 * P;
        j.mask = o->tile_mask ? o->tile_mask + (size_t)c0 * job_keys(one_channel(base)) : nullptr;
        r = grid_pass(ctx, {j}, o, P, prec, !o->accumulate, dstats, nu, st);
        if (r != MTI_OK) return r;
    }
    return stats_out(ctx, dstats, nu, stats, st);
}

extern "C" mti_status mti_grid_jobs_workspace_size(int32_t njobs, const mti_job *jobs, const mti_grid_opts *o,
                                                   size_t *bytes)
{
    if (!bytes || !o || check_opts(nullptr, o) != MTI_OK || njobs < 0 || njobs > 65535 || (njobs > 0 && !jobs))
        return MTI_ERR_INVALID;
    std::vector<PassJob> all;
    int P = 0;
    for (int i = 0; i < njobs; i++) {
        const mti_job &j = jobs[i];
        if (!j.kernels || check_sizes(&j.params, &j.vis, j.chan0, j.nchan) != MTI_OK) return MTI_ERR_INVALID;
        if (i == 0) P = j.vis.npol;
        if (j.vis.npol != P) return MTI_ERR_INVALID;
        all.push_back(PassJob{j.kernels, j.params, make_view(&j.vis), j.chan0, j.nchan, nullptr, o->grid_layout,
                              nullptr});
        all.back().mfs = o->mfs;
        if (channels_per_pass(all.back()) < (o->mfs ? j.nchan : 1)) return MTI_ERR_INVALID;
    }
    ScratchPlan pl;
    plan_grid_jobs(all, njobs, jobs, P, o->weighting != MTI_WEIGHT_NATURAL, pl);
    *bytes = pl.total();
    return MTI_OK;
}

extern "C" mti_status mti_grid_jobs(mti_ctx *ctx, int32_t njobs, const mti_job *jobs, const mti_grid_opts *o,
                                    mti_grid_stats *stats, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = check_opts(ctx, o);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, njobs >= 0 && njobs <= 65535 && (njobs == 0 || jobs), "njobs must be in [0, 65535]");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    int P = 0, prec = -1;
    std::vector<PassJob> all;
    int64_t sort_rows_total = 0;  // rows of the jobs whose order the call computes
    for (int i = 0; i < njobs; i++) {
        const mti_job &j = jobs[i];
        if (j.ant1) {
            MTI_CHECK(ctx, j.ant2 && !j.vis.row_index, "a job with ant1 needs ant2 and no row_index");
            MTI_CHECK(ctx, j.nant >= 1 && j.nant <= 46340, "job antenna count out of range");
            sort_rows_total += j.vis.nrows;
        }
        r = check_job_geometry(ctx, j.kernels, &j.params, &j.vis);
        if (r != MTI_OK) return r;
        MTI_CHECK(ctx, j.grids && j.norm, "null job outputs");
        MTI_CHECK(ctx, j.nchan >= 1 && j.nchan <= 65535, "job nchan must be in [1, 65535]");
        if (j.vis.inv_lambda_per_row)
            MTI_CHECK(ctx, j.nchan == 1, "per-row 1/lambda implies a single channel per job");
        else
            MTI_CHECK(ctx, j.chan0 >= 0 && j.chan0 + j.nchan <= j.vis.nchan, "job channel range out of bounds");
        if (i == 0) {
            P = j.vis.npol;
            prec = j.kernels->info.precision;
        }
        MTI_CHECK(ctx, j.vis.npol == P, "jobs of one call must share npol");
        MTI_CHECK(ctx, j.kernels->info.precision == prec, "jobs of one call must share the table precision");
        if (j.tile_mask) MTI_CHECK(ctx, !o->accumulate, "a job tile mask (sparse grids) needs accumulate = 0");
        MTI_CHECK(ctx, j.grids_c64 == 0 || j.grids_c64 == 1, "job grids_c64 must be 0 or 1");
        const PassJob pj{j.kernels, j.params, make_view(&j.vis), j.chan0, j.nchan, j.grids, o->grid_layout,
                         j.norm, j.tile_mask, j.grids_c64};
        MTI_CHECK(ctx, pj.vw.nidx < (int64_t(1) << 31), "more than 2^31-1 rows in a job");
        MTI_CHECK(ctx, channels_per_pass(pj) >= 1, "one channel has too many (record, tile) pairs");
        if (o->mfs)
            MTI_CHECK(ctx, channels_per_pass(pj) >= j.nchan, "MFS job above one pass's record/tile limits");
        all.push_back(pj);
        all.back().mfs = o->mfs;
    }
    CallScope cs(ctx);
    if (cs.planning()) {
        ScratchPlan pl;
        plan_grid_jobs(all, njobs, jobs, P, o->weighting != MTI_WEIGHT_NATURAL, pl);
        if ((r = cs.begin(pl)) != MTI_OK) return r;
    }
    void *sw = nullptr;
    r = scratch_get(ctx, SLOT_MISC2, 2 * BIN_STAT_STRIPES * sizeof(unsigned long long), &sw);
    if (r != MTI_OK) return r;
    unsigned long long *dstats = (unsigned long long *)sw;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (uint32_t *)dstats, 4 * BIN_STAT_STRIPES, 0u);
    MTI_HIP(ctx, hipGetLastError());
    // baseline-major row order of the jobs that ask for it, kept counts on the
    // device (index and counts live in SLOT_RECVIS for the whole call)
    if (sort_rows_total > 0) {
        StageTimer tm(ctx, MTI_STAGE_SORT_ROWS, st);
        size_t maxrows = 0;
        for (int i = 0; i < njobs; i++)
            if (jobs[i].ant1) maxrows = std::max<size_t>(maxrows, (size_t)jobs[i].vis.nrows);
        const size_t idx_bytes = align_up((size_t)sort_rows_total * 4) + align_up((size_t)njobs * 8);
        void *wi = nullptr, *wk = nullptr;
        if ((r = scratch_get(ctx, SLOT_RECVIS, idx_bytes, &wi)) != MTI_OK) return r;
        if ((r = scratch_get(ctx, SLOT_KEYS, maxrows * 12 + 4096, &wk)) != MTI_OK) return r;
        int32_t *ip = (int32_t *)wi;
        unsigned long long *kp = (unsigned long long *)((char *)wi + align_up((size_t)sort_rows_total * 4));
        for (int i = 0; i < njobs; i++) {
            if (!jobs[i].ant1) continue;
            const int64_t n = jobs[i].vis.nrows;
            r = sort_rows_device(ctx, jobs[i].ant1, jobs[i].ant2, n, jobs[i].nant, ip, (uint32_t *)wk, kp + i, st);
            if (r != MTI_OK) return r;
            all[i].vw.row_index = ip;
            all[i].vw.nidx = n;
            all[i].vw.nkept = kp + i;
            ip += n;
        }
    }
    uint64_t nu[2] = {0, 0};
    // zero the normalisers; the grids are zeroed by the passes (tiles not stored whole)
    if (!o->accumulate) {
        for (const PassJob &j : all)
            hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, st, j.norm, (int64_t)grids_of(j) * P, 0.0);
        MTI_HIP(ctx, hipGetLastError());
    }
    r = for_each_pass(all, P, [&](const std::vector<PassJob> &pass) {
        return grid_pass(ctx, pass, o, P, prec, !o->accumulate, dstats, nu, st);
    });
    if (r != MTI_OK) return r;
    return stats_out(ctx, dstats, nu, stats, st);
}
