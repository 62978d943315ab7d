// mti_grid.hpp -- structures shared by the binning (mti_bin.hip, built with
// -ffp-contract=off) and the gridder (mti_grid.hip).
#pragma once
#include "mti_internal.hpp"

namespace mti {

// Device view of an mti_vis descriptor.
struct VisView {
    const void *uvw;
    int uvw_f32;
    int64_t uvw_rs, uvw_cs;
    const void *invl;
    int invl_f32, invl_per_row;
    const double *dop;  // per-row factor on 1/lambda (TOPO -> LSRK), or null
    const float *vis;
    int64_t vis_rs, vis_cs;  // complex elements
    const void *flags;
    int flags_i32;
    int64_t fl_rs, fl_cs;
    const float *weight;
    int64_t w_rs;
    const int32_t *row_index;
    int64_t nidx;  // records per channel
    int npol;
    // device count of leading row_index entries to grid (the rest are rows a
    // device-side row sort put behind the kept ones), or null: all nidx
    const unsigned long long *nkept;
};

struct BinGeom {
    int nx, ny;
    double du, dv, dw, hx, hy;
    int mode, nplanes, s, T;
    const int2 *convinfo;  // {support, base}
};

// One imaging job of a gridding call (device-resident table).  Records are
// numbered job-major; within a job channel-major (c_local * nidx + i).
struct JobDev {
    VisView v;
    BinGeom g;
    int chan0, nb;           // first channel, channels gridded
    int64_t rec0;            // first record of the job
    int64_t key0;            // first (channel, tile) key of the job
    int ntx, nty;            // tiles per row / column
    const void *table;       // kernel samples (precision of the call)
    uint32_t table_bytes;    // its size (at most 2^32 - 16)
    const float *sums;       // [nplanes][T*T]
    int T2;
    void *grid;              // grid of channel chan0, complex128 (or complex64 if grid_c64);
                             // channel c at complex element c*chan_stride
    int grid_c64;
    int64_t cs, ps;          // complex strides of cell and correlation
    int64_t chan_stride;     // complex elements per channel grid
    double *norm;            // [nb][P]
    double *f2;              // [nb][P] Briggs factor
    int psf, weighting;
    uint8_t *tile_mask;      // [nb][nty][ntx]: 1 where a record reaches the tile; null: complete grids
    int mfs;                 // 1: every channel grids into grid 0 (one grid, normaliser, density)
};

// The jobs of a call and prefix tables for locating them.
struct JobSet {
    const JobDev *jobs;
    const int64_t *rec0;     // [njobs + 1] record prefix
    const int64_t *key0;     // [njobs + 1] key prefix
    const int64_t *blk0;     // [njobs + 1] binning-block prefix
    int njobs;
    int P;
    int masks;               // some job leaves untouched tiles unwritten (tile_mask)
};

// Largest j with prefix[j] <= x (prefix ascending, prefix[0] = 0).
__device__ __forceinline__ int job_of(const int64_t *prefix, int n, int64_t x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Record descriptor written by the binning pass (16 bytes):
//   x = sub-kernel (plane T^2 + oy T + ox: the sums index),
//   y = corner_y << 16 | corner_x, z = conv_idx,
//   w = plane << 16 | conj << 15 | support
// Every (job, channel, row) gets a record; one that is not gridded (all
// correlations flagged, or its support box off the grid) has support 0 and
// no visibilities or weights.
constexpr int BIN_STAT_STRIPES = 32;

struct BinPassArgs {
    JobSet js;
    int32_t *cjc;               // [nrec] job << 16 | local channel
    int4 *desc;                 // [nrec] record descriptors (layout at GridArgs::cdesc)
    float2 *wv;                 // [nrec][npol] visibilities of gridded records
    float *wt;                  // [nrec][npol] weights, 0 where flagged
    unsigned long long *stats;  // [2][BIN_STAT_STRIPES]: gridded records, gridded points
    uint32_t *tile_cnt;         // [keys] (record, tile) pairs per tile key, added to (zeroed by the caller)
};

mti_status launch_bin_pass(mti_ctx *ctx, const BinPassArgs &a, int64_t nblocks, hipStream_t st);

// ---- tiles of the gridder ------------------------------------------------------

constexpr int TG_T = 32;  // tile side in cells

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

// Adds the tiles a wave's records touch to the pair counts of their tile
// keys (base: key of the records' channel grid, ntx: tiles per row); records
// of support 0 touch none.  Every lane of the wave takes part: consecutive
// lanes hold consecutive records, so runs of equal keys share one atomic.
__device__ __forceinline__ void count_tile_pairs(int4 d, int base, int ntx, uint32_t *cnt)
{
    const TileRange t = tile_range(d);
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
        if (r.head && key >= 0) atomicAdd(&cnt[key], (uint32_t)r.len);
    }
}
int64_t bin_blocks(int64_t nidx, int nb);
VisView make_view(const mti_vis *v);
BinGeom make_geom(const mti_kernels *k, const mti_image_params *p);

__device__ __forceinline__ bool vis_flag(const VisView &v, int64_t row, int chan, int p)
{
    int64_t o = row * v.fl_rs + (int64_t)chan * v.fl_cs + p;
    return v.flags_i32 ? (((const int32_t *)v.flags)[o] != 0) : (((const uint8_t *)v.flags)[o] != 0);
}

}  // namespace mti
