// mti_bin.hip -- binning of visibility records.
//
// Restates op/WProjPreProcessUVW.cu:111-286 in FP64 (built with
// -ffp-contract=off so every operation rounds exactly as the FP64 oracle's):
//   to_grid  = any correlation unflagged (:65-83) and support box inside the
//              grid (:196-199);
//   w < 0    -> (u,v) negated, imaginary part negated later (:170-175);
//   plane    = round(sqrt(|w f| / dw)), clamped to W-1 (:181-182);
//   pos      = (u f)/du + nx/2 (:188-189), loc = round(pos) (half away from 0),
//   corner   = loc - S/2;
//   offset   = round((loc - pos) s) + s/2, clamped to [0,s-1] for odd s (:206-227);
//   conv_idx = base_k + oy S^2 T + ox S^2 (:235-239).
// The reference's per-1024-block "compress if equal to the previous record"
// flag (:270-324) is not produced by the gridding path: the tile gridder adds
// equal records one after another into the same FP64 register sums, which
// gives the same sum (mti_preprocess_reference writes the flag for callers
// of the reference's operator).
#include "mti_binning.hpp"

#include <algorithm>

namespace mti {

__device__ BinOut bin_record(const VisView &v, const BinGeom &g, int64_t row, int chan)
{
    bool any = false;
    for (int p = 0; p < v.npol; p++)
        if (!vis_flag(v, row, chan, p)) any = true;
    if (!any) {
        BinOut b;
        b.to_grid = b.plane = b.S = b.cx = b.cy = b.ox = b.oy = b.conj = b.conv_idx = 0;
        return b;
    }
    return bin_core(load_uvw(v, row, 0), load_uvw(v, row, 1), load_uvw(v, row, 2), load_invl(v, row, chan), g);
}

__global__ void k_bin_only(VisView v, BinGeom g, int chan, int32_t *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.nidx) return;
    const int64_t row = v.row_index ? (int64_t)v.row_index[i] : i;
    BinOut b = bin_record(v, g, row, chan);
    int32_t *o = out + 8 * i;
    o[0] = b.to_grid;
    o[1] = b.plane;
    o[2] = b.S;
    o[3] = b.cx;
    o[4] = b.cy;
    o[5] = b.ox;
    o[6] = b.oy;
    o[7] = b.conj;
}

// One binning block: BT_ROWS index entries of one job, all its channels in
// chunks of BT_CHANS, 256 threads.  The rows' (u, v, w) and weights are staged
// once; per chunk, phase 1 stages the flags and visibilities in LDS with loads
// that run along each row (MS order: correlations, then channels,
// contiguous), and phase 2 bins the records with the 64 lanes of a wave on 64
// consecutive index entries of one channel, so every store of descriptors,
// visibilities and weights is one contiguous run per channel (records are
// numbered e = rec0 + c_local * nidx + i).  Chunks of 8 channels keep the
// block's LDS at 20 KB for four correlations, so 8 blocks fit a CU: 16-channel
// chunks (40 KB, 4 blocks) took 8.8-9.1 ms on the 64-channel full-polarisation
// cube against 7.5-7.6 ms, and 4-channel chunks 7.7-7.9 ms.
constexpr int BT_THREADS = 256;
constexpr int BT_ROWS = 64;
constexpr int BT_CHANS = 8;

template <int P>
struct BinTile {
    float2 vis[BT_CHANS][BT_ROWS][P];
    uint8_t flag[BT_CHANS][BT_ROWS][P];
    double uvw[BT_ROWS][3];
    float wgt[BT_ROWS][P];
    int row[BT_ROWS];
    unsigned long long st[BT_THREADS / 64][2];
};

// Seven waves a SIMD (at most 72 VGPRs; the LDS tile allows seven blocks of
// four correlations a CU): 7.9 against 8.2-8.4 ms at six on the cube.
template <int P>
__global__ __launch_bounds__(BT_THREADS) __attribute__((amdgpu_waves_per_eu(7))) void k_bin_pass(BinPassArgs a)
{
    __shared__ BinTile<P> t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int job = job_of(a.js.blk0, a.js.njobs, (int64_t)blockIdx.x);
    const JobDev jd = a.js.jobs[job];  // uniform: scalar loads into registers
    const VisView &v = jd.v;
    const int64_t i0 = ((int64_t)blockIdx.x - a.js.blk0[job]) * BT_ROWS;
    const int64_t nvalid = v.nkept ? (int64_t)*v.nkept : v.nidx;
    const int nr = (int)max<int64_t>(0, min<int64_t>(BT_ROWS, nvalid - i0));  // rows of the tile to bin
    const int64_t i = i0 + lane;

    if (tid < nr) {
        const int64_t row = v.row_index ? (int64_t)v.row_index[i0 + tid] : i0 + tid;
        t.row[tid] = (int)row;
        t.uvw[tid][0] = load_uvw(v, row, 0);
        t.uvw[tid][1] = load_uvw(v, row, 1);
        t.uvw[tid][2] = load_uvw(v, row, 2);
#pragma unroll
        for (int p = 0; p < P; p++) t.wgt[tid][p] = v.weight[row * v.w_rs + p];
    }
    unsigned long long ngrid = 0, npts = 0;
    for (int c0 = 0; c0 < jd.nb; c0 += BT_CHANS) {
        const int nc = min(jd.nb - c0, BT_CHANS);
        __syncthreads();  // row data staged / previous chunk consumed
        // phase 1: the chunk's flags and visibilities, along each row
        constexpr int PER_ROW = BT_CHANS * P;
        for (int q = tid; q < nr * PER_ROW; q += BT_THREADS) {
            const int r = q / PER_ROW, rem = q - r * PER_ROW;
            const int c = rem / P, p = rem - c * P;
            if (c >= nc) continue;
            const int64_t row = t.row[r];
            const int chan = jd.chan0 + c0 + c;
            t.flag[c][r][p] = vis_flag(v, row, chan, p) ? 1 : 0;
            if (!jd.psf)
                t.vis[c][r][p] = *(const float2 *)(v.vis + 2 * (row * v.vis_rs + (int64_t)chan * v.vis_cs + p));
        }
        __syncthreads();
        // phase 2: wave w takes channels w, w + 4, ...; lane = index entry.
        // Every lane then adds its record's tiles to the pair counts (all
        // lanes take part in the count's shuffles; past the index: none).
        for (int c = wave; c < nc; c += BT_THREADS / 64) {
            const int cl = c0 + c;
            int4 d = make_int4(0, 0, 0, 0);
            if (i < v.nidx) {
                const int64_t e = jd.rec0 + (int64_t)cl * v.nidx + i;
                if (lane < nr) {
                    bool any = false;
#pragma unroll
                    for (int p = 0; p < P; p++) any |= !t.flag[c][lane][p];
                    if (any) {
                        const BinOut b = bin_core(t.uvw[lane][0], t.uvw[lane][1], t.uvw[lane][2],
                                                  load_invl(v, (int64_t)t.row[lane], jd.chan0 + cl), jd.g);
                        if (b.to_grid) {
                            d.x = (b.plane * jd.g.T + b.oy) * jd.g.T + b.ox;  // sub-kernel: sums index
                            d.y = (b.cy << 16) | b.cx;
                            d.z = b.conv_idx;
                            d.w = (b.plane << 16) | (b.conj << 15) | b.S;
                            ngrid += 1;
                            npts += (unsigned long long)b.S * b.S * P;
#pragma unroll
                            for (int p = 0; p < P; p++) {
                                float2 x = jd.psf ? make_float2(1.0f, 0.0f) : t.vis[c][lane][p];
                                if (b.conj) x.y = -x.y;
                                a.wv[e * P + p] = x;
                                a.wt[e * P + p] = t.flag[c][lane][p] ? 0.0f : t.wgt[lane][p];
                            }
                        }
                    }
                }
                // entries behind the kept rows (device row order) are not gridded;
                // the tag names the record's grid (channel 0 for MFS jobs)
                a.cjc[e] = (job << 16) | (jd.mfs ? 0 : cl);
                a.desc[e] = d;
            }
            if (a.tile_cnt)
                count_tile_pairs(d, (int)a.js.key0[job] + (jd.mfs ? 0 : cl) * jd.nty * jd.ntx, jd.ntx, a.tile_cnt);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        ngrid += __shfl_xor(ngrid, o, 64);
        npts += __shfl_xor(npts, o, 64);
    }
    if (lane == 0) {
        t.st[wave][0] = ngrid;
        t.st[wave][1] = npts;
    }
    __syncthreads();
    if (tid < 2) {
        unsigned long long s = 0;
        for (int k = 0; k < BT_THREADS / 64; k++) s += t.st[k][tid];
        // striped counters: blocks spread their adds over BIN_STAT_STRIPES addresses
        if (s) atomicAdd(&a.stats[tid * BIN_STAT_STRIPES + blockIdx.x % BIN_STAT_STRIPES], s);
    }
}

VisView make_view(const mti_vis *v)
{
    VisView w;
    w.uvw = v->uvw;
    w.uvw_f32 = v->uvw_f32;
    w.uvw_rs = v->uvw_row_stride;
    w.uvw_cs = v->uvw_comp_stride;
    w.invl = v->inv_lambda;
    w.invl_f32 = v->inv_lambda_f32;
    w.invl_per_row = v->inv_lambda_per_row;
    w.dop = v->row_doppler;
    w.vis = v->vis;
    w.vis_rs = v->vis_row_stride;
    w.vis_cs = v->vis_chan_stride;
    w.flags = v->flags;
    w.flags_i32 = v->flags_i32;
    w.fl_rs = v->flags_row_stride;
    w.fl_cs = v->flags_chan_stride;
    w.weight = v->weight;
    w.w_rs = v->weight_row_stride;
    w.row_index = v->row_index;
    w.nidx = v->row_index ? v->nindex : v->nrows;
    w.npol = v->npol;
    w.nkept = nullptr;
    return w;
}

BinGeom make_geom(const mti_kernels *k, const mti_image_params *p)
{
    BinGeom g;
    g.nx = p->nx;
    g.ny = p->ny;
    g.du = 1.0 / (p->cell_l * (double)p->nx);
    g.dv = 1.0 / (p->cell_m * (double)p->ny);
    g.dw = k->info.w_increment;
    g.hx = (double)(p->nx / 2);
    g.hy = (double)(p->ny / 2);
    g.mode = p->mode;
    g.nplanes = k->info.nplanes;
    g.s = k->info.sampling;
    g.T = k->info.T;
    g.convinfo = k->d_convinfo;
    return g;
}

mti_status launch_bin_pass(mti_ctx *ctx, const BinPassArgs &a, int64_t nblocks, hipStream_t st)
{
    if (nblocks <= 0) return MTI_OK;
    MTI_CHECK(ctx, nblocks < (int64_t(1) << 31), "binning grid above 2^31 blocks");
    switch (a.js.P) {
    case 1: hipLaunchKernelGGL(k_bin_pass<1>, dim3((unsigned)nblocks), dim3(BT_THREADS), 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_bin_pass<2>, dim3((unsigned)nblocks), dim3(BT_THREADS), 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_bin_pass<4>, dim3((unsigned)nblocks), dim3(BT_THREADS), 0, st, a); break;
    default: return set_error(ctx, MTI_ERR_INVALID, "npol must be 1, 2 or 4");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

// Binning blocks of a job with nidx rows and nb channels.
int64_t bin_blocks(int64_t nidx, int nb)
{
    (void)nb;
    return ceil_div(nidx, (int64_t)BT_ROWS);
}

// ---- PreProcessUVW with the reference's seven outputs -----------------------

struct PreRefOut {
    int4 *data;        // {corner_x, corner_y * columns + corner_x, -(corner_x % S), -(corner_y % S)}
    int *conv_index;
    int *to_grid;
    int *to_compress;
    int *support;
    float *take_conj;  // -1 where the record is conjugated (w < 0), else 1
    float *conv_sum;   // [P][N]
};

struct PreRefRec {
    int grid, S, conv_idx;
    int4 loc;
};

// The reference's per-record result before its compression step: zeros for a
// record that is not gridded (all correlations flagged or box off the grid).
__device__ __forceinline__ PreRefRec pre_ref_record(const VisView &v, const BinGeom &g, int64_t rec, int chan,
                                                    BinOut *bo)
{
    const int64_t row = v.row_index ? (int64_t)v.row_index[rec] : rec;
    const BinOut b = bin_record(v, g, row, chan);
    if (bo) *bo = b;
    PreRefRec r;
    r.grid = b.to_grid;
    r.S = b.to_grid ? b.S : 0;
    r.conv_idx = b.to_grid ? b.conv_idx : 0;
    r.loc = b.to_grid ? make_int4(b.cx, b.cy * g.nx + b.cx, -(b.cx % b.S), -(b.cy % b.S)) : make_int4(0, 0, 0, 0);
    return r;
}

// One thread per record.  A gridded record equal in convolution index and
// upper-corner position to the record before it (whatever that one's own
// flags) is marked compressed and not gridded: the rule of
// op/WProjPreProcessUVW.cu:270-285 and its block-boundary pass :287-324,
// which together compare every record with its predecessor.
__global__ void k_pre_ref(VisView v, BinGeom g, int chan, int P, const float *sums, PreRefOut o)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.nidx) return;
    BinOut b;
    const PreRefRec r = pre_ref_record(v, g, i, chan, &b);
    const int64_t row = v.row_index ? (int64_t)v.row_index[i] : i;
    int compress = 0;
    if (r.grid && i > 0) {
        const PreRefRec q = pre_ref_record(v, g, i - 1, chan, nullptr);
        compress = (q.conv_idx == r.conv_idx && q.loc.y == r.loc.y) ? 1 : 0;
    }
    o.data[i] = r.loc;
    o.conv_index[i] = r.conv_idx;
    o.to_grid[i] = r.grid && !compress;
    o.to_compress[i] = compress;
    o.support[i] = r.S;
    // conjugation follows w < 0 for every record with an unflagged
    // correlation, gridded or not (:160-175)
    bool any = false;
    for (int p = 0; p < P; p++)
        if (!vis_flag(v, row, chan, p)) any = true;
    const double w = v.uvw_f32 ? (double)((const float *)v.uvw)[row * v.uvw_rs + 2 * v.uvw_cs]
                               : ((const double *)v.uvw)[row * v.uvw_rs + 2 * v.uvw_cs];
    o.take_conj[i] = (any && w < 0.0) ? -1.0f : 1.0f;
    const double cs = r.grid ? (double)sums[(int64_t)b.plane * g.T * g.T + b.oy * g.T + b.ox] : 0.0;
    for (int p = 0; p < P; p++) {
        const double wt = (double)v.weight[row * v.w_rs + p];
        o.conv_sum[(int64_t)p * v.nidx + i] = vis_flag(v, row, chan, p) ? 0.0f : (float)(cs * wt);
    }
}

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_preprocess(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                                     const mti_vis *v, int32_t chan, int32_t *bins, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    r = validate_vis(ctx, v);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, k, "null kernels");
    MTI_CHECK(ctx, chan >= 0 && (v->inv_lambda_per_row || chan < v->nchan), "channel out of range");
    MTI_CHECK(ctx, k->info.sampling == p->sampling, "kernels built for another sampling");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    VisView vw = make_view(v);
    BinGeom g = make_geom(k, p);
    if (vw.nidx == 0) return MTI_OK;
    MTI_CHECK(ctx, bins, "null output");
    hipLaunchKernelGGL(k_bin_only, dim3((unsigned)ceil_div(vw.nidx, 256)), dim3(256), 0, to_stream(stream),
                       vw, g, chan, bins);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

extern "C" mti_status mti_preprocess_reference(mti_ctx *ctx, const mti_preprocess_ref_params *rp,
                                               const int32_t *convinfo, const float *sums, const mti_vis *v,
                                               int32_t chan, const mti_preprocess_ref_outputs *out, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, rp && out, "null parameters or outputs");
    mti_status r = validate_vis(ctx, v);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, rp->rows >= 2 && rp->columns >= 2, "uvImage.rows/uvImage.columns must be at least 2");
    MTI_CHECK(ctx, rp->u_increment > 0.0 && rp->v_increment > 0.0, "uvImage.u_increment/v_increment must be positive");
    MTI_CHECK(ctx, rp->sampling >= 1, "ConvolutionFunction.sampling must be at least 1");
    MTI_CHECK(ctx, rp->wplanes >= 1, "the convolution-function data needs at least one plane");
    MTI_CHECK(ctx, chan >= 0 && (v->inv_lambda_per_row || chan < v->nchan), "channel out of range");
    VisView vw = make_view(v);
    if (vw.nidx == 0) return MTI_OK;
    MTI_CHECK(ctx, convinfo && sums, "null convolution-function data");
    MTI_CHECK(ctx, out->gridding_data && out->conv_index && out->to_grid && out->to_compress && out->support &&
                       out->take_conjugate && out->conv_sum,
              "null output");
    MTI_CHECK(ctx, vw.nidx < (int64_t(1) << 31), "more than 2^31-1 records");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    BinGeom g;
    g.nx = rp->columns;
    g.ny = rp->rows;
    g.du = rp->u_increment;
    g.dv = rp->v_increment;
    g.dw = rp->wsquareincrement;
    g.hx = (double)(rp->columns / 2);
    g.hy = (double)(rp->rows / 2);
    g.mode = MTI_MODE_WPROJ;
    g.nplanes = rp->wplanes;
    g.s = rp->sampling;
    g.T = (rp->sampling % 2) ? rp->sampling : rp->sampling + 1;
    g.convinfo = (const int2 *)convinfo;
    PreRefOut o{(int4 *)out->gridding_data, out->conv_index, out->to_grid, out->to_compress, out->support,
                out->take_conjugate, out->conv_sum};
    hipLaunchKernelGGL(k_pre_ref, dim3((unsigned)ceil_div(vw.nidx, 256)), dim3(256), 0, to_stream(stream), vw, g,
                       chan, v->npol, sums, o);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}
