// mti_bin.hip -- binning of visibility records (row a4).
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
// flag (:270-324) is not produced: the tile gridder accumulates equal records
// in LDS, which gives the same sum.
#include "mti_grid.hpp"

#include <algorithm>

namespace mti {

struct BinOut {
    int to_grid, plane, S, cx, cy, ox, oy, conj, conv_idx;
};

__device__ __forceinline__ double load_uvw(const VisView &v, int64_t row, int comp)
{
    int64_t o = row * v.uvw_rs + (int64_t)comp * v.uvw_cs;
    return v.uvw_f32 ? (double)((const float *)v.uvw)[o] : ((const double *)v.uvw)[o];
}

__device__ __forceinline__ double load_invl(const VisView &v, int64_t row, int chan)
{
    int64_t o = v.invl_per_row ? row : (int64_t)chan;
    return v.invl_f32 ? (double)((const float *)v.invl)[o] : ((const double *)v.invl)[o];
}

// The FP64 binning of one record from its row's (u, v, w) and 1/lambda.
__device__ __forceinline__ BinOut bin_core(double u, double vv, double w, double f, const BinGeom &g)
{
    BinOut b;
    b.to_grid = 0;
    b.plane = b.S = b.cx = b.cy = b.ox = b.oy = b.conj = b.conv_idx = 0;
    int conj = 0;
    if (w < 0.0) {
        u = -u;
        vv = -vv;
        conj = 1;
    }
    int plane = 0;
    if (g.mode == MTI_MODE_WPROJ) {
        double rp = round(sqrt(fabs((w * f) / g.dw)));
        plane = (rp <= (double)(g.nplanes - 1)) ? (int)rp : g.nplanes - 1;
    }
    const int2 ci = g.convinfo[plane];
    const int S = ci.x;
    const double pos_v = ((vv * f) / g.dv) + g.hy;
    const double pos_u = ((u * f) / g.du) + g.hx;
    const double lx = round(pos_u), ly = round(pos_v);
    const double cxd = lx - (double)(S / 2), cyd = ly - (double)(S / 2);
    if (!(cyd >= 0.0 && cyd <= (double)((g.ny - 1) - S) && cxd >= 0.0 &&
          cxd <= (double)((g.nx - 1) - S)))
        return b;
    int ox = (int)round((lx - pos_u) * (double)g.s) + g.s / 2;
    int oy = (int)round((ly - pos_v) * (double)g.s) + g.s / 2;
    if (g.s % 2) {
        if (ox < 0) ox = 0;
        if (oy < 0) oy = 0;
        if (ox == g.s) ox--;
        if (oy == g.s) oy--;
    }
    const int S2 = S * S;
    b.to_grid = 1;
    b.plane = plane;
    b.S = S;
    b.cx = (int)cxd;
    b.cy = (int)cyd;
    b.ox = ox;
    b.oy = oy;
    b.conj = conj;
    b.conv_idx = ci.y + oy * S2 * g.T + ox * S2;
    return b;
}

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

// Rows and channels of one binning block: a thread owns one row of the
// job's index and walks a run of BIN_CHANS channels.
constexpr int BIN_ROWS = 256;
constexpr int BIN_CHANS = 64;

// One row of a job over channels [c0, c1): the row's (u, v, w) and weights
// are loaded once; record e = rec0 + c_local * nidx + i is written
// channel-major, so the lanes of a wave (consecutive i) store contiguous
// descriptors, visibilities and weights for each channel.
template <int P>
__device__ __forceinline__ void bin_row(const BinPassArgs &a, const JobDev &jd, int job, int64_t i, int c0, int c1,
                                        unsigned long long &ngrid, unsigned long long &npts)
{
    const VisView &v = jd.v;
    const int64_t row = v.row_index ? (int64_t)v.row_index[i] : i;
    const double u = load_uvw(v, row, 0), vv = load_uvw(v, row, 1), w = load_uvw(v, row, 2);
    float wgt[P];
#pragma unroll
    for (int p = 0; p < P; p++) wgt[p] = v.weight[row * v.w_rs + p];
    for (int c = c0; c < c1; c++) {
        const int chan = jd.chan0 + c;
        const int64_t e = jd.rec0 + (int64_t)c * v.nidx + i;
        bool fl[P];
        bool any = false;
#pragma unroll
        for (int p = 0; p < P; p++) {
            fl[p] = vis_flag(v, row, chan, p);
            any |= !fl[p];
        }
        int4 d = make_int4((int)row, 0, 0, 0);
        if (any) {
            const BinOut b = bin_core(u, vv, w, load_invl(v, row, chan), jd.g);
            if (b.to_grid) {
                d.y = (b.cy << 16) | b.cx;
                d.z = b.conv_idx;
                d.w = (b.plane << 16) | (b.conj << 15) | b.S;
                ngrid += 1;
                npts += (unsigned long long)b.S * b.S * P;
                const float *src = v.vis + 2 * (row * v.vis_rs + (int64_t)chan * v.vis_cs);
#pragma unroll
                for (int p = 0; p < P; p++) {
                    float2 x = make_float2(1.0f, 0.0f);
                    if (!jd.psf) x = *(const float2 *)(src + 2 * p);
                    if (b.conj) x.y = -x.y;
                    a.wv[e * P + p] = x;
                    a.wt[e * P + p] = fl[p] ? 0.0f : wgt[p];
                }
            }
        }
        a.cjc[e] = (job << 16) | c;
        a.desc[e] = d;
    }
}

// Blocks are assigned to jobs by the block prefix js.blk0 (so the job is
// uniform in a block); a job's blocks cover its rows in runs of BIN_ROWS
// (fastest) and its channels in runs of BIN_CHANS.  Gridded records also get
// their visibilities (PSF: 1+0i, imaginary part negated for w < 0) and
// flag-masked weights.  Counts are reduced per block and added once.
template <int P>
__global__ __launch_bounds__(BIN_ROWS) void k_bin_pass(BinPassArgs a)
{
    __shared__ unsigned long long s_st[BIN_ROWS / 64][2];
    const int job = job_of(a.js.blk0, a.js.njobs, (int64_t)blockIdx.x);
    const JobDev jd = a.js.jobs[job];  // uniform: scalar loads into registers
    const int64_t lb = (int64_t)blockIdx.x - a.js.blk0[job];
    const int64_t nrb = (jd.v.nidx + BIN_ROWS - 1) / BIN_ROWS;
    const int64_t rb = lb % nrb;
    const int c0 = (int)(lb / nrb) * BIN_CHANS;
    const int c1 = min(jd.nb, c0 + BIN_CHANS);
    const int64_t i = rb * BIN_ROWS + threadIdx.x;
    unsigned long long ngrid = 0, npts = 0;
    if (i < jd.v.nidx) bin_row<P>(a, jd, job, i, c0, c1, ngrid, npts);
    for (int o = 32; o >= 1; o >>= 1) {
        ngrid += __shfl_xor(ngrid, o, 64);
        npts += __shfl_xor(npts, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_st[threadIdx.x >> 6][0] = ngrid;
        s_st[threadIdx.x >> 6][1] = npts;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long t = 0;
        for (int k = 0; k < BIN_ROWS / 64; k++) t += s_st[k][threadIdx.x];
        if (t) atomicAdd(&a.stats[threadIdx.x], t);
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
    case 1: hipLaunchKernelGGL(k_bin_pass<1>, dim3((unsigned)nblocks), dim3(BIN_ROWS), 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_bin_pass<2>, dim3((unsigned)nblocks), dim3(BIN_ROWS), 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_bin_pass<4>, dim3((unsigned)nblocks), dim3(BIN_ROWS), 0, st, a); break;
    default: return set_error(ctx, MTI_ERR_INVALID, "npol must be 1, 2 or 4");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

// Binning blocks of a job with nidx rows and nb channels.
int64_t bin_blocks(int64_t nidx, int nb)
{
    return ceil_div(nidx, (int64_t)BIN_ROWS) * ceil_div((int64_t)nb, (int64_t)BIN_CHANS);
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
