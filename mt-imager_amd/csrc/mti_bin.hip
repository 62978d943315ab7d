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

__device__ BinOut bin_record(const VisView &v, const BinGeom &g, int64_t row, int chan)
{
    BinOut b;
    b.to_grid = 0;
    b.plane = b.S = b.cx = b.cy = b.ox = b.oy = b.conj = b.conv_idx = 0;
    bool any = false;
    for (int p = 0; p < v.npol; p++)
        if (!vis_flag(v, row, chan, p)) any = true;
    if (!any) return b;
    double u = load_uvw(v, row, 0), vv = load_uvw(v, row, 1), w = load_uvw(v, row, 2);
    const double f = load_invl(v, row, chan);
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

__device__ __forceinline__ void bin_one(const BinPassArgs &a, int64_t t, unsigned long long &ngrid,
                                        unsigned long long &npts)
{
    const int64_t i = t / a.nb;
    const int c_local = (int)(t - i * a.nb);
    const int64_t e = (int64_t)c_local * a.v.nidx + i;
    const int64_t row = a.v.row_index ? (int64_t)a.v.row_index[i] : i;
    const int chan = a.chan0 + c_local;
    BinOut b = bin_record(a.v, a.g, row, chan);
    int4 d = make_int4((int)row, 0, 0, 0);
    if (b.to_grid) {
        d.y = (b.cy << 16) | b.cx;
        d.z = b.conv_idx;
        d.w = (b.plane << 16) | (b.conj << 15) | b.S;
        ngrid += 1;
        npts += (unsigned long long)b.S * b.S * a.v.npol;
        const int P = a.v.npol;
        for (int p = 0; p < P; p++) {
            const bool fl = vis_flag(a.v, row, chan, p);
            a.wt[e * P + p] = fl ? 0.0f : a.v.weight[row * a.v.w_rs + p];
            float2 vv = make_float2(1.0f, 0.0f);
            if (!a.psf) vv = *(const float2 *)(a.v.vis + 2 * (row * a.v.vis_rs + (int64_t)chan * a.v.vis_cs + p));
            if (b.conj) vv.y = -vv.y;
            a.wv[e * P + p] = vv;
        }
    }
    a.valid[e] = b.to_grid ? 1u : 0u;
    a.desc[e] = d;
}

// Threads run over (row, channel) with the channel fastest, so a wave reads
// one row's flags and visibilities contiguously (MeasurementSet order); the
// record e = c_local * nidx + i is written channel-major (baseline-major
// within a channel).  Gridded records also get their visibilities (PSF: 1+0i,
// imaginary part negated for w < 0) and flag-masked weights, which the
// compaction moves next to the descriptors.  Grid-stride loop: the gridded
// record / point counts are reduced per block and added once.
__global__ __launch_bounds__(256) void k_bin_pass(BinPassArgs a, int64_t nrec)
{
    __shared__ unsigned long long s_st[4][2];
    unsigned long long ngrid = 0, npts = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrec; t += (int64_t)gridDim.x * blockDim.x)
        bin_one(a, t, ngrid, npts);
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
        const unsigned long long v = s_st[0][threadIdx.x] + s_st[1][threadIdx.x] + s_st[2][threadIdx.x] +
                                     s_st[3][threadIdx.x];
        if (v) atomicAdd(&a.stats[threadIdx.x], v);
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

mti_status launch_bin_pass(mti_ctx *ctx, const BinPassArgs &a, int64_t nrec, hipStream_t st)
{
    if (nrec <= 0) return MTI_OK;
    const int64_t blocks = std::min<int64_t>(ceil_div(nrec, 256), (int64_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_bin_pass, dim3((unsigned)blocks), dim3(256), 0, st, a, nrec);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
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
