// mti_binning.hpp -- the FP64 binning of one record (op/WProjPreProcessUVW.cu:
// 111-286), shared by the gridding path's binning pass (mti_bin.hip) and the
// degridder (mti_degrid.hip).  Every source including it is built with
// -ffp-contract=off, so that each operation rounds exactly as the FP64 oracle's.
#pragma once
#include "mti_grid.hpp"

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
    const double f = v.invl_f32 ? (double)((const float *)v.invl)[o] : ((const double *)v.invl)[o];
    return v.dop ? f * v.dop[row] : f;
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

}  // namespace mti
