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
};

struct BinGeom {
    int nx, ny;
    double du, dv, dw, hx, hy;
    int mode, nplanes, s, T;
    const int2 *convinfo;  // {support, base}
};

// Record descriptor written by the binning pass (16 bytes):
//   x = row, y = corner_y << 16 | corner_x, z = conv_idx,
//   w = plane << 16 | conj << 15 | support
struct BinPassArgs {
    VisView v;
    BinGeom g;
    int chan0, nb;
    int psf;
    uint32_t *valid;            // [nrec] 1 if the record is gridded
    int4 *desc;                 // [nrec]
    float2 *wv;                 // [nrec][npol] visibilities of gridded records
    float *wt;                  // [nrec][npol] weights, 0 where flagged
    unsigned long long *stats;  // [0] gridded records, [1] gridded points
};

mti_status launch_bin_pass(mti_ctx *ctx, const BinPassArgs &a, int64_t nrec, hipStream_t st);
VisView make_view(const mti_vis *v);
BinGeom make_geom(const mti_kernels *k, const mti_image_params *p);

__device__ __forceinline__ bool vis_flag(const VisView &v, int64_t row, int chan, int p)
{
    int64_t o = row * v.fl_rs + (int64_t)chan * v.fl_cs + p;
    return v.flags_i32 ? (((const int32_t *)v.flags)[o] != 0) : (((const uint8_t *)v.flags)[o] != 0);
}

}  // namespace mti
