// mti_degrid.hip -- W-projection prediction (degridding): model visibilities
// from model images, the step beside the imaging path in a major cycle.
//
// The reference has no degridder (it images only); this is built as the
// adjoint of its imaging steps:
//   * mti_model_grids: Stokes model images in the finalize's layout (m axis
//     flipped, op/PostIFFTStokesConverter.cu:172) -> correlation images (the
//     inverse of the Stokes formulas, op/PostIFFTStokesConverter.cu:51-80)
//     times the inverse of the attenuation the gridding kernel imposes at
//     each pixel's sky position (the casagridsf taper, op/CasaGridSF.cu:
//     151-199, and the nearest-sample sinc, op/SincCorrector.cu:29-71), then
//     the unnormalised forward 2-D FFT with the imaging's shifts (FP64
//     butterflies of mti_fft2d_inverse on conjugated planes);
//   * mti_degrid: every record binned exactly as the gridder bins it (FP64,
//     -ffp-contract=off, flags ignored), V_p = sum_taps conj(K) G_p over its
//     support box, divided by the sub-kernel's sum (the quantity the imaging
//     normaliser divides by); a record binned at (-u, -v, |w|) (w < 0) gets
//     the Hermitian image of that prediction, V_p(u) = conj(V_p'(-u)), with
//     p' = p except XY <-> YX (RL <-> LR) for four correlations.
// The degridder is a gather: one wave per record, the lanes over the box's
// taps, FP64 sums reduced across the wave -- no atomics.  It reads S^2 kernel
// samples and P S^2 grid cells per record, neighbouring records (baseline
// order) sharing grid cells in L2.
#include "mti_binning.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>

namespace mti {

constexpr int DG_THREADS = 256;
constexpr int MG_THREADS = 256;
constexpr double DG_PI = 3.14159265358979323846;

// ---- model grids -----------------------------------------------------------------

// Inverse kernel attenuation at transform pixel (jy, jx): the taper and
// sinc of the finalize's corrections evaluated at the pixel's sky offset.
__device__ __forceinline__ double model_correction(int jy, int jx, int nx, int ny, int s)
{
    const double nuy = fabs((double)(jy - ny / 2)) / (double)(ny / 2);
    const double nux = fabs((double)(jx - nx / 2)) / (double)(nx / 2);
    const double tap = mti_dev::gridsf(nuy) * mti_dev::gridsf(nux);
    const double xx = (double)(jx - nx / 2) / ((double)nx * s);
    const double yy = (double)(jy - ny / 2) / ((double)ny * s);
    const double sx = xx == 0.0 ? 1.0 : (DG_PI * xx) / sin(DG_PI * xx);
    const double sy = yy == 0.0 ? 1.0 : (DG_PI * yy) / sin(DG_PI * yy);
    return tap > 0.0 ? (sx * sy) / tap : 0.0;
}

__device__ __forceinline__ int64_t pmod64(int64_t a, int64_t m)
{
    const int64_t r = a % m;
    return r < 0 ? r + m : r;
}

// Correlation q of the Stokes values s[] (re, im).
__device__ __forceinline__ double2 stokes_corr(int P, int poltype, const double *s, int q)
{
    if (P == 1) return make_double2(s[0], 0.0);
    if (P == 2) return make_double2(q == 0 ? s[0] + s[1] : s[0] - s[1], 0.0);
    if (poltype == MTI_POL_LINEAR) {  // XX = I+Q, XY = U+iV, YX = U-iV, YY = I-Q
        switch (q) {
        case 0: return make_double2(s[0] + s[1], 0.0);
        case 1: return make_double2(s[2], s[3]);
        case 2: return make_double2(s[2], -s[3]);
        default: return make_double2(s[0] - s[1], 0.0);
        }
    }
    switch (q) {  // RR = I+V, RL = Q-iU, LR = Q+iU, LL = I-V
    case 0: return make_double2(s[0] + s[3], 0.0);
    case 1: return make_double2(s[1], -s[2]);
    case 2: return make_double2(s[1], s[2]);
    default: return make_double2(s[0] - s[3], 0.0);
    }
}

// One lane per transform pixel of plane batch [b0, b0 + nb) (plane = channel
// * P + correlation): the conjugated, corrected correlation value, placed at
// the transform's origin-at-corner position.
__global__ __launch_bounds__(MG_THREADS) void k_model_prep(const float *model, int P, int poltype, int nx, int ny,
                                                         int s, int64_t b0, int nb, float2 *planes)
{
    const int64_t plane_px = (int64_t)nx * ny;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane_px * nb) return;
    const int64_t b = i / plane_px, px = i - b * plane_px;
    const int64_t pl = b0 + b;
    const int chan = (int)(pl / P), q = (int)(pl - (int64_t)chan * P);
    const int jy = (int)(px / nx), jx = (int)(px - (int64_t)jy * nx);
    const int r = ny - 1 - jy;  // image rows are m-flipped
    double st[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < P; k++) st[k] = (double)model[(((int64_t)chan * P + k) * ny + r) * nx + jx];
    const double2 c = stokes_corr(P, poltype, st, q);
    const double f = model_correction(jy, jx, nx, ny, s);
    const int64_t o = pmod64(jy - ny / 2, ny) * nx + pmod64(jx - nx / 2, nx);
    // conjugated: the inverse transform of conj(x) is conj(forward transform of x)
    planes[b * plane_px + o] = make_float2((float)(c.x * f), (float)(-c.y * f));
}

// Grid cell (gy, gx) of every plane of the batch from the transformed planes:
// conjugated back and shifted to the grid's origin at (ny/2, nx/2).
__global__ __launch_bounds__(MG_THREADS) void k_model_post(const float2 *planes, int nx, int ny, int64_t b0, int nb,
                                                         float2 *grids)
{
    const int64_t plane_px = (int64_t)nx * ny;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane_px * nb) return;
    const int64_t b = i / plane_px, px = i - b * plane_px;
    const int gy = (int)(px / nx), gx = (int)(px - (int64_t)gy * nx);
    const int64_t o = pmod64(gy - ny / 2, ny) * nx + pmod64(gx - nx / 2, nx);
    const float2 v = planes[b * plane_px + o];
    grids[(b0 + b) * plane_px + px] = make_float2(v.x, -v.y);
}

// ---- degridder ---------------------------------------------------------------------

template <int PREC>
__device__ __forceinline__ float2 table_sample(const void *tab, int64_t k)
{
    if (PREC == MTI_TABLE_F32) return ((const float2 *)tab)[k];
    const uint32_t w = ((const uint32_t *)tab)[k];
    if (PREC == MTI_TABLE_F16)
        return make_float2(__half2float(__ushort_as_half((unsigned short)(w & 0xffffu))),
                           __half2float(__ushort_as_half((unsigned short)(w >> 16))));
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

struct DegridArgs {
    VisView v;
    BinGeom g;
    int chan0, nchan;
    const void *table;
    const float *sums;  // [nplanes][T*T]
    const float2 *grids;  // [nchan][P][ny][nx]
    float *out;           // complex64, the strides of v
    unsigned long long *count;  // predicted records
};

// One wave per record: lanes over the support box's taps, FP64 sums of
// conj(K) G per correlation reduced across the wave.
template <int P, int PREC>
__global__ __launch_bounds__(DG_THREADS) void k_degrid(DegridArgs a)
{
    const int lane = threadIdx.x & 63;
    const int64_t nrec = a.v.nidx * a.nchan;
    const int64_t waves = (int64_t)gridDim.x * (DG_THREADS / 64);
    const int64_t plane_px = (int64_t)a.g.nx * a.g.ny;
    unsigned long long npred = 0;
    for (int64_t rec = (int64_t)blockIdx.x * (DG_THREADS / 64) + (threadIdx.x >> 6); rec < nrec; rec += waves) {
        const int c = (int)(rec / a.v.nidx);
        const int64_t i = rec - (int64_t)c * a.v.nidx;
        const int64_t row = a.v.row_index ? (int64_t)a.v.row_index[i] : i;
        const int chan = a.chan0 + c;
        const BinOut b = bin_core(load_uvw(a.v, row, 0), load_uvw(a.v, row, 1), load_uvw(a.v, row, 2),
                                  load_invl(a.v, row, chan), a.g);
        double acc[2 * P];
#pragma unroll
        for (int q = 0; q < 2 * P; q++) acc[q] = 0.0;
        if (b.to_grid) {
            const int S = b.S, S2 = S * S;
            const float2 *gc = a.grids + (int64_t)c * P * plane_px;
            for (int t = lane; t < S2; t += 64) {
                const int sy = t / S, sx = t - sy * S;
                const float2 k = table_sample<PREC>(a.table, (int64_t)b.conv_idx + t);
                const double kr = k.x, ki = k.y;
                const int64_t cell = (int64_t)(b.cy + sy) * a.g.nx + (b.cx + sx);
#pragma unroll
                for (int q = 0; q < P; q++) {
                    const float2 gv = gc[(int64_t)q * plane_px + cell];
                    const double gr = gv.x, gi = gv.y;
                    // conj(k) * g
                    acc[2 * q] = fma(kr, gr, fma(ki, gi, acc[2 * q]));
                    acc[2 * q + 1] = fma(kr, gi, fma(-ki, gr, acc[2 * q + 1]));
                }
            }
#pragma unroll
            for (int q = 0; q < 2 * P; q++)
                for (int o = 32; o >= 1; o >>= 1) acc[q] += __shfl_xor(acc[q], o, 64);
        }
        if (lane == 0) {
            float *o = a.out + 2 * (row * a.v.vis_rs + (int64_t)chan * a.v.vis_cs);
            const double sum = b.to_grid ? (double)a.sums[b.plane * a.g.T * a.g.T + b.oy * a.g.T + b.ox] : 1.0;
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int src = (b.conj && P == 4 && (q == 1 || q == 2)) ? 3 - q : q;
                const double re = acc[2 * src] / sum, im = acc[2 * src + 1] / sum;
                o[2 * q] = (float)re;
                o[2 * q + 1] = (float)(b.conj ? -im : im);
            }
            npred += b.to_grid ? 1ull : 0ull;
        }
    }
    if (lane == 0 && npred) atomicAdd(a.count, npred);
}

template <int P>
static void launch_degrid(const DegridArgs &a, int prec, int blocks, hipStream_t st)
{
    if (prec == MTI_TABLE_F32) hipLaunchKernelGGL((k_degrid<P, MTI_TABLE_F32>), dim3(blocks), dim3(DG_THREADS), 0, st, a);
    else if (prec == MTI_TABLE_F16)
        hipLaunchKernelGGL((k_degrid<P, MTI_TABLE_F16>), dim3(blocks), dim3(DG_THREADS), 0, st, a);
    else hipLaunchKernelGGL((k_degrid<P, MTI_TABLE_BF16>), dim3(blocks), dim3(DG_THREADS), 0, st, a);
}

__global__ void k_zero_count(unsigned long long *c) { *c = 0ull; }

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_model_grids(mti_ctx *ctx, const mti_image_params *p, int32_t npol, int32_t poltype,
                                      int32_t nchan, const float *model, void *grids, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, npol == 1 || npol == 2 || npol == 4, "npol must be 1, 2 or 4");
    MTI_CHECK(ctx, poltype == MTI_POL_LINEAR || poltype == MTI_POL_CIRCULAR, "bad poltype");
    MTI_CHECK(ctx, nchan >= 1 && model && grids, "null model/grids or nchan < 1");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    const int64_t plane_px = (int64_t)p->nx * p->ny, nplanes = (int64_t)nchan * npol;
    // planes transformed per batch: at most 512 MiB of complex64 scratch
    const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(nplanes, (int64_t(512) << 20) / (plane_px * 8)));
    void *w = nullptr;
    if ((r = scratch_get(ctx, SLOT_CORR, (size_t)(batch * plane_px * 8), &w)) != MTI_OK) return r;
    float2 *planes = (float2 *)w;
    for (int64_t b0 = 0; b0 < nplanes; b0 += batch) {
        const int nb = (int)std::min(batch, nplanes - b0);
        const int64_t n = plane_px * nb;
        const unsigned blocks = (unsigned)ceil_div(n, MG_THREADS);
        hipLaunchKernelGGL(k_model_prep, dim3(blocks), dim3(MG_THREADS), 0, st, model, (int)npol, (int)poltype,
                           p->nx, p->ny, p->sampling, b0, nb, planes);
        MTI_HIP(ctx, hipGetLastError());
        if ((r = mti_fft2d_inverse(ctx, (float *)planes, p->nx, p->ny, nb, stream)) != MTI_OK) return r;
        hipLaunchKernelGGL(k_model_post, dim3(blocks), dim3(MG_THREADS), 0, st, (const float2 *)planes, p->nx, p->ny,
                           b0, nb, (float2 *)grids);
        MTI_HIP(ctx, hipGetLastError());
    }
    return MTI_OK;
}

extern "C" mti_status mti_degrid(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p, const mti_vis *v,
                                 int32_t chan0, int32_t nchan, const void *grids, float *vis_out,
                                 int64_t *predicted, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    r = validate_vis(ctx, v);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, k && k->info.sampling == p->sampling, "null kernels or kernels built for another sampling");
    MTI_CHECK(ctx, grids && vis_out, "null grids or output");
    MTI_CHECK(ctx, nchan >= 1, "nchan < 1");
    if (v->inv_lambda_per_row) MTI_CHECK(ctx, nchan == 1, "per-row 1/lambda implies a single channel per call");
    else MTI_CHECK(ctx, chan0 >= 0 && chan0 + nchan <= v->nchan, "channel range out of bounds");
    MTI_CHECK(ctx, p->nx <= 65535 && p->ny <= 65535, "grid side above 65535");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    void *w = nullptr;
    if ((r = scratch_get(ctx, SLOT_MISC2, 256, &w)) != MTI_OK) return r;
    DegridArgs a;
    a.v = make_view(v);
    a.g = make_geom(k, p);
    a.chan0 = chan0;
    a.nchan = nchan;
    a.table = k->d_table;
    a.sums = k->d_sums;
    a.grids = (const float2 *)grids;
    a.out = vis_out;
    a.count = (unsigned long long *)w;
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, st, a.count);
    const int64_t nrec = a.v.nidx * nchan;
    if (nrec > 0) {
        const int blocks = (int)std::min<int64_t>(ceil_div(nrec, DG_THREADS / 64), (int64_t)ctx->num_cus * 16);
        switch (v->npol) {
        case 1: launch_degrid<1>(a, k->info.precision, blocks, st); break;
        case 2: launch_degrid<2>(a, k->info.precision, blocks, st); break;
        default: launch_degrid<4>(a, k->info.precision, blocks, st); break;
        }
        MTI_HIP(ctx, hipGetLastError());
    }
    if (predicted) {
        unsigned long long h = 0;
        MTI_HIP(ctx, hipMemcpyAsync(&h, a.count, sizeof(h), hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipStreamSynchronize(st));
        *predicted = (int64_t)h;
    }
    return MTI_OK;
}
