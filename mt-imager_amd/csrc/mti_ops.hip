// mti_ops.hip -- the reference's finalizer operators with their own array
// contracts, for callers that keep the reference's operator DAG
// (PreProcessUVW with its seven outputs is mti_preprocess_reference, in
// mti_bin.hip beside the binning it shares):
//   PreIFFTUVCorrection  op/PreIFFTUVCorrection.cpp:36-62, .cu:28-90
//   ZeroToCentre2DShift  gafw/StandardOperators/ZeroToCentre2DShift.cpp:36-62, .cu:31-119
//   StokesConverter      op/PostIFFTStokesConverter.cpp:37-67, .cu:42-234
//   casagridsf           op/CasaGridSF.cpp:38-128, .cu:107-199 (its three modes)
//   SincCorrector        op/SincCorrector.cpp:37-50, .cu:29-71
// plus the f32 normaliser the reference's WImager outputs (comp/WImager.cpp:196-197).
// Each is one elementwise HBM pass; the arithmetic is FP64 where the
// reference's is FP32 (psi, sinc, the normalisation), rounded once to the
// output type.
#include "mti_internal.hpp"

#include <algorithm>
#include <cmath>

namespace mti {

// ---- the finalizer's operators ---------------------------------------------

// PreIFFTUVCorrection: [rows][cols][P] -> [P][rows][cols], each axis moved by
// OriginToCorner (op/common.hcu:122-136): output index o holds input
// (o + n/2) mod n.
template <typename T>
__global__ void k_uv_reorder(const T *in, T *out, int rows, int cols, int P)
{
    const int64_t n = (int64_t)rows * cols * P;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t plane = (int64_t)rows * cols;
        const int p = (int)(e / plane);
        const int64_t r = e - (int64_t)p * plane;
        const int oy = (int)(r / cols), ox = (int)(r - (int64_t)oy * cols);
        int y = oy + rows / 2, x = ox + cols / 2;
        if (y >= rows) y -= rows;
        if (x >= cols) x -= cols;
        out[e] = in[((int64_t)y * cols + x) * P + p];
    }
}

// ZeroToCentre2DShift: per [rows][cols] plane, input index i goes to
// (i + n/2) mod n (OriginToCentre1D, ZeroToCentre2DShift.cu:31-45); output
// index o reads input (o + n - n/2) mod n.
template <typename T>
__global__ void k_centre_shift(const T *in, T *out, int64_t planes, int rows, int cols)
{
    const int64_t plane = (int64_t)rows * cols, n = plane * planes;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pl = e / plane;
        const int64_t r = e - pl * plane;
        const int oy = (int)(r / cols), ox = (int)(r - (int64_t)oy * cols);
        int y = oy + rows - rows / 2, x = ox + cols - cols / 2;
        if (y >= rows) y -= rows;
        if (x >= cols) x -= cols;
        out[e] = in[pl * plane + (int64_t)y * cols + x];
    }
}

// StokesConverter: normalise the used parts, form Stokes, write rows flipped
// (op/PostIFFTStokesConverter.cu:51-59, 73-80, 150-180).
__global__ void k_stokes(const float2 *in, const float *norm, float *out, int P, int rows, int cols, int circular)
{
    const int64_t plane = (int64_t)rows * cols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < plane; e += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(e / cols), x = (int)(e - (int64_t)y * cols);
        double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0}, s[4] = {0, 0, 0, 0};
        for (int p = 0; p < P; p++) {
            const float2 v = in[(int64_t)p * plane + e];
            const double inv = 1.0 / (double)norm[p];
            re[p] = (double)v.x * inv;
            im[p] = (double)v.y * inv;
        }
        if (P == 1) {
            s[0] = re[0];
        } else if (P == 2) {
            s[0] = 0.5 * (re[0] + re[1]);
            s[1] = 0.5 * (re[0] - re[1]);
        } else if (!circular) {
            s[0] = 0.5 * (re[0] + re[3]);
            s[1] = 0.5 * (re[0] - re[3]);
            s[2] = 0.5 * (re[1] + re[2]);
            s[3] = 0.5 * (im[1] - im[2]);
        } else {
            s[0] = 0.5 * (re[0] + re[3]);
            s[1] = 0.5 * (re[1] + re[2]);
            s[2] = 0.5 * (im[2] - im[1]);
            s[3] = 0.5 * (re[0] - re[3]);
        }
        const int64_t o = (int64_t)(rows - 1 - y) * cols + x;
        for (int p = 0; p < P; p++) out[(int64_t)p * plane + o] = (float)s[p];
    }
}

__device__ __forceinline__ double taper_at(int j, int n)
{
    const int h = n / 2;
    return mti_dev::gridsf(fabs((double)(j - h)) / (double)h);
}

// casagridsf with lmPlanePostCorrection = true: divide each plane by
// psi(|j - rows/2|/(rows/2)) psi(|i - cols/2|/(cols/2)) (CasaGridSF.cu:174-199),
// or (in == null) write that product (the generate mode, :151-173).
template <typename T>
__global__ void k_taper(const T *in, T *out, int64_t planes, int rows, int cols)
{
    const int64_t plane = (int64_t)rows * cols, n = plane * planes;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % plane;
        const int y = (int)(r / cols), x = (int)(r - (int64_t)y * cols);
        const double t = taper_at(y, rows) * taper_at(x, cols);
        out[e] = in ? (T)((double)in[e] / t) : (T)t;
    }
}

// casagridsf with lmPlanePostCorrection = false: the oversampled taper-only
// gridding function (1 - nu^2) psi(nu) per axis, nu = |i - dim/2| / (3 s),
// zero from 3 s on, of dim = S s (odd s) or S s + 1 (CasaGridSF.cu:107-149);
// complex outputs get a zero imaginary part.
template <typename R>
__global__ void k_taper_kernel(R *out, int dim, int s, int complex_out)
{
    const int64_t n = (int64_t)dim * dim;
    const int c = dim / 2, fz = 3 * s;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = abs((int)(e / dim) - c), i = abs((int)(e % dim) - c);
        double v = 0.0;
        if (i < fz && j < fz) {
            const double ni = (double)i / fz, nj = (double)j / fz;
            v = (1.0 - ni * ni) * mti_dev::gridsf(ni) * (1.0 - nj * nj) * mti_dev::gridsf(nj);
        }
        if (complex_out) {
            out[2 * e] = (R)v;
            out[2 * e + 1] = (R)0;
        } else {
            out[e] = (R)v;
        }
    }
}

// SincCorrector: multiply by the inverse sinc of the oversampled pixel,
// pi x / sin(pi x), x = (i - n/2) / (n s) per axis (SincCorrector.cu:29-58).
__global__ void k_sinc(const float *in, float *out, int64_t planes, int rows, int cols, int s)
{
    const int64_t plane = (int64_t)rows * cols, n = plane * planes;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % plane;
        const int y = (int)(r / cols), x = (int)(r - (int64_t)y * cols);
        const double xx = (double)(x - cols / 2) / ((double)cols * s), yy = (double)(y - rows / 2) / ((double)rows * s);
        const double sx = xx == 0.0 ? 1.0 : M_PI * xx / sin(M_PI * xx);
        const double sy = yy == 0.0 ? 1.0 : M_PI * yy / sin(M_PI * yy);
        out[e] = (float)((double)in[e] * (sx * sy));
    }
}

__global__ void k_norm_f32(const double *a, float *b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

static dim3 ew_blocks(int64_t n)
{
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 1 << 20)));
}

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_uv_reorder(const void *in, void *out, int32_t rows, int32_t cols, int32_t npol,
                                     int32_t c128, void *stream)
{
    if (!in || !out || rows < 1 || cols < 1 || npol < 1) return MTI_ERR_INVALID;
    const int64_t n = (int64_t)rows * cols * npol;
    if (c128)
        hipLaunchKernelGGL(k_uv_reorder<double2>, ew_blocks(n), dim3(256), 0, to_stream(stream), (const double2 *)in,
                           (double2 *)out, rows, cols, npol);
    else
        hipLaunchKernelGGL(k_uv_reorder<float2>, ew_blocks(n), dim3(256), 0, to_stream(stream), (const float2 *)in,
                           (float2 *)out, rows, cols, npol);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}

extern "C" mti_status mti_centre_shift(const void *in, void *out, int64_t planes, int32_t rows, int32_t cols,
                                       int32_t c128, void *stream)
{
    if (!in || !out || rows < 1 || cols < 1 || planes < 1) return MTI_ERR_INVALID;
    const int64_t n = (int64_t)rows * cols * planes;
    if (c128)
        hipLaunchKernelGGL(k_centre_shift<double2>, ew_blocks(n), dim3(256), 0, to_stream(stream),
                           (const double2 *)in, (double2 *)out, planes, rows, cols);
    else
        hipLaunchKernelGGL(k_centre_shift<float2>, ew_blocks(n), dim3(256), 0, to_stream(stream),
                           (const float2 *)in, (float2 *)out, planes, rows, cols);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}

extern "C" mti_status mti_stokes(const float *in, const float *norm, float *out, int32_t npol, int32_t rows,
                                 int32_t cols, int32_t poltype, void *stream)
{
    if (!in || !norm || !out || rows < 1 || cols < 1 || !(npol == 1 || npol == 2 || npol == 4) ||
        !(poltype == MTI_POL_LINEAR || poltype == MTI_POL_CIRCULAR))
        return MTI_ERR_INVALID;
    hipLaunchKernelGGL(k_stokes, ew_blocks((int64_t)rows * cols), dim3(256), 0, to_stream(stream),
                       (const float2 *)in, norm, out, npol, rows, cols, poltype == MTI_POL_CIRCULAR ? 1 : 0);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}

extern "C" mti_status mti_taper_correct(const void *in, void *out, int64_t planes, int32_t rows, int32_t cols,
                                        int32_t f64, void *stream)
{
    if (!out || rows < 2 || cols < 2 || planes < 1) return MTI_ERR_INVALID;
    const int64_t n = (int64_t)rows * cols * planes;
    if (f64)
        hipLaunchKernelGGL(k_taper<double>, ew_blocks(n), dim3(256), 0, to_stream(stream), (const double *)in,
                           (double *)out, planes, rows, cols);
    else
        hipLaunchKernelGGL(k_taper<float>, ew_blocks(n), dim3(256), 0, to_stream(stream), (const float *)in,
                           (float *)out, planes, rows, cols);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}

extern "C" mti_status mti_taper_kernel(void *out, int32_t support, int32_t sampling, int32_t complex_out,
                                       int32_t f64, void *stream)
{
    if (!out || support < 1 || support % 2 == 0 || sampling < 1) return MTI_ERR_INVALID;
    const int dim = (sampling % 2) ? support * sampling : support * sampling + 1;
    const int64_t n = (int64_t)dim * dim;
    if (f64)
        hipLaunchKernelGGL(k_taper_kernel<double>, ew_blocks(n), dim3(256), 0, to_stream(stream), (double *)out, dim,
                           sampling, complex_out ? 1 : 0);
    else
        hipLaunchKernelGGL(k_taper_kernel<float>, ew_blocks(n), dim3(256), 0, to_stream(stream), (float *)out, dim,
                           sampling, complex_out ? 1 : 0);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}

extern "C" mti_status mti_sinc_correct(const float *in, float *out, int64_t planes, int32_t rows, int32_t cols,
                                       int32_t sampling, void *stream)
{
    if (!in || !out || rows < 1 || cols < 1 || planes < 1 || sampling < 1) return MTI_ERR_INVALID;
    hipLaunchKernelGGL(k_sinc, ew_blocks((int64_t)rows * cols * planes), dim3(256), 0, to_stream(stream), in, out,
                       planes, rows, cols, sampling);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}

extern "C" mti_status mti_normalizer_f32(const double *norm, float *out, int64_t n, void *stream)
{
    if (!norm || !out || n < 1) return MTI_ERR_INVALID;
    hipLaunchKernelGGL(k_norm_f32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, to_stream(stream), norm, out, n);
    return hipGetLastError() == hipSuccess ? MTI_OK : MTI_ERR_HIP;
}
