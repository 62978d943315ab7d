// mti_fft.hip -- 2-D inverse FFT and the image finalizer (rows a16-a19).
//
// Reference (comp/ImageFinalizer.cpp:31-44): PreIFFTUVCorrection
// ([v][u][P] -> [P][v][u] plus ifftshift, op/PreIFFTUVCorrection.cu:28-90),
// cuFFT C2C inverse per polarisation, unnormalised (gafw/StandardOperators/
// IFFT2D.cpp:86-111), fftshift (ZeroToCentre2DShift.cu:66-91),
// StokesConverter (divide by the normaliser, Stokes I/Q/U/V, m-axis flip,
// op/PostIFFTStokesConverter.cu:42-180), casagridsf taper correction
// (op/CasaGridSF.cu:174-199) and SincCorrector (op/SincCorrector.cu:29-58),
// the last two evaluated on the flipped array exactly as the reference does.
//
// MI355X design, three HBM passes per channel:
//   A  row transforms in LDS (radix-4/2 Stockham), fused with the uv-plane
//      reorder/ifftshift;                          grid   -> X[p][v][u]
//   B1 column transforms, four-step first half (length N1 over rows strided
//      by N2, times the inter-step twiddle), in place;  tiles of 16 columns
//      so every row segment read is 128 contiguous bytes;
//   B2 four-step second half (length N2 over N2 contiguous rows) fused with
//      fftshift + normalise + Stokes + flip + taper and sinc correction,
//      all P correlations of a pixel in one workgroup  -> f32 image.
// Twiddles come from an FP64-computed table exp(+2 pi i t / NT).
#include "mti_internal.hpp"

#include <algorithm>
#include <cmath>

namespace mti {

constexpr int FFT_THREADS = 256;
constexpr int COLS = 16;  // columns per tile in the column passes

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Batch of nb inverse FFTs of length n (power of two) stored contiguously in
// LDS buffer a (stride n); b is a scratch buffer of the same size.  Returns
// the buffer holding the result.  tw[t] = exp(+2 pi i t / NT), tw_mul = NT/n.
__device__ float2 *lds_ifft(float2 *a, float2 *b, int n, int logn, int nb, const float2 *tw,
                            int tw_mul)
{
    int Ns = 1;
    while (Ns < n) {
        const int R = (n / Ns >= 4) ? 4 : 2;
        const int per = n / R;
        const int logper = logn - (R == 4 ? 2 : 1);
        const int twstep = (n / (Ns * R)) * tw_mul;
        for (int idx = threadIdx.x; idx < per * nb; idx += blockDim.x) {
            const int bt = idx >> logper;
            const int j = idx & (per - 1);
            const float2 *src = a + bt * n;
            float2 *dst = b + bt * n;
            const int k = j & (Ns - 1);
            const int idxD = (j - k) * R + k;
            if (R == 4) {
                float2 v0 = src[j], v1 = src[j + per], v2 = src[j + 2 * per], v3 = src[j + 3 * per];
                if (Ns > 1) {
                    v1 = cmulf(v1, tw[k * twstep]);
                    v2 = cmulf(v2, tw[2 * k * twstep]);
                    v3 = cmulf(v3, tw[3 * k * twstep]);
                }
                const float2 s02 = make_float2(v0.x + v2.x, v0.y + v2.y);
                const float2 d02 = make_float2(v0.x - v2.x, v0.y - v2.y);
                const float2 s13 = make_float2(v1.x + v3.x, v1.y + v3.y);
                const float2 d13 = make_float2(v1.x - v3.x, v1.y - v3.y);
                // i * d13 = (-d13.y, d13.x)
                dst[idxD] = make_float2(s02.x + s13.x, s02.y + s13.y);
                dst[idxD + Ns] = make_float2(d02.x - d13.y, d02.y + d13.x);
                dst[idxD + 2 * Ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
                dst[idxD + 3 * Ns] = make_float2(d02.x + d13.y, d02.y - d13.x);
            } else {
                float2 v0 = src[j], v1 = src[j + per];
                if (Ns > 1) v1 = cmulf(v1, tw[k * twstep]);
                dst[idxD] = make_float2(v0.x + v1.x, v0.y + v1.y);
                dst[idxD + Ns] = make_float2(v0.x - v1.x, v0.y - v1.y);
            }
        }
        __syncthreads();
        float2 *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// Pass A.  mode 0: from the gridder layout [c][gy][gx][P] with the
// PreIFFTUVCorrection reorder (ix = (gx - nx/2) mod nx, iy = (gy - ny/2) mod ny)
// into X[c][p][iy][kx].  mode 1: plain in-place rows of [c][y][x].
template <int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows(const float2 *in, float2 *out, int nx,
                                                         int lognx, int ny, int P,
                                                         const float2 *tw, int tw_mul)
{
    extern __shared__ float2 sm[];
    float2 *a = sm, *b = sm + nx;
    const int y = blockIdx.x, p = blockIdx.y, c = blockIdx.z;
    const int hx = nx >> 1;
    if (MODE == 0) {
        const float2 *src = in + (((size_t)c * ny + y) * nx) * P + p;
        for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
            const int gx = (ix + hx) & (nx - 1);
            a[ix] = src[(size_t)gx * P];
        }
    } else {
        const float2 *src = in + ((size_t)c * ny + y) * nx;
        for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) a[ix] = src[ix];
    }
    __syncthreads();
    float2 *r = lds_ifft(a, b, nx, lognx, 1, tw, tw_mul);
    float2 *dst;
    if (MODE == 0) {
        const int iy = (y - (ny >> 1)) & (ny - 1);
        dst = out + (((size_t)c * P + p) * ny + iy) * nx;
    } else {
        dst = out + ((size_t)c * ny + y) * nx;
    }
    for (int kx = threadIdx.x; kx < nx; kx += blockDim.x) dst[kx] = r[kx];
}

// Pass B1: for plane pl (= c*P + p), columns [x0, x0+cols), index n2:
// rows N2*n1 + n2 -> length-N1 FFT -> * exp(+2 pi i n2 k1 / ny) -> rows N2*k1 + n2.
__global__ __launch_bounds__(FFT_THREADS) void k_fft_cols1(float2 *X, int nx, int ny, int N1, int logN1,
                                                          int N2, const float2 *tw, int NT)
{
    extern __shared__ float2 sm[];
    const int cols = min(COLS, nx);
    float2 *a = sm, *b = sm + cols * N1;
    const int x0 = blockIdx.x * cols, n2 = blockIdx.y;
    float2 *plane = X + (size_t)blockIdx.z * nx * ny;
    for (int i = threadIdx.x; i < N1 * cols; i += blockDim.x) {
        const int n1 = i / cols, cx = i - n1 * cols;
        a[cx * N1 + n1] = plane[((size_t)N2 * n1 + n2) * nx + x0 + cx];
    }
    __syncthreads();
    float2 *r = lds_ifft(a, b, N1, logN1, cols, tw, NT / N1);
    const int mul = NT / ny;
    for (int i = threadIdx.x; i < N1 * cols; i += blockDim.x) {
        const int k1 = i / cols, cx = i - k1 * cols;
        float2 v = r[cx * N1 + k1];
        const int t = ((n2 * k1) & (ny - 1)) * mul;
        v = cmulf(v, tw[t]);
        plane[((size_t)N2 * k1 + n2) * nx + x0 + cx] = v;
    }
}

// Pass B2 (+ finalize): for channel c, index k1, columns [x0, x0+cols):
// rows N2*k1 + n2 -> length-N2 FFT -> ky = k1 + N1*k2.  mode 0 writes the
// finalized image, mode 1 writes the raw transform to out[c][ky][kx].
template <int MODE, int P>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_cols2(const float2 *X, int nx, int ny, int N1,
                                                          int N2, int logN2, const float2 *tw, int NT,
                                                          const double *norm, int poltype,
                                                          const float *corr_x, const float *corr_y,
                                                          float *image, float2 *out)
{
    extern __shared__ float2 sm[];
    const int cols = min(COLS, nx);
    float2 *a = sm, *b = sm + P * cols * N2;
    const int x0 = blockIdx.x * cols, k1 = blockIdx.y, c = blockIdx.z;
    for (int i = threadIdx.x; i < P * N2 * cols; i += blockDim.x) {
        const int p = i / (N2 * cols);
        const int rem = i - p * N2 * cols;
        const int n2 = rem / cols, cx = rem - n2 * cols;
        const float2 *plane = X + ((size_t)c * P + p) * nx * ny;
        a[(p * cols + cx) * N2 + n2] = plane[((size_t)N2 * k1 + n2) * nx + x0 + cx];
    }
    __syncthreads();
    float2 *r = lds_ifft(a, b, N2, logN2, P * cols, tw, NT / N2);
    if (MODE == 1) {
        for (int i = threadIdx.x; i < N2 * cols; i += blockDim.x) {
            const int k2 = i / cols, cx = i - k2 * cols;
            const int ky = k1 + N1 * k2;
            out[((size_t)c * ny + ky) * nx + x0 + cx] = r[cx * N2 + k2];
        }
        return;
    }
    double nrm[P];
#pragma unroll
    for (int p = 0; p < P; p++) nrm[p] = norm[c * P + p];
    float inv[P];
#pragma unroll
    for (int p = 0; p < P; p++) inv[p] = (float)(1.0 / nrm[p]);
    for (int i = threadIdx.x; i < N2 * cols; i += blockDim.x) {
        const int k2 = i / cols, cx = i - k2 * cols;
        const int ky = k1 + N1 * k2, kx = x0 + cx;
        const int jy = (ky + (ny >> 1)) & (ny - 1), jx = (kx + (nx >> 1)) & (nx - 1);
        const int row = ny - 1 - jy;  // m axis flipped for FITS
        float re[P], im[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            const float2 v = r[(p * cols + cx) * N2 + k2];
            re[p] = v.x * inv[p];
            im[p] = v.y * inv[p];
        }
        float o[P];
        if (P == 1) {
            o[0] = re[0];
        } else if (P == 2) {
            o[0] = (re[0] + re[1]) * 0.5f;
            o[1] = (re[0] - re[1]) * 0.5f;
        } else if (poltype == MTI_POL_LINEAR) {
            o[0] = (re[0] + re[3]) * 0.5f;
            o[1] = (re[0] - re[3]) * 0.5f;
            o[2] = (re[1] + re[2]) * 0.5f;
            o[3] = (im[1] - im[2]) * 0.5f;
        } else {
            o[0] = (re[0] + re[3]) * 0.5f;
            o[1] = (re[1] + re[2]) * 0.5f;
            o[2] = (im[2] - im[1]) * 0.5f;
            o[3] = (re[0] - re[3]) * 0.5f;
        }
        const float cf = corr_x[jx] * corr_y[row];
        float *img = image + (size_t)c * P * nx * ny;
#pragma unroll
        for (int q = 0; q < P; q++) img[((size_t)q * ny + row) * nx + jx] = o[q] * cf;
    }
}

// 1-D correction vectors: sinc(x)/psi(|x-n/2|/(n/2)), x = (i-n/2)/(n s)
__global__ void k_corr(float *corr, int n, int s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int h = n / 2;
    double nu = fabs((double)(i - h)) / (double)h;
    double tap = mti_dev::gridsf(nu);
    double x = (double)(i - h) / ((double)n * s);
    double sc = (x == 0.0) ? 1.0 : (M_PI * x) / sin(M_PI * x);
    corr[i] = (float)(sc / tap);
}

static bool is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }
static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

static mti_status twiddle_table(mti_ctx *ctx, int NT, const float2 **out, hipStream_t st)
{
    for (auto &t : ctx->twiddles)
        if (t.first == NT) {
            *out = t.second;
            return MTI_OK;
        }
    std::vector<float2> h(NT);
    for (int t = 0; t < NT; t++) {
        double a = 2.0 * M_PI * (double)t / (double)NT;
        h[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2 *d = nullptr;
    MTI_HIP(ctx, hipMalloc(&d, NT * sizeof(float2)));
    MTI_HIP(ctx, hipMemcpyAsync(d, h.data(), NT * sizeof(float2), hipMemcpyHostToDevice, st));
    MTI_HIP(ctx, hipStreamSynchronize(st));
    ctx->twiddles.push_back({NT, d});
    *out = d;
    return MTI_OK;
}

template <typename K>
static mti_status set_lds(mti_ctx *ctx, K kern, size_t bytes)
{
    if (bytes > 64 * 1024)
        MTI_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)bytes));
    return MTI_OK;
}

struct FftGeom {
    int nx, ny, lognx, N1, N2, logN1, logN2, NT, cols;
};

static FftGeom geom(int nx, int ny)
{
    FftGeom g;
    g.nx = nx;
    g.ny = ny;
    g.lognx = ilog2(nx);
    int m = ilog2(ny);
    g.logN1 = (m + 1) / 2;
    g.logN2 = m / 2;
    g.N1 = 1 << g.logN1;
    g.N2 = 1 << g.logN2;
    g.NT = std::max(nx, ny);
    g.cols = std::min(COLS, nx);
    return g;
}

static mti_status cols1(mti_ctx *ctx, const FftGeom &g, float2 *X, int planes, const float2 *tw,
                        hipStream_t st)
{
    size_t lds = (size_t)2 * g.cols * g.N1 * sizeof(float2);
    mti_status r = set_lds(ctx, k_fft_cols1, lds);
    if (r != MTI_OK) return r;
    dim3 gr((unsigned)(g.nx / g.cols), (unsigned)g.N2, (unsigned)planes);
    hipLaunchKernelGGL(k_fft_cols1, gr, dim3(FFT_THREADS), lds, st, X, g.nx, g.ny, g.N1, g.logN1, g.N2, tw,
                       g.NT);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <int MODE, int P>
static mti_status cols2(mti_ctx *ctx, const FftGeom &g, const float2 *X, int nch, const float2 *tw,
                        const double *norm, int poltype, const float *cx, const float *cy, float *image,
                        float2 *out, hipStream_t st)
{
    auto kern = k_fft_cols2<MODE, P>;
    size_t lds = (size_t)2 * P * g.cols * g.N2 * sizeof(float2);
    mti_status r = set_lds(ctx, kern, lds);
    if (r != MTI_OK) return r;
    dim3 gr((unsigned)(g.nx / g.cols), (unsigned)g.N1, (unsigned)nch);
    hipLaunchKernelGGL(kern, gr, dim3(FFT_THREADS), lds, st, X, g.nx, g.ny, g.N1, g.N2, g.logN2, tw, g.NT,
                       norm, poltype, cx, cy, image, out);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_fft2d_inverse(mti_ctx *ctx, float *data, int32_t nx, int32_t ny, int32_t nbatch,
                                        void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    if (!is_pow2(nx) || !is_pow2(ny) || nx > 8192 || ny > 8192)
        return set_error(ctx, MTI_ERR_UNSUPPORTED, "FFT sizes must be powers of two in [2, 8192]");
    MTI_CHECK(ctx, data && nbatch >= 1, "null data or nbatch < 1");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    FftGeom g = geom(nx, ny);
    const float2 *tw = nullptr;
    mti_status r = twiddle_table(ctx, g.NT, &tw, st);
    if (r != MTI_OK) return r;
    float2 *X = (float2 *)data;
    size_t lds = (size_t)2 * nx * sizeof(float2);
    r = set_lds(ctx, k_fft_rows<1>, lds);
    if (r != MTI_OK) return r;
    hipLaunchKernelGGL(k_fft_rows<1>, dim3((unsigned)ny, 1, (unsigned)nbatch), dim3(FFT_THREADS), lds, st, X,
                       X, nx, g.lognx, ny, 1, tw, g.NT / nx);
    MTI_HIP(ctx, hipGetLastError());
    r = cols1(ctx, g, X, nbatch, tw, st);
    if (r != MTI_OK) return r;
    void *w = nullptr;
    size_t bytes = (size_t)nbatch * nx * ny * sizeof(float2);
    r = scratch_get(ctx, SLOT_FFT, bytes, &w);
    if (r != MTI_OK) return r;
    r = cols2<1, 1>(ctx, g, X, nbatch, tw, nullptr, 0, nullptr, nullptr, nullptr, (float2 *)w, st);
    if (r != MTI_OK) return r;
    MTI_HIP(ctx, hipMemcpyAsync(data, w, bytes, hipMemcpyDeviceToDevice, st));
    return MTI_OK;
}

extern "C" mti_status mti_finalize(mti_ctx *ctx, const mti_image_params *p, int32_t npol, int32_t poltype,
                                   int32_t nchan, float *grids, const double *norm, float *image,
                                   void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, npol == 1 || npol == 2 || npol == 4, "npol must be 1, 2 or 4");
    MTI_CHECK(ctx, poltype == MTI_POL_LINEAR || poltype == MTI_POL_CIRCULAR, "bad PolarizationType");
    MTI_CHECK(ctx, nchan >= 1 && grids && norm && image, "null pointer or nchan < 1");
    if (!is_pow2(p->nx) || !is_pow2(p->ny) || p->nx > 8192 || p->ny > 8192)
        return set_error(ctx, MTI_ERR_UNSUPPORTED, "image.nx/ny must be powers of two in [2, 8192]");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    FftGeom g = geom(p->nx, p->ny);
    const float2 *tw = nullptr;
    r = twiddle_table(ctx, g.NT, &tw, st);
    if (r != MTI_OK) return r;
    // scratch: transform buffer for a batch of channels + correction vectors
    const size_t plane_bytes = (size_t)p->nx * p->ny * sizeof(float2);
    int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nchan, (size_t(8) << 30) / (plane_bytes * npol)));
    void *w = nullptr;
    r = scratch_get(ctx, SLOT_FFT, plane_bytes * npol * batch + (size_t)(p->nx + p->ny + 128) * 4, &w);
    if (r != MTI_OK) return r;
    float2 *X = (float2 *)w;
    float *corr_x = (float *)((char *)w + plane_bytes * npol * batch);
    float *corr_y = corr_x + p->nx + 64;
    hipLaunchKernelGGL(k_corr, dim3((unsigned)ceil_div(p->nx, 256)), dim3(256), 0, st, corr_x, p->nx, p->sampling);
    hipLaunchKernelGGL(k_corr, dim3((unsigned)ceil_div(p->ny, 256)), dim3(256), 0, st, corr_y, p->ny, p->sampling);
    MTI_HIP(ctx, hipGetLastError());
    size_t lds = (size_t)2 * p->nx * sizeof(float2);
    r = set_lds(ctx, k_fft_rows<0>, lds);
    if (r != MTI_OK) return r;
    for (int c0 = 0; c0 < nchan; c0 += batch) {
        const int nb = std::min(batch, nchan - c0);
        const float2 *G = (const float2 *)grids + (size_t)c0 * p->nx * p->ny * npol;
        hipLaunchKernelGGL(k_fft_rows<0>, dim3((unsigned)p->ny, (unsigned)npol, (unsigned)nb), dim3(FFT_THREADS),
                           lds, st, G, X, p->nx, g.lognx, p->ny, npol, tw, g.NT / p->nx);
        MTI_HIP(ctx, hipGetLastError());
        r = cols1(ctx, g, X, nb * npol, tw, st);
        if (r != MTI_OK) return r;
        float *img = image + (size_t)c0 * npol * p->nx * p->ny;
        const double *nr = norm + (size_t)c0 * npol;
        if (npol == 1)
            r = cols2<0, 1>(ctx, g, X, nb, tw, nr, poltype, corr_x, corr_y, img, nullptr, st);
        else if (npol == 2)
            r = cols2<0, 2>(ctx, g, X, nb, tw, nr, poltype, corr_x, corr_y, img, nullptr, st);
        else
            r = cols2<0, 4>(ctx, g, X, nb, tw, nr, poltype, corr_x, corr_y, img, nullptr, st);
        if (r != MTI_OK) return r;
    }
    return MTI_OK;
}
