// mti_wkernel.hip -- convolution-function tables on the device.
//
// W-projection kernels (cfg/WProjectionConvFunction.cpp:186-291,
// cfg/WPlaneConvFunctionCalculator.cpp:31-148):
//   screen_k(c) = psi(|cy|/h) psi(|cx|/h) * exp(i pi 2 w_k (sqrt(1-l^2-m^2)-1))
//     on the S_t x S_t centre of an (S_t s)^2 zero-padded plane
//     (op/WProjPreFFTconvFunc.cu:34-105, op/CasaGridSF.cu:151-173);
//   C_k(q) = FFT(screen_k) fftshifted, divided by the w=0 DC value
//     (MultiFFT2D + ZeroToCentre2DShift + SingleMultiply);
//   support from the outermost centre-row/column sample with |C|^2 > 1e-6
//     (op/ConvFunctionSupportFind.cu:28-77), clamped to S_t-2;
//   table = S*T x S*T window reordered to [oy][ox][sy][sx]
//     (op/ConvFunctionReOrganize.cu:30-71), sums of Re per sub-kernel
//     (op/SumConvolutionFunctions.cu:28-56).
//
// MI355X design: the FFT of a zero-padded plane followed by a crop is
// evaluated as a separable DFT restricted to the window actually kept (the
// same numbers; the padded FFT is never materialised): two tiled FP64 complex
// matrix products per plane, all planes of a batch in one launch, with the
// DFT matrix generated from an FP64 twiddle table while tiles are staged.
// The support search uses the identity "centre row of the 2-D DFT = 1-D DFT
// of the column sums".  One host sync (support read-back) per build.
#include "mti_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

namespace {

struct ScreenGeom {
    int St, h, s, L;
    double linc, minc;
};

// twiddle table exp(-2 pi i m / L), FP64
__global__ void k_twiddle(double2 *tw, int L)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)L, &sn, &cs);
    tw[m] = make_double2(cs, -sn);
}

// screens for planes [k0, k0+nb): scr[b][iy][ix], c = i - h
__global__ void k_screen(double2 *scr, ScreenGeom g, double dw, int k0)
{
    const int b = blockIdx.z;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= g.St || iy >= g.St) return;
    const int k = k0 + b;
    const double w = (double)k * (double)k * dw;
    const int cx = ix - g.h, cy = iy - g.h;
    double taper = mti_dev::gridsf(fabs((double)cy) / (double)g.h) *
                   mti_dev::gridsf(fabs((double)cx) / (double)g.h);
    double msq = g.minc * (double)cy;
    msq *= msq;
    double lsq = g.linc * (double)cx;
    lsq *= lsq;
    double rsq = lsq + msq;
    double2 v;
    if (rsq < 1.0) {
        double ph = 2.0 * w * (sqrt(1.0 - rsq) - 1.0);
        double sn, cs;
        sincospi(ph, &sn, &cs);
        v = make_double2(taper * cs, taper * sn);
    } else {
        v = make_double2(taper, 0.0);
    }
    scr[((size_t)b * g.St + iy) * g.St + ix] = v;
}

__device__ __forceinline__ int mod_pos(long long a, int m)
{
    long long r = a % m;
    return (int)(r < 0 ? r + m : r);
}

// Projections of each screen of a batch: rowp[ix] = sum over iy of S[iy][ix]
// and colp[iy] = sum over ix of S[iy][ix], each summed in ascending index
// order (the oracle's order, so the support decision sees the same values).
// Blocks (plane, y): y < PJ_RB(St) takes 64 rows' sums, staging 64 x 64 tiles
// in LDS so the reads run along rows; the others take 256 columns' sums, a
// thread a column walking the rows.  Many blocks a plane and reads that run
// along rows, where one block a plane summing rows thread by thread read
// 16 bytes a lane at a stride of St x 16 bytes.
constexpr int PJ_THREADS = 256, PJ_T = 64;
__host__ __device__ inline int pj_row_blocks(int St) { return (St + PJ_T - 1) / PJ_T; }
__host__ __device__ inline int pj_col_blocks(int St) { return (St + PJ_THREADS - 1) / PJ_THREADS; }
__global__ __launch_bounds__(PJ_THREADS) void k_proj(const double2 *scr, int St, double2 *proj /* [nb][2][St] */)
{
    __shared__ double2 tile[PJ_T][PJ_T + 1];
    const int b = blockIdx.x, rb = pj_row_blocks(St);
    const double2 *S = scr + (size_t)b * St * St;
    double2 *rowp = proj + (size_t)b * 2 * St;
    double2 *colp = rowp + St;
    if ((int)blockIdx.y >= rb) {
        const int ix = ((int)blockIdx.y - rb) * PJ_THREADS + threadIdx.x;
        if (ix >= St) return;
        double2 r = make_double2(0, 0);
        for (int iy = 0; iy < St; iy++) {
            const double2 a = S[(size_t)iy * St + ix];
            r.x += a.x;
            r.y += a.y;
        }
        rowp[ix] = r;
        return;
    }
    const int y0 = (int)blockIdx.y * PJ_T, t = threadIdx.x;
    double2 c = make_double2(0, 0);  // thread t < 64: row y0 + t
    for (int x0 = 0; x0 < St; x0 += PJ_T) {
        for (int e = t; e < PJ_T * PJ_T; e += PJ_THREADS) {
            const int ry = e / PJ_T, rx = e % PJ_T;
            tile[ry][rx] = (y0 + ry < St && x0 + rx < St) ? S[(size_t)(y0 + ry) * St + x0 + rx] : make_double2(0, 0);
        }
        __syncthreads();
        if (t < PJ_T) {
            const int n = min(PJ_T, St - x0);
            for (int rx = 0; rx < n; rx++) {
                c.x += tile[t][rx].x;
                c.y += tile[t][rx].y;
            }
        }
        __syncthreads();
    }
    if (t < PJ_T && y0 + t < St) colp[y0 + t] = c;
}

// Support find on the centre row / column (the DFTs of the projections
// k_proj wrote): the support comes from the largest q whose row or column
// sample exceeds the threshold, so every q is tried at once -- thread
// (blockIdx.x, threadIdx.x) of plane blockIdx.y takes q = qmax - 256 x - t
// -- and found[k0 + plane] keeps the largest hit (filled with -1 before).
// Each sum runs in ascending c as in the oracle, so the threshold sees the
// same values.
constexpr int SUP_THREADS = 256;
__global__ __launch_bounds__(SUP_THREADS) void k_support(ScreenGeom g, const double2 *tw, double inv_div,
                                                         int k0, int *found,
                                                         const double2 *proj /* [nb][2][St] */)
{
    const int b = blockIdx.y;
    const int qmax = g.L - 1 - g.L / 2;
    const int q = qmax - (int)blockIdx.x * SUP_THREADS - (int)threadIdx.x;
    if (q < 0) return;
    const double2 *rowp = proj + (size_t)b * 2 * g.St;
    const double2 *colp = rowp + g.St;
    for (int pass = 0; pass < 2; pass++) {
        const double2 *pr = pass ? colp : rowp;
        int m = mod_pos(-(long long)g.h * q, g.L);
        double re = 0, im = 0;
        // unrolled so the twiddle gathers go out ahead of the sums; the
        // sums stay one chain in ascending c (the oracle's order)
#pragma unroll 8
        for (int c = 0; c < g.St; c++) {
            double2 t = tw[m];
            double2 v = pr[c];
            re += v.x * t.x - v.y * t.y;
            im += v.x * t.y + v.y * t.x;
            m += q;
            if (m >= g.L) m -= g.L;
        }
        re *= inv_div;
        im *= inv_div;
        if (re * re + im * im > 1e-6) {
            atomicMax(&found[k0 + b], q);
            return;
        }
    }
}

// found[k] = -1 for the planes of a build, before the support search.
__global__ void k_support_init(int *found, int W)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < W) found[k] = -1;
}

// Support of each plane from its largest hit (odd, at most St - 2; 0: none).
__global__ void k_support_size(ScreenGeom g, const int *found, int W, int *support_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= W) return;
    int S_ = 0;
    if (found[k] >= 0) {
        S_ = ((found[k] + (g.s + 1) / 2) / g.s) * 2 + 1;
        if (S_ > g.St - 2) S_ = g.St - 2;
    }
    support_out[k] = S_;
}

template <int PREC>
__device__ __forceinline__ void store_sample(void *table, int64_t idx, double re, double im)
{
    if (PREC == MTI_TABLE_F32) {
        ((float2 *)table)[idx] = make_float2((float)re, (float)im);
    } else if (PREC == MTI_TABLE_F16) {
        // direct FP64 -> FP16 rounding (no intermediate FP32 rounding)
        _Float16 *h = (_Float16 *)table;
        h[2 * idx] = (_Float16)re;
        h[2 * idx + 1] = (_Float16)im;
    } else {
        __hip_bfloat16 a = __double2bfloat16(re), b = __double2bfloat16(im);
        ((__hip_bfloat162 *)table)[idx] = __hip_bfloat162{a, b};
    }
}

template <int PREC>
__device__ __forceinline__ float2 load_sample(const void *table, int64_t idx)
{
    if (PREC == MTI_TABLE_F32) return ((const float2 *)table)[idx];
    if (PREC == MTI_TABLE_F16) return __half22float2(((const __half2 *)table)[idx]);
    __hip_bfloat162 v = ((const __hip_bfloat162 *)table)[idx];
    return make_float2(__bfloat162float(v.x), __bfloat162float(v.y));
}

// taper-only table (mode=normal): op/CasaGridSF.cu:107-149 + reorganize
template <int PREC>
__global__ void k_taper_table(int S, int s, int T, void *table)
{
    const int64_t S2 = (int64_t)S * S;
    const int64_t n = (int64_t)T * T * S2;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int o = (int)(e / S2);
    const int t = (int)(e - (int64_t)o * S2);
    const int oy = o / T, ox = o - oy * T;
    const int sy = t / S, sx = t - sy * S;
    const int dim = (s % 2) ? S * s : S * s + 1;
    const int c = dim / 2, fz = 3 * s;
    auto g1 = [&](int i) -> double {
        int a = abs(i - c);
        if (a >= fz) return 0.0;
        double nu = (double)a / (double)fz;
        return (1.0 - nu * nu) * mti_dev::gridsf(nu);
    };
    store_sample<PREC>(table, e, g1(sy * s + oy) * g1(sx * s + ox), 0.0);
}

template <int PREC>
__global__ void k_sums(const void *table, const int2 *convinfo, int nplanes, int T2, float *sums)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nplanes * T2) return;
    const int k = id / T2, o = id - k * T2;
    const int S = convinfo[k].x;
    const int64_t S2 = (int64_t)S * S;
    const int64_t b = (int64_t)convinfo[k].y + (int64_t)o * S2;
    double acc = 0;
    for (int64_t i = 0; i < S2; i++) acc += (double)load_sample<PREC>(table, b + i).x;
    sums[id] = (float)acc;
}

// The two windowed-DFT stages as tiled complex FP64 matrix products on the
// FP64 matrix cores, all planes of a batch in one launch (blockIdx.z =
// plane).  E[c][j] = tw[((c - h) q_j) mod L] is generated while a tile is
// staged.
//   stage 1: A[b][iy][j] = sum_cx scr[b][iy][cx] E[cx][j]            (St x Wk)
//   stage 2: T[b][jy][jx] = sum_iy E[iy][jy] A[b][iy][jx]            (Wk x Wk)
// Plane b uses the centred window j < Wk(S_b) of the widest window Wmax
// (offset (Wmax - Wk)/2), qmin = -(S/2) s - s/2.
//
// A block of 4 waves computes a GT x GT output tile, wave (wy, wx) its
// 32 x 32 quarter as 2 x 2 blocks of 16 x 16, with v_mfma_f64_16x16x4f64:
// operand a of lane l is X[l % 16][l / 16], operand b is Y[l / 16][l % 16],
// and result register r of lane l is C[4 r + l / 16][l % 16].  A complex
// product takes four real ones: Cr += Xr Yr - Xi Yi, Ci += Xr Yi + Xi Yr.
// The X and Y tiles of a k chunk are staged in LDS as complex FP64, so each
// lane reads one complex element per operand and k step.
constexpr int GT = 64;   // output tile
constexpr int GK = 16;   // k chunk (4 matrix steps of 4)

typedef double mti_f64x4 __attribute__((ext_vector_type(4)));

// Twiddle index ((c - h) q) mod L advanced by c += step: idx + (step q mod L),
// wrapped once (both terms below L).
__device__ __forceinline__ int tw_step(int idx, int inc, int L)
{
    idx += inc;
    return idx >= L ? idx - L : idx;
}

template <int STAGE>
__global__ __launch_bounds__(256) void k_wdft(const double2 *src, ScreenGeom g, const double2 *tw,
                                              const int *support, int k0, int Wmax, int T, double2 *dst)
{
    __shared__ double2 Xs[GT][GK + 1];
    __shared__ double2 Ys[GK][GT + 1];
    const int b = blockIdx.z;
    const int S = support[k0 + b];
    if (S == 0) return;
    const int Wk = (S - 1) * g.s + T;
    const int qmin = -(S / 2) * g.s - g.s / 2;
    const int M = STAGE == 1 ? g.St : Wk, N = Wk, K = g.St;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    if (m0 >= M || n0 >= N) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wy = (wave >> 1) * 32, wx = (wave & 1) * 32;  // the wave's quarter of the tile
    const int li = lane & 15, lk = lane >> 4;
    const double2 *Sb = src + (size_t)b * g.St * (STAGE == 1 ? g.St : Wmax);
    // A thread stages elements t = 0..3 of each k chunk: X[(tid >> 4) + 16 t][tid & 15]
    // and Y[(tid >> 6) + 4 t][tid & 63].  The generated operand (Y in stage 1,
    // X in stage 2) comes from the twiddle table through running indices
    // (one wrapped add per element instead of a modulo).
    const int xm = tid >> 4, xk = tid & 15, yk = tid >> 6, yn = tid & 63;
    // stage 1: ti[0] (element 0), toff[t] = 4 t q mod L, tinc[0] = 16 q mod L;
    // stage 2: ti[t] and tinc[t] = 16 q_t mod L per element t
    int ti[4], tinc[4], toff[4];
    if (STAGE == 1) {
        const int q = qmin + n0 + yn;
        ti[0] = mod_pos((long long)(yk - g.h) * q, g.L);
#pragma unroll
        for (int t = 0; t < 4; t++) toff[t] = mod_pos(4LL * t * q, g.L);
        tinc[0] = mod_pos(16LL * q, g.L);
    } else {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int q = qmin + m0 + xm + 16 * t;
            ti[t] = mod_pos((long long)(xk - g.h) * q, g.L);
            tinc[t] = mod_pos(16LL * q, g.L);
        }
    }
    double2 xv[4], yv[4];
    auto fetch = [&](int k0c) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int m = m0 + xm + 16 * t, kx = k0c + xk;
            const int ky = k0c + yk + 4 * t, n = n0 + yn;
            if (STAGE == 1) {
                xv[t] = m < M && kx < K ? Sb[(size_t)m * g.St + kx] : make_double2(0.0, 0.0);
                yv[t] = n < N && ky < K ? tw[tw_step(ti[0], toff[t], g.L)] : make_double2(0.0, 0.0);
            } else {
                xv[t] = m < M && kx < K ? tw[ti[t]] : make_double2(0.0, 0.0);
                yv[t] = n < N && ky < K ? Sb[(size_t)ky * Wmax + n] : make_double2(0.0, 0.0);
            }
        }
        // the next chunk is 16 further along k
        if (STAGE == 1) {
            ti[0] = tw_step(ti[0], tinc[0], g.L);
        } else {
#pragma unroll
            for (int t = 0; t < 4; t++) ti[t] = tw_step(ti[t], tinc[t], g.L);
        }
    };
    mti_f64x4 cr[2][2], ci[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) cr[i][j] = ci[i][j] = mti_f64x4{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int k0c = 0; k0c < K; k0c += GK) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            Xs[xm + 16 * t][xk] = xv[t];
            Ys[yk + 4 * t][yn] = yv[t];
        }
        __syncthreads();
        if (k0c + GK < K) fetch(k0c + GK);  // in flight during the products
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            double2 x[2], y[2];
#pragma unroll
            for (int i = 0; i < 2; i++) x[i] = Xs[wy + 16 * i + li][ks + lk];
#pragma unroll
            for (int j = 0; j < 2; j++) y[j] = Ys[ks + lk][wx + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i].x, y[j].x, cr[i][j], 0, 0, 0);
                    cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-x[i].y, y[j].y, cr[i][j], 0, 0, 0);
                    ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i].x, y[j].y, ci[i][j], 0, 0, 0);
                    ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i].y, y[j].x, ci[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wy + 16 * i + 4 * r + lk, n = n0 + wx + 16 * j + li;
                if (m < M && n < N)
                    dst[((size_t)b * (STAGE == 1 ? g.St : Wmax) + m) * Wmax + n] =
                        make_double2(cr[i][j][r], ci[i][j][r]);
            }
}

// Stage-2 products (T[jy][jx], jy, jx < Wk) scaled by 1/divisor into the
// [oy][ox][sy][sx] layout of plane k0 + b.
template <int PREC>
__global__ void k_wstore(const double2 *Tm, ScreenGeom g, const int *support, const int64_t *base, int k0,
                         int Wmax, int T, double inv_div, void *table)
{
    const int b = blockIdx.y;
    const int S = support[k0 + b];
    const int64_t S2 = (int64_t)S * S;
    const int64_t n = (int64_t)T * T * S2;
    const double2 *Tb = Tm + (size_t)b * Wmax * Wmax;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(e / S2);
        const int t = (int)(e - (int64_t)o * S2);
        const int oy = o / T, ox = o - oy * T;
        const int sy = t / S, sx = t - sy * S;
        const double2 v = Tb[(size_t)(sy * g.s + oy) * Wmax + sx * g.s + ox];
        store_sample<PREC>(table, base[k0 + b] + e, v.x * inv_div, v.y * inv_div);
    }
}

double host_gridsf(double nu)
{
    static const double P[2][5] = {
        {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
        {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
    static const double Q[2][3] = {{1.0, 8.212018e-1, 2.078043e-1},
                                   {1.0, 9.599102e-1, 2.918724e-1}};
    int part = (nu < 0.75) ? 0 : 1;
    double d = nu * nu - (part ? 1.0 : 0.5625);
    double top = P[part][0], bot = Q[part][0], dk = 1.0;
    for (int k = 1; k < 5; k++) {
        dk *= d;
        top += P[part][k] * dk;
    }
    dk = 1.0;
    for (int k = 1; k < 3; k++) {
        dk *= d;
        bot += Q[part][k] * dk;
    }
    return bot != 0.0 ? top / bot : 0.0;
}

size_t elem_bytes(int prec) { return prec == MTI_TABLE_F32 ? 8 : 4; }

}  // namespace

using namespace mti;

// The build's outputs in one device allocation: the sample table (total
// samples of eb bytes), the {support, base} pairs and the sub-kernel sums.
static mti_status alloc_tables(mti_ctx *ctx, mti_kernels *K, size_t eb)
{
    const int W = K->info.nplanes, T2 = K->info.T * K->info.T;
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t tb = up(std::max<size_t>(1, (size_t)K->info.total) * eb), cb = up((size_t)W * sizeof(int2));
    const size_t sb = up((size_t)W * T2 * sizeof(float));
    MTI_HIP(ctx, hipMalloc(&K->d_mem, tb + cb + sb));
    K->d_table = K->d_mem;
    K->d_convinfo = (int2 *)((char *)K->d_mem + tb);
    K->d_sums = (float *)((char *)K->d_mem + tb + cb);
    return MTI_OK;
}

static mti_status build_impl(mti_ctx *ctx, const mti_image_params *p, int precision,
                             mti_kernels *K, hipStream_t st)
{
    const int s = p->sampling;
    const int T = (s % 2) ? s : s + 1;
    K->info.sampling = s;
    K->info.T = T;
    K->info.precision = precision;
    K->info.w_increment = w_increment(p);
    const size_t eb = elem_bytes(precision);

    if (p->mode == MTI_MODE_NORMAL) {
        const int S = p->taper_support;
        K->info.nplanes = 1;
        K->info.trial_support = S;
        K->info.max_support = S;
        K->info.divisor = 1.0;
        K->support = {S};
        K->base = {0};
        K->info.total = (int64_t)T * T * S * S;
        // the gridder addresses samples with 32-bit byte offsets, and offsets
        // from 2^32 - 16 up must lie beyond the table
        if ((uint64_t)K->info.total * (uint64_t)eb > (uint64_t(1) << 32) - 16)
            return set_error(ctx, MTI_ERR_UNSUPPORTED, "convolution table above 4 GiB");
    } else {
        const int W = p->wplanes;
        int mn = std::min(p->nx, p->ny);
        int St = mn / s;
        if (St % 2 == 0) St++;
        ScreenGeom g;
        g.St = St;
        g.h = St / 2;
        g.s = s;
        g.L = St * s;
        g.linc = (p->cell_l * (double)p->nx) / (double)St;
        g.minc = (p->cell_m * (double)p->ny) / (double)St;
        K->info.nplanes = W;
        K->info.trial_support = St;
        // divisor: DC term of the w=0 screen = sum of the lm taper
        double div = 0;
        for (int cy = -g.h; cy <= g.h; cy++) {
            double ty = host_gridsf(std::fabs((double)cy) / (double)g.h);
            for (int cx = -g.h; cx <= g.h; cx++)
                div += ty * host_gridsf(std::fabs((double)cx) / (double)g.h);
        }
        K->info.divisor = div;
        const double inv = 1.0 / div;

        // planes per batch bounded by ~2 GB of FP64 screens; when one batch
        // holds every plane, the DFT stage reuses the screens of the support
        // search instead of computing them again
        const size_t plane_bytes = (size_t)St * St * sizeof(double2);
        int nb = (int)std::max<size_t>(1, std::min<size_t>((size_t)W, (size_t(2) << 30) / plane_bytes));
        const size_t tw_bytes = (size_t)g.L * sizeof(double2);
        const size_t proj_bytes = (size_t)nb * 2 * St * sizeof(double2);
        void *w0 = nullptr;
        const size_t sup_bytes = ((size_t)W * sizeof(int) + 255) / 256 * 256;
        mti_status r = scratch_get(ctx, SLOT_SORTED,
                                   tw_bytes + (size_t)nb * plane_bytes + proj_bytes + 2 * sup_bytes + 512, &w0);
        if (r != MTI_OK) return r;
        char *wp = (char *)w0;
        double2 *d_tw = (double2 *)wp;
        wp += (tw_bytes + 255) / 256 * 256;
        double2 *d_scr = (double2 *)wp;
        wp += (size_t)nb * plane_bytes;
        double2 *d_proj = (double2 *)wp;
        wp += proj_bytes;
        wp = (char *)w0 + (((size_t)(wp - (char *)w0) + 255) / 256 * 256);
        int *d_sup = (int *)wp;
        int *d_found = (int *)(wp + sup_bytes);

        hipLaunchKernelGGL(k_twiddle, dim3((unsigned)ceil_div(g.L, 256)), dim3(256), 0, st, d_tw, g.L);
        hipLaunchKernelGGL(k_support_init, dim3((unsigned)ceil_div(W, 256)), dim3(256), 0, st, d_found, W);
        MTI_HIP(ctx, hipGetLastError());
        const int qchunks = ceil_div(g.L - g.L / 2, SUP_THREADS);  // q = 0 .. L - 1 - L/2
        dim3 sb(16, 16);
        for (int k0 = 0; k0 < W; k0 += nb) {
            int n = std::min(nb, W - k0);
            dim3 sg((unsigned)ceil_div(St, 16), (unsigned)ceil_div(St, 16), (unsigned)n);
            hipLaunchKernelGGL(k_screen, sg, sb, 0, st, d_scr, g, K->info.w_increment, k0);
            MTI_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_proj, dim3((unsigned)n, (unsigned)(pj_row_blocks(St) + pj_col_blocks(St))),
                               dim3(PJ_THREADS), 0, st, d_scr, St, d_proj);
            MTI_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_support, dim3((unsigned)qchunks, (unsigned)n), dim3(SUP_THREADS), 0, st, g, d_tw,
                               inv, k0, d_found, d_proj);
            MTI_HIP(ctx, hipGetLastError());
        }
        hipLaunchKernelGGL(k_support_size, dim3((unsigned)ceil_div(W, 256)), dim3(256), 0, st, g, d_found, W, d_sup);
        MTI_HIP(ctx, hipGetLastError());
        K->support.assign(W, 0);
        MTI_HIP(ctx, hipMemcpyAsync(K->support.data(), d_sup, W * sizeof(int),
                                    hipMemcpyDeviceToHost, st));
        MTI_HIP(ctx, hipStreamSynchronize(st));
        K->base.assign(W, 0);
        int64_t tot = 0;
        int smax = 0;
        for (int k = 0; k < W; k++) {
            K->base[k] = tot;
            tot += (int64_t)T * T * K->support[k] * K->support[k];
            smax = std::max(smax, K->support[k]);
        }
        K->info.total = tot;
        K->info.max_support = smax;
        // the gridder addresses samples with 32-bit byte offsets, and offsets
        // from 2^32 - 16 up must lie beyond the table
        if ((uint64_t)tot * (uint64_t)eb > (uint64_t(1) << 32) - 16)
            return set_error(ctx, MTI_ERR_UNSUPPORTED, "convolution table above 4 GiB");
        if ((r = alloc_tables(ctx, K, eb)) != MTI_OK) return r;

        // windowed DFT as two batched complex products (all planes of a batch)
        const int Wmax = (std::max(smax, 1) - 1) * s + T;
        const size_t a_plane = (size_t)St * Wmax * sizeof(double2), t_plane = (size_t)Wmax * Wmax * sizeof(double2);
        void *aw = nullptr;
        r = scratch_get(ctx, SLOT_ITEMS, (size_t)nb * (a_plane + t_plane) + (size_t)W * sizeof(int64_t) + 256, &aw);
        if (r != MTI_OK) return r;
        double2 *d_A = (double2 *)aw;
        double2 *d_T = (double2 *)((char *)aw + (size_t)nb * a_plane);
        // plane bases, staged through the host vector (the synchronisation at
        // the end of the build keeps it alive until the copy is done)
        int64_t *d_base = (int64_t *)((char *)aw + (size_t)nb * (a_plane + t_plane));
        MTI_HIP(ctx, hipMemcpyAsync(d_base, K->base.data(), (size_t)W * sizeof(int64_t), hipMemcpyHostToDevice, st));
        for (int k0 = 0; k0 < W; k0 += nb) {
            int n = std::min(nb, W - k0);
            if (nb < W) {
                dim3 sg((unsigned)ceil_div(St, 16), (unsigned)ceil_div(St, 16), (unsigned)n);
                hipLaunchKernelGGL(k_screen, sg, sb, 0, st, d_scr, g, K->info.w_increment, k0);
            }
            dim3 g1((unsigned)ceil_div(Wmax, GT), (unsigned)ceil_div(St, GT), (unsigned)n);
            hipLaunchKernelGGL(k_wdft<1>, g1, dim3(256), 0, st, d_scr, g, d_tw, d_sup, k0, Wmax, T, d_A);
            dim3 g2((unsigned)ceil_div(Wmax, GT), (unsigned)ceil_div(Wmax, GT), (unsigned)n);
            hipLaunchKernelGGL(k_wdft<2>, g2, dim3(256), 0, st, d_A, g, d_tw, d_sup, k0, Wmax, T, d_T);
            dim3 g3(64, (unsigned)n);
            if (precision == MTI_TABLE_F32)
                hipLaunchKernelGGL(k_wstore<MTI_TABLE_F32>, g3, dim3(256), 0, st, d_T, g, d_sup, d_base, k0, Wmax, T,
                                   inv, K->d_table);
            else if (precision == MTI_TABLE_F16)
                hipLaunchKernelGGL(k_wstore<MTI_TABLE_F16>, g3, dim3(256), 0, st, d_T, g, d_sup, d_base, k0, Wmax, T,
                                   inv, K->d_table);
            else
                hipLaunchKernelGGL(k_wstore<MTI_TABLE_BF16>, g3, dim3(256), 0, st, d_T, g, d_sup, d_base, k0, Wmax,
                                   T, inv, K->d_table);
            MTI_HIP(ctx, hipGetLastError());
        }
    }

    const int W = K->info.nplanes;
    if (!K->d_mem) {
        mti_status r = alloc_tables(ctx, K, eb);
        if (r != MTI_OK) return r;
    }
    std::vector<int2> ci(W);
    for (int k = 0; k < W; k++) ci[k] = make_int2(K->support[k], (int)K->base[k]);
    MTI_HIP(ctx, hipMemcpyAsync(K->d_convinfo, ci.data(), W * sizeof(int2), hipMemcpyHostToDevice, st));
    if (p->mode == MTI_MODE_NORMAL) {
        const int S = p->taper_support;
        dim3 gr((unsigned)ceil_div(K->info.total, 256)), bl(256);
        if (precision == MTI_TABLE_F32)
            hipLaunchKernelGGL(k_taper_table<MTI_TABLE_F32>, gr, bl, 0, st, S, s, T, K->d_table);
        else if (precision == MTI_TABLE_F16)
            hipLaunchKernelGGL(k_taper_table<MTI_TABLE_F16>, gr, bl, 0, st, S, s, T, K->d_table);
        else
            hipLaunchKernelGGL(k_taper_table<MTI_TABLE_BF16>, gr, bl, 0, st, S, s, T, K->d_table);
        MTI_HIP(ctx, hipGetLastError());
    }
    const int T2 = T * T;
    dim3 gr((unsigned)ceil_div((int64_t)W * T2, 256)), bl(256);
    if (precision == MTI_TABLE_F32)
        hipLaunchKernelGGL(k_sums<MTI_TABLE_F32>, gr, bl, 0, st, K->d_table, K->d_convinfo, W, T2, K->d_sums);
    else if (precision == MTI_TABLE_F16)
        hipLaunchKernelGGL(k_sums<MTI_TABLE_F16>, gr, bl, 0, st, K->d_table, K->d_convinfo, W, T2, K->d_sums);
    else
        hipLaunchKernelGGL(k_sums<MTI_TABLE_BF16>, gr, bl, 0, st, K->d_table, K->d_convinfo, W, T2, K->d_sums);
    MTI_HIP(ctx, hipGetLastError());
    MTI_HIP(ctx, hipStreamSynchronize(st));
    return MTI_OK;
}

extern "C" mti_status mti_kernels_build(mti_ctx *ctx, const mti_image_params *p, int32_t precision,
                                        mti_kernels **out, void *stream)
{
    if (!ctx || !out) return MTI_ERR_INVALID;
    *out = nullptr;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, precision >= MTI_TABLE_F32 && precision <= MTI_TABLE_BF16, "bad table precision");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    mti_kernels *K = new mti_kernels();
    K->device = ctx->device;
    {
        StageTimer tm(ctx, MTI_STAGE_KERNELS, to_stream(stream));
        r = build_impl(ctx, p, precision, K, to_stream(stream));
    }
    if (r != MTI_OK) {
        mti_kernels_free(K);
        return r;
    }
    *out = K;
    return MTI_OK;
}

extern "C" mti_status mti_kernels_info(const mti_kernels *k, mti_kernel_info *info)
{
    if (!k || !info) return MTI_ERR_INVALID;
    *info = k->info;
    return MTI_OK;
}

extern "C" mti_status mti_kernels_supports(const mti_kernels *k, int32_t *support, int64_t *base)
{
    if (!k) return MTI_ERR_INVALID;
    for (int i = 0; i < k->info.nplanes; i++) {
        if (support) support[i] = k->support[i];
        if (base) base[i] = k->base[i];
    }
    return MTI_OK;
}

extern "C" const void *mti_kernels_table(const mti_kernels *k) { return k ? k->d_table : nullptr; }
extern "C" const float *mti_kernels_sums(const mti_kernels *k) { return k ? k->d_sums : nullptr; }
extern "C" const int32_t *mti_kernels_position(const mti_kernels *k)
{
    return k ? (const int32_t *)k->d_convinfo : nullptr;
}

extern "C" void mti_kernels_free(mti_kernels *k)
{
    if (!k) return;
    (void)hipSetDevice(k->device);
    if (k->d_mem) (void)hipFree(k->d_mem);
    delete k;
}
