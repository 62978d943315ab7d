// mti_fft.hip -- 2-D inverse FFT and the image finalizer.
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
// MI355X design.  Every Stokes image is real, so two of them share one
// complex FP64 transform ("packed planes": the Hermitian parts of two Stokes
// combinations of correlation planes, one as the real and one as the
// imaginary part); a full-polarisation channel takes two transforms, and two
// Stokes-I channels share one.  Per packed plane:
//   A  row transforms on 4096/8192-element LDS blocks (radix-16 Stockham),
//      rows y and -y together, fused with the uv-plane reorder/ifftshift,
//      normalisation and Stokes coefficients;     grid -> X[pack][v][u]
//   B  columns whole in one block when ny <= 256, else the four-step split
//      B1 (length N1 over rows strided by N2, times the inter-step twiddle,
//      in place) and B2 (length N2 over N2 contiguous rows); the last pass is
//      fused with fftshift + m flip + taper and sinc correction and writes
//      the real and imaginary parts as two f32 images.
// Other sizes (a side that is not a power of two, rows above 4096 or a side
// below 16) take the general path: the same packed planes built in one HBM
// pass, each axis transformed by mixed-radix Stockham passes through HBM
// (radices 16, 8, 4, 2, 3, 5, 7, 11, 13) or, for a length with a larger prime
// factor, by Bluestein's chirp-z algorithm over a power-of-two length, and a
// last pass that writes the images.  Twiddles come from FP64-computed tables
// exp(+2 pi i t / L).
#include "mti_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>

namespace mti {


// 1-D correction vectors: sinc(x)/psi(|i-n/2|/(n/2)), x = (i-n/2)/(n s)
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

// exp(+2 pi i t / NT) in FP64, cached per length on the context.
static mti_status twiddle_table(mti_ctx *ctx, int NT, const double2 **out, hipStream_t st)
{
    for (auto &t : ctx->twiddles)
        if (t.first == NT) {
            *out = t.second;
            return MTI_OK;
        }
    std::vector<double2> h(NT);
    for (int t = 0; t < NT; t++) {
        double a = 2.0 * M_PI * (double)t / (double)NT;
        h[t] = make_double2(std::cos(a), std::sin(a));
    }
    if (ctx->ws_active) {
        // a call in a caller workspace allocates nothing: the table goes to
        // its twiddle slot for this call (same values as the cached tables)
        void *w = nullptr;
        mti_status r = scratch_get(ctx, SLOT_TWIDDLE, (size_t)NT * sizeof(double2), &w);
        if (r != MTI_OK) return r;
        if ((r = upload(ctx, w, h.data(), NT * sizeof(double2), st)) != MTI_OK) return r;
        if ((r = host_sync(ctx, st)) != MTI_OK) return r;
        *out = (const double2 *)w;
        return MTI_OK;
    }
    double2 *d = nullptr;
    MTI_HIP(ctx, hipMalloc(&d, NT * sizeof(double2)));
    MTI_HIP(ctx, hipMemcpyAsync(d, h.data(), NT * sizeof(double2), hipMemcpyHostToDevice, st));
    MTI_HIP(ctx, hipStreamSynchronize(st));
    ctx->twiddles.push_back({NT, d});
    *out = d;
    return MTI_OK;
}

struct FftGeom {
    int nx, ny, lognx, N1, N2, logN1, logN2, NT;
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
    return g;
}

// ---- FP64 transform on 4096-element blocks (finalize path) --------------------
//
// A block holds NF = 4096 / N independent inverse transforms of length N in one
// LDS buffer (padded by one element every 16, so radix-16 writes are
// conflict-free).  Stockham passes of radix 16 (then 8, 4 or 2) keep R
// elements per butterfly in registers: every pass reads its inputs, waits,
// and overwrites in place.  The first pass reads HBM through a loader and
// the last pass hands results to a storer, so each block touches HBM once
// each way.  256 threads; twiddles from the FP64 table exp(+2 pi i t / NT).

constexpr int FB_E = 4096;
constexpr int FB_THREADS = 256;
constexpr int FB_LDS = FB_E + FB_E / 16;

// Packed Stokes planes.  Every output image is real (the Stokes formulas take
// real parts), and Re IDFT(S) = IDFT(H) with H(y, x) = (S(y, x) +
// conj(S(-y, -x))) / 2, the Hermitian part of S; two Hermitian planes then
// share one complex transform, IDFT(H0 + i H1) = IDFT(H0) + i IDFT(H1), whose
// real and imaginary parts are the two images.  A packed plane is built in
// the row pass from two correlation planes A and B of one channel (or, for
// one correlation, the planes of two channels): S_s = ca[s] A / na +
// cb[s] B / nb for s = 0, 1.  A full-polarisation channel takes two
// transforms instead of four; two Stokes-I channels share one.
struct PackDesc {
    const void *a, *b;        // element (0, 0) of planes A and B (b null: none)
    const double *na, *nb;    // their normalisers (device)
    double2 ca[2], cb[2];     // Stokes coefficients of s = 0 (real part) and 1 (imaginary part)
    float *img[2];            // image planes [ny][nx] of s = 0, 1 (img[1] null: none)
    const float *corr_x;      // [nx] sinc / taper correction
    const float *corr_y;      // [ny]
    const uint8_t *ma, *mb;   // tile masks of planes A and B ([ny/32][nx/32]; null: complete grids)
};

__device__ __forceinline__ int fb_pad(int i) { return i + (i >> 4); }

__device__ __forceinline__ double2 c2add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c2sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c2mul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 c2muli(double2 a) { return make_double2(-a.y, a.x); }  // * (+i)

// Packed planes X between the passes are stored as complex128 or complex64
// (XT); the arithmetic is FP64 either way.
__device__ __forceinline__ double2 x_ld(double2 v) { return v; }
__device__ __forceinline__ double2 x_ld(float2 v) { return make_double2(v.x, v.y); }
template <typename XT>
__device__ __forceinline__ XT x_st(double2 v);
template <>
__device__ __forceinline__ double2 x_st<double2>(double2 v)
{
    return v;
}
template <>
__device__ __forceinline__ float2 x_st<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}

// Inverse DFT of length 4 in place: X[k] = sum_n x[n] i^(nk).
__device__ __forceinline__ void idft4(double2 &x0, double2 &x1, double2 &x2, double2 &x3)
{
    const double2 a0 = c2add(x0, x2), a1 = c2sub(x0, x2), a2 = c2add(x1, x3), a3 = c2muli(c2sub(x1, x3));
    x0 = c2add(a0, a2);
    x2 = c2sub(a0, a2);
    x1 = c2add(a1, a3);
    x3 = c2sub(a1, a3);
}

// exp(+2 pi i m / 16) for m = 0..15, as double constants.
__device__ __forceinline__ double2 w16(int m)
{
    constexpr double C1 = 0.92387953251128675613, S1 = 0.38268343236508977173, H = 0.70710678118654752440;
    switch (m & 15) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(C1, S1);
    case 2: return make_double2(H, H);
    case 3: return make_double2(S1, C1);
    case 4: return make_double2(0.0, 1.0);
    case 5: return make_double2(-S1, C1);
    case 6: return make_double2(-H, H);
    case 7: return make_double2(-C1, S1);
    case 8: return make_double2(-1.0, 0.0);
    case 9: return make_double2(-C1, -S1);
    case 10: return make_double2(-H, -H);
    case 11: return make_double2(-S1, -C1);
    case 12: return make_double2(0.0, -1.0);
    case 13: return make_double2(S1, -C1);
    case 14: return make_double2(H, -H);
    default: return make_double2(C1, -S1);
    }
}

// v * exp(+2 pi i M / 16) for a constant M: the multiples of 1/8 turn need
// no general complex multiply (a product with 0.0 or 1.0 is not folded away
// by the compiler under IEEE rules).
template <int M>
__device__ __forceinline__ double2 c2mul_w16(double2 v)
{
    constexpr int m = M & 15;
    constexpr double H = 0.70710678118654752440;
    if constexpr (m == 0) return v;
    else if constexpr (m == 4) return make_double2(-v.y, v.x);
    else if constexpr (m == 8) return make_double2(-v.x, -v.y);
    else if constexpr (m == 12) return make_double2(v.y, -v.x);
    else if constexpr (m == 2) return make_double2(H * (v.x - v.y), H * (v.x + v.y));
    else if constexpr (m == 6) return make_double2(-H * (v.x + v.y), H * (v.x - v.y));
    else if constexpr (m == 10) return make_double2(H * (v.y - v.x), -H * (v.x + v.y));
    else if constexpr (m == 14) return make_double2(H * (v.x + v.y), H * (v.y - v.x));
    else return c2mul(v, w16(m));
}

// Inverse DFT of length R in place (natural order in and out).
template <int R>
__device__ __forceinline__ void idft(double2 (&v)[R])
{
    if constexpr (R == 2) {
        const double2 a = v[0];
        v[0] = c2add(a, v[1]);
        v[1] = c2sub(a, v[1]);
    } else if constexpr (R == 4) {
        idft4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // n = 2 n1 + n2, k = k1 + 4 k2
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        idft4(e[0], e[1], e[2], e[3]);
        idft4(o[0], o[1], o[2], o[3]);
        const double2 t[4] = {o[0], c2mul_w16<2>(o[1]), c2mul_w16<4>(o[2]), c2mul_w16<6>(o[3])};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = c2add(e[k], t[k]);
            v[k + 4] = c2sub(e[k], t[k]);
        }
    } else {
        static_assert(R == 16, "radix");
        // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1, twiddle w16^(n2 k1), DFT4 over n2
        double2 y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            y[n2][0] = v[n2];
            y[n2][1] = v[4 + n2];
            y[n2][2] = v[8 + n2];
            y[n2][3] = v[12 + n2];
            idft4(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
        }
        // twiddles w16^(n2 k1)
        y[1][1] = c2mul_w16<1>(y[1][1]);
        y[1][2] = c2mul_w16<2>(y[1][2]);
        y[1][3] = c2mul_w16<3>(y[1][3]);
        y[2][1] = c2mul_w16<2>(y[2][1]);
        y[2][2] = c2mul_w16<4>(y[2][2]);
        y[2][3] = c2mul_w16<6>(y[2][3]);
        y[3][1] = c2mul_w16<3>(y[3][1]);
        y[3][2] = c2mul_w16<6>(y[3][2]);
        y[3][3] = c2mul_w16<9>(y[3][3]);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            idft4(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
            v[k1] = y[0][k1];
            v[k1 + 4] = y[1][k1];
            v[k1 + 8] = y[2][k1];
            v[k1 + 12] = y[3][k1];
        }
    }
}

// LDS index of element n of transform f: row layout f*N + n (transforms of a
// row block, lanes along n) or column layout n*NF + f (column blocks, lanes
// along f so that HBM rows are read and written contiguously).
template <int E, int N, bool COL>
__device__ __forceinline__ int fb_idx(int f, int n)
{
    return COL ? n * (E / N) + f : fb_pad(f * N + n);
}

// A loader with issue(f, n) -> token and finish(token) -> element (and a
// GROUP constant) has the first pass's reads split in two: the reads of GROUP
// elements are all issued before the first is finished, so they are in flight
// together instead of one element's reads at a time.
template <class L, class = void>
struct HasIssue : std::false_type {};
template <class L>
struct HasIssue<L, std::void_t<decltype(std::declval<const L &>().issue(0, 0))>> : std::true_type {};

// One Stockham pass of radix R at stride NS over the block.  load(f, n) gives
// element n of transform f (FIRST); store(f, k, v) takes output k (LAST).
template <int E, int TH, int N, bool COL, int R, int NS, bool FIRST, bool LAST, class Load, class Store>
__device__ __forceinline__ void fb_pass(double2 *lds, const double2 *tw, int NT, const Load &load,
                                        const Store &store)
{
    constexpr int PER = N / R;
    constexpr int NF = E / N;
    constexpr int B = E / (R * TH);
    static_assert(B >= 1, "block too small for the radix");
    double2 v[B][R];
    // the pass's twiddles are read before its LDS reads and barrier, so their
    // round trip overlaps the wait for the other waves
    double2 t[B][R > 1 ? R - 1 : 1];
    if constexpr (NS > 1) {
        const int step = NT / (NS * R);
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int b = threadIdx.x + TH * s;
            const int j = COL ? b / NF : b % PER, k = j % NS;
#pragma unroll
            for (int r = 1; r < R; r++) t[s][r - 1] = tw[r * k * step];
        }
    }
    if constexpr (FIRST && HasIssue<Load>::value) {
        constexpr int G = Load::GROUP < B * R ? Load::GROUP : B * R;
        static_assert((B * R) % G == 0, "group");
#pragma unroll
        for (int e0 = 0; e0 < B * R; e0 += G) {
            typename Load::Token t[G];
#pragma unroll
            for (int e = 0; e < G; e++) {
                const int s = (e0 + e) / R, r = (e0 + e) % R, b = threadIdx.x + TH * s;
                const int f = COL ? b % NF : b / PER, j = COL ? b / NF : b % PER;
                t[e] = load.issue(f, j + PER * r);
            }
            __builtin_amdgcn_sched_barrier(0);  // all of the group's reads before the first use
#pragma unroll
            for (int e = 0; e < G; e++) v[(e0 + e) / R][(e0 + e) % R] = load.finish(t[e]);
        }
    } else {
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int b = threadIdx.x + TH * s;
            const int f = COL ? b % NF : b / PER, j = COL ? b / NF : b % PER;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if constexpr (FIRST) v[s][r] = load(f, j + PER * r);
                else v[s][r] = lds[fb_idx<E, N, COL>(f, j + PER * r)];
            }
        }
    }
    if (!FIRST) __syncthreads();
#pragma unroll
    for (int s = 0; s < B; s++) {
        const int b = threadIdx.x + TH * s;
        const int f = COL ? b % NF : b / PER, j = COL ? b / NF : b % PER;
        const int k = j % NS;
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[s][r] = c2mul(v[s][r], t[s][r - 1]);
        }
        idft<R>(v[s]);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (LAST) store(f, base + NS * r, v[s][r]);
            else lds[fb_idx<E, N, COL>(f, base + NS * r)] = v[s][r];
        }
    }
    if (!LAST) __syncthreads();
}

template <int E, int TH, int N, bool COL, int NS, bool FIRST, int RMAX, class Load, class Store>
__device__ __forceinline__ void fb_passes(double2 *lds, const double2 *tw, int NT, const Load &load,
                                          const Store &store)
{
    constexpr int REM = N / NS;
    constexpr int R = REM >= RMAX ? RMAX : REM;
    fb_pass<E, TH, N, COL, R, NS, FIRST, REM == R>(lds, tw, NT, load, store);
    if constexpr (REM > R) fb_passes<E, TH, N, COL, NS * R, false, RMAX>(lds, tw, NT, load, store);
}

// Full inverse transforms of a block of E elements (E / N transforms of
// length N) on TH threads, passes of radix up to RMAX: load -> passes -> store.
template <int N, bool COL, int E = FB_E, int TH = FB_THREADS, int RMAX = 16, class Load, class Store>
__device__ __forceinline__ void fb_ifft(double2 *lds, const double2 *tw, int NT, const Load &load,
                                        const Store &store)
{
    static_assert(N >= 16 && N <= E, "length");
    fb_passes<E, TH, N, COL, 1, true, RMAX>(lds, tw, NT, load, store);
}

// Row pass of packed planes.  Blocks of E = 4096 elements on 256 threads
// (68 KB of LDS: two blocks per CU).  Below N = 4096 a block transforms
// E / (2N) pairs of rows (y, -y mod ny) of packed plane pk; at N = 4096 it
// transforms one row, y or its mirror.  Each element is built on the fly from
// the grid planes with the PreIFFTUVCorrection reorder (transform row y,
// column x holds grid cell ((y + ny/2) mod ny, (x + N/2) mod N)) from the
// cells at (y, x) and (-y, -x); rows 0 and ny/2 are their own mirrors and
// are transformed once.
constexpr int PK_E = 4096;
constexpr int PK_TH = 256;
constexpr int PK_RMAX = 16;
template <int N>
struct PkGeom {
    static constexpr bool ROW = N == PK_E;             // one row a block (y or its mirror)
    static constexpr int NF = PK_E / N;                // transform rows of a block
    static constexpr int PB = ROW ? 1 : PK_E / (2 * N);  // row pairs of a block
    static constexpr int M = ROW ? 2 : 1;              // blocks per row pair
    static constexpr int NTX = (N + 31) / 32;          // 32-cell tiles of a row
    static constexpr int MR = ROW ? 2 : NF;            // tile-mask rows staged: own and mirror rows
};

// Grid elements are read through buffer resources over the grid planes: the
// byte offset of a cell is one 32-bit register, and the buffer's range check
// makes an offset past the plane's end (PK_OUTSIDE) read as zero bits without
// a memory access -- the read of a cell in a tile the mask leaves out, of a
// row past the last pair, or of an absent plane B needs no branch, so the
// compiler is free to issue the reads of all of a thread's elements before
// the arithmetic waits for the first.  Zero bits are +0 in either precision.
constexpr uint32_t PK_OUTSIDE = 0xfffffff0u;
constexpr int PK_BUFFER_DWORD3 = 0x00020000;  // raw buffer, gfx9-family dword 3

// Unit queues of the persistent row pass: one counter per XCD, each in its
// own 128-byte line (PK_QSTRIDE words apart), reset by a one-block kernel
// before every launch (a fill kernel rather than a memset, so that a
// captured graph replays it).
constexpr int PK_QUEUES = 8;
constexpr int PK_QSTRIDE = 32;
constexpr size_t PK_QUEUE_BYTES = (size_t)PK_QUEUES * PK_QSTRIDE * sizeof(uint32_t);

__global__ void k_pk_queue_reset(uint32_t *q)
{
    if (threadIdx.x < PK_QUEUES) q[threadIdx.x * PK_QSTRIDE] = 0u;
}

// Bytes of a descriptor array in the descriptor slot, rounded up so that
// what follows it starts 256-byte aligned.
static inline size_t desc_region(size_t n, size_t size)
{
    return (std::max<size_t>(1, n) * size + 255) / 256 * 256;
}
typedef unsigned int pk_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int pk_u32x4 __attribute__((ext_vector_type(4)));

template <typename TI>
struct PkRaw;
template <>
struct PkRaw<float2> {  // complex64 grid cell
    typedef pk_u32x2 T;
    static __device__ __forceinline__ T ld(__amdgpu_buffer_rsrc_t r, uint32_t off)
    {
        return __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    }
    static __device__ __forceinline__ double2 wide(T t)
    {
        return make_double2((double)__uint_as_float(t.x), (double)__uint_as_float(t.y));
    }
};
template <>
struct PkRaw<double2> {  // complex128 grid cell
    typedef pk_u32x4 T;
    static __device__ __forceinline__ T ld(__amdgpu_buffer_rsrc_t r, uint32_t off)
    {
        return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    }
    static __device__ __forceinline__ double2 wide(T t)
    {
        return make_double2(__hiloint2double((int)t.y, (int)t.x), __hiloint2double((int)t.w, (int)t.z));
    }
};

// Bytes a grid plane spans (cells [ny][N], cs elements apart).
__host__ __device__ inline uint64_t pk_span_bytes(int N, int ny, int64_t cs, size_t elem)
{
    return (((uint64_t)ny * N - 1) * (uint64_t)cs + 1) * elem;
}

// With S_s = ka_s A + kb_s B (ka_s = ca[s] / na, kb_s = cb[s] / nb) and
// H_s = (S_s + conj(S_s(-y, -x))) / 2, the packed element Z = H_0 + i H_1 is
//   Z = cA A + cB B + cAm conj(A(-y, -x)) + cBm conj(B(-y, -x)),
//   cA = (ka_0 + i ka_1) / 2,  cAm = (conj(ka_0) + i conj(ka_1)) / 2,
// and cB, cBm alike from kb: four complex multiply-adds an element.
__device__ __forceinline__ void pk_coefs(double2 k0, double2 k1, double2 &c, double2 &cm)
{
    c = make_double2(0.5 * (k0.x - k1.y), 0.5 * (k0.y + k1.x));
    cm = make_double2(0.5 * (k0.x + k1.y), 0.5 * (k1.x - k0.y));
}

// The row pass's element loader: element x of transform row f of the block
// is Z = H_0 + i H_1 built from four grid cells (planes A and B at (y, x) and
// at (-y, -x)).  issue() starts the four reads, finish() widens and packs.
// GROUP elements (64 reads of complex64 cells, 32 of complex128) are issued
// before the first is finished, 128 VGPRs of raw cells either way.
template <int N, typename TI>
struct PkRowLoad {
    using G = PkGeom<N>;
    static constexpr int GROUP = 128 / (int)sizeof(TI);
    struct Token {
        typename PkRaw<TI>::T a0, a1, b0, b1;
    };
    __amdgpu_buffer_rsrc_t rA, rB;
    double2 cA, cB, cAm, cBm;  // packing coefficients (pk_coefs)
    const uint8_t (*mask)[G::MR][G::NTX];  // LDS tile-mask rows [2][MR][NTX]
    int bx, ny, y1;                       // y1: the row of a one-row block
    uint32_t cb;                          // bytes between cells of a grid row
    // transform row of block row f, and the mask rows of it and of its mirror
    __device__ __forceinline__ int row(int f) const
    {
        if (G::ROW) return y1;
        const int pair = bx * G::PB + (f >> 1);
        return (f & 1) ? ((ny - pair) & (ny - 1)) : pair;
    }
    __device__ __forceinline__ uint32_t off(int ab, int mr, int y, int x) const
    {
        const int gx = (x + N / 2) & (N - 1);
        const uint32_t o = ((uint32_t)((y + (ny >> 1)) & (ny - 1)) * N + gx) * cb;
        return mask[ab][mr][gx >> 5] ? o : PK_OUTSIDE;
    }
    __device__ __forceinline__ Token issue(int f, int x) const
    {
        const int y = row(f), mr = G::ROW ? 0 : f, mm = G::ROW ? 1 : f ^ 1;
        const int ym = (ny - y) & (ny - 1), xm = (N - x) & (N - 1);
        Token t;
        t.a0 = PkRaw<TI>::ld(rA, off(0, mr, y, x));
        t.a1 = PkRaw<TI>::ld(rA, off(0, mm, ym, xm));
        t.b0 = PkRaw<TI>::ld(rB, off(1, mr, y, x));
        t.b1 = PkRaw<TI>::ld(rB, off(1, mm, ym, xm));
        return t;
    }
    __device__ __forceinline__ double2 finish(const Token &t) const
    {
        const double2 a = PkRaw<TI>::wide(t.a0), am = PkRaw<TI>::wide(t.a1);
        const double2 b = PkRaw<TI>::wide(t.b0), bm = PkRaw<TI>::wide(t.b1);
        // Z = cA A + cB B + cAm conj(A(-y,-x)) + cBm conj(B(-y,-x))
        double zr = cA.x * a.x;
        zr = fma(-cA.y, a.y, zr);
        zr = fma(cB.x, b.x, zr);
        zr = fma(-cB.y, b.y, zr);
        zr = fma(cAm.x, am.x, zr);
        zr = fma(cAm.y, am.y, zr);
        zr = fma(cBm.x, bm.x, zr);
        zr = fma(cBm.y, bm.y, zr);
        double zi = cA.x * a.y;
        zi = fma(cA.y, a.x, zi);
        zi = fma(cB.x, b.y, zi);
        zi = fma(cB.y, b.x, zi);
        zi = fma(cAm.y, am.x, zi);
        zi = fma(-cAm.x, am.y, zi);
        zi = fma(cBm.y, bm.x, zi);
        zi = fma(-cBm.x, bm.y, zi);
        return make_double2(zr, zi);
    }
};

// Block -> (row block bx, mirror bit mir, pack pk) of the row pass, a
// bijection over the npk * nbx * M blocks of a 1-D launch.  The two packs 2c,
// 2c+1 of a full-polarisation channel read the same grid cache lines when the
// correlations are interleaved per cell (each uses two of the four), and a
// one-row block reads the lines of its mirror row's block too.  Blocks go
// round-robin to the 8 XCDs, each with its own L2, so the 2M blocks that
// share lines are made L, L + 8, ...: they start together on one XCD and the
// later ones find the lines in L2.  An unpaired last pack takes the tail.
template <int M>
__device__ __forceinline__ void pk_block(int L, int nbx, int npk, int &bx, int &mir, int &pk)
{
    constexpr int G = 2 * M;
    const int paired = G * (npk >> 1) * nbx;
    if (L >= paired) {
        pk = npk - 1;
        bx = (L - paired) / M;
        mir = (L - paired) % M;
        return;
    }
    const int full = paired - paired % (8 * G);
    int u, k;
    if (L < full) {
        u = (L / (8 * G)) * 8 + (L & 7);
        k = (L >> 3) % G;
    } else {
        u = full / G + (L - full) / G;
        k = (L - full) % G;
    }
    pk = 2 * (u / nbx) + (k & 1);
    bx = u % nbx;
    mir = k >> 1;
}

template <int N, typename TI, typename XT>
__global__ __launch_bounds__(PK_TH) void k_rows_pk(const PackDesc *pd, int npk, XT *X, int ny, int64_t cs,
                                                  const double2 *tw, int NT)
{
    using G = PkGeom<N>;
    constexpr int E = PK_E, NTX = G::NTX;
    __shared__ double2 lds[E + E / 16];
    __shared__ uint8_t s_mask[2][G::MR][NTX];  // tile masks of the block's rows (sparse grids)
    const int npairs = ny / 2 + 1, nbx = (npairs + G::PB - 1) / G::PB;
    int bx, mir, pk;
    pk_block<G::M>(blockIdx.x, nbx, npk, bx, mir, pk);
    // a one-row block: row bx, or its mirror; a self-mirrored row has one block
    const int y1 = mir ? ((ny - bx) & (ny - 1)) : bx;
    if (G::ROW && mir && y1 == bx) return;
    const PackDesc &d = pd[pk];
    const void *B = d.b;
    const double ia = 1.0 / *d.na, ib = B ? 1.0 / *d.nb : 0.0;
    XT *dst = X + (size_t)pk * N * ny;
    PkRowLoad<N, TI> load;
    load.bx = bx;
    load.ny = ny;
    load.y1 = y1;
    // the tile-mask rows of the block's transform rows and of their mirrors
    // (all 1 for complete grids): cells of tiles no record reached, and of
    // rows past the last pair, are read as zeros, never loaded
    for (int i = threadIdx.x; i < 2 * G::MR * NTX; i += PK_TH) {
        const int ab = i / (G::MR * NTX), f = (i / NTX) % G::MR, t = i % NTX;
        const uint8_t *m = ab ? d.mb : d.ma;
        uint8_t v;
        int y;
        if (G::ROW) {
            v = 1;
            y = f ? (ny - y1) & (ny - 1) : y1;
        } else {
            v = bx * G::PB + (f >> 1) < npairs;
            y = load.row(f);
        }
        if (m && v) v = m[(((y + (ny >> 1)) & (ny - 1)) >> 5) * NTX + t];
        s_mask[ab][f][t] = v;
    }
    __syncthreads();
    const uint32_t span = (uint32_t)pk_span_bytes(N, ny, cs, sizeof(TI));
    load.rA = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(d.a), (short)0, (int)span, PK_BUFFER_DWORD3);
    load.rB = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(d.b), (short)0, B ? (int)span : 0,
                                                PK_BUFFER_DWORD3);
    pk_coefs(make_double2(d.ca[0].x * ia, d.ca[0].y * ia), make_double2(d.ca[1].x * ia, d.ca[1].y * ia), load.cA,
             load.cAm);
    pk_coefs(make_double2(d.cb[0].x * ib, d.cb[0].y * ib), make_double2(d.cb[1].x * ib, d.cb[1].y * ib), load.cB,
             load.cBm);
    load.mask = s_mask;
    load.cb = (uint32_t)cs * (uint32_t)sizeof(TI);
    auto store = [&](int f, int kx, double2 v) {
        if (G::ROW) {
            dst[(size_t)y1 * N + kx] = x_st<XT>(v);
            return;
        }
        const int pair = bx * G::PB + (f >> 1);
        if (pair >= npairs) return;
        const int y = load.row(f);
        if ((f & 1) && y == pair) return;  // a self-paired row is written once
        dst[(size_t)y * N + kx] = x_st<XT>(v);
    };
    fb_ifft<N, false, E, PK_TH, PK_RMAX>(lds, tw, NT, load, store);
}

// Values the same in every lane, moved to scalar registers: in a loop after
// global stores the compiler cannot prove a descriptor load unclobbered and
// loads it per lane; a buffer resource built from per-lane registers turns
// every read through it into a waterfall loop.
__device__ __forceinline__ const void *pk_uniform_ptr(const void *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const void *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double pk_uniform_f64(double x)
{
    const uint64_t v = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// v[r] * w^r for r = 1..R-1 from w and w^4 (powers w^2, w^3, w^8, w^12 and
// w^(4a + b) = w^(4a) w^b): the twiddles of a pass from two table values.
template <int R>
__device__ __forceinline__ void pk_twiddle(double2 (&v)[R], double2 w1, double2 w4)
{
    if constexpr (R > 1) {
        const double2 one = make_double2(1.0, 0.0);
        const double2 w2 = R > 2 ? c2mul(w1, w1) : one, w3 = R > 3 ? c2mul(w2, w1) : one;
        const double2 w8 = R > 8 ? c2mul(w4, w4) : one, w12 = R > 8 ? c2mul(w8, w4) : one;
        const double2 lo[4] = {one, w1, w2, w3}, hi[4] = {one, w4, w8, w12};
#pragma unroll
        for (int r = 1; r < R; r++) {
            const int a = r >> 2, b = r & 3;
            const double2 w = a == 0 ? lo[b] : (b == 0 ? hi[a] : c2mul(hi[a], lo[b]));
            v[r] = c2mul(v[r], w);
        }
    }
}

// One Stockham pass (radix R, stride NS > 1) over the E = 4096 elements of a
// row-pass block in LDS, as fb_pass, with the twiddles of transform-local
// index k from the bases tb[2k] = w^k, tb[2k + 1] = w^4k in LDS.
template <int N, int R, int NS, bool LAST, class Store>
__device__ __forceinline__ void pk_pipe_pass(double2 *lds, const double2 *tb, const Store &store)
{
    constexpr int PER = N / R, B = PK_E / (R * PK_TH);
    double2 v[B][R];
#pragma unroll
    for (int s = 0; s < B; s++) {
        const int b = threadIdx.x + PK_TH * s, f = b / PER, jj = b % PER;
#pragma unroll
        for (int r = 0; r < R; r++) v[s][r] = lds[fb_pad(f * N + jj + PER * r)];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < B; s++) {
        const int b = threadIdx.x + PK_TH * s, f = b / PER, jj = b % PER, k = jj % NS;
        pk_twiddle<R>(v[s], tb[2 * k], tb[2 * k + 1]);
        idft<R>(v[s]);
        const int base = (jj - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (LAST) store(f, base + NS * r, v[s][r]);
            else lds[fb_pad(f * N + base + NS * r)] = v[s][r];
        }
    }
    if (!LAST) __syncthreads();
}

// The passes after the first (strides 16 and 256 at most, rows <= 4096).
template <int N, class Store>
__device__ __forceinline__ void pk_pipe_rest(double2 *lds, const double2 *tb2, const double2 *tb3, const Store &store)
{
    constexpr int R2 = N / 16 >= 16 ? 16 : N / 16;
    if constexpr (N / 16 == R2) {
        pk_pipe_pass<N, R2, 16, true>(lds, tb2, store);
    } else {
        pk_pipe_pass<N, R2, 16, false>(lds, tb2, store);
        pk_pipe_pass<N, N / 256, 256, true>(lds, tb3, store);
    }
}

// Row pass from complex64 grids, software-pipelined: a persistent block
// walks the units of k_rows_pk (a row pair, or at N = 4096 one row; the same
// unit order, so the blocks that share grid lines run on one XCD) and issues
// the grid reads of the first AHEAD elements of its next unit before the
// later passes of the current one, so those reads are in flight during the
// passes instead of after them.  A global load waits for every load issued before it, so no
// global load follows the early reads until they are used: the passes'
// twiddles come from bases w^k, w^4k staged in LDS once per block (the
// others as products, a few ulp from the table values), and the next unit's
// tile-mask rows and normalisers are read during the current unit's first
// pass.
template <int N, typename XT>
__global__ __launch_bounds__(PK_TH) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_rows_pk_pipe(
    const PackDesc *__restrict__ pd, int npk, XT *X, int ny, int64_t cs, const double2 *tw, int NT,
    uint32_t *queues)
{
    using TI = float2;
    using G = PkGeom<N>;
    using Load = PkRowLoad<N, TI>;
    constexpr int E = PK_E, NTX = G::NTX, PER = N / 16, NQ = 2 * G::MR * NTX / PK_TH;
    static_assert(N >= 256 && PK_TH == 256 && E / (16 * PK_TH) == 1 && NQ >= 1 && NQ <= 2, "row geometry");
    __shared__ double2 lds[E + E / 16];
    __shared__ double2 s_tw2[2 * 16], s_tw3[2 * 256];   // [k][w^k, w^4k] of the strides 16 and 256
    __shared__ uint8_t s_mask[2][2][G::MR][NTX];         // [buffer][plane A, B][row][tile]
    const int npairs = ny / 2 + 1, nbx = (npairs + G::PB - 1) / G::PB, U = G::M * npk * nbx;
    const uint32_t span = (uint32_t)pk_span_bytes(N, ny, cs, sizeof(TI));
    const uint32_t cb = (uint32_t)cs * (uint32_t)sizeof(TI);
    const int j = threadIdx.x;
    // unit L -> pack, row block and (one-row blocks) row; the mirror unit of
    // a self-mirrored row is empty
    struct Unit {
        int pk, bx, y1;
    };
    auto unit = [&](int L, Unit &u) {
        int mir;
        pk_block<G::M>(L, nbx, npk, u.bx, mir, u.pk);
        u.y1 = mir ? ((ny - u.bx) & (ny - 1)) : u.bx;
        return !(G::ROW && mir && u.y1 == u.bx);
    };
    // Rows of 4096 (one-row units) come from one queue per XCD (workgroups
    // go round-robin to the 8 XCDs): queue q holds L = q + nq k in order, k
    // from the queue's atomic counter, so a block that finishes early takes
    // more units instead of idling while others finish theirs.  Thread 0
    // takes the unit after the next one while the block works on the current
    // one; a queue's empty (self-mirrored) units are skipped.  Shorter rows
    // walk L = blockIdx.x + k * gridDim.x: their launches are small, and the
    // queue measured slower there.
    constexpr bool QUEUE = G::ROW;
    auto next_static = [&](int L, Unit &u) {
        while (L < U && !unit(L, u)) L += gridDim.x;
        return L;
    };
    __shared__ int s_take[2];
    const int nq = min((int)gridDim.x, PK_QUEUES), q = (int)(blockIdx.x % (unsigned)nq);
    uint32_t *qc = queues + q * PK_QSTRIDE;
    auto taken = [&](uint32_t k) {  // thread 0: queue entry k, or the next non-empty one
        Unit t;
        int L = q + nq * (int)k;
        while (L < U && !unit(L, t)) L = q + nq * (int)atomicAdd(qc, 1u);
        return L;
    };
    // transform row of block row f
    auto row_of = [&](const Unit &u, int f) {
        if (G::ROW) return f ? ((ny - u.y1) & (ny - 1)) : u.y1;
        const int pair = u.bx * G::PB + (f >> 1);
        return (f & 1) ? ((ny - pair) & (ny - 1)) : pair;
    };
    // a unit's tile-mask bytes (entries j, j + 256 of [plane][row][tile]) and
    // normalisers, through global-address-space pointers: a FLAT load would
    // also hold up every LDS wait until it returns
    typedef const __attribute__((address_space(1))) uint8_t *GU8;
    typedef const __attribute__((address_space(1))) double *GF64;
    auto prefetch = [&](const Unit &u, uint8_t (&m)[NQ], double &na, double &nb) {
        const PackDesc &d = pd[u.pk];
        const uint8_t *ma = (const uint8_t *)pk_uniform_ptr(d.ma), *mb = (const uint8_t *)pk_uniform_ptr(d.mb);
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int i = j + PK_TH * q, ab = i / (G::MR * NTX), f = (i / NTX) % G::MR, t = i % NTX;
            const uint8_t *mk = ab ? mb : ma;
            const bool valid = G::ROW || u.bx * G::PB + (f >> 1) < npairs;  // rows past the last pair: zeros
            const int y = row_of(u, f);
            m[q] = !valid ? 0 : (mk ? ((GU8)mk)[(((y + (ny >> 1)) & (ny - 1)) >> 5) * NTX + t] : 1);
        }
        na = *(GF64)pk_uniform_ptr(d.na);
        nb = d.b ? *(GF64)pk_uniform_ptr(d.nb) : 1.0;
    };
    auto stage = [&](const uint8_t (&m)[NQ], int buf) {
#pragma unroll
        for (int q = 0; q < NQ; q++) (&s_mask[buf][0][0][0])[j + PK_TH * q] = m[q];
    };
    // the reads of a unit need its planes and mask rows only
    auto reader = [&](const Unit &u, int buf) {
        const PackDesc &d = pd[u.pk];
        const void *a = pk_uniform_ptr(d.a), *b = pk_uniform_ptr(d.b);
        Load l;
        l.rA = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a), (short)0, (int)span, PK_BUFFER_DWORD3);
        l.rB = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(b), (short)0, b ? (int)span : 0,
                                                 PK_BUFFER_DWORD3);
        l.mask = s_mask[buf];
        l.bx = u.bx;
        l.ny = ny;
        l.y1 = u.y1;
        l.cb = cb;
        return l;
    };
    auto coefs = [&](const Unit &u, double na, double nb, Load &l) {
        const PackDesc &d = pd[u.pk];
        const double ia = 1.0 / pk_uniform_f64(na), ib = pk_uniform_ptr(d.b) ? 1.0 / pk_uniform_f64(nb) : 0.0;
        pk_coefs(make_double2(d.ca[0].x * ia, d.ca[0].y * ia), make_double2(d.ca[1].x * ia, d.ca[1].y * ia), l.cA,
                 l.cAm);
        pk_coefs(make_double2(d.cb[0].x * ib, d.cb[0].y * ib), make_double2(d.cb[1].x * ib, d.cb[1].y * ib), l.cB,
                 l.cBm);
    };
    Unit u;
    int L, Ln = 0;  // the current unit; with the queue also the next (U: none)
    if constexpr (QUEUE) {
        if (j == 0) {
            const int L0 = taken(atomicAdd(qc, 1u));
            s_take[0] = L0;
            s_take[1] = L0 < U ? taken(atomicAdd(qc, 1u)) : U;
        }
        __syncthreads();
        L = __builtin_amdgcn_readfirstlane(s_take[0]);
        Ln = __builtin_amdgcn_readfirstlane(s_take[1]);
        if (L >= U) return;
        unit(L, u);
    } else {
        L = next_static(blockIdx.x, u);
        if (L >= U) return;
    }
    // twiddle bases of the passes of stride 16 (radix R2) and 256 (radix
    // N / 256): fb_pass's twiddle tw[r k step] with step = NT / (NS R)
    {
        constexpr int R2 = N / 16 >= 16 ? 16 : N / 16;
        const int s3 = N > 256 ? NT / N : 0, s2 = NT / (16 * R2);
        s_tw3[2 * j] = tw[(j * s3) & (NT - 1)];
        s_tw3[2 * j + 1] = tw[(4 * j * s3) & (NT - 1)];
        if (j < 16) {
            s_tw2[2 * j] = tw[(j * s2) & (NT - 1)];
            s_tw2[2 * j + 1] = tw[(4 * j * s2) & (NT - 1)];
        }
    }
    // elements 0..AHEAD-1 of a unit are read during the previous unit's
    // passes, the rest when its own first pass starts (more ahead does not
    // fit two waves per SIMD without spilling: the radix-8 and radix-4 last
    // passes of 2048 and 1024 hold more twiddles)
    constexpr int AHEAD = (N == 1024 || N == 2048) ? 4 : N == 4096 ? 7 : 8;
    const int f0 = j / PER, jj = j % PER;  // first pass: transform row f0, elements jj + PER r
    int buf = 0;
    uint8_t m[NQ];
    double na, nb;
    prefetch(u, m, na, nb);
    stage(m, buf);
    __syncthreads();
    typename Load::Token tok[16];
    {
        const Load rd = reader(u, buf);
#pragma unroll
        for (int r = 0; r < AHEAD; r++) tok[r] = rd.issue(f0, jj + PER * r);
    }
    while (true) {
        {
            const Load rd = reader(u, buf);
#pragma unroll
            for (int r = AHEAD; r < 16; r++) tok[r] = rd.issue(f0, jj + PER * r);
        }
        Unit un;
        uint32_t kt = 0;
        if constexpr (QUEUE) {
            if (Ln < U) unit(Ln, un);
            // the unit after next (wave 0 holds the queue entry in a scalar
            // register: a vector one would spill at the first pass's peak)
            if (j < 64) kt = __builtin_amdgcn_readfirstlane(j == 0 && Ln < U ? atomicAdd(qc, 1u) : 0u);
        } else {
            Ln = next_static(L + gridDim.x, un);
        }
        const bool more = Ln < U;
        uint8_t mn[NQ];
        double nan_ = 1.0, nbn = 1.0;
        if (more) prefetch(un, mn, nan_, nbn);
        // first pass: finish the elements, radix 16 (stride 1), into LDS
        {
            Load cur;
            coefs(u, na, nb, cur);
            double2 v[16];
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = cur.finish(tok[r]);
            idft<16>(v);
#pragma unroll
            for (int r = 0; r < 16; r++) lds[fb_pad(f0 * N + 16 * jj + r)] = v[r];
        }
        if (more) stage(mn, buf ^ 1);
        if constexpr (QUEUE)
            if (j == 0) s_take[buf] = more ? taken(kt) : U;
        __syncthreads();
        const int Lnn = QUEUE ? __builtin_amdgcn_readfirstlane(s_take[buf]) : 0;
        if (more) {
            const Load rd = reader(un, buf ^ 1);
#pragma unroll
            for (int r = 0; r < AHEAD; r++) tok[r] = rd.issue(f0, jj + PER * r);
        }
        XT *dst = X + (size_t)u.pk * N * ny;
        pk_pipe_rest<N>(lds, s_tw2, s_tw3, [&](int f, int kx, double2 v) {
            if (G::ROW) {
                dst[(size_t)u.y1 * N + kx] = x_st<XT>(v);
                return;
            }
            const int pair = u.bx * G::PB + (f >> 1);
            if (pair >= npairs) return;
            const int y = row_of(u, f);
            if ((f & 1) && y == pair) return;  // a self-paired row is written once
            dst[(size_t)y * N + kx] = x_st<XT>(v);
        });
        if (!more) break;
        if constexpr (QUEUE)
            Ln = Lnn;
        else
            L = Ln;
        u = un;
        na = nan_;
        nb = nbn;
        buf ^= 1;
    }
}

// Pass B1: length-N1 transforms down the columns (rows N2*n1 + n2) of NF =
// 4096/N1 columns starting at x0, times exp(+2 pi i n2 k1 / ny), in place.
template <int N1, typename XT>
__global__ __launch_bounds__(FB_THREADS) void k_cols1_fb(XT *X, int nx, int ny, int N2, const double2 *tw, int NT)
{
    __shared__ double2 lds[FB_LDS];
    constexpr int NF = FB_E / N1;
    const int x0 = blockIdx.x * NF, n2 = blockIdx.y;
    XT *plane = X + (size_t)blockIdx.z * nx * ny;
    const int mul = NT / ny;
    // transform f = column x0 + f; loads with the column index fastest across lanes
    auto load = [&](int f, int n1) { return x_ld(plane[((size_t)N2 * n1 + n2) * nx + x0 + f]); };
    auto store = [&](int f, int k1, double2 v) {
        plane[((size_t)N2 * k1 + n2) * nx + x0 + f] = x_st<XT>(c2mul(v, tw[((n2 * k1) & (ny - 1)) * mul]));
    };
    fb_ifft<N1, true>(lds, tw, NT, load, store);
}

// Finalize of packed column transforms held in LDS: output k of column
// x0 + cx at lds[k * NF + cx] is image row ky = ky_of(k); fftshift, the m flip
// and both corrections; the real part is image s = 0, the imaginary part s = 1.
template <int NF, int TH, class KyOf>
__device__ __forceinline__ void pack_store(const double2 *lds, int nk, int x0, int nx, int ny, const PackDesc &d,
                                           const KyOf &ky_of)
{
    float *i0 = d.img[0], *i1 = d.img[1];
    for (int i = threadIdx.x; i < nk * NF; i += TH) {
        const int k = i / NF, cx = i - k * NF;
        const int ky = ky_of(k), kx = x0 + cx;
        const int jy = (ky + (ny >> 1)) & (ny - 1), jx = (kx + (nx >> 1)) & (nx - 1);
        const int row = ny - 1 - jy;  // m axis flipped for FITS
        const double2 v = lds[k * NF + cx];
        const double cf = (double)d.corr_x[jx] * (double)d.corr_y[row];
        const size_t o = (size_t)row * nx + jx;
        i0[o] = (float)(v.x * cf);
        if (i1) i1[o] = (float)(v.y * cf);
    }
}

// Pass B2 + finalize: packed plane blockIdx.z, index k1, NF = 4096 / N2
// columns from x0: rows N2*k1 + n2 -> length-N2 transforms -> ky = k1 + N1*k2.
template <int N2, typename XT>
__global__ __launch_bounds__(FB_THREADS) void k_cols2_pk(const XT *X, int nx, int ny, int N1, const double2 *tw,
                                                        int NT, const PackDesc *pd)
{
    __shared__ double2 lds[FB_LDS];
    constexpr int NF = FB_E / N2;
    const int x0 = blockIdx.x * NF, k1 = blockIdx.y;
    const XT *plane = X + (size_t)blockIdx.z * nx * ny;
    auto load = [&](int f, int n2) { return x_ld(plane[((size_t)N2 * k1 + n2) * nx + x0 + f]); };
    auto store = [&](int f, int k2, double2 v) { lds[k2 * NF + f] = v; };
    fb_ifft<N2, true>(lds, tw, NT, load, store);
    __syncthreads();
    pack_store<NF, FB_THREADS>(lds, N2, x0, nx, ny, pd[blockIdx.z], [&](int k2) { return k1 + N1 * k2; });
}

// Whole-column pass + finalize (ny <= FBC_E / 32): a block holds NF >= 32
// columns of packed plane blockIdx.z (128 KB of LDS, rows read in runs of NF
// elements) and transforms them completely, so the transform takes two HBM
// passes instead of three.
constexpr int FBC_E = 8192;
constexpr int FBC_THREADS = 512;

template <int N, typename XT>
__global__ __launch_bounds__(FBC_THREADS) void k_cols_pk(const XT *X, int nx, int ny, const double2 *tw, int NT,
                                                        const PackDesc *pd)
{
    __shared__ double2 lds[FBC_E];
    constexpr int NF = FBC_E / N;
    static_assert(NF >= 32, "columns per block");
    const int x0 = blockIdx.x * NF;
    const XT *plane = X + (size_t)blockIdx.z * nx * ny;
    auto load = [&](int f, int n) { return x_ld(plane[(size_t)n * nx + x0 + f]); };
    auto store = [&](int f, int k, double2 v) { lds[k * NF + f] = v; };
    fb_ifft<N, true, FBC_E, FBC_THREADS>(lds, tw, NT, load, store);
    __syncthreads();
    pack_store<NF, FBC_THREADS>(lds, N, x0, nx, ny, pd[blockIdx.z], [](int k) { return k; });
}

// Host dispatch of the packed block-transform path: rows of 16-4096 elements;
// columns whole when a block spans at least 32 of them (ny <= 256: image rows
// written in runs of >= 128 bytes; narrower blocks write the f32 image in
// pieces too small for HBM and lose to the two-pass split), else four-step
// (N1, N2 >= 16, a 4096-element block of columns fits the image width).
static bool fbc_ok(const FftGeom &g) { return g.ny >= 16 && g.ny * 32 <= FBC_E && g.nx >= FBC_E / g.ny; }

static bool split_ok(const FftGeom &g)
{
    return g.N1 >= 16 && g.N2 >= 16 && FB_E / g.N1 <= g.nx && FB_E / g.N2 <= g.nx;
}

static bool pk_ok(const FftGeom &g)
{
    return is_pow2(g.nx) && is_pow2(g.ny) && g.nx >= 16 && g.nx <= 4096 && g.ny >= 16 && (fbc_ok(g) || split_ok(g));
}

// Whole columns in one pass, or the four-step split: where both apply (ny =
// 256) the context's tuning picks.
static bool whole_columns(const mti_ctx *ctx, const FftGeom &g)
{
    return fbc_ok(g) && (!split_ok(g) || g.ny <= ctx->whole_columns_ny);
}

template <typename TI, typename XT>
static mti_status pk_rows(mti_ctx *ctx, const FftGeom &g, const PackDesc *pd, int npk, XT *X, int64_t cs,
                          const double2 *tw, uint32_t *queues, hipStream_t st)
{
    const int npairs = g.ny / 2 + 1;
    if (pk_span_bytes(g.nx, g.ny, cs, sizeof(TI)) > PK_OUTSIDE)
        return set_error(ctx, MTI_ERR_UNSUPPORTED, "grid plane spans more than 4 GiB");
    if constexpr (std::is_same<TI, float2>::value) {
        // complex64 grids, rows of 256 or more: persistent and pipelined, two
        // blocks per CU (LDS), a multiple of 32 blocks so each keeps the XCD
        // grouping of its units
        const int64_t cap = (int64_t)2 * ((ctx->num_cus + 15) / 16) * 16;
        switch (g.nx) {
#define MTI_PP(NN)                                                                                                \
    case NN: {                                                                                                    \
        using G = PkGeom<NN>;                                                                                     \
        const int64_t units = (int64_t)G::M * npk * ceil_div(npairs, G::PB);                                      \
        if (NN == PK_E) hipLaunchKernelGGL(k_pk_queue_reset, dim3(1), dim3(PK_QUEUES), 0, st, queues);           \
        hipLaunchKernelGGL((k_rows_pk_pipe<NN, XT>), dim3((unsigned)std::min(units, cap)), dim3(PK_TH), 0, st, pd, \
                           npk, X, g.ny, cs, tw, g.NT, queues);                                                   \
        MTI_HIP(ctx, hipGetLastError());                                                                          \
        return MTI_OK;                                                                                            \
    }
            MTI_PP(256) MTI_PP(512) MTI_PP(1024) MTI_PP(2048) MTI_PP(4096)
#undef MTI_PP
        default: break;
        }
    }
    switch (g.nx) {
#define MTI_PR(NN)                                                                                          \
    case NN: {                                                                                              \
        using G = PkGeom<NN>;                                                                               \
        const dim3 gr((unsigned)(ceil_div(npairs, G::PB) * G::M * npk));                                    \
        hipLaunchKernelGGL((k_rows_pk<NN, TI, XT>), gr, dim3(PK_TH), 0, st, pd, npk, X, g.ny, cs, tw, g.NT); \
        break;                                                                                              \
    }
        MTI_PR(16) MTI_PR(32) MTI_PR(64) MTI_PR(128) MTI_PR(256) MTI_PR(512) MTI_PR(1024) MTI_PR(2048)
        MTI_PR(4096)
#undef MTI_PR
    default: return set_error(ctx, MTI_ERR_UNSUPPORTED, "row length");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <typename XT>
static mti_status fb_cols1(mti_ctx *ctx, const FftGeom &g, XT *X, int planes, const double2 *tw, hipStream_t st)
{
    const dim3 gr((unsigned)(g.nx / (FB_E / g.N1)), (unsigned)g.N2, (unsigned)planes);
    switch (g.N1) {
#define MTI_C1(NN) \
    case NN: hipLaunchKernelGGL((k_cols1_fb<NN, XT>), gr, dim3(FB_THREADS), 0, st, X, g.nx, g.ny, g.N2, tw, g.NT); break;
        MTI_C1(16) MTI_C1(32) MTI_C1(64) MTI_C1(128)
#undef MTI_C1
    default: return set_error(ctx, MTI_ERR_UNSUPPORTED, "column split");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <typename XT>
static mti_status pk_cols2(mti_ctx *ctx, const FftGeom &g, const XT *X, int npk, const double2 *tw,
                           const PackDesc *pd, hipStream_t st)
{
    const dim3 gr((unsigned)(g.nx / (FB_E / g.N2)), (unsigned)g.N1, (unsigned)npk);
    switch (g.N2) {
#define MTI_C2(NN) \
    case NN:                                                                                                \
        hipLaunchKernelGGL((k_cols2_pk<NN, XT>), gr, dim3(FB_THREADS), 0, st, X, g.nx, g.ny, g.N1, tw, g.NT, pd); \
        break;
        MTI_C2(16) MTI_C2(32) MTI_C2(64)
#undef MTI_C2
    default: return set_error(ctx, MTI_ERR_UNSUPPORTED, "column split");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

template <typename XT>
static mti_status pk_cols(mti_ctx *ctx, const FftGeom &g, const XT *X, int npk, const double2 *tw,
                          const PackDesc *pd, hipStream_t st)
{
    const dim3 gr((unsigned)(g.nx / (FBC_E / g.ny)), 1u, (unsigned)npk);
    switch (g.ny) {
#define MTI_CF(NN) \
    case NN: hipLaunchKernelGGL((k_cols_pk<NN, XT>), gr, dim3(FBC_THREADS), 0, st, X, g.nx, g.ny, tw, g.NT, pd); break;
        MTI_CF(16) MTI_CF(32) MTI_CF(64) MTI_CF(128) MTI_CF(256)
#undef MTI_CF
    default: return set_error(ctx, MTI_ERR_UNSUPPORTED, "column length");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}


// One channel of a block-transform finalize.
struct ChanRef {
    const void *grid;   // the channel's grid (layout strides cs, ps)
    const double *norm; // [P]
    float *image;       // [P][ny][nx]
    const float *corr_x, *corr_y;
    const uint8_t *mask = nullptr;  // tile mask of sparse grids (null: complete)
};

// Packed planes of a list of channels of one image size (the Stokes formulas
// of op/PostIFFTStokesConverter.cu:51-59, 73-80 as coefficients of the
// correlation planes: linear I = Re(XX+YY)/2, Q = Re(XX-YY)/2,
// U = Re(XY+YX)/2, V = Im(XY-YX)/2 = Re(-i(XY-YX))/2; circular
// I = Re(RR+LL)/2, Q = Re(RL+LR)/2, U = Im(LR-RL)/2, V = Re(RR-LL)/2;
// two correlations (a+b)/2, (a-b)/2; one correlation pairs channels).
static void make_packs(const std::vector<ChanRef> &ch, int P, int poltype, int layout, int nx, int ny, size_t elem,
                       std::vector<PackDesc> &out)
{
    const int64_t plane = (int64_t)nx * ny;
    const int64_t ps = layout == MTI_GRID_VUP ? 1 : plane;  // complex elements between correlation planes
    auto pl = [&](const ChanRef &c, int p) { return (const void *)((const char *)c.grid + (size_t)(p * ps) * elem); };
    const double2 h = make_double2(0.5, 0.0), mh = make_double2(-0.5, 0.0);
    const double2 ih = make_double2(0.0, 0.5), mih = make_double2(0.0, -0.5), one = make_double2(1.0, 0.0);
    const double2 zero = make_double2(0.0, 0.0);
    if (P == 1) {
        // a pack has one pair of correction vectors: only channels that share
        // them (same size and oversampling) are paired; the stable order keeps
        // the channels of a job together
        std::vector<size_t> ord(ch.size());
        for (size_t i = 0; i < ch.size(); i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
            return std::less<const float *>()(ch[x].corr_x, ch[y].corr_x);
        });
        for (size_t i = 0; i < ord.size();) {
            const ChanRef &c0 = ch[ord[i]];
            const bool two = i + 1 < ord.size() && ch[ord[i + 1]].corr_x == c0.corr_x &&
                             ch[ord[i + 1]].corr_y == c0.corr_y;
            const ChanRef &c1 = two ? ch[ord[i + 1]] : c0;
            i += two ? 2 : 1;
            PackDesc d{pl(c0, 0), two ? pl(c1, 0) : nullptr, c0.norm, two ? c1.norm : nullptr,
                       {one, zero}, {zero, one}, {c0.image, two ? c1.image : nullptr}, c0.corr_x, c0.corr_y,
                       c0.mask, two ? c1.mask : nullptr};
            out.push_back(d);
        }
        return;
    }
    for (const ChanRef &c : ch) {
        float *im[4] = {c.image, c.image + plane, c.image + 2 * plane, c.image + 3 * plane};
        if (P == 2) {
            out.push_back(PackDesc{pl(c, 0), pl(c, 1), c.norm, c.norm + 1, {h, h}, {h, mh}, {im[0], im[1]},
                                   c.corr_x, c.corr_y, c.mask, c.mask});
        } else if (poltype == MTI_POL_LINEAR) {
            out.push_back(PackDesc{pl(c, 0), pl(c, 3), c.norm, c.norm + 3, {h, h}, {h, mh}, {im[0], im[1]},
                                   c.corr_x, c.corr_y, c.mask, c.mask});  // I, Q from XX, YY
            out.push_back(PackDesc{pl(c, 1), pl(c, 2), c.norm + 1, c.norm + 2, {h, mih}, {h, ih}, {im[2], im[3]},
                                   c.corr_x, c.corr_y, c.mask, c.mask});  // U, V from XY, YX
        } else {
            out.push_back(PackDesc{pl(c, 0), pl(c, 3), c.norm, c.norm + 3, {h, h}, {h, mh}, {im[0], im[3]},
                                   c.corr_x, c.corr_y, c.mask, c.mask});  // I, V from RR, LL
            out.push_back(PackDesc{pl(c, 1), pl(c, 2), c.norm + 1, c.norm + 2, {h, ih}, {h, mih}, {im[1], im[2]},
                                   c.corr_x, c.corr_y, c.mask, c.mask});  // Q, U from RL, LR
        }
    }
}

// Transform + finalize of npk packed planes (device descriptors pd) through
// the planes X (complex128 or complex64, XT).
template <typename TI, typename XT>
static mti_status finalize_packs(mti_ctx *ctx, const FftGeom &g, const PackDesc *pd, int npk, XT *X, int64_t cs,
                                 const double2 *tw, uint32_t *queues, hipStream_t st)
{
    mti_status r;
    {
        StageTimer tm(ctx, MTI_STAGE_FFT_ROWS, st);
        r = pk_rows<TI, XT>(ctx, g, pd, npk, X, cs, tw, queues, st);
        if (r != MTI_OK) return r;
    }
    if (whole_columns(ctx, g)) {
        StageTimer tm(ctx, MTI_STAGE_FINALIZE, st);
        return pk_cols(ctx, g, X, npk, tw, pd, st);
    }
    {
        StageTimer tm(ctx, MTI_STAGE_FFT_COLS, st);
        r = fb_cols1(ctx, g, X, npk, tw, st);
        if (r != MTI_OK) return r;
    }
    StageTimer tm(ctx, MTI_STAGE_FINALIZE, st);
    return pk_cols2(ctx, g, X, npk, tw, pd, st);
}

// Planes per transform batch (FP64 scratch of at most 16 GB).
static int transform_batch(const FftGeom &g, int planes)
{
    const size_t plane_bytes = (size_t)g.nx * g.ny * sizeof(double2);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)planes, (size_t(16) << 30) / plane_bytes));
}

// ---- general sizes: HBM passes, mixed radix and Bluestein ------------------
//
// Element (b, i, v) of a batch of transforms along one axis sits at
// b * bs + i * is + v: rows of a plane are (b = plane row, i = x, V = 1,
// is = 1, bs = nx); columns are (b = plane, i = y, v = x, V = nx, is = nx,
// bs = nx * ny), so a column pass reads and writes along image rows.
// Per pass each thread runs one radix-R butterfly of a Stockham step: inputs
// i = j + r L / R (r < R), twiddles exp(+2 pi i r k / (Ns R)) with
// k = j mod Ns (Ns = product of the radices before this pass), outputs
// (j - k) R + k + r Ns -- the self-sorting form, so the result needs no
// digit reversal.

// Inverse DFT of length R (any R up to 16; the powers of two take the fixed
// butterflies above).  tw holds exp(+2 pi i t / L) with R dividing L.
template <int R>
__device__ __forceinline__ void idft_any(double2 (&v)[R], const double2 *__restrict__ tw, int L)
{
    if constexpr (R == 2 || R == 4 || R == 8 || R == 16) {
        idft<R>(v);
    } else {
        const int st = L / R;
        double2 w[R];
        w[0] = make_double2(1.0, 0.0);
#pragma unroll
        for (int m = 1; m < R; m++) w[m] = tw[m * st];
        double2 o[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            double2 acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                const double2 t = w[(r * q) % R];
                acc.x = fma(v[r].x, t.x, fma(-v[r].y, t.y, acc.x));
                acc.y = fma(v[r].x, t.y, fma(v[r].y, t.x, acc.y));
            }
            o[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = o[q];
    }
}

constexpr int GEN_THREADS = 256;

// One Stockham pass of radix R over nb * V transforms of length L.  Every
// index fits 32 bits: a batch holds at most 2^30 elements per buffer.
template <int R>
__global__ __launch_bounds__(GEN_THREADS) void k_gen_pass(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                          int L, int Ns, int V, int64_t is, int64_t bs, int nb,
                                                          const double2 *__restrict__ tw)
{
    const int per = L / R;
    const uint32_t total = (uint32_t)nb * (uint32_t)V * (uint32_t)per;
    const int step = L / (Ns * R);
    for (uint32_t t = blockIdx.x * GEN_THREADS + threadIdx.x; t < total; t += gridDim.x * GEN_THREADS) {
        const uint32_t v = t % (uint32_t)V, rest = t / (uint32_t)V;
        const uint32_t j = rest % (uint32_t)per, b = rest / (uint32_t)per;
        const double2 *s = src + (int64_t)b * bs + v;
        double2 *d = dst + (int64_t)b * bs + v;
        double2 x[R];
#pragma unroll
        for (int r = 0; r < R; r++) x[r] = s[(int64_t)(j + r * per) * is];
        const int k = (int)(j % (uint32_t)Ns);
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) x[r] = c2mul(x[r], tw[r * k * step]);
        }
        idft_any<R>(x, tw, L);
        const int64_t base = (int64_t)(j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) d[(base + (int64_t)r * Ns) * is] = x[r];
    }
}

// Bluestein, for a length N with a prime factor above 13: with
// c_n = exp(+i pi n^2 / N), X_k = c_k sum_n (x_n c_n) conj(c_(k-n)), a
// circular convolution of length M (a power of two >= 2N - 1).  With only
// inverse transforms I: T = I(a); T_k *= D_k (D_k = B_(-k) / M, B the
// forward DFT of b_m = conj(c_m), built on the host); T = I(T); the
// convolution at n is T_((-n) mod M).  Layout of T: (b, m, v) at
// b * M * V + m * V + v.
__global__ __launch_bounds__(GEN_THREADS) void k_blue_pre(const double2 *__restrict__ X, double2 *__restrict__ T,
                                                          int N, int M, int V, int64_t is, int64_t bs, int nb,
                                                          const double2 *__restrict__ chirp)
{
    const uint32_t total = (uint32_t)nb * (uint32_t)M * (uint32_t)V;
    for (uint32_t t = blockIdx.x * GEN_THREADS + threadIdx.x; t < total; t += gridDim.x * GEN_THREADS) {
        const uint32_t v = t % (uint32_t)V, rest = t / (uint32_t)V;
        const uint32_t m = rest % (uint32_t)M, b = rest / (uint32_t)M;
        double2 a = make_double2(0.0, 0.0);
        if ((int)m < N) a = c2mul(X[(int64_t)b * bs + (int64_t)m * is + v], chirp[m]);
        T[t] = a;
    }
}

__global__ __launch_bounds__(GEN_THREADS) void k_blue_mid(double2 *__restrict__ T, int M, int V, int nb,
                                                          const double2 *__restrict__ D)
{
    const uint32_t total = (uint32_t)nb * (uint32_t)M * (uint32_t)V;
    for (uint32_t t = blockIdx.x * GEN_THREADS + threadIdx.x; t < total; t += gridDim.x * GEN_THREADS)
        T[t] = c2mul(T[t], D[(t / (uint32_t)V) % (uint32_t)M]);
}

__global__ __launch_bounds__(GEN_THREADS) void k_blue_post(const double2 *__restrict__ T, double2 *__restrict__ X,
                                                           int N, int M, int V, int64_t is, int64_t bs, int nb,
                                                           const double2 *__restrict__ chirp)
{
    const uint32_t total = (uint32_t)nb * (uint32_t)N * (uint32_t)V;
    for (uint32_t t = blockIdx.x * GEN_THREADS + threadIdx.x; t < total; t += gridDim.x * GEN_THREADS) {
        const uint32_t v = t % (uint32_t)V, rest = t / (uint32_t)V;
        const uint32_t n = rest % (uint32_t)N, b = rest / (uint32_t)N;
        const uint32_t m = n ? (uint32_t)M - n : 0u;
        X[(int64_t)b * bs + (int64_t)n * is + v] = c2mul(T[((int64_t)b * M + m) * V + v], chirp[n]);
    }
}

// Packed plane Z = H0 + i H1 of pack blockIdx.y at every (y, x) of the
// transform, from the grids with the PreIFFTUVCorrection reorder (transform
// index i holds grid cell (i + n/2) mod n on each axis, op/common.hcu:122-136
// for odd and even n) -- the load of k_rows_pk for any size.
template <typename TI>
__global__ __launch_bounds__(GEN_THREADS) void k_gen_pack(const PackDesc *pd, double2 *X, int nx, int ny, int64_t cs)
{
    const PackDesc &d = pd[blockIdx.y];
    const TI *A = (const TI *)d.a, *B = (const TI *)d.b;
    const double ia = 1.0 / *d.na, ib = B ? 1.0 / *d.nb : 0.0;
    const double2 ka0 = make_double2(d.ca[0].x * ia, d.ca[0].y * ia), ka1 = make_double2(d.ca[1].x * ia, d.ca[1].y * ia);
    const double2 kb0 = make_double2(d.cb[0].x * ib, d.cb[0].y * ib), kb1 = make_double2(d.cb[1].x * ib, d.cb[1].y * ib);
    const int ntx = (nx + 31) / 32, hx = nx / 2, hy = ny / 2;
    const int64_t n = (int64_t)nx * ny;
    double2 *dst = X + (int64_t)blockIdx.y * n;
    auto elem = [&](const TI *P0, const uint8_t *m, int gy, int gx) {
        if (m && !m[(gy >> 5) * ntx + (gx >> 5)]) return make_double2(0.0, 0.0);
        const TI v = P0[((int64_t)gy * nx + gx) * cs];
        return make_double2((double)v.x, (double)v.y);
    };
    for (int64_t i = (int64_t)blockIdx.x * GEN_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * GEN_THREADS) {
        const int y = (int)(i / nx), x = (int)(i - (int64_t)y * nx);
        const int ym = y ? ny - y : 0, xm = x ? nx - x : 0;
        int gy = y + hy, gx = x + hx, gym = ym + hy, gxm = xm + hx;
        if (gy >= ny) gy -= ny;
        if (gx >= nx) gx -= nx;
        if (gym >= ny) gym -= ny;
        if (gxm >= nx) gxm -= nx;
        const double2 a0 = elem(A, d.ma, gy, gx), a1 = elem(A, d.ma, gym, gxm);
        double2 b0 = make_double2(0.0, 0.0), b1 = b0;
        if (B) {
            b0 = elem(B, d.mb, gy, gx);
            b1 = elem(B, d.mb, gym, gxm);
        }
        const double2 s0 = c2add(c2mul(ka0, a0), c2mul(kb0, b0)), s0m = c2add(c2mul(ka0, a1), c2mul(kb0, b1));
        const double2 s1 = c2add(c2mul(ka1, a0), c2mul(kb1, b0)), s1m = c2add(c2mul(ka1, a1), c2mul(kb1, b1));
        const double h0r = 0.5 * (s0.x + s0m.x), h0i = 0.5 * (s0.y - s0m.y);
        const double h1r = 0.5 * (s1.x + s1m.x), h1i = 0.5 * (s1.y - s1m.y);
        dst[i] = make_double2(h0r - h1i, h0i + h1r);
    }
}

// Images of pack blockIdx.y: fftshift (output index (k + n/2) mod n,
// ZeroToCentre2DShift.cu:31-91 for odd and even n), the m flip and both
// corrections; the real part is image 0, the imaginary part image 1.
__global__ __launch_bounds__(GEN_THREADS) void k_gen_store(const PackDesc *pd, const double2 *X, int nx, int ny)
{
    const PackDesc &d = pd[blockIdx.y];
    const int64_t n = (int64_t)nx * ny;
    const double2 *src = X + (int64_t)blockIdx.y * n;
    float *i0 = d.img[0], *i1 = d.img[1];
    const int hx = nx / 2, hy = ny / 2;
    for (int64_t i = (int64_t)blockIdx.x * GEN_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * GEN_THREADS) {
        const int ky = (int)(i / nx), kx = (int)(i - (int64_t)ky * nx);
        int jy = ky + hy, jx = kx + hx;
        if (jy >= ny) jy -= ny;
        if (jx >= nx) jx -= nx;
        const int row = ny - 1 - jy;  // m axis flipped for FITS
        const double2 v = src[i];
        const double cf = (double)d.corr_x[jx] * (double)d.corr_y[row];
        const int64_t o = (int64_t)row * nx + jx;
        i0[o] = (float)(v.x * cf);
        if (i1) i1[o] = (float)(v.y * cf);
    }
}

// complex64 <-> complex128 planes of mti_fft2d_inverse.
__global__ __launch_bounds__(GEN_THREADS) void k_gen_widen(const float2 *a, double2 *X, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * GEN_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * GEN_THREADS)
        X[i] = make_double2((double)a[i].x, (double)a[i].y);
}
__global__ __launch_bounds__(GEN_THREADS) void k_gen_narrow(const double2 *X, float2 *a, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * GEN_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * GEN_THREADS)
        a[i] = make_float2((float)X[i].x, (float)X[i].y);
}

// One axis of the general transform (host side).
struct GenAxis {
    int N = 1;                 // transform length
    int L = 1;                 // length of the Stockham passes: N, or M for Bluestein
    bool blue = false;
    std::vector<int> radices;  // factors of L in pass order
    size_t tw = 0, chirp = 0, dk = 0;  // offsets (elements) in the table blob: tw[L], chirp[N], D[L]
};

struct GenPlan {
    GenAxis x, y;
    size_t elems = 0;  // double2 elements of the table blob
};

// Radices 16, 8, 4, 2, 3, 5, 7, 11, 13; false if a larger prime factor is left.
static bool gen_factor(int L, std::vector<int> &out)
{
    out.clear();
    while (L % 16 == 0) {
        out.push_back(16);
        L /= 16;
    }
    for (int r : {8, 4, 2})
        if (L % r == 0) {
            out.push_back(r);
            L /= r;
        }
    for (int r : {3, 5, 7, 11, 13})
        while (L % r == 0) {
            out.push_back(r);
            L /= r;
        }
    return L == 1;
}

static GenAxis gen_axis(int N, size_t &off)
{
    GenAxis a;
    a.N = N;
    if (N <= 1) return a;
    if (gen_factor(N, a.radices)) {
        a.L = N;
    } else {
        a.blue = true;
        int M = 1;
        while (M < 2 * N - 1) M <<= 1;
        a.L = M;
        gen_factor(M, a.radices);
    }
    a.tw = off;
    off += a.L;
    if (a.blue) {
        a.chirp = off;
        off += N;
        a.dk = off;
        off += a.L;
    }
    return a;
}

static GenPlan gen_plan(int nx, int ny)
{
    GenPlan g;
    size_t off = 0;
    g.x = gen_axis(nx, off);
    g.y = gen_axis(ny, off);
    g.elems = std::max<size_t>(off, 1);
    return g;
}

// Forward DFT of a power-of-two length on the host (FP64, radix 2, twiddles
// from cos/sin of the exact angle).
static void host_fft_fwd(std::vector<double2> &a)
{
    const int n = (int)a.size();
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int h = len / 2;
        for (int k = 0; k < h; k++) {
            const double ang = -2.0 * M_PI * (double)k / (double)len;
            const double wr = std::cos(ang), wi = std::sin(ang);
            for (int i = 0; i < n; i += len) {
                double2 &u = a[i + k], &v = a[i + k + h];
                const double tr = v.x * wr - v.y * wi, ti = v.x * wi + v.y * wr;
                v = make_double2(u.x - tr, u.y - ti);
                u = make_double2(u.x + tr, u.y + ti);
            }
        }
    }
}

static void gen_fill(const GenAxis &a, std::vector<double2> &h)
{
    if (a.N <= 1) return;
    for (int t = 0; t < a.L; t++) {
        const double ang = 2.0 * M_PI * (double)t / (double)a.L;
        h[a.tw + t] = make_double2(std::cos(ang), std::sin(ang));
    }
    if (!a.blue) return;
    const int N = a.N, M = a.L;
    for (int n = 0; n < N; n++) {
        const int64_t q = ((int64_t)n * n) % (2 * (int64_t)N);  // exact argument
        const double ang = M_PI * (double)q / (double)N;
        h[a.chirp + n] = make_double2(std::cos(ang), std::sin(ang));
    }
    std::vector<double2> b(M, make_double2(0.0, 0.0));
    for (int n = 0; n < N; n++) {
        const double2 c = h[a.chirp + n];
        b[n] = make_double2(c.x, -c.y);
        if (n) b[M - n] = b[n];
    }
    host_fft_fwd(b);
    for (int k = 0; k < M; k++) {
        const double2 v = b[k ? M - k : 0];
        h[a.dk + k] = make_double2(v.x / M, v.y / M);
    }
}

// The table blob of (nx, ny), cached per size on the context (or, in a
// caller workspace, uploaded into its slot for the call).
static mti_status gen_tables(mti_ctx *ctx, const GenPlan &g, const double2 **out, hipStream_t st)
{
    for (auto &t : ctx->gen_tables)
        if (t.nx == g.x.N && t.ny == g.y.N) {
            *out = t.ptr;
            return MTI_OK;
        }
    std::vector<double2> h(g.elems, make_double2(0.0, 0.0));
    gen_fill(g.x, h);
    gen_fill(g.y, h);
    const size_t bytes = g.elems * sizeof(double2);
    if (ctx->ws_active) {
        void *w = nullptr;
        mti_status r = scratch_get(ctx, SLOT_GTAB, bytes, &w);
        if (r != MTI_OK) return r;
        if ((r = upload(ctx, w, h.data(), bytes, st)) != MTI_OK) return r;
        if ((r = host_sync(ctx, st)) != MTI_OK) return r;
        *out = (const double2 *)w;
        return MTI_OK;
    }
    double2 *d = nullptr;
    MTI_HIP(ctx, hipMalloc(&d, bytes));
    MTI_HIP(ctx, hipMemcpyAsync(d, h.data(), bytes, hipMemcpyHostToDevice, st));
    MTI_HIP(ctx, hipStreamSynchronize(st));
    ctx->gen_tables.push_back({g.x.N, g.y.N, d, bytes});
    *out = d;
    return MTI_OK;
}

static unsigned gen_blocks(uint64_t threads)
{
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((threads + GEN_THREADS - 1) / GEN_THREADS, 1u << 20));
}

static mti_status gen_pass(mti_ctx *ctx, int R, const double2 *src, double2 *dst, int L, int Ns, int V, int64_t is,
                           int64_t bs, int nb, const double2 *tw, hipStream_t st)
{
    const dim3 gr(gen_blocks((uint64_t)nb * V * (L / R))), bl(GEN_THREADS);
    switch (R) {
#define MTI_GP(RR) \
    case RR: hipLaunchKernelGGL(k_gen_pass<RR>, gr, bl, 0, st, src, dst, L, Ns, V, is, bs, nb, tw); break;
        MTI_GP(2) MTI_GP(3) MTI_GP(4) MTI_GP(5) MTI_GP(7) MTI_GP(8) MTI_GP(11) MTI_GP(13) MTI_GP(16)
#undef MTI_GP
    default: return set_error(ctx, MTI_ERR_UNSUPPORTED, "FFT radix");
    }
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

// Stockham passes over a length-L layout; the data moves between *a and *b
// and ends in *a.
static mti_status gen_passes(mti_ctx *ctx, const std::vector<int> &radices, double2 *&a, double2 *&b, int L, int V,
                             int64_t is, int64_t bs, int nb, const double2 *tw, hipStream_t st)
{
    int Ns = 1;
    for (int R : radices) {
        mti_status r = gen_pass(ctx, R, a, b, L, Ns, V, is, bs, nb, tw, st);
        if (r != MTI_OK) return r;
        std::swap(a, b);
        Ns *= R;
    }
    return MTI_OK;
}

// One axis of nb * V transforms held in *X (ping-pong buffer *Y; Bluestein
// buffers T, U); the result ends in *X.
static mti_status gen_axis_run(mti_ctx *ctx, const GenAxis &ax, const double2 *tab, double2 *&X, double2 *&Y,
                               double2 *T, double2 *U, int V, int64_t is, int64_t bs, int nb, hipStream_t st)
{
    if (ax.N <= 1) return MTI_OK;
    const double2 *tw = tab + ax.tw;
    if (!ax.blue) return gen_passes(ctx, ax.radices, X, Y, ax.L, V, is, bs, nb, tw, st);
    const int N = ax.N, M = ax.L;
    const uint64_t tot = (uint64_t)nb * M * V;
    hipLaunchKernelGGL(k_blue_pre, dim3(gen_blocks(tot)), dim3(GEN_THREADS), 0, st, X, T, N, M, V, is, bs, nb,
                       tab + ax.chirp);
    MTI_HIP(ctx, hipGetLastError());
    double2 *a = T, *b = U;
    mti_status r = gen_passes(ctx, ax.radices, a, b, M, V, (int64_t)V, (int64_t)M * V, nb, tw, st);
    if (r != MTI_OK) return r;
    hipLaunchKernelGGL(k_blue_mid, dim3(gen_blocks(tot)), dim3(GEN_THREADS), 0, st, a, M, V, nb, tab + ax.dk);
    MTI_HIP(ctx, hipGetLastError());
    if ((r = gen_passes(ctx, ax.radices, a, b, M, V, (int64_t)V, (int64_t)M * V, nb, tw, st)) != MTI_OK) return r;
    hipLaunchKernelGGL(k_blue_post, dim3(gen_blocks((uint64_t)nb * N * V)), dim3(GEN_THREADS), 0, st, a, X, N, M, V,
                       is, bs, nb, tab + ax.chirp);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

// Unnormalised 2-D inverse transform of nb planes [nb][ny][nx] held in *X.
static mti_status gen_transform(mti_ctx *ctx, const GenPlan &g, const double2 *tab, double2 *&X, double2 *&Y,
                                double2 *T, double2 *U, int nb, hipStream_t st)
{
    const int nx = g.x.N, ny = g.y.N;
    mti_status r = gen_axis_run(ctx, g.x, tab, X, Y, T, U, 1, 1, nx, nb * ny, st);
    if (r != MTI_OK) return r;
    return gen_axis_run(ctx, g.y, tab, X, Y, T, U, nx, nx, (int64_t)nx * ny, nb, st);
}

// Scratch of one general-path batch of nb planes: X, Y and (Bluestein) T, U.
static size_t gen_plane_elems(const GenPlan &g) { return (size_t)g.x.N * g.y.N; }
static size_t gen_blue_elems(const GenPlan &g)
{
    size_t t = 0;
    if (g.x.blue) t = std::max(t, (size_t)g.x.L * g.y.N);
    if (g.y.blue) t = std::max(t, (size_t)g.y.L * g.x.N);
    return t;
}
static size_t gen_batch_bytes(const GenPlan &g, int nb)
{
    return (2 * gen_plane_elems(g) + 2 * gen_blue_elems(g)) * sizeof(double2) * (size_t)nb;
}
// Planes per general-path batch: at most 16 GB of scratch (2^30 elements).
static int gen_batch(const GenPlan &g, int planes)
{
    const size_t per = gen_batch_bytes(g, 1);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)planes, (size_t(16) << 30) / per));
}
static void gen_buffers(const GenPlan &g, int nb, void *w, double2 **X, double2 **Y, double2 **T, double2 **U)
{
    double2 *p = (double2 *)w;
    const size_t pe = gen_plane_elems(g) * nb, be = gen_blue_elems(g) * nb;
    *X = p;
    *Y = p + pe;
    *T = p + 2 * pe;
    *U = p + 2 * pe + be;
}

// General-path finalize of npk packs (device descriptors pd).
template <typename TI>
static mti_status gen_finalize_packs(mti_ctx *ctx, const GenPlan &g, const double2 *tab, const PackDesc *pd,
                                     int npk, void *w, int64_t cs, hipStream_t st)
{
    const int nx = g.x.N, ny = g.y.N;
    double2 *X, *Y, *T, *U;
    gen_buffers(g, npk, w, &X, &Y, &T, &U);
    const dim3 gr(gen_blocks((uint64_t)nx * ny / 4 + 1), (unsigned)npk);
    {
        StageTimer tm(ctx, MTI_STAGE_FFT_ROWS, st);
        hipLaunchKernelGGL(k_gen_pack<TI>, gr, dim3(GEN_THREADS), 0, st, pd, X, nx, ny, cs);
        MTI_HIP(ctx, hipGetLastError());
        mti_status r = gen_axis_run(ctx, g.x, tab, X, Y, T, U, 1, 1, nx, npk * ny, st);
        if (r != MTI_OK) return r;
    }
    {
        StageTimer tm(ctx, MTI_STAGE_FFT_COLS, st);
        mti_status r = gen_axis_run(ctx, g.y, tab, X, Y, T, U, nx, nx, (int64_t)nx * ny, npk, st);
        if (r != MTI_OK) return r;
    }
    StageTimer tm(ctx, MTI_STAGE_FINALIZE, st);
    hipLaunchKernelGGL(k_gen_store, gr, dim3(GEN_THREADS), 0, st, pd, X, nx, ny);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

// Uploads the descriptors of a call (once; the copy completes before any
// batch is enqueued, so later batches never overwrite what earlier ones read).
template <typename D>
static mti_status upload_desc(mti_ctx *ctx, const std::vector<D> &h, D **d, hipStream_t st, size_t extra = 0)
{
    void *w = nullptr;
    mti_status r = scratch_get(ctx, SLOT_CHANDESC, desc_region(h.size(), sizeof(D)) + extra, &w);
    if (r != MTI_OK) return r;
    if ((r = upload(ctx, w, h.data(), h.size() * sizeof(D), st)) != MTI_OK) return r;
    if ((r = host_sync(ctx, st)) != MTI_OK) return r;
    *d = (D *)w;
    return MTI_OK;
}

// Transform + finalize of a channel list in batches of planes.
static mti_status finalize_channels(mti_ctx *ctx, const FftGeom &g, const std::vector<ChanRef> &ch, int P,
                                    int poltype, int layout, bool c64, hipStream_t st)
{
    std::vector<PackDesc> h;
    make_packs(ch, P, poltype, layout, g.nx, g.ny, c64 ? sizeof(float2) : sizeof(double2), h);
    PackDesc *pd = nullptr;
    mti_status r = upload_desc(ctx, h, &pd, st, PK_QUEUE_BYTES);
    if (r != MTI_OK) return r;
    const int npk = (int)h.size();
    // the row pass's unit queues sit behind the descriptors
    uint32_t *queues = (uint32_t *)((char *)pd + desc_region(h.size(), sizeof(PackDesc)));
    const int64_t cs = layout == MTI_GRID_VUP ? P : 1;
    if (!pk_ok(g)) {
        // general sizes: complex128 planes, HBM passes
        const GenPlan gp = gen_plan(g.nx, g.ny);
        const double2 *tab = nullptr;
        if ((r = gen_tables(ctx, gp, &tab, st)) != MTI_OK) return r;
        const int batch = gen_batch(gp, npk);
        void *w = nullptr;
        if ((r = scratch_get(ctx, SLOT_FFT, gen_batch_bytes(gp, batch), &w)) != MTI_OK) return r;
        for (int k0 = 0; k0 < npk; k0 += batch) {
            const int nk = std::min(batch, npk - k0);
            r = c64 ? gen_finalize_packs<float2>(ctx, gp, tab, pd + k0, nk, w, cs, st)
                    : gen_finalize_packs<double2>(ctx, gp, tab, pd + k0, nk, w, cs, st);
            if (r != MTI_OK) return r;
        }
        return MTI_OK;
    }
    const double2 *tw = nullptr;
    r = twiddle_table(ctx, g.NT, &tw, st);
    if (r != MTI_OK) return r;
    // the batch is sized for complex128 planes (the workspace plans do the
    // same), so complex64 planes take half of what was planned
    const int batch = transform_batch(g, npk);
    const bool x64 = ctx->xform_c64 != 0;
    void *w = nullptr;
    r = scratch_get(ctx, SLOT_FFT, (size_t)g.nx * g.ny * (x64 ? sizeof(float2) : sizeof(double2)) * batch, &w);
    if (r != MTI_OK) return r;
    for (int k0 = 0; k0 < npk; k0 += batch) {
        const int nk = std::min(batch, npk - k0);
        if (x64)
            r = c64 ? finalize_packs<float2>(ctx, g, pd + k0, nk, (float2 *)w, cs, tw, queues, st)
                    : finalize_packs<double2>(ctx, g, pd + k0, nk, (float2 *)w, cs, tw, queues, st);
        else
            r = c64 ? finalize_packs<float2>(ctx, g, pd + k0, nk, (double2 *)w, cs, tw, queues, st)
                    : finalize_packs<double2>(ctx, g, pd + k0, nk, (double2 *)w, cs, tw, queues, st);
        if (r != MTI_OK) return r;
    }
    return MTI_OK;
}

// ---- scratch plans (the requests of each entry point, largest per slot) ----
static void plan_twiddles(int NT, ScratchPlan &pl) { pl.need(SLOT_TWIDDLE, (size_t)NT * sizeof(double2)); }

static void plan_gen(const GenPlan &gp, int planes, ScratchPlan &pl)
{
    pl.need(SLOT_GTAB, gp.elems * sizeof(double2));
    pl.need(SLOT_FFT, gen_batch_bytes(gp, gen_batch(gp, planes)));
}

static void plan_finalize_channels(const FftGeom &g, int npk, ScratchPlan &pl)
{
    pl.need(SLOT_CHANDESC, desc_region((size_t)npk, sizeof(PackDesc)) + PK_QUEUE_BYTES);
    if (!pk_ok(g)) {
        plan_gen(gen_plan(g.nx, g.ny), npk, pl);
        return;
    }
    plan_twiddles(g.NT, pl);
    pl.need(SLOT_FFT, (size_t)g.nx * g.ny * sizeof(double2) * transform_batch(g, npk));
}

// packed planes of nchan channels of P correlations (one correlation: two
// channels per pack, all of them sharing one correction here)
static int pack_count(int P, int nchan) { return P == 1 ? (nchan + 1) / 2 : P == 2 ? nchan : 2 * nchan; }

static void plan_finalize(const mti_image_params *p, int npol, int nchan, ScratchPlan &pl)
{
    const FftGeom g = geom(p->nx, p->ny);
    pl.need(SLOT_FCORR, (size_t)(p->nx + p->ny + 128) * 4);
    plan_finalize_channels(g, pack_count(npol, nchan), pl);
}

static void plan_fft2d(int nx, int ny, int nbatch, ScratchPlan &pl) { plan_gen(gen_plan(nx, ny), nbatch, pl); }

static void plan_finalize_jobs(int32_t njobs, const mti_job *jobs, int P, ScratchPlan &pl)
{
    std::vector<std::array<int, 3>> ckey;
    size_t ctot = 0;
    std::vector<std::pair<int, int>> sizes;
    for (int i = 0; i < njobs; i++) {
        const mti_image_params &p = jobs[i].params;
        const std::array<int, 3> key{p.nx, p.ny, p.sampling};
        if (std::find(ckey.begin(), ckey.end(), key) == ckey.end()) {
            ckey.push_back(key);
            ctot += (size_t)p.nx + p.ny + 64;
        }
        const std::pair<int, int> sz{p.nx, p.ny};
        if (std::find(sizes.begin(), sizes.end(), sz) == sizes.end()) sizes.push_back(sz);
    }
    pl.need(SLOT_CORR, ctot * sizeof(float));
    for (auto sz : sizes) {
        const FftGeom g = geom(sz.first, sz.second);
        int nch = 0;
        for (int i = 0; i < njobs; i++)
            if (jobs[i].params.nx == sz.first && jobs[i].params.ny == sz.second) nch += jobs[i].nchan;
        // one correlation: at most one pack per channel (channels pair only
        // within a correction group)
        plan_finalize_channels(g, P == 1 ? nch : pack_count(P, nch), pl);
    }
}

}  // namespace mti

using namespace mti;

extern "C" mti_status mti_fft2d_inverse(mti_ctx *ctx, float *data, int32_t nx, int32_t ny, int32_t nbatch,
                                        void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, nx >= 1 && ny >= 1 && nx <= 32768 && ny <= 32768, "FFT sizes must be in [1, 32768]");
    MTI_CHECK(ctx, data && nbatch >= 1, "null data or nbatch < 1");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    const GenPlan gp = gen_plan(nx, ny);
    mti_status r;
    CallScope cs(ctx);
    if (cs.planning()) {
        ScratchPlan pl;
        plan_fft2d(nx, ny, nbatch, pl);
        if ((r = cs.begin(pl)) != MTI_OK) return r;
    }
    const double2 *tab = nullptr;
    if ((r = gen_tables(ctx, gp, &tab, st)) != MTI_OK) return r;
    const int batch = gen_batch(gp, nbatch);
    void *w = nullptr;
    if ((r = scratch_get(ctx, SLOT_FFT, gen_batch_bytes(gp, batch), &w)) != MTI_OK) return r;
    const int64_t plane = (int64_t)nx * ny;
    for (int b0 = 0; b0 < nbatch; b0 += batch) {
        const int nb = std::min(batch, nbatch - b0);
        double2 *X, *Y, *T, *U;
        gen_buffers(gp, nb, w, &X, &Y, &T, &U);
        float2 *a = (float2 *)data + (int64_t)b0 * plane;
        const int64_t n = plane * nb;
        hipLaunchKernelGGL(k_gen_widen, dim3(gen_blocks((uint64_t)n / 4 + 1)), dim3(GEN_THREADS), 0, st, a, X, n);
        MTI_HIP(ctx, hipGetLastError());
        if ((r = gen_transform(ctx, gp, tab, X, Y, T, U, nb, st)) != MTI_OK) return r;
        hipLaunchKernelGGL(k_gen_narrow, dim3(gen_blocks((uint64_t)n / 4 + 1)), dim3(GEN_THREADS), 0, st, X, a, n);
        MTI_HIP(ctx, hipGetLastError());
    }
    return MTI_OK;
}

extern "C" mti_status mti_finalize(mti_ctx *ctx, const mti_image_params *p, int32_t npol, int32_t poltype,
                                   int32_t nchan, int32_t grid_layout, int32_t grid_precision,
                                   const void *grids, const double *norm, float *image, void *stream)
{
    return mti_finalize_sparse(ctx, p, npol, poltype, nchan, grid_layout, grid_precision, grids, nullptr, norm,
                               image, stream);
}

extern "C" mti_status mti_finalize_sparse(mti_ctx *ctx, const mti_image_params *p, int32_t npol, int32_t poltype,
                                          int32_t nchan, int32_t grid_layout, int32_t grid_precision,
                                          const void *grids, const uint8_t *tile_mask, const double *norm,
                                          float *image, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = validate_params(ctx, p);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, npol == 1 || npol == 2 || npol == 4, "npol must be 1, 2 or 4");
    MTI_CHECK(ctx, poltype == MTI_POL_LINEAR || poltype == MTI_POL_CIRCULAR, "bad PolarizationType");
    MTI_CHECK(ctx, nchan >= 1 && grids && norm && image, "null pointer or nchan < 1");
    MTI_CHECK(ctx, grid_layout == MTI_GRID_VUP || grid_layout == MTI_GRID_PVU, "bad grid_layout");
    MTI_CHECK(ctx, grid_precision == MTI_GRID_C64 || grid_precision == MTI_GRID_C128, "bad grid_precision");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    FftGeom g = geom(p->nx, p->ny);
    CallScope cs(ctx);
    if (cs.planning()) {
        ScratchPlan pl;
        plan_finalize(p, npol, nchan, pl);
        if ((r = cs.begin(pl)) != MTI_OK) return r;
    }
    // correction vectors
    void *wc = nullptr;
    r = scratch_get(ctx, SLOT_FCORR, (size_t)(p->nx + p->ny + 128) * 4, &wc);
    if (r != MTI_OK) return r;
    float *corr_x = (float *)wc;
    float *corr_y = corr_x + p->nx + 64;
    hipLaunchKernelGGL(k_corr, dim3((unsigned)ceil_div(p->nx, 256)), dim3(256), 0, st, corr_x, p->nx, p->sampling);
    hipLaunchKernelGGL(k_corr, dim3((unsigned)ceil_div(p->ny, 256)), dim3(256), 0, st, corr_y, p->ny, p->sampling);
    MTI_HIP(ctx, hipGetLastError());
    const size_t elem = grid_precision == MTI_GRID_C64 ? sizeof(float2) : sizeof(double2);
    const size_t chan_grid = (size_t)p->nx * p->ny * npol;
    const size_t tiles = (size_t)((p->nx + 31) / 32) * ((p->ny + 31) / 32);
    std::vector<ChanRef> ch(nchan);
    for (int c = 0; c < nchan; c++)
        ch[c] = ChanRef{(const char *)grids + elem * chan_grid * c, norm + (size_t)c * npol, image + chan_grid * c,
                        corr_x, corr_y, tile_mask ? tile_mask + tiles * c : nullptr};
    return finalize_channels(ctx, g, ch, npol, poltype, grid_layout, grid_precision == MTI_GRID_C64, st);
}

extern "C" mti_status mti_finalize_jobs(mti_ctx *ctx, int32_t njobs, const mti_job *jobs, int32_t poltype,
                                        int32_t grid_layout, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, njobs >= 0 && (njobs == 0 || jobs), "bad job list");
    MTI_CHECK(ctx, poltype == MTI_POL_LINEAR || poltype == MTI_POL_CIRCULAR, "bad PolarizationType");
    MTI_CHECK(ctx, grid_layout == MTI_GRID_VUP || grid_layout == MTI_GRID_PVU, "bad grid_layout");
    if (njobs == 0) return MTI_OK;
    const int P = jobs[0].vis.npol;
    for (int i = 0; i < njobs; i++) {
        mti_status r = validate_params(ctx, &jobs[i].params);
        if (r != MTI_OK) return r;
        MTI_CHECK(ctx, jobs[i].vis.npol == P, "jobs of one call must share npol");
        MTI_CHECK(ctx, jobs[i].grids && jobs[i].norm && jobs[i].image && jobs[i].nchan >= 1, "null job outputs");
        MTI_CHECK(ctx, jobs[i].grids_c64 == 0 || jobs[i].grids_c64 == 1, "job grids_c64 must be 0 or 1");
    }
    MTI_CHECK(ctx, P == 1 || P == 2 || P == 4, "npol must be 1, 2 or 4");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = to_stream(stream);
    CallScope cs(ctx);
    if (cs.planning()) {
        ScratchPlan pl;
        plan_finalize_jobs(njobs, jobs, P, pl);
        mti_status rp = cs.begin(pl);
        if (rp != MTI_OK) return rp;
    }
    // correction vectors, once per distinct (nx, ny, sampling): they depend on
    // nothing else, and jobs that share them can share packed planes
    std::vector<std::array<int, 3>> ckey;
    std::vector<size_t> cbase, coff(njobs);
    size_t ctot = 0;
    for (int i = 0; i < njobs; i++) {
        const mti_image_params &p = jobs[i].params;
        const std::array<int, 3> key{p.nx, p.ny, p.sampling};
        size_t k = std::find(ckey.begin(), ckey.end(), key) - ckey.begin();
        if (k == ckey.size()) {
            ckey.push_back(key);
            cbase.push_back(ctot);
            ctot += (size_t)p.nx + p.ny + 64;
        }
        coff[i] = cbase[k];
    }
    void *wc = nullptr;
    mti_status r = scratch_get(ctx, SLOT_CORR, ctot * sizeof(float), &wc);
    if (r != MTI_OK) return r;
    float *corr = (float *)wc;
    for (size_t k = 0; k < ckey.size(); k++) {
        const int nx = ckey[k][0], ny = ckey[k][1], smp = ckey[k][2];
        hipLaunchKernelGGL(k_corr, dim3((unsigned)ceil_div(nx, 256)), dim3(256), 0, st, corr + cbase[k], nx, smp);
        hipLaunchKernelGGL(k_corr, dim3((unsigned)ceil_div(ny, 256)), dim3(256), 0, st, corr + cbase[k] + nx, ny,
                           smp);
    }
    MTI_HIP(ctx, hipGetLastError());
    // one packed-plane transform per image size (and grid element type) over
    // the channels of all its jobs
    std::vector<std::array<int, 3>> sizes;
    for (int i = 0; i < njobs; i++) {
        const std::array<int, 3> sz{jobs[i].params.nx, jobs[i].params.ny, jobs[i].grids_c64 ? 1 : 0};
        if (std::find(sizes.begin(), sizes.end(), sz) == sizes.end()) sizes.push_back(sz);
    }
    for (const auto &sz : sizes) {
        const FftGeom g = geom(sz[0], sz[1]);
        const bool c64 = sz[2] != 0;
        auto member = [&](const mti_job &j) {
            return j.params.nx == sz[0] && j.params.ny == sz[1] && (j.grids_c64 != 0) == c64;
        };
        const size_t chan_grid = (size_t)sz[0] * sz[1] * P;
        const size_t elem = c64 ? sizeof(float2) : sizeof(double2);
        std::vector<ChanRef> ch;
        for (int i = 0; i < njobs; i++) {
            const mti_job &j = jobs[i];
            if (!member(j)) continue;
            const size_t tiles = (size_t)((sz[0] + 31) / 32) * ((sz[1] + 31) / 32);
            for (int c = 0; c < j.nchan; c++)
                ch.push_back(ChanRef{(const char *)j.grids + chan_grid * c * elem, j.norm + (size_t)c * P,
                                     j.image + chan_grid * c, corr + coff[i], corr + coff[i] + sz[0],
                                     j.tile_mask ? j.tile_mask + tiles * c : nullptr});
        }
        r = finalize_channels(ctx, g, ch, P, poltype, grid_layout, c64, st);
        if (r != MTI_OK) return r;
    }
    return MTI_OK;
}

extern "C" mti_status mti_fft2d_workspace_size(int32_t nx, int32_t ny, int32_t nbatch, size_t *bytes)
{
    if (!bytes || nx < 1 || ny < 1 || nx > 32768 || ny > 32768 || nbatch < 1) return MTI_ERR_INVALID;
    ScratchPlan pl;
    plan_fft2d(nx, ny, nbatch, pl);
    *bytes = pl.total();
    return MTI_OK;
}

extern "C" mti_status mti_finalize_workspace_size(const mti_image_params *p, int32_t npol, int32_t nchan,
                                                  size_t *bytes)
{
    if (!bytes || !p || validate_params(nullptr, p) != MTI_OK ||
        !(npol == 1 || npol == 2 || npol == 4) || nchan < 1)
        return MTI_ERR_INVALID;
    ScratchPlan pl;
    plan_finalize(p, npol, nchan, pl);
    *bytes = pl.total();
    return MTI_OK;
}

extern "C" mti_status mti_finalize_jobs_workspace_size(int32_t njobs, const mti_job *jobs, size_t *bytes)
{
    if (!bytes || njobs < 0 || (njobs > 0 && !jobs)) return MTI_ERR_INVALID;
    if (njobs == 0) {
        *bytes = ScratchPlan().total();
        return MTI_OK;
    }
    const int P = jobs[0].vis.npol;
    if (!(P == 1 || P == 2 || P == 4)) return MTI_ERR_INVALID;
    for (int i = 0; i < njobs; i++)
        if (validate_params(nullptr, &jobs[i].params) != MTI_OK ||
            jobs[i].vis.npol != P || jobs[i].nchan < 1)
            return MTI_ERR_INVALID;
    ScratchPlan pl;
    plan_finalize_jobs(njobs, jobs, P, pl);
    *bytes = pl.total();
    return MTI_OK;
}

extern "C" mti_status mti_finalize_out_shape(const mti_image_params *p, int32_t npol, int32_t nchan,
                                             int64_t image_dims[4])
{
    if (!p || !image_dims || validate_params(nullptr, p) != MTI_OK ||
        !(npol == 1 || npol == 2 || npol == 4) || nchan < 1)
        return MTI_ERR_INVALID;
    image_dims[0] = nchan;
    image_dims[1] = npol;
    image_dims[2] = p->ny;
    image_dims[3] = p->nx;
    return MTI_OK;
}
