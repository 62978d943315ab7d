/*
 * mti_oracle.c -- FP64 CPU restatement of mt-imager's imaging path.
 *
 * TEST INFRASTRUCTURE ONLY (see mti_oracle.h).  Not part of the product.
 *
 * Reference paths below are relative to code/mt-imager/lib/ ("op/" =
 * operators/, "cfg/" = components/ConvolutionFunctionsGenerator/, "comp/" =
 * components/, "vis/" = components/VisibilityManager/) and to
 * code/gafw/Implimentation/StandardOperators/ ("std/").
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_gridsf          closed-form values, branch continuity at 0.75 and the
 *                       spheroidal eigen-property (finite Fourier transform of
 *                       (1-x^2) psi(x) is proportional to psi).
 *   orc_kernels_build   numpy.fft.fft2 of the full zero-padded, ifftshifted
 *                       screen followed by fftshift (the reference's cuFFT
 *                       pipeline) on small configs; w=0 kernel vs the
 *                       analytic gridding function; brute-force support find.
 *   orc_sort_index      brute-force stable sort by baseline key.
 *   orc_preprocess      hand-derived known-answer records: pixel centres,
 *                       exact .5 ties, w<0 conjugation, w = k^2 dw plane
 *                       boundaries and (k+1/2)^2 dw ties, clamp to the last
 *                       plane, edge exclusion, odd-sampling clamp.
 *   orc_grid            brute-force complex scatter of single records, and
 *                       end-to-end against a direct DFT image with w-term.
 *   orc_weights         textbook Briggs/uniform definitions on tiny inputs,
 *                       and the R->+inf (natural) / R->-inf (uniform) limits.
 *   orc_fft1d           numpy.fft on random input, powers of two and not.
 *   orc_finalize        numpy.fft.ifft2 with shifts; per-pixel Stokes closed
 *                       forms; point-source position and PSF peak.
 */
#include "mti_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* psi(nu): Schwab's rational approximation of the M=6, alpha=1 prolate      */
/* spheroidal function.  op/CasaGridSF.cu:41-85 (coefficients :44-50,        */
/* breakpoint 0.75).                                                         */
/* ------------------------------------------------------------------------ */
double orc_gridsf(double nu)
{
    static const double P[2][5] = {
        {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
        {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
    static const double Q[2][3] = {{1.0, 8.212018e-1, 2.078043e-1},
                                   {1.0, 9.599102e-1, 2.918724e-1}};
    int part = (nu < 0.75) ? 0 : 1;
    double nuend2 = part ? 1.0 : 0.5625;
    double d = nu * nu - nuend2;
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

/* dw: cfg/WProjectionConvFunction.cpp:105-112 (w_max = 0.25/cell_l,
 * dw = w_max/(W-1)^2, W==1 -> 1.0).  Normal mode has one plane; dw=1.0 there
 * (the reference leaves it uninitialised, cfg/OnlyTaperConvFunction.h:62). */
double orc_w_increment(const orc_params *p)
{
    if (p->mode != 1 || p->wplanes <= 1) return 1.0;
    double wmax = 0.25 / p->cell_l;
    double wm1 = (double)(p->wplanes - 1);
    return wmax / (wm1 * wm1);
}

/* Trial support S_t = min(nx,ny)/s made odd: cfg/WProjectionConvFunction.cpp:194-200 */
int orc_trial_support(const orc_params *p)
{
    int mn = p->nx < p->ny ? p->nx : p->ny;
    int st = mn / p->sampling;
    if (st % 2 == 0) st++;
    return st;
}

static int imod(int64_t a, int64_t m)
{
    int64_t r = a % m;
    return (int)(r < 0 ? r + m : r);
}

/* Screen of one w-plane: op/WProjPreFFTconvFunc.cu:34-105 with the lm taper of
 * op/CasaGridSF.cu:151-173 (n = S_t, nu = |c|/(n/2)).  Only the S_t x S_t
 * central samples are non-zero; screen[(cy+h)*St+(cx+h)], c in [-h,h]. */
static void build_screen(const orc_params *p, int St, double w, double *scr)
{
    int h = St / 2;
    double tl = p->cell_l * (double)p->nx, tm = p->cell_m * (double)p->ny;
    double linc = tl / (double)St, minc = tm / (double)St;
    for (int cy = -h; cy <= h; cy++) {
        double ty = orc_gridsf(fabs((double)cy) / (double)h);
        double msq = minc * (double)cy;
        msq *= msq;
        for (int cx = -h; cx <= h; cx++) {
            double taper = ty * orc_gridsf(fabs((double)cx) / (double)h);
            double lsq = linc * (double)cx;
            lsq *= lsq;
            double rsq = lsq + msq;
            double re, im;
            if (rsq < 1.0) {
                double ph = 2.0 * w * (sqrt(1.0 - rsq) - 1.0); /* units of pi */
                re = taper * cos(M_PI * ph);
                im = taper * sin(M_PI * ph);
            } else {
                re = taper;
                im = 0.0;
            }
            double *o = scr + 2 * ((int64_t)(cy + h) * St + (cx + h));
            o[0] = re;
            o[1] = im;
        }
    }
}

static double *twiddles(int L)
{
    /* tw[m] = exp(-2 pi i m / L) */
    double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)L);
    for (int m = 0; m < L; m++) {
        double a = 2.0 * M_PI * (double)m / (double)L;
        tw[2 * m] = cos(a);
        tw[2 * m + 1] = -sin(a);
    }
    return tw;
}

/* Outermost centre-row / centre-column sample with |C|^2 > 1e-6 of the
 * fftshifted, normalised uv kernel; support S=((q+(s+1)/2)/s)*2+1
 * (op/ConvFunctionSupportFind.cu:28-77, threshold 1e-3 from
 * cfg/WPlaneConvFunctionCalculator.cpp:61), clamped to S_t-2
 * (cfg/WProjectionConvFunction.cpp:248-249).
 * The centre row (q_y = 0) of the 2-D DFT is the 1-D DFT of the column sums
 * of the screen; likewise the centre column. */
static int find_support(const double *scr, int St, int s, const double *tw,
                        double inv_div)
{
    int h = St / 2, L = St * s;
    double *rowp = (double *)calloc(2 * (size_t)St, sizeof(double));
    double *colp = (double *)calloc(2 * (size_t)St, sizeof(double));
    for (int iy = 0; iy < St; iy++)
        for (int ix = 0; ix < St; ix++) {
            const double *v = scr + 2 * ((int64_t)iy * St + ix);
            rowp[2 * ix] += v[0];
            rowp[2 * ix + 1] += v[1];
            colp[2 * iy] += v[0];
            colp[2 * iy + 1] += v[1];
        }
    int qmax = L - 1 - L / 2, found = -1;
    for (int q = qmax; q >= 0 && found < 0; q--) {
        for (int pass = 0; pass < 2; pass++) {
            const double *pr = pass ? colp : rowp;
            double re = 0, im = 0;
            for (int c = -h; c <= h; c++) {
                const double *t = tw + 2 * imod((int64_t)c * q, L);
                const double *v = pr + 2 * (c + h);
                re += v[0] * t[0] - v[1] * t[1];
                im += v[0] * t[1] + v[1] * t[0];
            }
            re *= inv_div;
            im *= inv_div;
            if (re * re + im * im > 1e-6) {
                found = q;
                break;
            }
        }
    }
    free(rowp);
    free(colp);
    if (found < 0) return 0;
    int S = ((found + (s + 1) / 2) / s) * 2 + 1;
    if (S > St - 2) S = St - 2;
    return S;
}

/* Cut the S*T x S*T window around the kernel centre out of the (virtual)
 * fftshifted FFT of the screen and reorder to [oy][ox][sy][sx]
 * (op/ConvFunctionReOrganize.cu:30-71): sample (o, s_) sits at uv offset
 * q = (s_ - S/2)*s + o - s/2 from the centre.  Evaluated as a separable DFT
 * restricted to the window (identical to FFT-then-crop). */
static void fill_plane(const double *scr, int St, int s, int T, int S,
                       const double *tw, double inv_div, double *out)
{
    int h = St / 2, L = St * s;
    int qmin = -(S / 2) * s - s / 2;
    int W = (S - 1) * s + T; /* window length covering every q used */
    double *A = (double *)malloc(sizeof(double) * 2 * (size_t)St * W);
    for (int iy = 0; iy < St; iy++)
        for (int j = 0; j < W; j++) {
            int qx = qmin + j;
            double re = 0, im = 0;
            for (int cx = -h; cx <= h; cx++) {
                const double *t = tw + 2 * imod((int64_t)cx * qx, L);
                const double *v = scr + 2 * ((int64_t)iy * St + (cx + h));
                re += v[0] * t[0] - v[1] * t[1];
                im += v[0] * t[1] + v[1] * t[0];
            }
            A[2 * ((int64_t)iy * W + j)] = re;
            A[2 * ((int64_t)iy * W + j) + 1] = im;
        }
    double *Y = (double *)malloc(sizeof(double) * 2 * (size_t)W * W);
    for (int jy = 0; jy < W; jy++) {
        int qy = qmin + jy;
        for (int jx = 0; jx < W; jx++) {
            double re = 0, im = 0;
            for (int cy = -h; cy <= h; cy++) {
                const double *t = tw + 2 * imod((int64_t)cy * qy, L);
                const double *v = A + 2 * ((int64_t)(cy + h) * W + jx);
                re += v[0] * t[0] - v[1] * t[1];
                im += v[0] * t[1] + v[1] * t[0];
            }
            Y[2 * ((int64_t)jy * W + jx)] = re * inv_div;
            Y[2 * ((int64_t)jy * W + jx) + 1] = im * inv_div;
        }
    }
    int64_t S2 = (int64_t)S * S;
    for (int oy = 0; oy < T; oy++)
        for (int ox = 0; ox < T; ox++)
            for (int sy = 0; sy < S; sy++)
                for (int sx = 0; sx < S; sx++) {
                    int jy = sy * s + oy, jx = sx * s + ox;
                    int64_t o = (int64_t)oy * T * S2 + (int64_t)ox * S2 + sy * S + sx;
                    out[2 * o] = Y[2 * ((int64_t)jy * W + jx)];
                    out[2 * o + 1] = Y[2 * ((int64_t)jy * W + jx) + 1];
                }
    free(A);
    free(Y);
}

/* Sums of Re over every sub-kernel: op/SumConvolutionFunctions.cu:28-56 */
static void fill_sums(orc_kernels *K)
{
    int T2 = K->T * K->T;
    for (int k = 0; k < K->nplanes; k++) {
        int64_t S2 = (int64_t)K->support[k] * K->support[k];
        for (int o = 0; o < T2; o++) {
            double acc = 0;
            const double *t = K->table + 2 * (K->base[k] + o * S2);
            for (int64_t i = 0; i < S2; i++) acc += t[2 * i];
            K->sums[(int64_t)k * T2 + o] = acc;
        }
    }
}

static int alloc_tables(orc_kernels *K)
{
    int T2 = K->T * K->T;
    K->base = (int64_t *)malloc(sizeof(int64_t) * K->nplanes);
    int64_t tot = 0;
    for (int k = 0; k < K->nplanes; k++) {
        K->base[k] = tot;
        tot += (int64_t)T2 * K->support[k] * K->support[k];
    }
    K->total = tot;
    K->table = (double *)calloc(2 * (size_t)(tot > 0 ? tot : 1), sizeof(double));
    K->sums = (double *)calloc((size_t)K->nplanes * T2, sizeof(double));
    return (K->table && K->sums) ? 0 : -1;
}

int orc_wkernel_supports(const orc_params *p, int *support_out)
{
    int s = p->sampling, St = orc_trial_support(p), L = St * s;
    if (St < 3) return -1;
    double dw = orc_w_increment(p);
    double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)St * St);
    double *tw = twiddles(L);
    build_screen(p, St, 0.0, scr);
    double div = 0;
    for (int64_t i = 0; i < (int64_t)St * St; i++) div += scr[2 * i];
    double inv = 1.0 / div;
    for (int k = 0; k < p->wplanes; k++) {
        build_screen(p, St, (double)k * (double)k * dw, scr);
        support_out[k] = find_support(scr, St, s, tw, inv);
    }
    free(scr);
    free(tw);
    return 0;
}

/* W-projection kernels: cfg/WProjectionConvFunction.cpp:186-291 and
 * cfg/WPlaneConvFunctionCalculator.cpp:31-148; taper-only kernel for
 * mode=normal: cfg/OnlyTaperConvFunction.cpp:29-99 with op/CasaGridSF.cu:107-149. */
int orc_kernels_build(const orc_params *p, orc_kernels *K)
{
    memset(K, 0, sizeof(*K));
    int s = p->sampling;
    if (s < 1) return -1;
    K->sampling = s;
    K->T = (s % 2) ? s : s + 1;
    K->w_increment = orc_w_increment(p);
    if (p->mode == 1) {
        int St = orc_trial_support(p), L = St * s;
        if (St < 3 || p->wplanes < 1) return -1;
        K->nplanes = p->wplanes;
        K->trial_support = St;
        K->support = (int *)malloc(sizeof(int) * K->nplanes);
        double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)St * St);
        double *tw = twiddles(L);
        build_screen(p, St, 0.0, scr);
        double div = 0;
        for (int64_t i = 0; i < (int64_t)St * St; i++) div += scr[2 * i];
        K->divisor = div;
        double inv = 1.0 / div;
        for (int k = 0; k < K->nplanes; k++) {
            build_screen(p, St, (double)k * (double)k * K->w_increment, scr);
            K->support[k] = find_support(scr, St, s, tw, inv);
        }
        if (alloc_tables(K)) return -1;
        for (int k = 0; k < K->nplanes; k++) {
            if (K->support[k] == 0) continue;
            build_screen(p, St, (double)k * (double)k * K->w_increment, scr);
            fill_plane(scr, St, s, K->T, K->support[k], tw, inv,
                       K->table + 2 * K->base[k]);
        }
        free(scr);
        free(tw);
    } else {
        int S = p->taper_support;
        if (S < 1 || S % 2 == 0) return -1;
        K->nplanes = 1;
        K->trial_support = S;
        K->divisor = 1.0;
        K->support = (int *)malloc(sizeof(int));
        K->support[0] = S;
        if (alloc_tables(K)) return -1;
        int dim = (s % 2) ? S * s : S * s + 1, c = dim / 2, fz = 3 * s;
        double *g = (double *)malloc(sizeof(double) * dim);
        for (int i = 0; i < dim; i++) {
            int a = abs(i - c);
            if (a < fz) {
                double nu = (double)a / (double)fz;
                g[i] = (1.0 - nu * nu) * orc_gridsf(nu);
            } else
                g[i] = 0.0;
        }
        int T = K->T;
        int64_t S2 = (int64_t)S * S;
        for (int oy = 0; oy < T; oy++)
            for (int ox = 0; ox < T; ox++)
                for (int sy = 0; sy < S; sy++)
                    for (int sx = 0; sx < S; sx++) {
                        int64_t o = (int64_t)oy * T * S2 + (int64_t)ox * S2 + sy * S + sx;
                        K->table[2 * o] = g[sy * s + oy] * g[sx * s + ox];
                        K->table[2 * o + 1] = 0.0;
                    }
        free(g);
    }
    fill_sums(K);
    return 0;
}

void orc_kernels_free(orc_kernels *K)
{
    free(K->support);
    free(K->base);
    free(K->table);
    free(K->sums);
    memset(K, 0, sizeof(*K));
}

/* Baseline-major stable order, autocorrelations dropped:
 * vis/SortIndex.cpp:46-87 (key = Nant^2 - a1*Nant - a2 - 1, ascending). */
int64_t orc_sort_index(const int32_t *ant1, const int32_t *ant2, int64_t nrows,
                       int nant, int32_t *index_out)
{
    int64_t nk = (int64_t)nant * nant;
    int64_t *cnt = (int64_t *)calloc((size_t)nk + 1, sizeof(int64_t));
    for (int64_t i = 0; i < nrows; i++) {
        if (ant1[i] == ant2[i]) continue;
        int64_t key = nk - (int64_t)ant1[i] * nant - ant2[i] - 1;
        cnt[key + 1]++;
    }
    for (int64_t k = 0; k < nk; k++) cnt[k + 1] += cnt[k];
    int64_t kept = cnt[nk];
    for (int64_t i = 0; i < nrows; i++) {
        if (ant1[i] == ant2[i]) continue;
        int64_t key = nk - (int64_t)ant1[i] * nant - ant2[i] - 1;
        index_out[cnt[key]++] = (int32_t)i;
    }
    free(cnt);
    return kept;
}

/* Per-record factor on 1/lambda (TOPO -> LSRK), set by orc_set_row_doppler. */
static const double *g_row_doppler = NULL;
void orc_set_row_doppler(const double *factor) { g_row_doppler = factor; }

/* Binning: op/WProjPreProcessUVW.cu:111-286 restated in FP64.
 *  to_grid  = any correlation unflagged (:65-83) and support box inside the
 *             grid (:196-199);
 *  w<0      -> (u,v) negated, imaginary part negated (:170-175);
 *  plane    = round(sqrt(|w f| / dw)) clamped to W-1 (:181-182);
 *  pos      = (u f)/du + nx/2 (:188-189), loc = round(pos),
 *  corner   = loc - S/2;
 *  offset   = round((loc-pos)*s) + s/2, clamped to [0,s-1] for odd s (:206-227);
 *  conv_idx = base_k + oy*S^2*T + ox*S^2 (:235-239). */
void orc_preprocess(const orc_params *p, const orc_kernels *K,
                    int64_t nrows, int nchan, int npol,
                    const double *uvw, const double *inv_lambda,
                    const uint8_t *flags, int chan, orc_bin *out)
{
    double dw = K->w_increment;
    double du = 1.0 / (p->cell_l * (double)p->nx);
    double dv = 1.0 / (p->cell_m * (double)p->ny);
    int s = K->sampling, T = K->T, W = K->nplanes;
    const double fc = inv_lambda[chan];
    double hx = (double)(p->nx / 2), hy = (double)(p->ny / 2);
    for (int64_t r = 0; r < nrows; r++) {
        orc_bin b;
        memset(&b, 0, sizeof(b));
        /* per-record LSRK factor (vis/FrequencyConverter.cpp:112-141) */
        const double f = g_row_doppler ? fc * g_row_doppler[r] : fc;
        const uint8_t *fl = flags + ((int64_t)r * nchan + chan) * npol;
        int any = 0;
        for (int q = 0; q < npol; q++)
            if (fl[q] == 0) any = 1;
        if (any) {
            double u = uvw[3 * r], v = uvw[3 * r + 1], w = uvw[3 * r + 2];
            int conj = 0;
            if (w < 0.0) {
                u = -u;
                v = -v;
                conj = 1;
            }
            int plane = 0;
            if (p->mode == 1) {
                double rp = round(sqrt(fabs((w * f) / dw)));
                plane = (rp <= (double)(W - 1)) ? (int)rp : W - 1;
            }
            int S = K->support[plane];
            double pos_v = ((v * f) / dv) + hy;
            double pos_u = ((u * f) / du) + hx;
            double lx = round(pos_u), ly = round(pos_v);
            double cxd = lx - (double)(S / 2), cyd = ly - (double)(S / 2);
            if (cyd >= 0.0 && cyd <= (double)((p->ny - 1) - S) && cxd >= 0.0 &&
                cxd <= (double)((p->nx - 1) - S)) {
                int ox = (int)round((lx - pos_u) * (double)s) + s / 2;
                int oy = (int)round((ly - pos_v) * (double)s) + s / 2;
                if (s % 2) {
                    if (ox < 0) ox = 0;
                    if (oy < 0) oy = 0;
                    if (ox == s) ox--;
                    if (oy == s) oy--;
                }
                int64_t S2 = (int64_t)S * S;
                b.to_grid = 1;
                b.plane = plane;
                b.support = S;
                b.corner_x = (int)cxd;
                b.corner_y = (int)cyd;
                b.off_x = ox;
                b.off_y = oy;
                b.conj = conj;
                b.conv_idx = K->base[plane] + (int64_t)oy * S2 * T + (int64_t)ox * S2;
            }
        }
        out[r] = b;
    }
}

/* Imaging weights (not in the reference; textbook definitions).
 *  density W_p(cell) = sum of w_p (1-flag_p) of gridded records whose nearest
 *  grid cell (loc) is that cell;
 *  uniform: w' = w / W(cell);
 *  briggs(R): f^2 = (5*10^-R)^2 / (sum_cells W^2 / sum_i w_i),
 *             w' = w / (1 + f^2 W(cell))  (Briggs 1995).
 *  weighting: 0 natural, 1 uniform, 2 briggs. */
int orc_weights(const orc_params *p, const orc_kernels *K,
                int64_t nrows, int nchan, int npol,
                const double *uvw, const double *inv_lambda,
                const uint8_t *flags, const float *weight, int chan,
                int weighting, double robust, double *wout)
{
    return orc_weights_mfs(p, K, nrows, nchan, npol, uvw, inv_lambda, flags, weight, chan, 1,
                           weighting, robust, wout);
}

/* The same definitions with one density over the records of channels
 * chan0 .. chan0 + nch - 1 (multi-frequency synthesis: those channels share
 * one grid).  wout [nch][nrows][npol]. */
int orc_weights_mfs(const orc_params *p, const orc_kernels *K,
                    int64_t nrows, int nchan, int npol,
                    const double *uvw, const double *inv_lambda,
                    const uint8_t *flags, const float *weight, int chan0, int nch,
                    int weighting, double robust, double *wout)
{
    const int64_t nb = nrows > 0 ? nrows : 1;
    orc_bin *b = (orc_bin *)malloc(sizeof(orc_bin) * (size_t)nb * (size_t)nch);
    for (int c = 0; c < nch; c++)
        orc_preprocess(p, K, nrows, nchan, npol, uvw, inv_lambda, flags, chan0 + c, b + (int64_t)c * nb);
    for (int c = 0; c < nch; c++)
        for (int64_t r = 0; r < nrows; r++)
            for (int q = 0; q < npol; q++) wout[((int64_t)c * nrows + r) * npol + q] = weight[r * npol + q];
    if (weighting == 0) {
        free(b);
        return 0;
    }
    int64_t ncell = (int64_t)p->nx * p->ny;
    double *dens = (double *)calloc((size_t)ncell * npol, sizeof(double));
    for (int c = 0; c < nch; c++)
        for (int64_t r = 0; r < nrows; r++) {
            const orc_bin *br = b + (int64_t)c * nb + r;
            if (!br->to_grid) continue;
            const uint8_t *fl = flags + ((int64_t)r * nchan + chan0 + c) * npol;
            int64_t cell = (int64_t)(br->corner_y + br->support / 2) * p->nx + (br->corner_x + br->support / 2);
            for (int q = 0; q < npol; q++)
                dens[q * ncell + cell] += (double)weight[r * npol + q] * (fl[q] ? 0.0 : 1.0);
        }
    for (int q = 0; q < npol; q++) {
        double f2 = 0.0;
        if (weighting == 2) {
            double sw = 0, sw2 = 0;
            for (int64_t c = 0; c < ncell; c++) {
                sw += dens[q * ncell + c];
                sw2 += dens[q * ncell + c] * dens[q * ncell + c];
            }
            double k = 5.0 * pow(10.0, -robust);
            f2 = (sw2 > 0.0) ? (k * k) / (sw2 / sw) : 0.0;
        }
        for (int c = 0; c < nch; c++)
            for (int64_t r = 0; r < nrows; r++) {
                const orc_bin *br = b + (int64_t)c * nb + r;
                if (!br->to_grid) continue;
                int64_t cell = (int64_t)(br->corner_y + br->support / 2) * p->nx + (br->corner_x + br->support / 2);
                double d = dens[q * ncell + cell];
                double w = weight[r * npol + q];
                double *o = &wout[((int64_t)c * nrows + r) * npol + q];
                if (weighting == 1)
                    *o = d > 0.0 ? w / d : 0.0;
                else
                    *o = w / (1.0 + f2 * d);
            }
    }
    free(dens);
    free(b);
    return 0;
}

/* Gridding: op/ConvolutionGridder.cu:130-181 (full complex product, kernel
 * not conjugated) over op/ConvolutionGridder.cu:394-441's footprint
 * grid[corner_y+sy][corner_x+sx] += K[conv_idx + sy*S + sx] * V; visibility
 * weighting op/common.hcu:80-104 (V*w*(1-flag), imag negated for w<0);
 * normaliser op/WProjPreProcessUVW.cu:84-108, 229 with comp/WImager.cpp:196-197
 * (sum of sums[k][oy*T+ox] * w * (1-flag) per correlation).  Records that
 * the reference "compresses" (op/CompressVisibilties.cu:40-74) are gridded
 * individually: the sum is the same by linearity.  psf: all visibilities 1+0i
 * (vis/VisibilityForPSFStorageManager.cpp:63-100). */
int orc_grid(const orc_params *p, const orc_kernels *K,
             int64_t nrows, int nchan, int npol,
             const double *uvw, const double *inv_lambda,
             const float *vis, const uint8_t *flags, const float *weight,
             const double *wts, int chan, int psf,
             double *grid, double *norm, int64_t *gridded_records)
{
    orc_bin *b = (orc_bin *)malloc(sizeof(orc_bin) * (size_t)(nrows > 0 ? nrows : 1));
    orc_preprocess(p, K, nrows, nchan, npol, uvw, inv_lambda, flags, chan, b);
    int T = K->T, nx = p->nx;
    int64_t ng = 0;
    for (int64_t r = 0; r < nrows; r++) {
        if (!b[r].to_grid) continue;
        ng++;
        const uint8_t *fl = flags + ((int64_t)r * nchan + chan) * npol;
        const float *vv = vis + 2 * ((int64_t)r * nchan + chan) * npol;
        int S = b[r].support;
        double vre[4], vim[4];
        for (int q = 0; q < npol; q++) {
            double w = wts ? wts[r * npol + q] : (double)weight[r * npol + q];
            double wf = w * (fl[q] ? 0.0 : 1.0);
            double re = psf ? 1.0 : (double)vv[2 * q];
            double im = psf ? 0.0 : (double)vv[2 * q + 1];
            vre[q] = re * wf;
            vim[q] = im * wf * (b[r].conj ? -1.0 : 1.0);
            norm[q] += K->sums[(int64_t)b[r].plane * T * T + b[r].off_y * T + b[r].off_x] * wf;
        }
        const double *kk = K->table + 2 * b[r].conv_idx;
        for (int sy = 0; sy < S; sy++)
            for (int sx = 0; sx < S; sx++) {
                double kre = kk[2 * (sy * S + sx)], kim = kk[2 * (sy * S + sx) + 1];
                double *g = grid + 2 * (((int64_t)(b[r].corner_y + sy) * nx +
                                         (b[r].corner_x + sx)) * npol);
                for (int q = 0; q < npol; q++) {
                    g[2 * q] += vre[q] * kre - vim[q] * kim;
                    g[2 * q + 1] += vim[q] * kre + vre[q] * kim;
                }
            }
    }
    if (gridded_records) *gridded_records = ng;
    free(b);
    return 0;
}

/* ---------------- FP64 FFT (radix-2 + Bluestein), test/oracle use -------- */
static void fft_pow2(double *d, int n, int sign)
{
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t0 = d[2 * i], t1 = d[2 * i + 1];
            d[2 * i] = d[2 * j];
            d[2 * i + 1] = d[2 * j + 1];
            d[2 * j] = t0;
            d[2 * j + 1] = t1;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            double a = sign * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(a), wi = sin(a);
            for (int i = 0; i < n; i += len) {
                double *x = d + 2 * (i + k), *y = d + 2 * (i + k + half);
                double tr = y[0] * wr - y[1] * wi, ti = y[0] * wi + y[1] * wr;
                y[0] = x[0] - tr;
                y[1] = x[1] - ti;
                x[0] += tr;
                x[1] += ti;
            }
        }
    }
}

void orc_fft1d(double *d, int n, int sign)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        fft_pow2(d, n, sign);
        return;
    }
    /* Bluestein: jk = (j^2 + k^2 - (k-j)^2)/2, so with c_j = e^{i pi s j^2/n}
     * X_k = c_k sum_j (x_j c_j) conj(c_{k-j}). */
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    double *a = (double *)calloc(2 * (size_t)m, sizeof(double));
    double *bq = (double *)calloc(2 * (size_t)m, sizeof(double));
    double *c = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int j = 0; j < n; j++) {
        int64_t j2 = ((int64_t)j * j) % (2 * (int64_t)n);
        double ang = sign * M_PI * (double)j2 / (double)n;
        c[2 * j] = cos(ang);
        c[2 * j + 1] = sin(ang);
    }
    for (int j = 0; j < n; j++) {
        /* a_j = x_j * c_j */
        a[2 * j] = d[2 * j] * c[2 * j] - d[2 * j + 1] * c[2 * j + 1];
        a[2 * j + 1] = d[2 * j + 1] * c[2 * j] + d[2 * j] * c[2 * j + 1];
    }
    /* b_j = conj(c_j), wrapped for the circular convolution */
    bq[0] = c[0];
    bq[1] = -c[1];
    for (int j = 1; j < n; j++) {
        bq[2 * j] = bq[2 * (m - j)] = c[2 * j];
        bq[2 * j + 1] = bq[2 * (m - j) + 1] = -c[2 * j + 1];
    }
    fft_pow2(a, m, -1);
    fft_pow2(bq, m, -1);
    for (int i = 0; i < m; i++) {
        double re = a[2 * i] * bq[2 * i] - a[2 * i + 1] * bq[2 * i + 1];
        double im = a[2 * i] * bq[2 * i + 1] + a[2 * i + 1] * bq[2 * i];
        a[2 * i] = re;
        a[2 * i + 1] = im;
    }
    fft_pow2(a, m, +1);
    for (int k = 0; k < n; k++) {
        /* X_k = c_k * conv_k / m */
        double re = a[2 * k] / m, im = a[2 * k + 1] / m;
        d[2 * k] = re * c[2 * k] - im * c[2 * k + 1];
        d[2 * k + 1] = im * c[2 * k] + re * c[2 * k + 1];
    }
    free(a);
    free(bq);
    free(c);
}

/* Finalize: comp/ImageFinalizer.cpp:31-44.
 *  PreIFFTUVCorrection ([v][u][P] -> [P][v][u] + ifftshift,
 *  op/PreIFFTUVCorrection.cu:28-90), unnormalised inverse FFT
 *  (std/IFFT2D.cpp:86-111), fftshift (std/ZeroToCentre2DShift.cu:66-91),
 *  normalise + Stokes + m-flip (op/PostIFFTStokesConverter.cu:51-180),
 *  taper correction (op/CasaGridSF.cu:174-199) and sinc correction
 *  (op/SincCorrector.cu:29-58), the last two on the flipped array.
 *  poltype: 0 linear, 1 circular. */
int orc_finalize(const orc_params *p, int npol, int poltype,
                 const double *grid, const double *norm, double *image)
{
    int nx = p->nx, ny = p->ny, s = p->sampling;
    int64_t np = (int64_t)nx * ny;
    double *x = (double *)malloc(sizeof(double) * 2 * (size_t)np * npol);
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny));
    for (int q = 0; q < npol; q++) {
        double *xq = x + 2 * np * q;
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                int iy = imod(gy - ny / 2, ny), ix = imod(gx - nx / 2, nx);
                const double *g = grid + 2 * (((int64_t)gy * nx + gx) * npol + q);
                xq[2 * ((int64_t)iy * nx + ix)] = g[0];
                xq[2 * ((int64_t)iy * nx + ix) + 1] = g[1];
            }
        for (int iy = 0; iy < ny; iy++) orc_fft1d(xq + 2 * (int64_t)iy * nx, nx, +1);
        for (int ix = 0; ix < nx; ix++) {
            for (int iy = 0; iy < ny; iy++) {
                line[2 * iy] = xq[2 * ((int64_t)iy * nx + ix)];
                line[2 * iy + 1] = xq[2 * ((int64_t)iy * nx + ix) + 1];
            }
            orc_fft1d(line, ny, +1);
            for (int iy = 0; iy < ny; iy++) {
                xq[2 * ((int64_t)iy * nx + ix)] = line[2 * iy];
                xq[2 * ((int64_t)iy * nx + ix) + 1] = line[2 * iy + 1];
            }
        }
    }
    for (int jy = 0; jy < ny; jy++)
        for (int jx = 0; jx < nx; jx++) {
            int ky = imod(jy - ny / 2, ny), kx = imod(jx - nx / 2, nx);
            int64_t src = (int64_t)ky * nx + kx;
            double re[4] = {0}, im[4] = {0}, out[4] = {0};
            for (int q = 0; q < npol; q++) {
                re[q] = x[2 * (np * q + src)] / norm[q];
                im[q] = x[2 * (np * q + src) + 1] / norm[q];
            }
            if (npol == 1) {
                out[0] = re[0];
            } else if (npol == 2) {
                out[0] = (re[0] + re[1]) * 0.5;
                out[1] = (re[0] - re[1]) * 0.5;
            } else if (poltype == 0) {
                out[0] = (re[0] + re[3]) * 0.5;
                out[1] = (re[0] - re[3]) * 0.5;
                out[2] = (re[1] + re[2]) * 0.5;
                out[3] = (im[1] - im[2]) * 0.5;
            } else {
                out[0] = (re[0] + re[3]) * 0.5;
                out[1] = (re[1] + re[2]) * 0.5;
                out[2] = (im[2] - im[1]) * 0.5;
                out[3] = (re[0] - re[3]) * 0.5;
            }
            int r = ny - 1 - jy; /* m axis flipped for FITS */
            double nuy = fabs((double)(r - ny / 2)) / (double)(ny / 2);
            double nux = fabs((double)(jx - nx / 2)) / (double)(nx / 2);
            double tap = orc_gridsf(nuy) * orc_gridsf(nux);
            double xx = (double)(jx - nx / 2) / ((double)nx * s);
            double yy = (double)(r - ny / 2) / ((double)ny * s);
            double sx = xx == 0.0 ? 1.0 : M_PI * xx / sin(M_PI * xx);
            double sy = yy == 0.0 ? 1.0 : M_PI * yy / sin(M_PI * yy);
            for (int q = 0; q < npol; q++)
                image[np * q + (int64_t)r * nx + jx] = out[q] / tap * (sx * sy);
        }
    free(x);
    free(line);
    return 0;
}
