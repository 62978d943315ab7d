/*
 * mti_oracle.h -- FP64 CPU restatement of mt-imager's imaging path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the parity checker for the HIP
 * imaging library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline -- never as a product code path.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to code/mt-imager/lib/ and code/gafw/Implimentation/).
 * All arithmetic is IEEE FP64, built with -ffp-contract=off so that the
 * uvw->pixel / w-plane / oversample-offset arithmetic is the exact op sequence
 * the GPU binning kernel performs.
 *
 * Pinning status of each function (how the tests check it) is listed in the
 * header comment of mti_oracle.c.
 */
#ifndef MTI_ORACLE_H
#define MTI_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Imaging parameters; names follow conf/mtimager.properties keys. */
typedef struct {
    int    nx, ny;          /* image.nx, image.ny (pixels)                  */
    double cell_l, cell_m;  /* image.lintervalsize/mintervalsize in radians */
    int    mode;            /* 0 = normal (taper only), 1 = wproj           */
    int    wplanes;         /* wproj.wplanes                                */
    int    sampling;        /* wproj.convFunction.sampling (wproj) or
                               taper.sampling (normal)                      */
    int    taper_support;   /* taper.support (normal mode)                  */
} orc_params;

/* Convolution-function tables (the WProjectionConvFunction outputs). */
typedef struct {
    int      nplanes;
    int      sampling;
    int      T;              /* samples per cell axis: s (odd) or s+1 (even) */
    int      trial_support;  /* S_t (wproj) or taper support (normal)        */
    double   w_increment;    /* dw : w_k = k^2 dw                            */
    double   divisor;        /* Re of the w=0 screen DC term                 */
    int     *support;        /* [nplanes]                                    */
    int64_t *base;           /* [nplanes] complex offset of plane k          */
    int64_t  total;          /* complex elements in table                    */
    double  *table;          /* [total][2]  layout [plane][oy][ox][sy][sx]   */
    double  *sums;           /* [nplanes][T*T] sum of Re over each sub-kernel */
} orc_kernels;

/* Per-record binning result (WProjPreProcessUVW outputs). */
typedef struct {
    int32_t to_grid;
    int32_t plane;
    int32_t support;
    int32_t corner_x, corner_y;
    int32_t off_x, off_y;
    int32_t conj;            /* 1 when w<0: (u,v)->(-u,-v), imag negated     */
    int64_t conv_idx;        /* complex offset into table                    */
} orc_bin;

double orc_gridsf(double nu);
double orc_w_increment(const orc_params *p);
int    orc_trial_support(const orc_params *p);

int  orc_kernels_build(const orc_params *p, orc_kernels *K);
void orc_kernels_free(orc_kernels *K);
/* Support-find only (no tables): fills support[nplanes]. */
int  orc_wkernel_supports(const orc_params *p, int *support_out);

int64_t orc_sort_index(const int32_t *ant1, const int32_t *ant2, int64_t nrows,
                       int nant, int32_t *index_out);

void orc_preprocess(const orc_params *p, const orc_kernels *K,
                    int64_t nrows, int nchan, int npol,
                    const double *uvw, const double *inv_lambda,
                    const uint8_t *flags, int chan, orc_bin *out);

/* Imaging weights (natural/uniform/briggs) for one channel.
 * weight [nrows][npol] in, wout [nrows][npol] out (effective weights). */
int orc_weights(const orc_params *p, const orc_kernels *K,
                int64_t nrows, int nchan, int npol,
                const double *uvw, const double *inv_lambda,
                const uint8_t *flags, const float *weight, int chan,
                int weighting, double robust, double *wout);

/* Weights with one density over channels chan0 .. chan0 + nch - 1 (MFS:
 * those channels share one grid); wout [nch][nrows][npol]. */
int orc_weights_mfs(const orc_params *p, const orc_kernels *K,
                    int64_t nrows, int nchan, int npol,
                    const double *uvw, const double *inv_lambda,
                    const uint8_t *flags, const float *weight, int chan0, int nch,
                    int weighting, double robust, double *wout);

/* Grid one channel.  vis complex64 [nrows][nchan][npol]; flags u8 same
 * layout; weight f32 [nrows][npol] or wts f64 [nrows][npol] (if non-NULL,
 * overrides weight).  psf!=0 replaces every visibility by 1+0i.
 * grid [ny][nx][npol][2] and norm [npol] are ACCUMULATED into. */
int orc_grid(const orc_params *p, const orc_kernels *K,
             int64_t nrows, int nchan, int npol,
             const double *uvw, const double *inv_lambda,
             const float *vis, const uint8_t *flags, const float *weight,
             const double *wts, int chan, int psf,
             double *grid, double *norm, int64_t *gridded_records);

/* 1-D FFT of any length, FP64, in place: sign=-1 forward, +1 inverse,
 * unnormalised. */
void orc_fft1d(double *data, int n, int sign);

/* Finalize: shift + unnormalised inverse 2-D FFT + shift + normalise +
 * Stokes + m-flip + taper correction + sinc correction.
 * grid [ny][nx][npol][2], norm [npol], image out [npol][ny][nx]. */
int orc_finalize(const orc_params *p, int npol, int poltype,
                 const double *grid, const double *norm, double *image);

/* TOPO -> LSRK (vis/FrequencyConverter.cpp:112-141; casacore Measures
 * restated in mti_oracle_lsrk.c). */
typedef struct {
    double obs_itrf[3];   /* observatory geocentric ITRF position, metres    */
    double ra, dec;       /* phase centre, J2000, radians                    */
    double tt_minus_utc;  /* TT - UTC in seconds (69.184 since 2017)         */
} orc_lsrk_frame;
/* Velocities in m/s, J2000 equatorial axes: observatory rotation, Earth
 * barycentric, solar motion relative to the LSR.  Time: MJD seconds (UTC). */
void orc_lsrk_velocity(const orc_lsrk_frame *f, double mjd_utc_s, double v_rot[3],
                       double v_bary[3], double v_lsr[3]);
/* The Earth's velocity relative to the Sun and the Sun's relative to the
 * solar-system barycentre, m/s, J2000 equatorial axes; time: MJD seconds (TT). */
void orc_earth_velocity(double mjd_tt_s, double v_helio[3], double v_sun[3]);
/* f_LSRK / f_TOPO at one time, and per record (one evaluation per run of
 * equal times). */
double orc_lsrk_factor(const orc_lsrk_frame *f, double mjd_utc_s);
void   orc_lsrk_factors(const orc_lsrk_frame *f, const double *mjd_utc_s, int64_t n, double *out);
/* Per-record factor on 1/lambda used by orc_preprocess/orc_weights/orc_grid
 * (f = inv_lambda[chan] * factor[row]); NULL (the default) = no factor. */
void   orc_set_row_doppler(const double *factor);

/* Prediction (degridding; mti_oracle_degrid.c).  Model grids [ny][nx][npol][2]
 * from Stokes model images [npol][ny][nx] in the finalize's layout: */
int    orc_model_grid(const orc_params *p, int npol, int poltype, const double *image, double *grid);
/* inverse kernel attenuation (sinc / taper) at transform pixel (jy, jx) */
double orc_model_correction(const orc_params *p, int jy, int jx);
/* Predicted visibilities vis [nrows][npol][2] of one channel (1/lambda
 * inv_lambda) from model grids; normalise = 1 divides by the sub-kernel sums
 * (prediction), 0 gives the adjoint of orc_grid's scatter.  Records whose
 * support box is off the grid get 0; *predicted counts the others. */
int    orc_degrid(const orc_params *p, const orc_kernels *K, int64_t nrows, int npol, const double *uvw,
                  double inv_lambda, const double *grid, int normalise, double *vis, int64_t *predicted);

#ifdef __cplusplus
}
#endif
#endif
