/*
 * mtimager_amd.h -- C-ABI of the MI355X (gfx950) W-projection imaging path.
 *
 * Plain C types only: device pointers, element counts, strides, POD parameter
 * structs and a hipStream_t passed as void*.  No C++ exceptions cross this
 * boundary; every entry point returns an mti_status and the message of the
 * last failure is available from mti_last_error().
 *
 * The caller owns every input and output device buffer (as the GAFW engine
 * does for operator I/O, gafw/Implimentation/GPUImplimentation/
 * CUDADeviceManager.cpp:1361-1530).  Scratch space is either a caller's
 * workspace (mti_set_workspace, sized by the mti_*_workspace_size queries)
 * or, without one, held by the context and grown on demand.  All work is
 * enqueued on the stream given to each call;
 * calls on one context must not overlap (one context per host thread, as the
 * reference runs one submit thread per device, CUDADeviceManager.cpp:410-489).
 *
 * Reference citations are relative to code/mt-imager/lib/ unless noted
 * ("gafw/" = code/gafw/Implimentation/).
 */
#ifndef MTIMAGER_AMD_H
#define MTIMAGER_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MTI_OK = 0,
    MTI_ERR_INVALID = 1,     /* validation failure (ValidationException)     */
    MTI_ERR_HIP = 2,         /* HIP runtime failure (CudaException)          */
    MTI_ERR_UNSUPPORTED = 3, /* shape/feature not supported by this build    */
    MTI_ERR_NOMEM = 4
} mti_status;

typedef enum { MTI_MODE_NORMAL = 0, MTI_MODE_WPROJ = 1 } mti_mode;
typedef enum { MTI_TABLE_F32 = 0, MTI_TABLE_F16 = 1, MTI_TABLE_BF16 = 2 } mti_table_precision;
typedef enum { MTI_WEIGHT_NATURAL = 0, MTI_WEIGHT_UNIFORM = 1, MTI_WEIGHT_BRIGGS = 2 } mti_weighting;
typedef enum { MTI_POL_LINEAR = 0, MTI_POL_CIRCULAR = 1 } mti_poltype;
/* Grid layouts: the reference's WImager Grid [v][u][P] (comp/WImager.cpp:231-232)
 * or planar [P][v][u] (what the reference's PreIFFTUVCorrection reorders it
 * to before the transform, op/PreIFFTUVCorrection.cu:28-90). */
typedef enum { MTI_GRID_VUP = 0, MTI_GRID_PVU = 1 } mti_grid_layout;
/* Grid element type.  Gridding accumulates in FP64 registers; a complex64
 * grid (the reference's complex<float> Grid) receives each work unit's FP64
 * cell sums rounded once -- stored, read-modify-written, or, in a tile split
 * over several work units, added with float atomics whose order (and so the
 * last-bit rounding) depends on the order units finish. */
typedef enum { MTI_GRID_C64 = 0, MTI_GRID_C128 = 1 } mti_grid_precision;
/* Gridder accumulation (a precision form): FP64 registers (default), or FP32
 * registers per four records with those partial sums added by Kahan's
 * compensated summation in FP32 (binning, oversample indices and the w-plane
 * choice stay FP64 either way). */
typedef enum { MTI_ACC_F64 = 0, MTI_ACC_F32_KAHAN = 1 } mti_accumulation;

/* Imaging parameters; the fields carry the mt-imager property keys
 * (comp/ConfigurationManager.cpp:33-62, conf/mtimager.properties:1-78). */
typedef struct {
    int32_t nx;            /* image.nx                                       */
    int32_t ny;            /* image.ny                                       */
    double  cell_l;        /* image.lintervalsize, converted to radians      */
    double  cell_m;        /* image.mintervalsize, converted to radians      */
    int32_t mode;          /* mode: MTI_MODE_NORMAL | MTI_MODE_WPROJ         */
    int32_t wplanes;       /* wproj.wplanes                                  */
    int32_t sampling;      /* wproj.convFunction.sampling (wproj) or
                              taper.sampling (normal)                        */
    int32_t taper_support; /* taper.support (normal mode; casagridsf)        */
} mti_image_params;

/* Host-side description of a built convolution-function set. */
typedef struct {
    int32_t nplanes;
    int32_t sampling;
    int32_t T;             /* samples per cell axis: s (odd s) or s+1        */
    int32_t trial_support; /* S_t                                            */
    int32_t max_support;
    int32_t precision;     /* mti_table_precision                            */
    double  w_increment;   /* ConvolutionFunction.wsquareincrement (w_k = k^2 dw) */
    double  divisor;       /* w=0 origin value the tables are divided by     */
    int64_t total;         /* complex samples in the table                   */
} mti_kernel_info;

typedef struct mti_ctx mti_ctx;
typedef struct mti_kernels mti_kernels;

/* Visibility set in HBM.  Defaults describe MeasurementSet order
 * (row, channel, correlation); strides let the reference's chunk layout
 * (vis/StorageManager.h:147-150: uvw f32 [3][N], vis [N][2P], flags int [N][P],
 * weights f32 [N][P], per-record f32 1/lambda) be passed without a copy. */
typedef struct {
    int64_t nrows;            /* rows in the arrays                           */
    int32_t nchan;            /* channels in the arrays                       */
    int32_t npol;             /* correlations: 1, 2 or 4                      */
    /* uvw in metres */
    const void *uvw;
    int32_t uvw_f32;          /* 0: f64, 1: f32                               */
    int64_t uvw_row_stride;   /* elements between rows (3 for [N][3])         */
    int64_t uvw_comp_stride;  /* elements between u,v,w (1 for [N][3])        */
    /* 1/lambda (= f/c, 1/m) */
    const void *inv_lambda;
    int32_t inv_lambda_f32;   /* 0: f64, 1: f32                               */
    int32_t inv_lambda_per_row; /* 0: indexed [chan]; 1: indexed [row] (one channel) */
    /* complex64 visibilities */
    const float *vis;         /* interleaved re,im                            */
    int64_t vis_row_stride;   /* complex elements between rows                */
    int64_t vis_chan_stride;  /* complex elements between channels            */
    /* flags (non-zero = flagged) */
    const void *flags;
    int32_t flags_i32;        /* 0: u8, 1: int32                              */
    int64_t flags_row_stride; /* elements between rows                        */
    int64_t flags_chan_stride;/* elements between channels                    */
    /* weights f32, one per (row, correlation) */
    const float *weight;
    int64_t weight_row_stride;
    /* optional row order (baseline-major from mti_sort_rows); NULL = 0..nrows-1 */
    const int32_t *row_index;
    int64_t nindex;           /* entries in row_index (ignored if NULL)       */
    /* optional f64 [nrows]: a per-row factor on 1/lambda, the TOPO -> LSRK
     * Doppler factor of the row's time (mti_lsrk_factors); the binning uses
     * 1/lambda = inv_lambda * row_doppler[row] in FP64.  NULL = 1. */
    const double *row_doppler;
} mti_vis;

/* Gridding options for one call (WImager::calculate for one chunk,
 * comp/WImager.cpp:271-290). */
typedef struct {
    int32_t chan0;            /* first channel gridded by this call           */
    int32_t nchan;            /* channels gridded; one grid per channel       */
    int32_t accumulate;       /* 0: grids/normalisers zeroed first
                                 (initialize_output=true); 1: add in place    */
    int32_t psf;              /* psf: every visibility replaced by 1+0i       */
    int32_t weighting;        /* mti_weighting                                */
    double  robust;           /* Briggs robustness R                          */
    int32_t grid_layout;      /* mti_grid_layout of `grids`                   */
    int32_t grid_precision;   /* mti_grid_precision of `grids`                */
    /* Optional (NULL: complete grids).  Sparse output for callers that only
     * finalize the grids: u8 [nchan][ceil(ny/32)][ceil(nx/32)] (device),
     * written by the call -- 1 where some record's support box reaches the
     * 32 x 32-cell tile, 0 where none does; tiles marked 0 are then left
     * unwritten (undefined) instead of zero-filled.  Needs accumulate = 0;
     * mti_finalize_sparse reads those tiles as zeros.
     * No counterpart in the reference (its grids are always complete). */
    uint8_t *tile_mask;
    /* 1: multi-frequency synthesis -- the nchan channels are gridded into ONE
     * grid and ONE normaliser (grids, norm and tile_mask then hold a single
     * channel), uniform/Briggs densities summed over those channels.  The
     * reference images one cube plane per channel (app/mtimager.cpp:151-170);
     * feeding a WImager several channels' per-record frequencies is its MFS.
     * The channel range must fit one gridding pass (< 2^31 records). */
    int32_t mfs;
    int32_t accumulation;     /* mti_accumulation (0: FP64)                      */
} mti_grid_opts;

typedef struct {
    int64_t gridded_records;  /* Statistic:TotalGriddedRecords                */
    int64_t gridded_points;   /* Statistic:TotalGriddedPoints (sum S^2 * P)   */
    int64_t work_items;       /* gridder work units: runs of <= unit_pairs
                                 (record, tile) pairs of one 32x32-cell tile   */
    int64_t flushes;          /* (record, tile) pairs: each record counted once
                                 per tile its support box touches             */
} mti_grid_stats;

/* Pipeline stages timed with HIP events when profiling is on (the
 * equivalent of the reference's per-operator start/end events,
 * gafw/GPUImplimentation/CUDADeviceManager.cpp:435-560 -> engine.csv). */
typedef enum {
    MTI_STAGE_SORT_ROWS = 0,  /* mti_sort_rows                               */
    MTI_STAGE_KERNELS = 1,    /* mti_kernels_build                           */
    MTI_STAGE_BIN = 2,        /* binning pass (and weight density)           */
    MTI_STAGE_SORT_RECORDS = 3,/* radix sort of records by (channel, tile)   */
    MTI_STAGE_PLAN = 4,       /* bucket scans                                */
    MTI_STAGE_GRID = 5,       /* tile gridder                                */
    MTI_STAGE_FFT_ROWS = 6,   /* row transforms (+ uv reorder)               */
    MTI_STAGE_FFT_COLS = 7,   /* first column pass                           */
    MTI_STAGE_FINALIZE = 8,   /* second column pass + Stokes + corrections   */
    MTI_STAGE_COUNT = 9
} mti_stage;

/* ---- context ------------------------------------------------------------ */
const char *mti_version(void);
mti_status  mti_create(int device, mti_ctx **out);
void        mti_destroy(mti_ctx *ctx);
const char *mti_last_error(const mti_ctx *ctx);
/* Profiling: when enabled every stage launch is bracketed by HIP events on
 * the caller's stream.  mti_stage_times waits for the recorded events, adds
 * each stage's elapsed milliseconds and launch count to ms[MTI_STAGE_COUNT] /
 * launches[MTI_STAGE_COUNT] (either may be NULL) and clears the record. */
mti_status  mti_set_profiling(mti_ctx *ctx, int32_t enable);
mti_status  mti_stage_times(mti_ctx *ctx, double *ms, int64_t *launches);

/* ---- convolution functions ----------------------------------------------
 * Replaces WProjectionConvFunction::calculate (cfg/WProjectionConvFunction.cpp:186-291)
 * with its operators casagridsf (op/CasaGridSF.cu:151-173), WProjPreFFTConvFunc
 * (op/WProjPreFFTconvFunc.cu:34-105), multiFFT2D + ZeroToCentre2DShift +
 * SingleMultiply (gafw/StandardOperators/MultiFFT2D.cpp:86-111,
 * ZeroToCentre2DShift.cu:66-91, SingleMultiply.cu:69-100),
 * ConvFunctionSupportFind (op/ConvFunctionSupportFind.cu:28-77),
 * ConvFunctionReOrganize (op/ConvFunctionReOrganize.cu:30-71), ArraysToVector
 * and SumConvolutionFunctions (op/SumConvolutionFunctions.cu:28-56); for
 * mode=normal OnlyTaperConvFunction (cfg/OnlyTaperConvFunction.cpp:29-99).
 * Blocks once (support read-back) like the reference's host loop. */
mti_status mti_kernels_build(mti_ctx *ctx, const mti_image_params *p,
                             int32_t precision, mti_kernels **out, void *stream);
mti_status mti_kernels_info(const mti_kernels *k, mti_kernel_info *info);
/* Host copies: support[nplanes], base[nplanes] (complex offsets). */
mti_status mti_kernels_supports(const mti_kernels *k, int32_t *support, int64_t *base);
/* Device pointers: table (complex, precision-dependent element size),
 * sums f32 [nplanes][T*T]. */
const void  *mti_kernels_table(const mti_kernels *k);
const float *mti_kernels_sums(const mti_kernels *k);
/* Device {support, base} int32 pairs per plane: the convolution position data
 * (getConvFunctionPositionData, cfg/WProjectionConvFunction.cpp:258-265) that
 * PreProcessUVW takes as input 0 (mti_preprocess_reference). */
const int32_t *mti_kernels_position(const mti_kernels *k);
void         mti_kernels_free(mti_kernels *k);

/* ---- TOPO -> LSRK per-record frequency ----------------------------------------
 * Replaces FrequencyConverter::calculate (vis/FrequencyConverter.cpp:112-141,
 * casacore MFrequency::Convert TOPO -> LSRK per distinct time).  FP64 on the
 * device: observatory rotation (mean sidereal time, IAU 1976 precession to
 * J2000), the Earth's barycentric velocity (JPL Keplerian elements 1800-2050,
 * lunar offset), solar motion 20 km/s towards 18h03m50.29s +30d00m16.8s
 * (J2000); factor = sqrt(1 - b.b) / (1 + b.n).  Agrees with casacore to the
 * ephemeris accuracy (~10 m/s, 3e-8 in frequency), not bit for bit. */
typedef struct {
    double obs_itrf[3];   /* observatory geocentric ITRF position, metres     */
    double ra, dec;       /* phase centre (J2000), radians                    */
    double tt_minus_utc;  /* TT - UTC, seconds (69.184 from 2017)             */
} mti_lsrk_frame;
/* factor_out[r] = f_LSRK / f_TOPO at time_mjd_s[r] (MJD seconds, UTC, the
 * MeasurementSet TIME column), f64, for mti_vis.row_doppler. */
mti_status mti_lsrk_factors(mti_ctx *ctx, const mti_lsrk_frame *f, const double *time_mjd_s,
                            int64_t nrows, double *factor_out, void *stream);
/* The reference's per-record frequency store: f32 [nchan][nrows] of
 * f_LSRK / c = freq_hz[c] * factor / c (1/lambda in 1/m), for an mti_vis with
 * inv_lambda_f32 = inv_lambda_per_row = 1 (one channel per call). */
mti_status mti_lsrk_inv_lambda(mti_ctx *ctx, const mti_lsrk_frame *f, const double *time_mjd_s,
                               int64_t nrows, const double *freq_hz, int32_t nchan,
                               float *inv_lambda_out, void *stream);

/* ---- baseline-major row order --------------------------------------------
 * Replaces SortIndex::sortFuncThread (vis/SortIndex.cpp:46-87): stable order by
 * key Nant^2 - a1*Nant - a2 - 1, autocorrelations dropped.
 * index_out must hold nrows entries; *kept_out receives the count (blocks). */
mti_status mti_sort_rows(mti_ctx *ctx, const int32_t *ant1, const int32_t *ant2,
                         int64_t nrows, int32_t nant, int32_t *index_out,
                         int64_t *kept_out, void *stream);

/* Allocation-free mti_sort_rows.  No counterpart in the reference: for
 * callers that cannot allocate device memory during the call (code under HIP
 * graph capture, tight serving loops, hosts that forbid allocation).  Its
 * scratch is a 52 MiB device array of the library (one per GPU, held until
 * the process ends, looked up by mti_create), so it calls no hipMalloc,
 * hipFree or allocating routine; nrows up to MTI_SORT_STATIC_MAX_ROWS
 * (MTI_ERR_INVALID above).  Every call on one GPU shares that array: calls
 * that overlap on one GPU (several streams, contexts or threads) race.
 * Results equal mti_sort_rows's. */
#define MTI_SORT_STATIC_MAX_ROWS 4194304
mti_status mti_sort_rows_static(mti_ctx *ctx, const int32_t *ant1, const int32_t *ant2,
                                int64_t nrows, int32_t nant, int32_t *index_out,
                                int64_t *kept_out, void *stream);

/* ---- binning (PreProcessUVW) ----------------------------------------------
 * Replaces WProjPreProcessUVW (op/WProjPreProcessUVW.cu:111-324) for one
 * channel: per record (row order of vis->row_index) writes
 *   bins[r] = {to_grid, plane, support, corner_x, corner_y, off_x, off_y, conj}
 * (int32 x8) computed in FP64.  Exposed for parity checks and for callers
 * that keep the reference's operator split. */
mti_status mti_preprocess(mti_ctx *ctx, const mti_kernels *k,
                          const mti_image_params *p, const mti_vis *v,
                          int32_t chan, int32_t *bins, void *stream);

/* ---- grid channels --------------------------------------------------------
 * Replaces the WImager DAG (comp/WImager.cpp:167-234: PreProcessUVW, the
 * AccumulatedSum/CreateIndexAndReorder/CreateBlockDataIndex/
 * NotSameSupportManipulate work plan, CreateCompressionPlan/
 * CompressVisibilities, ConvolutionGridder op/ConvolutionGridder.cu:183-550,
 * and the NormalizerValue Sum) for opts->nchan channels at once.
 *   grids : per channel [ny][nx][npol] (MTI_GRID_VUP, the reference's Grid,
 *           comp/WImager.cpp:231-232) or [npol][ny][nx] (MTI_GRID_PVU), of
 *           complex64 or complex128 elements (opts->grid_precision)
 *   norm  : f64 [nchan][npol]               (NormalizerValue)
 * Weighting (uniform/Briggs, not in the reference) is applied per channel.
 * Accumulation is FP64 in registers; complex128 grids add the sums in FP64,
 * complex64 grids are written by the gridder directly (each work unit's sum
 * rounded once; float atomics where a tile is split over work units). */
mti_status mti_grid(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p,
                    const mti_vis *v, const mti_grid_opts *opts, void *grids,
                    double *norm, mti_grid_stats *stats, void *stream);

/* ---- batched jobs (snapshot / multi-field imaging) ------------------------
 * No counterpart in the reference (it images one MeasurementSet per process,
 * app/mtimager.cpp:130-170): many independent jobs, each with its own
 * parameters, convolution functions, visibilities and outputs, gridded by one
 * persistent launch per pass (passes split only at the record/tile limits)
 * and finalized with batched transforms per image size.  All jobs of a call
 * share npol and the table precision; each job's grids are complex128, or
 * complex64 with grids_c64. */
typedef struct {
    const mti_kernels *kernels;  /* tables built for `params`                  */
    mti_image_params params;
    mti_vis vis;
    int32_t chan0;               /* first channel of `vis` imaged              */
    int32_t nchan;               /* channels imaged (one grid/image each)      */
    void   *grids;               /* [nchan] grids, opts.grid_layout; complex128,
                                    or complex64 if grids_c64                  */
    double *norm;                /* f64 [nchan][npol]                          */
    float  *image;               /* f32 [nchan][npol][ny][nx] (finalize)       */
    /* Optional: the antenna columns of vis's rows (int32 [vis.nrows], device)
     * and the antenna count.  When ant1 is set, mti_grid_jobs puts the rows in
     * baseline-major order itself, exactly as mti_sort_rows does
     * (vis/SortIndex.cpp:46-87, autocorrelations dropped), keeping the kept
     * count on the device: no host synchronisation per job.  vis.row_index
     * must then be null.  Null ant1: vis.row_index (or MS order) is used. */
    const int32_t *ant1;
    const int32_t *ant2;
    int32_t nant;
    /* Optional tile mask of the job's grids, as mti_grid_opts.tile_mask
     * (u8 [nchan][ceil(ny/32)][ceil(nx/32)]): written by mti_grid_jobs
     * (accumulate = 0), read by mti_finalize_jobs. */
    uint8_t *tile_mask;
    /* 1: grids are complex64 (FP64 sums rounded once as a cell is added, float
     * atomics where a tile is split over work units); 0: complex128. */
    int32_t grids_c64;
} mti_job;

/* Grids every job (opts: accumulate, psf, weighting, robust, grid_layout;
 * opts.grid_precision is not used -- each job's grids_c64 gives its grids'
 * element type; chan0/nchan come from the jobs). */
mti_status mti_grid_jobs(mti_ctx *ctx, int32_t njobs, const mti_job *jobs,
                         const mti_grid_opts *opts, mti_grid_stats *stats, void *stream);

/* mti_finalize of every job's grids into its image. */
mti_status mti_finalize_jobs(mti_ctx *ctx, int32_t njobs, const mti_job *jobs,
                             int32_t poltype, int32_t grid_layout, void *stream);

/* ---- 2-D inverse FFT (IFFT2D, gafw/StandardOperators/IFFT2D.cpp:86-111) --
 * Unnormalised C2C inverse of nbatch planes complex64 [nbatch][ny][nx], in
 * place, any nx, ny in [1, 32768] (the reference plans cuFFT for any
 * rows x cols).  FP64 arithmetic: mixed-radix passes (radices 2-16, 3, 5, 7,
 * 11, 13), Bluestein for a length with a larger prime factor. */
mti_status mti_fft2d_inverse(mti_ctx *ctx, float *data, int32_t nx, int32_t ny,
                             int32_t nbatch, void *stream);

/* ---- finalize (ImageFinalizer, comp/ImageFinalizer.cpp:31-44) -----
 * PreIFFTUVCorrection + IFFT2D + ZeroToCentre2DShift + StokesConverter
 * (normalise, Stokes, m-flip) + casagridsf taper correction + SincCorrector.
 * The transform runs in FP64; the image is FP32.
 *   grids : per channel in grid_layout (mti_grid_layout), complex64 or
 *           complex128 (grid_precision); read only
 *   norm  : f64 [nchan][npol]
 *   image : f32 [nchan][npol][ny][nx]  (Stokes I,Q,U,V / I,Q / I,V / I)
 * Any image size that validates (2..32768 per side): powers of two with
 * 16 <= nx <= 4096 and ny >= 16 take LDS-blocked passes, every other size
 * mixed-radix HBM passes (Bluestein for a prime factor above 13); odd sizes
 * shift as the reference's OriginToCorner/OriginToCentre (op/common.hcu:105-136). */
mti_status mti_finalize(mti_ctx *ctx, const mti_image_params *p, int32_t npol,
                        int32_t poltype, int32_t nchan, int32_t grid_layout,
                        int32_t grid_precision, const void *grids,
                        const double *norm, float *image, void *stream);
/* mti_finalize of sparse grids: tiles that tile_mask (from mti_grid with
 * opts.tile_mask) marks 0 are read as zeros and never loaded. */
mti_status mti_finalize_sparse(mti_ctx *ctx, const mti_image_params *p, int32_t npol,
                               int32_t poltype, int32_t nchan, int32_t grid_layout,
                               int32_t grid_precision, const void *grids,
                               const uint8_t *tile_mask, const double *norm, float *image,
                               void *stream);

/* ---- the reference's operators with their own array contracts ---------------
 * For callers that keep the reference's operator DAG by name
 * (ImagerFactoryHelper, op/ImagerFactoryHelper.cpp:45-87, and the GAFW
 * standard operators).  Device pointers, no scratch, enqueued on `stream`;
 * FP64 arithmetic rounded once to the reference's output types. */

/* PreProcessUVW parameters (op/WProjPreProcessUVW.cpp:49-56). */
typedef struct {
    int32_t rows, columns;           /* uvImage.rows / uvImage.columns                */
    double  u_increment;             /* uvImage.u_increment, 1 / (cell_l * columns)   */
    double  v_increment;             /* uvImage.v_increment, 1 / (cell_m * rows)      */
    int32_t sampling;                /* ConvolutionFunction.sampling                  */
    double  wsquareincrement;        /* ConvolutionFunction.wsquareincrement          */
    int32_t wplanes;                 /* planes of the convolution-function data       */
} mti_preprocess_ref_params;

/* PreProcessUVW outputs 0-6 (op/WProjPreProcessUVW.cpp:70-84), N records. */
typedef struct {
    int32_t *gridding_data;   /* int4 [N]: corner_x, corner_y * columns + corner_x,
                                 -(corner_x % S), -(corner_y % S); zeros if not gridded */
    int32_t *conv_index;      /* [N] base + oy S^2 T + ox S^2 (0 if not gridded)  */
    int32_t *to_grid;         /* [N] 1: grid this record                          */
    int32_t *to_compress;     /* [N] 1: same kernel and corner as the record before */
    int32_t *support;         /* [N] S (0 if not gridded)                         */
    float   *take_conjugate;  /* [N] -1 where w < 0 (some correlation unflagged), else 1 */
    float   *conv_sum;        /* [P][N] sums[plane][oy T + ox] * weight * (1 - flag) */
} mti_preprocess_ref_outputs;

/* Replaces WProjPreProcessUVW (op/WProjPreProcessUVW.cu:111-324) with its own
 * inputs and outputs: convinfo = input 0 ({support, base} int32 pairs per
 * plane), sums = input 1 (f32 [wplanes][T*T]), v = inputs 2-5 (any layout
 * mti_vis describes; the reference's is uvw f32 [3][N], per-record f32 1/lambda,
 * weights f32 [N][P], int flags [N][P]).  Binning in FP64 (bins equal
 * mti_preprocess's); each correlation's sum uses its own flag. */
mti_status mti_preprocess_reference(mti_ctx *ctx, const mti_preprocess_ref_params *rp,
                                    const int32_t *convinfo, const float *sums, const mti_vis *v,
                                    int32_t chan, const mti_preprocess_ref_outputs *out, void *stream);

/* PreIFFTUVCorrection (op/PreIFFTUVCorrection.cu:28-90): complex [rows][cols][npol]
 * -> [npol][rows][cols], each axis moved by OriginToCorner (op/common.hcu:122-136).
 * c128: complex128 elements, else complex64.  Out of place. */
mti_status mti_uv_reorder(const void *in, void *out, int32_t rows, int32_t cols, int32_t npol,
                          int32_t c128, void *stream);
/* ZeroToCentre2DShift (gafw/StandardOperators/ZeroToCentre2DShift.cu:66-119):
 * every [rows][cols] plane moved by OriginToCentre on both axes.  Out of place. */
mti_status mti_centre_shift(const void *in, void *out, int64_t planes, int32_t rows, int32_t cols,
                            int32_t c128, void *stream);
/* StokesConverter (op/PostIFFTStokesConverter.cu:42-234): complex64 [npol][rows][cols]
 * and the f32 normaliser [npol] -> f32 Stokes [npol][rows][cols], rows flipped. */
mti_status mti_stokes(const float *in, const float *norm, float *out, int32_t npol, int32_t rows,
                      int32_t cols, int32_t poltype, void *stream);
/* casagridsf, lmPlanePostCorrection=true (op/CasaGridSF.cu:151-199): every
 * [rows][cols] plane of `in` divided by psi(|j-rows/2|/(rows/2)) psi(|i-cols/2|/(cols/2))
 * (generate=false), or, with in = NULL, that product written (generate=true). f64: real_double. */
mti_status mti_taper_correct(const void *in, void *out, int64_t planes, int32_t rows, int32_t cols,
                             int32_t f64, void *stream);
/* casagridsf, lmPlanePostCorrection=false (op/CasaGridSF.cu:107-149): the oversampled
 * taper-only gridding function of support S and sampling s, [dim][dim] with dim = S s
 * (odd s) or S s + 1, real or complex (complex_out), f32 or f64. */
mti_status mti_taper_kernel(void *out, int32_t support, int32_t sampling, int32_t complex_out,
                            int32_t f64, void *stream);
/* SincCorrector (op/SincCorrector.cu:29-71): f32 planes times pi x / sin(pi x) per axis. */
mti_status mti_sinc_correct(const float *in, float *out, int64_t planes, int32_t rows, int32_t cols,
                            int32_t sampling, void *stream);
/* The f32 NormalizerValue of the reference's WImager (comp/WImager.cpp:196-197)
 * from the library's FP64 normalisers: n values rounded once. */
mti_status mti_normalizer_f32(const double *norm, float *out, int64_t n, void *stream);

/* ---- prediction (degridding) ---------------------------------------------------
 * No counterpart in the reference (it only images): the adjoint of its imaging
 * path, for the prediction step of a major cycle.
 * mti_model_grids: Stokes model images f32 [nchan][npol][ny][nx] in the
 * finalize's layout (m axis flipped; I,Q,U,V / I,Q / I,V / I as mti_finalize
 * writes them) -> complex64 model grids [nchan][npol][ny][nx] (origin at
 * (ny/2, nx/2), as the gridder's): correlation images from the Stokes
 * formulas' inverse (linear XX=I+Q, XY=U+iV, YX=U-iV, YY=I-Q; circular RR=I+V,
 * RL=Q-iU, LR=Q+iU, LL=I-V), times sinc/taper at each pixel's sky offset (the
 * inverse of the kernel's attenuation), unnormalised forward FFT in FP64.
 * Uses context scratch (at most 512 MiB) even with a caller workspace. */
mti_status mti_model_grids(mti_ctx *ctx, const mti_image_params *p, int32_t npol, int32_t poltype,
                           int32_t nchan, const float *model, void *grids, void *stream);
/* mti_degrid: model visibilities of channels chan0 .. chan0 + nchan - 1 of
 * v's rows (v->row_index order; flags ignored) written as complex64 into
 * vis_out with v's visibility strides (vis_row_stride, vis_chan_stride):
 * V_p = sum conj(K) G_p over the record's support box / the sub-kernel sum,
 * the record binned in FP64 exactly as mti_grid bins it; w < 0 records get
 * the Hermitian image of the mirrored prediction (XY <-> YX for P = 4).
 * Records whose box is off the grid get 0; rows outside row_index are not
 * written.  *predicted (may be NULL; then the call does not block) counts the
 * predicted records. */
mti_status mti_degrid(mti_ctx *ctx, const mti_kernels *k, const mti_image_params *p, const mti_vis *v,
                      int32_t chan0, int32_t nchan, const void *grids, float *vis_out,
                      int64_t *predicted, void *stream);

/* ---- captured batched imaging (HIP graphs) -----------------------------------
 * No counterpart in the reference (it plans each cuFFT call afresh,
 * gafw/StandardOperators/IFFT2D.cpp:86-111): one mti_grid_jobs call (and,
 * with finalize != 0, one mti_finalize_jobs call with poltype) of a fixed job
 * list captured once into a HIP graph.  The graph owns its workspace and the
 * pinned copies of its host tables; every mti_graph_launch images the current
 * contents of the jobs' input buffers into their grids/normalisers/images
 * with no host synchronisation and no allocation (no statistics are read
 * back).  The jobs' buffers, kernel tables and the context must outlive the
 * graph; launches of one graph must be ordered. */
typedef struct mti_graph mti_graph;
mti_status mti_graph_create(mti_ctx *ctx, int32_t njobs, const mti_job *jobs,
                            const mti_grid_opts *opts, int32_t finalize, int32_t poltype,
                            mti_graph **out);
mti_status mti_graph_launch(mti_graph *g, void *stream);
void       mti_graph_destroy(mti_graph *g);

/* ---- launch-shape tuning -----------------------------------------------------
 * Choices of how a context lays work onto the GPU, picked per workload by the
 * autotuner (mtimager_amd/autotune.py).  The reference fixes its gridder
 * blocks at <= 440 / <= 110 records (op/common.hcu:26-27, used by
 * NotSameSupportManipulate, op/NotSameSupportManipulate.cu:32-97); results do
 * not depend on these beyond FP64 summation order.  The tile itself is fixed
 * at 32 x 32 cells (one lane per cell of a 1024-lane workgroup); how the 16
 * waves split it into regions is tuned. */
typedef struct {
    int32_t unit_pairs;       /* (record, tile) pairs per gridder work unit:
                                 power of two in [256, 16384]; 0 = 2048      */
    int32_t whole_columns_ny; /* for ny = 256, where both transforms apply:
                                 column passes whole in one block when
                                 ny <= this, else the four-step split; 0 or a
                                 power of two in [16, 256] (default 256)     */
    int32_t xcd_chunk;        /* gridder work units dealt to each of the 8
                                 XCDs in runs of this many consecutive units
                                 (neighbouring tiles share w-planes, so one
                                 XCD's L2 serves their kernel samples); 0 =
                                 plain round-robin, else a power of two in
                                 [1, 4096]; default 0                         */
    int32_t region_width;     /* gridder: each wave of a tile's workgroup owns
                                 a region of region_width x (64 / region_width)
                                 cells: 4, 8 or 16; 0 = 8 (appended field)   */
} mti_tuning;
mti_status mti_set_tuning(mti_ctx *ctx, const mti_tuning *t);
mti_status mti_get_tuning(const mti_ctx *ctx, mti_tuning *t);

/* ---- transform storage (a precision form) -----------------------------------
 * The packed planes that the finalize's transform passes hand to each other
 * (mti_finalize, mti_finalize_sparse, mti_finalize_jobs, and graphs captured
 * while the setting holds) are stored as complex128 (MTI_GRID_C128, the
 * default) or complex64 (MTI_GRID_C64); butterflies, twiddles and
 * corrections are FP64 either way.  complex64 is the element type of the
 * reference's transform arrays (IFFT2D on complex64,
 * gafw/StandardOperators/IFFT2D.cpp:86-111) and moves half the bytes between
 * passes; its images differ from the complex128 form by the rounding of two
 * intermediate stores (relative RMS ~1e-7).  The small-image path (nx or ny
 * below 16) always stores complex128. */
mti_status mti_set_transform_storage(mti_ctx *ctx, int32_t grid_precision);
mti_status mti_get_transform_storage(const mti_ctx *ctx, int32_t *grid_precision);

/* ---- output shapes and scratch sizes (validate) ----------------------------
 * The validate() half of the reference's operator contract
 * (GPUArrayOperator::validate, gpu/GPUArrayOperator.h:27-43): output
 * dimensions, and the scratch bytes a call takes (ValidationData.bufferSize,
 * gpu/ValidationData.h:25-34), from parameters and sizes only -- no context
 * and no device.  Dimensions are row-major, last fastest; grid elements are
 * complex.  MTI_ERR_INVALID when the same arguments would fail validation. */
mti_status mti_grid_out_shape(const mti_image_params *p, int32_t npol, int32_t nchan,
                              int32_t grid_layout, int64_t grid_dims[4], int64_t norm_dims[2]);
mti_status mti_finalize_out_shape(const mti_image_params *p, int32_t npol, int32_t nchan,
                                  int64_t image_dims[4]);
mti_status mti_preprocess_out_shape(int64_t nrecords, int64_t bins_dims[2]);

/* Device workspace bytes a call with the same arguments takes.  Pointers in
 * the argument structs are not read (only counts, sizes and parameters; the
 * job variants read each job's kernel-table description). */
mti_status mti_sort_rows_workspace_size(int64_t nrows, size_t *bytes);
mti_status mti_grid_workspace_size(const mti_kernel_info *kernels, const mti_image_params *p,
                                   const mti_vis *v, const mti_grid_opts *opts, size_t *bytes);
mti_status mti_grid_jobs_workspace_size(int32_t njobs, const mti_job *jobs,
                                        const mti_grid_opts *opts, size_t *bytes);
mti_status mti_finalize_workspace_size(const mti_image_params *p, int32_t npol, int32_t nchan,
                                       size_t *bytes);
mti_status mti_finalize_jobs_workspace_size(int32_t njobs, const mti_job *jobs, size_t *bytes);
mti_status mti_fft2d_workspace_size(int32_t nx, int32_t ny, int32_t nbatch, size_t *bytes);

/* Caller-owned scratch, as the reference's engine hands each operator the
 * buffer its validate() asked for (GPUSubmissionData.bufferGPUPointer,
 * gpu/GPUSubmissionData.h:27-47).  After mti_set_workspace(ctx, ptr, bytes),
 * every mti_sort_rows, mti_grid, mti_grid_jobs, mti_finalize,
 * mti_finalize_jobs and mti_fft2d_inverse call on ctx lays its scratch out in
 * [ptr, ptr + bytes) and allocates no device memory; a call that needs more
 * than `bytes` (its *_workspace_size) fails with MTI_ERR_NOMEM before any
 * work is enqueued.  ptr = NULL with bytes = 0 returns to scratch owned by the
 * context.  Calls sharing a workspace must be ordered (one stream, or
 * synchronised).  mti_kernels_build stays outside: its tables and scratch are
 * sized by the supports it finds on the device (the reference's UnknownSize
 * buffer mode) and come from the context. */
mti_status mti_set_workspace(mti_ctx *ctx, void *ptr, size_t bytes);
/* Device bytes the context itself holds: scratch it allocated and its cached
 * FFT twiddle tables. */
mti_status mti_context_device_bytes(const mti_ctx *ctx, size_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
