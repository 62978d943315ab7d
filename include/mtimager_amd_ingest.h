/* mtimager_amd_ingest.h -- visibility container and chunked loader.
 *
 * Replaces the reference's MeasurementSet ingest: MSVisibilityDataSetAsync
 * (vis/MSVisibilityDataSetAsync.cpp:44-347) reads the MS columns (UVW, DATA,
 * FLAG, WEIGHT, TIME, ANTENNA1/2, the spectral window's CHAN_FREQ, the field's
 * PHASE_DIR, the polarisation type, the telescope position), and the
 * StorageManager typedefs (vis/StorageManager.h:147-150) sort them
 * baseline-major per channel into chunks of `records` rows that feed one
 * WImager::calculate each.  casacore is not part of this build, so the
 * columns live in a container file of this library's own ("MTIVIS01",
 * below) or in host arrays the caller describes; the loader streams them to
 * the device in chunks (a block of channels x a run of rows) through pinned
 * staging buffers on a copy stream, overlapped with the gridding of the
 * previous chunk, sorts each chunk's rows baseline-major on the device
 * (mti_sort_rows), optionally applies per-record TOPO -> LSRK factors
 * (mti_lsrk_factors, the reference's FrequencyConverter) and grids it with
 * mti_grid, accumulating the run of row chunks of one channel block.
 *
 * Container layout (little-endian): a 256-byte header
 *     char magic[8] = "MTIVIS01"; uint32 version = 1; uint32 header_bytes = 256;
 *     int64 nrows; int32 nchan, npol, nant, poltype;
 *     double ra, dec, obs_itrf[3];
 *     uint64 offset[8]  -- byte offset of each column below (256-aligned)
 * then the columns, indexed by mti_column:
 *     TIME f64 [nrows] (MJD seconds, UTC), ANTENNA1 / ANTENNA2 int32 [nrows],
 *     UVW f64 [nrows][3] (metres), WEIGHT f32 [nrows][npol],
 *     CHAN_FREQ f64 [nchan] (Hz, TOPO),
 *     DATA complex64 [nchan][nrows][npol] and FLAG u8 [nchan][nrows][npol]
 *     (channel-major: a block of channels x a run of rows is nchan contiguous
 *     pieces, and Briggs/uniform densities, which are per channel, need only
 *     the channel's own rows).
 */
#ifndef MTIMAGER_AMD_INGEST_H
#define MTIMAGER_AMD_INGEST_H
#include <stddef.h>
#include <stdint.h>

#include "mtimager_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MTI_COL_TIME = 0,
    MTI_COL_ANTENNA1 = 1,
    MTI_COL_ANTENNA2 = 2,
    MTI_COL_UVW = 3,
    MTI_COL_WEIGHT = 4,
    MTI_COL_CHAN_FREQ = 5,
    MTI_COL_DATA = 6,
    MTI_COL_FLAG = 7,
    MTI_COL_COUNT = 8
} mti_column;

typedef struct {
    int64_t nrows;
    int32_t nchan, npol, nant;
    int32_t poltype;        /* mti_poltype (feed POLARIZATION_TYPE X/Y or R/L) */
    double ra, dec;         /* FIELD PHASE_DIR, J2000 radians                  */
    double obs_itrf[3];     /* telescope position, ITRF metres                 */
} mti_dataset_info;

typedef struct mti_dataset mti_dataset;

/* Writes a container file.  Row columns as laid out above; vis and flags in
 * MeasurementSet order [nrows][nchan][npol] (ms_order = 1) or channel-major
 * (ms_order = 0).  MTI_ERR_INVALID for bad sizes, MTI_ERR_NOMEM for I/O
 * failures (message in *err when err is not NULL). */
mti_status mti_dataset_write(const char *path, const mti_dataset_info *info, const double *time,
                             const int32_t *ant1, const int32_t *ant2, const double *uvw,
                             const float *weight, const double *chan_freq, const float *vis,
                             const uint8_t *flags, int32_t ms_order, const char **err);
/* Maps a container file read-only (validated header and sizes). */
mti_status mti_dataset_open(const char *path, mti_dataset **out, const char **err);
/* Describes host arrays already in the container's column layout (no copy).
 * pinned = 1 when DATA/FLAG and the row columns are pinned host memory
 * (hipHostMalloc): the loader then copies them to the device directly instead
 * of through its staging buffers.  The arrays must outlive the dataset. */
mti_status mti_dataset_wrap(const mti_dataset_info *info, const double *time, const int32_t *ant1,
                            const int32_t *ant2, const double *uvw, const float *weight,
                            const double *chan_freq, const float *vis_chan_major,
                            const uint8_t *flags_chan_major, int32_t pinned, mti_dataset **out);
mti_status mti_dataset_get_info(const mti_dataset *ds, mti_dataset_info *info);
/* Host pointer to a column (in the mapping or the wrapped array). */
const void *mti_dataset_column(const mti_dataset *ds, int32_t column);
void mti_dataset_close(mti_dataset *ds);

/* How a loader cuts and grids a dataset. */
typedef struct {
    int32_t chan0, nchan;     /* channels imaged (one grid each)                  */
    int32_t channel_block;    /* channels per chunk (0: 8)                        */
    int64_t chunk_rows;       /* rows per chunk, the reference's `records`
                                 (0: all rows); more than one row chunk per block
                                 needs natural weighting and no tile mask         */
    int32_t lsrk;             /* 1: per-record TOPO -> LSRK factors from TIME and
                                 the dataset's frame                              */
    double  tt_minus_utc;     /* TT - UTC for the LSRK factors (s)                */
    int32_t psf;              /* visibilities replaced by 1+0i                    */
    int32_t weighting;        /* mti_weighting                                    */
    double  robust;
    int32_t grid_layout;      /* mti_grid_layout                                  */
    int32_t grid_precision;   /* mti_grid_precision                               */
    int32_t reader_threads;   /* host threads filling the staging buffers (0: 8)  */
} mti_loader_opts;

typedef struct {
    int64_t chunks;           /* chunks streamed                                  */
    int64_t h2d_bytes;        /* bytes copied to the device                       */
    double  seconds;          /* wall time of the call                            */
    double  reader_seconds;   /* host time spent filling staging buffers          */
    mti_grid_stats grid;      /* summed gridding statistics                       */
} mti_loader_stats;

typedef struct mti_loader mti_loader;

/* Allocates the loader's pinned staging and device chunk buffers for `o`.
 * The dataset and context must outlive the loader. */
mti_status mti_loader_create(mti_ctx *ctx, const mti_dataset *ds, const mti_loader_opts *o,
                             mti_loader **out);
/* Streams every chunk and grids it: grids/norm as mti_grid's for channels
 * o.chan0 .. o.chan0 + o.nchan - 1 (the first chunk of each channel block
 * overwrites, the rest accumulate); tile_mask as mti_grid_opts.tile_mask
 * (NULL: complete grids).  Returns after the last chunk's gridding (the
 * stream is synchronised). */
mti_status mti_loader_grid(mti_loader *l, const mti_kernels *k, const mti_image_params *p,
                           void *grids, double *norm, uint8_t *tile_mask, mti_loader_stats *stats,
                           void *stream);
const char *mti_loader_last_error(const mti_loader *l);
void mti_loader_destroy(mti_loader *l);

#ifdef __cplusplus
}
#endif
#endif
