/* mtimager_amd_fits.h -- FITS image-cube writer of the imaging path's output.
 *
 * Replaces ImageManager::createFITS / saveChannel (comp/ImageManager.cpp:68-215),
 * which builds the header from casacore coordinates (a SIN direction
 * coordinate with increments (-l_increment, m_increment) and reference pixel
 * (lpixels/2, mpixels/2), a Stokes coordinate, a spectral coordinate) and
 * writes float32 channels of [stokes][m][l] with cfitsio.  This writer needs
 * neither library: it formats the same cards itself and writes the primary
 * HDU (BITPIX -32, NAXIS = 4: l, m, Stokes, frequency) channel by channel, in
 * the image layout mti_finalize produces (f32 [npol][ny][nx], m already
 * flipped), big-endian and padded to 2880-byte records.
 */
#ifndef MTIMAGER_AMD_FITS_H
#define MTIMAGER_AMD_FITS_H
#include <stdint.h>

#include "mtimager_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mti_fits mti_fits;

typedef struct {
    int32_t nx, ny;          /* image.nx, image.ny (NAXIS1, NAXIS2)                */
    int32_t nstokes;         /* NAXIS3: 1, 2 or 4                                  */
    int32_t nchan;           /* NAXIS4                                             */
    int32_t stokes[4];       /* FITS Stokes codes of the planes (I=1 Q=2 U=3 V=4),
                                equally spaced (e.g. I,Q,U,V or I,V)               */
    double cell_l, cell_m;   /* image.lintervalsize / mintervalsize, radians       */
    double ra, dec;          /* phase centre, radians                              */
    double freq0, dfreq;     /* Hz: channel 0 and the channel spacing              */
    double equinox;          /* 2000 for J2000 (RADESYS FK5)                       */
    const char *specsys;     /* spectral frame, e.g. "LSRK" (NULL: omitted)        */
    const char *telescope;   /* TELESCOP (NULL: omitted)                           */
    const char *observer;    /* OBSERVER (NULL: omitted)                           */
    const char *date_obs;    /* DATE-OBS, ISO 8601 (NULL: omitted)                 */
} mti_fits_desc;

/* Creates the file and writes the header; the data area is written as
 * channels arrive (MTI_ERR_INVALID for a bad description, MTI_ERR_HIP never;
 * I/O failures return MTI_ERR_NOMEM with errno's message in *err if err). */
mti_status mti_fits_create(const char *path, const mti_fits_desc *d, mti_fits **out, const char **err);
/* Writes channel `chan` from host memory, f32 [nstokes][ny][nx]. */
mti_status mti_fits_write_channel(mti_fits *f, int32_t chan, const float *image);
/* Pads the data to a whole 2880-byte record and closes the file. */
mti_status mti_fits_close(mti_fits *f);

#ifdef __cplusplus
}
#endif
#endif
