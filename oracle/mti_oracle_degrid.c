/*
 * mti_oracle_degrid.c -- FP64 CPU definition of W-projection prediction
 * (degridding), the step next to the imaging path in a major cycle.
 *
 * TEST INFRASTRUCTURE ONLY (see mti_oracle.h).  Not part of the product.
 *
 * The reference has no degridder; this is the new oracle for one, built as
 * the inverse of the reference's imaging steps restated in mti_oracle.c:
 *   - orc_model_grid: Stokes model images in the finalize's layout (m axis
 *     flipped, op/PostIFFTStokesConverter.cu:172) -> correlation images
 *     (the inverse of the Stokes formulas of orc_finalize), multiplied by
 *     the inverse of the attenuation the gridding kernel imposes at each
 *     pixel's sky position (the casagridsf taper, op/CasaGridSF.cu:151-199,
 *     and the nearest-sample sinc, op/SincCorrector.cu:29-71), placed with
 *     the shifts of orc_finalize and transformed with the unnormalised
 *     FORWARD FFT (the adjoint of the imaging's inverse FFT, std/IFFT2D.cpp);
 *   - orc_degrid: each record binned exactly as the gridder bins it
 *     (orc_preprocess, op/WProjPreProcessUVW.cu:111-286, flags ignored),
 *     V_p = sum_taps conj(K[conv_idx + sy S + sx]) G_p[cy + sy][cx + sx]
 *     divided by the sub-kernel's sum sums[plane][oy T + ox] (the quantity
 *     the imaging normaliser divides by, comp/WImager.cpp:196-197), and for
 *     w < 0 (binned at (-u, -v, |w|)) the Hermitian image of the mirrored
 *     prediction: V_p(u) = conj(V_p'(-u)) with p' = p except XY <-> YX (RL <->
 *     LR) for four correlations.
 * Pins (tests/test_oracle_degrid.py): a point source of the model predicts
 * S exp(-2 pi i (u l + v m + w (n - 1))) (the convention the imaging's
 * point-source tests pin) to the kernels' accuracy; the unnormalised
 * degridder is the adjoint of orc_grid in the real inner product; model
 * grid -> orc_finalize returns the model where the taper is not small.
 */
#include "mti_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int64_t pmod(int64_t a, int64_t m)
{
    const int64_t r = a % m;
    return r < 0 ? r + m : r;
}

/* correlation value c[q] (re, im) of Stokes s[] at one pixel */
static void stokes_to_corr(int npol, int poltype, const double *s, double c[4][2])
{
    memset(c, 0, sizeof(double) * 8);
    if (npol == 1) {
        c[0][0] = s[0];
    } else if (npol == 2) {
        c[0][0] = s[0] + s[1];
        c[1][0] = s[0] - s[1];
    } else if (poltype == 0) { /* XX = I+Q, XY = U+iV, YX = U-iV, YY = I-Q */
        c[0][0] = s[0] + s[1];
        c[1][0] = s[2];
        c[1][1] = s[3];
        c[2][0] = s[2];
        c[2][1] = -s[3];
        c[3][0] = s[0] - s[1];
    } else { /* RR = I+V, RL = Q-iU, LR = Q+iU, LL = I-V */
        c[0][0] = s[0] + s[3];
        c[1][0] = s[1];
        c[1][1] = -s[2];
        c[2][0] = s[1];
        c[2][1] = s[2];
        c[3][0] = s[0] - s[3];
    }
}

/* Inverse attenuation of the gridding kernel at the pixel (jy, jx) of the
 * transform output (sky offsets jx - nx/2, jy - ny/2 pixels). */
double orc_model_correction(const orc_params *p, int jy, int jx)
{
    const int nx = p->nx, ny = p->ny, s = p->sampling;
    const double nuy = fabs((double)(jy - ny / 2)) / (double)(ny / 2);
    const double nux = fabs((double)(jx - nx / 2)) / (double)(nx / 2);
    const double tap = orc_gridsf(nuy) * orc_gridsf(nux);
    const double xx = (double)(jx - nx / 2) / ((double)nx * s);
    const double yy = (double)(jy - ny / 2) / ((double)ny * s);
    const double sx = xx == 0.0 ? 1.0 : M_PI * xx / sin(M_PI * xx);
    const double sy = yy == 0.0 ? 1.0 : M_PI * yy / sin(M_PI * yy);
    return tap > 0.0 ? (sx * sy) / tap : 0.0;
}

int orc_model_grid(const orc_params *p, int npol, int poltype, const double *image, double *grid)
{
    const int nx = p->nx, ny = p->ny;
    const int64_t np = (int64_t)nx * ny;
    double *x = (double *)calloc((size_t)(2 * np * npol), sizeof(double));
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny));
    if (!x || !line) {
        free(x);
        free(line);
        return -1;
    }
    for (int jy = 0; jy < ny; jy++)
        for (int jx = 0; jx < nx; jx++) {
            const int r = ny - 1 - jy; /* image rows are m-flipped */
            double s[4] = {0, 0, 0, 0}, c[4][2];
            for (int q = 0; q < npol; q++) s[q] = image[np * q + (int64_t)r * nx + jx];
            stokes_to_corr(npol, poltype, s, c);
            const double f = orc_model_correction(p, jy, jx);
            const int64_t k = pmod(jy - ny / 2, ny) * nx + pmod(jx - nx / 2, nx);
            for (int q = 0; q < npol; q++) {
                x[2 * (np * q + k)] = c[q][0] * f;
                x[2 * (np * q + k) + 1] = c[q][1] * f;
            }
        }
    for (int q = 0; q < npol; q++) {
        double *xq = x + 2 * np * q;
        for (int iy = 0; iy < ny; iy++) orc_fft1d(xq + 2 * (int64_t)iy * nx, nx, -1);
        for (int ix = 0; ix < nx; ix++) {
            for (int iy = 0; iy < ny; iy++) {
                line[2 * iy] = xq[2 * ((int64_t)iy * nx + ix)];
                line[2 * iy + 1] = xq[2 * ((int64_t)iy * nx + ix) + 1];
            }
            orc_fft1d(line, ny, -1);
            for (int iy = 0; iy < ny; iy++) {
                xq[2 * ((int64_t)iy * nx + ix)] = line[2 * iy];
                xq[2 * ((int64_t)iy * nx + ix) + 1] = line[2 * iy + 1];
            }
        }
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                const int64_t k = pmod(gy - ny / 2, ny) * nx + pmod(gx - nx / 2, nx);
                double *g = grid + 2 * (((int64_t)gy * nx + gx) * npol + q);
                g[0] = xq[2 * k];
                g[1] = xq[2 * k + 1];
            }
    }
    free(x);
    free(line);
    return 0;
}

int orc_degrid(const orc_params *p, const orc_kernels *K, int64_t nrows, int npol, const double *uvw,
               double inv_lambda, const double *grid, int normalise, double *vis, int64_t *predicted)
{
    const int64_t nb = nrows > 0 ? nrows : 1;
    orc_bin *b = (orc_bin *)malloc(sizeof(orc_bin) * (size_t)nb);
    uint8_t *fl = (uint8_t *)calloc((size_t)nb * npol, 1);
    if (!b || !fl) {
        free(b);
        free(fl);
        return -1;
    }
    orc_preprocess(p, K, nrows, 1, npol, uvw, &inv_lambda, fl, 0, b);
    const int nx = p->nx, T = K->T;
    int64_t n = 0;
    for (int64_t r = 0; r < nrows; r++) {
        double v[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        if (b[r].to_grid) {
            const int S = b[r].support;
            for (int sy = 0; sy < S; sy++)
                for (int sx = 0; sx < S; sx++) {
                    const double *k = K->table + 2 * (b[r].conv_idx + (int64_t)sy * S + sx);
                    const double *g = grid + 2 * (((int64_t)(b[r].corner_y + sy) * nx + b[r].corner_x + sx) * npol);
                    for (int q = 0; q < npol; q++) { /* conj(k) * g */
                        v[q][0] += k[0] * g[2 * q] + k[1] * g[2 * q + 1];
                        v[q][1] += k[0] * g[2 * q + 1] - k[1] * g[2 * q];
                    }
                }
            const double sum = normalise ? K->sums[(int64_t)b[r].plane * T * T + b[r].off_y * T + b[r].off_x] : 1.0;
            for (int q = 0; q < npol; q++) {
                v[q][0] /= sum;
                v[q][1] /= sum;
            }
            n++;
        }
        for (int q = 0; q < npol; q++) {
            int src = q;
            if (b[r].conj && npol == 4 && (q == 1 || q == 2)) src = 3 - q;
            vis[2 * (r * npol + q)] = v[src][0];
            vis[2 * (r * npol + q) + 1] = b[r].conj ? -v[src][1] : v[src][1];
        }
    }
    if (predicted) *predicted = n;
    free(b);
    free(fl);
    return 0;
}
