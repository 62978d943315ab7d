// mti_lsrk.hip -- TOPO -> LSRK frequency factors per record.
//
// Replaces FrequencyConverter::calculate (vis/FrequencyConverter.cpp:112-141),
// which converts every channel's TOPO frequency to LSRK with casacore Measures
// once per distinct TIME value and stores f_LSRK / c as float per (channel,
// record).  Here one lane per record computes, in FP64, the Doppler factor
// f_LSRK / f_TOPO = sqrt(1 - b.b) / (1 + b.n) of an observer moving with
// velocity b c relative to the local standard of rest, towards the phase
// centre n, where the velocity is the sum of
//   * the observatory's rotation about the Earth's axis: ITRF position turned
//     by Greenwich mean sidereal time (Meeus, Astronomical Algorithms, eq.
//     12.4), omega x r, precessed from the mean equator of date to J2000 with
//     the IAU 1976 angles (Lieske et al. 1977);
//   * the Earth's barycentric velocity: central differences of the heliocentric
//     Earth-Moon barycentre and giant planets from JPL's Keplerian elements for
//     1800-2050 (Standish, table 1), the Sun's barycentric offset from their
//     masses, and the Earth's share of the lunar offset (Astronomical Almanac
//     low-precision lunar series);
//   * the standard solar motion: 20 km/s towards RA 18h03m50.29s, Dec
//     +30d00m16.8s (J2000; 18h, +30 in B1900).
// Lanes of equal time compute equal factors (the reference caches the last
// time's answer instead).  Cost: ~10^4 FP64 operations per record, well below
// the record's binning work.
#include "mti_internal.hpp"

namespace mti {
namespace lsrk {

constexpr double PI = 3.14159265358979323846;
constexpr double D2R = PI / 180.0;
constexpr double AS2R = PI / (180.0 * 3600.0);
constexpr double AU_M = 149597870700.0;
constexpr double DAY_S = 86400.0;
constexpr double C_MS = 299792458.0;

// a [au], e, I [deg], L [deg], longitude of perihelion [deg], node [deg] at
// J2000 and their rates per Julian century (ecliptic and equinox of J2000):
// Earth-Moon barycentre, Jupiter, Saturn, Uranus, Neptune.
__constant__ double kElements[5][12] = {
    {1.00000261, 0.01671123, -0.00001531, 100.46457166, 102.93768193, 0.0,
     0.00000562, -0.00004392, -0.01294668, 35999.37244981, 0.32327364, 0.0},
    {5.20288700, 0.04838624, 1.30439695, 34.39644051, 14.72847983, 100.47390909,
     -0.00011607, -0.00013253, -0.00183714, 3034.74612775, 0.21252668, 0.20469106},
    {9.53667594, 0.05386179, 2.48599187, 49.95424423, 92.59887831, 113.66242448,
     -0.00125060, -0.00050991, 0.00193609, 1222.49362201, -0.41897216, -0.28867794},
    {19.18916464, 0.04725744, 0.77263783, 313.23810451, 170.95427630, 74.01692503,
     -0.00196176, -0.00004397, -0.00242939, 428.48202785, 0.40805281, 0.04240589},
    {30.06992276, 0.00859048, 1.77004347, -55.12002969, 44.96476227, 131.78422574,
     0.00026291, 0.00005105, 0.00035372, 218.45945325, -0.32241464, -0.00508664}};
// body mass / solar mass, same order
__constant__ double kMassRatio[5] = {1.0 / 328900.56, 1.0 / 1047.3486, 1.0 / 3497.898,
                                     1.0 / 22902.98, 1.0 / 19412.24};
constexpr double MOON_EARTH_MASS = 0.0123000371;
constexpr double OBLIQUITY = 84381.448 * AS2R;

struct V3 {
    double x, y, z;
};

__device__ V3 orbit_position(int b, double T)
{
    const double *e0 = kElements[b], *de = kElements[b] + 6;
    const double a = e0[0] + de[0] * T, e = e0[1] + de[1] * T;
    const double inc = (e0[2] + de[2] * T) * D2R, L = e0[3] + de[3] * T;
    const double peri = e0[4] + de[4] * T, node = (e0[5] + de[5] * T) * D2R;
    const double argp = peri * D2R - node;
    double M = fmod(L - peri, 360.0);
    if (M > 180.0) M -= 360.0;
    if (M < -180.0) M += 360.0;
    M *= D2R;
    // Kepler's equation by Newton steps
    double E = M + e * sin(M);
    for (int it = 0; it < 30; it++) {
        const double dE = (E - e * sin(E) - M) / (1.0 - e * cos(E));
        E -= dE;
        if (fabs(dE) < 1e-15) break;
    }
    const double px = a * (cos(E) - e), py = a * sqrt(1.0 - e * e) * sin(E);
    double sw, cw, sn, cn, si, ci;
    sincos(argp, &sw, &cw);
    sincos(node, &sn, &cn);
    sincos(inc, &si, &ci);
    return V3{(cw * cn - sw * sn * ci) * px + (-sw * cn - cw * sn * ci) * py,
              (cw * sn + sw * cn * ci) * px + (-sw * sn + cw * cn * ci) * py,
              (sw * si) * px + (cw * si) * py};
}

__device__ double sind(double deg) { return sin(deg * D2R); }
__device__ double cosd(double deg) { return cos(deg * D2R); }

// Geocentric ecliptic Moon [au].
__device__ V3 moon_position(double T)
{
    const double lon = (218.32 + 481267.881 * T + 6.29 * sind(135.0 + 477198.87 * T) -
                        1.27 * sind(259.3 - 413335.36 * T) + 0.66 * sind(235.7 + 890534.22 * T) +
                        0.21 * sind(269.9 + 954397.74 * T) - 0.19 * sind(357.5 + 35999.05 * T) -
                        0.11 * sind(186.5 + 966404.03 * T)) * D2R;
    const double lat = (5.13 * sind(93.3 + 483202.02 * T) + 0.28 * sind(228.2 + 960400.89 * T) -
                        0.28 * sind(318.3 + 6003.15 * T) - 0.17 * sind(217.6 - 407332.21 * T)) * D2R;
    const double par = (0.9508 + 0.0518 * cosd(135.0 + 477198.87 * T) + 0.0095 * cosd(259.3 - 413335.36 * T) +
                        0.0078 * cosd(235.7 + 890534.22 * T) + 0.0028 * cosd(269.9 + 954397.74 * T)) * D2R;
    const double dist = (6378137.0 / sin(par)) / AU_M;
    return V3{dist * cos(lat) * cos(lon), dist * cos(lat) * sin(lon), dist * sin(lat)};
}

// Barycentric ecliptic Earth [au]: the Sun's barycentric offset plus the
// heliocentric Earth-Moon barycentre minus the Earth's share of the Moon.
__device__ V3 earth_position(double T)
{
    const V3 emb = orbit_position(0, T);
    V3 sun{-kMassRatio[0] * emb.x, -kMassRatio[0] * emb.y, -kMassRatio[0] * emb.z};
    double mtot = 1.0 + kMassRatio[0];
    for (int b = 1; b < 5; b++) {
        const V3 p = orbit_position(b, T);
        sun.x -= kMassRatio[b] * p.x;
        sun.y -= kMassRatio[b] * p.y;
        sun.z -= kMassRatio[b] * p.z;
        mtot += kMassRatio[b];
    }
    const V3 m = moon_position(T);
    const double fm = MOON_EARTH_MASS / (1.0 + MOON_EARTH_MASS);
    return V3{sun.x / mtot + emb.x - fm * m.x, sun.y / mtot + emb.y - fm * m.y,
              sun.z / mtot + emb.z - fm * m.z};
}

__device__ V3 rot_z(double a, V3 v)  // frame rotation R3(a)
{
    double s, c;
    sincos(a, &s, &c);
    return V3{c * v.x + s * v.y, -s * v.x + c * v.y, v.z};
}

__device__ V3 rot_y(double a, V3 v)  // frame rotation R2(a)
{
    double s, c;
    sincos(a, &s, &c);
    return V3{c * v.x - s * v.z, v.y, s * v.x + c * v.z};
}

// f_LSRK / f_TOPO at MJD seconds t (UTC).
__device__ double doppler(const mti_lsrk_frame &f, double t)
{
    const double d_ut = (t / DAY_S + 2400000.5) - 2451545.0;
    const double T = (d_ut + f.tt_minus_utc / DAY_S) / 36525.0;
    // rotation: mean sidereal time, omega x r in the mean equator of date
    double gmst = 280.46061837 + 360.98564736629 * d_ut + 0.000387933 * T * T - T * T * T / 38710000.0;
    gmst = fmod(gmst, 360.0) * D2R;
    // the Earth's rotation rate (IERS: 1.00273781191135448 turns per UT1 day)
    const double omega = 2.0 * PI * 1.00273781191135448 / DAY_S;
    double sg, cg;
    sincos(gmst, &sg, &cg);
    const double X = f.obs_itrf[0] * cg - f.obs_itrf[1] * sg;
    const double Y = f.obs_itrf[0] * sg + f.obs_itrf[1] * cg;
    const double zeta = (2306.2181 * T + 0.30188 * T * T + 0.017998 * T * T * T) * AS2R;
    const double zz = (2306.2181 * T + 1.09468 * T * T + 0.018203 * T * T * T) * AS2R;
    const double th = (2004.3109 * T - 0.42665 * T * T - 0.041833 * T * T * T) * AS2R;
    const V3 vrot = rot_z(zeta, rot_y(-th, rot_z(zz, V3{-omega * Y, omega * X, 0.0})));
    // Earth's barycentric velocity (central difference over +-0.01 day)
    const double h = 0.01 / 36525.0;
    const V3 p1 = earth_position(T + h), p0 = earth_position(T - h);
    const double sc = AU_M / (2.0 * 0.01 * DAY_S);
    const V3 ve{(p1.x - p0.x) * sc, (p1.y - p0.y) * sc, (p1.z - p0.z) * sc};
    double se, ce;
    sincos(OBLIQUITY, &se, &ce);
    const V3 vb{ve.x, ce * ve.y - se * ve.z, se * ve.y + ce * ve.z};
    // solar motion
    const double ra_apex = (270.0 + (3.0 * 60.0 + 50.29) * 15.0 / 3600.0) * D2R;
    const double de_apex = (30.0 + 16.8 / 3600.0) * D2R;
    const V3 vl{20000.0 * cos(de_apex) * cos(ra_apex), 20000.0 * cos(de_apex) * sin(ra_apex),
                20000.0 * sin(de_apex)};
    const double bx = (vrot.x + vb.x + vl.x) / C_MS, by = (vrot.y + vb.y + vl.y) / C_MS,
                 bz = (vrot.z + vb.z + vl.z) / C_MS;
    const double nx = cos(f.dec) * cos(f.ra), ny = cos(f.dec) * sin(f.ra), nz = sin(f.dec);
    return sqrt(1.0 - (bx * bx + by * by + bz * bz)) / (1.0 + (bx * nx + by * ny + bz * nz));
}

__global__ __launch_bounds__(256) void k_lsrk_factors(mti_lsrk_frame f, const double *time, int64_t n,
                                                    double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = doppler(f, time[i]);
}

// The reference's per-(channel, record) float 1/lambda: (float)(f_c d / c).
__global__ __launch_bounds__(256) void k_lsrk_inv_lambda(mti_lsrk_frame f, const double *time, int64_t n,
                                                       const double *freq, int nchan, float *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = doppler(f, time[i]);
    for (int c = 0; c < nchan; c++) out[(int64_t)c * n + i] = (float)((freq[c] * d) / C_MS);
}

static mti_status check_frame(mti_ctx *ctx, const mti_lsrk_frame *f)
{
    MTI_CHECK(ctx, f, "null LSRK frame");
    bool fin = std::isfinite(f->ra) && std::isfinite(f->dec) && std::isfinite(f->tt_minus_utc);
    for (int k = 0; k < 3; k++) fin = fin && std::isfinite(f->obs_itrf[k]);
    MTI_CHECK(ctx, fin, "non-finite LSRK frame");
    MTI_CHECK(ctx, f->dec >= -PI / 2 && f->dec <= PI / 2, "phase-centre declination outside [-pi/2, pi/2]");
    const double r = sqrt(f->obs_itrf[0] * f->obs_itrf[0] + f->obs_itrf[1] * f->obs_itrf[1] +
                          f->obs_itrf[2] * f->obs_itrf[2]);
    MTI_CHECK(ctx, r < 1.0e8, "observatory position above 10^8 m from the geocentre");
    return MTI_OK;
}

}  // namespace lsrk
}  // namespace mti

using namespace mti;

extern "C" mti_status mti_lsrk_factors(mti_ctx *ctx, const mti_lsrk_frame *f, const double *time_mjd_s,
                                       int64_t nrows, double *factor_out, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = lsrk::check_frame(ctx, f);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, nrows >= 0, "negative row count");
    if (nrows == 0) return MTI_OK;
    MTI_CHECK(ctx, time_mjd_s && factor_out, "null pointer");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(lsrk::k_lsrk_factors, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0,
                       to_stream(stream), *f, time_mjd_s, nrows, factor_out);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}

extern "C" mti_status mti_lsrk_inv_lambda(mti_ctx *ctx, const mti_lsrk_frame *f, const double *time_mjd_s,
                                          int64_t nrows, const double *freq_hz, int32_t nchan,
                                          float *inv_lambda_out, void *stream)
{
    if (!ctx) return MTI_ERR_INVALID;
    mti_status r = lsrk::check_frame(ctx, f);
    if (r != MTI_OK) return r;
    MTI_CHECK(ctx, nrows >= 0 && nchan >= 1, "bad row or channel count");
    if (nrows == 0) return MTI_OK;
    MTI_CHECK(ctx, time_mjd_s && freq_hz && inv_lambda_out, "null pointer");
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(lsrk::k_lsrk_inv_lambda, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0,
                       to_stream(stream), *f, time_mjd_s, nrows, freq_hz, nchan, inv_lambda_out);
    MTI_HIP(ctx, hipGetLastError());
    return MTI_OK;
}
