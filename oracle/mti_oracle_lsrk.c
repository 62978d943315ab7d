/*
 * mti_oracle_lsrk.c -- FP64 CPU restatement of the TOPO -> LSRK frequency
 * conversion that mt-imager applies per record.
 *
 * TEST INFRASTRUCTURE ONLY (see mti_oracle.h).  Not part of the product.
 *
 * The reference (vis/FrequencyConverter.cpp:112-141) converts each channel's
 * TOPO frequency to LSRK once per distinct TIME value with casacore Measures
 * (casacore 1.5: MFrequency::Convert TOPO -> LSRK in a MeasFrame of epoch,
 * observatory position and phase centre) and stores f_LSRK / c as float per
 * (channel, record).  casacore is not in this image, so its algorithm is
 * restated from the published models it composes:
 *   - observatory velocity from the Earth's rotation: the ITRF position turned
 *     by Greenwich mean sidereal time (Meeus, Astronomical Algorithms, 2nd
 *     ed., eq. 12.4), v = omega x r, then taken from the mean equator of date
 *     to J2000 by the IAU 1976 precession angles (Lieske et al. 1977);
 *   - the Earth's barycentric velocity: heliocentric Earth-Moon barycentre
 *     and giant planets from the JPL Keplerian elements valid 1800-2050
 *     (Standish, "Keplerian Elements for Approximate Positions of the Major
 *     Planets", table 1), the Sun's barycentric offset from the planets'
 *     masses, the Earth's offset from the Earth-Moon barycentre from the
 *     low-precision lunar series of the Astronomical Almanac (section D);
 *     velocities are central differences of those positions;
 *   - the solar motion defining LSRK: 20 km/s towards RA 18h, Dec +30 (B1900),
 *     i.e. RA 18h03m50.29s, Dec +30d00m16.8s (J2000);
 *   - the relativistic Doppler factor of an observer moving with beta towards
 *     a source in direction n: f_LSRK = f_TOPO sqrt(1 - beta^2) / (1 + beta.n).
 * Parity with casacore's own numbers is unpinned (no casacore here).  What
 * pins this restatement (tests/test_oracle_lsrk.py): the Earth's barycentric
 * velocity at J2000.0 against the JPL ephemeris value, the diurnal velocity
 * of an equatorial site against omega * R at a known sidereal time, the solar
 * apex against its galactic coordinates, and the sign and first-order size
 * of the shift along and across the total velocity.
 */
#include "mti_oracle.h"

#include <math.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define D2R (M_PI / 180.0)
#define AS2R (M_PI / (180.0 * 3600.0))
#define AU_M 149597870700.0
#define DAY_S 86400.0
#define C_MS 299792458.0

/* Standish table 1 (1800-2050): a [au], e, I [deg], L [deg], varpi [deg],
 * Omega [deg] at J2000 and their rates per Julian century. */
static const double kElem[5][12] = {
    /* Earth-Moon barycentre */
    {1.00000261, 0.01671123, -0.00001531, 100.46457166, 102.93768193, 0.0,
     0.00000562, -0.00004392, -0.01294668, 35999.37244981, 0.32327364, 0.0},
    /* Jupiter */
    {5.20288700, 0.04838624, 1.30439695, 34.39644051, 14.72847983, 100.47390909,
     -0.00011607, -0.00013253, -0.00183714, 3034.74612775, 0.21252668, 0.20469106},
    /* Saturn */
    {9.53667594, 0.05386179, 2.48599187, 49.95424423, 92.59887831, 113.66242448,
     -0.00125060, -0.00050991, 0.00193609, 1222.49362201, -0.41897216, -0.28867794},
    /* Uranus */
    {19.18916464, 0.04725744, 0.77263783, 313.23810451, 170.95427630, 74.01692503,
     -0.00196176, -0.00004397, -0.00242939, 428.48202785, 0.40805281, 0.04240589},
    /* Neptune */
    {30.06992276, 0.00859048, 1.77004347, -55.12002969, 44.96476227, 131.78422574,
     0.00026291, 0.00005105, 0.00035372, 218.45945325, -0.32241464, -0.00508664}};
/* mass / mass of the Sun: Earth+Moon, Jupiter, Saturn, Uranus, Neptune */
static const double kMu[5] = {1.0 / 328900.56, 1.0 / 1047.3486, 1.0 / 3497.898,
                              1.0 / 22902.98, 1.0 / 19412.24};
#define MOON_EARTH_MASS 0.0123000371 /* m_Moon / m_Earth */
#define OBLIQUITY_J2000 (84381.448 * AS2R)

/* Heliocentric J2000 ecliptic position [au] of body b at T (TT centuries). */
static void kepler_position(int b, double T, double r[3])
{
    const double *e0 = kElem[b], *de = kElem[b] + 6;
    const double a = e0[0] + de[0] * T, e = e0[1] + de[1] * T;
    const double I = (e0[2] + de[2] * T) * D2R, L = e0[3] + de[3] * T;
    const double varpi = e0[4] + de[4] * T, Om = (e0[5] + de[5] * T) * D2R;
    const double w = varpi * D2R - Om;
    double M = fmod(L - varpi, 360.0);
    if (M > 180.0) M -= 360.0;
    if (M < -180.0) M += 360.0;
    M *= D2R;
    double E = M + e * sin(M);
    for (int it = 0; it < 30; it++) {
        const double dE = (E - e * sin(E) - M) / (1.0 - e * cos(E));
        E -= dE;
        if (fabs(dE) < 1e-15) break;
    }
    const double xp = a * (cos(E) - e), yp = a * sqrt(1.0 - e * e) * sin(E);
    const double cw = cos(w), sw = sin(w), cO = cos(Om), sO = sin(Om), cI = cos(I), sI = sin(I);
    r[0] = (cw * cO - sw * sO * cI) * xp + (-sw * cO - cw * sO * cI) * yp;
    r[1] = (cw * sO + sw * cO * cI) * xp + (-sw * sO + cw * cO * cI) * yp;
    r[2] = (sw * sI) * xp + (cw * sI) * yp;
}

/* Geocentric ecliptic position of the Moon [au] (Astronomical Almanac
 * low-precision series; T in centuries from J2000). */
static void moon_position(double T, double r[3])
{
    const double lam = (218.32 + 481267.881 * T + 6.29 * sin((135.0 + 477198.87 * T) * D2R) -
                        1.27 * sin((259.3 - 413335.36 * T) * D2R) +
                        0.66 * sin((235.7 + 890534.22 * T) * D2R) +
                        0.21 * sin((269.9 + 954397.74 * T) * D2R) -
                        0.19 * sin((357.5 + 35999.05 * T) * D2R) -
                        0.11 * sin((186.5 + 966404.03 * T) * D2R)) * D2R;
    const double bet = (5.13 * sin((93.3 + 483202.02 * T) * D2R) +
                        0.28 * sin((228.2 + 960400.89 * T) * D2R) -
                        0.28 * sin((318.3 + 6003.15 * T) * D2R) -
                        0.17 * sin((217.6 - 407332.21 * T) * D2R)) * D2R;
    const double par = (0.9508 + 0.0518 * cos((135.0 + 477198.87 * T) * D2R) +
                        0.0095 * cos((259.3 - 413335.36 * T) * D2R) +
                        0.0078 * cos((235.7 + 890534.22 * T) * D2R) +
                        0.0028 * cos((269.9 + 954397.74 * T) * D2R)) * D2R;
    const double dist = (6378137.0 / sin(par)) / AU_M;
    r[0] = dist * cos(bet) * cos(lam);
    r[1] = dist * cos(bet) * sin(lam);
    r[2] = dist * sin(bet);
}

/* J2000 ecliptic positions [au]: the Earth relative to the Sun (Earth-Moon
 * barycentre minus the Earth's share of the lunar offset) and the Sun relative
 * to the solar-system barycentre (sum of mu_b r_b over the bodies, divided by
 * the total mass). */
static void earth_positions(double T, double helio[3], double sun[3])
{
    double emb[3], p[3], moon[3];
    double mtot = 1.0;
    sun[0] = sun[1] = sun[2] = 0.0;
    kepler_position(0, T, emb);
    for (int b = 0; b < 5; b++) {
        if (b == 0) {
            p[0] = emb[0]; p[1] = emb[1]; p[2] = emb[2];
        } else {
            kepler_position(b, T, p);
        }
        for (int k = 0; k < 3; k++) sun[k] -= kMu[b] * p[k];
        mtot += kMu[b];
    }
    for (int k = 0; k < 3; k++) sun[k] /= mtot;
    moon_position(T, moon);
    const double fm = MOON_EARTH_MASS / (1.0 + MOON_EARTH_MASS);
    for (int k = 0; k < 3; k++) helio[k] = emb[k] - fm * moon[k];
}

void orc_earth_velocity(double mjd_tt_s, double v_helio[3], double v_sun[3])
{
    /* central differences of the positions over +-0.01 day, ecliptic ->
     * equator of J2000 */
    const double T = (mjd_tt_s / DAY_S + 2400000.5 - 2451545.0) / 36525.0;
    const double h = 0.01 / 36525.0;
    double hp[3], hm[3], sp[3], sm[3], a[3], b[3];
    earth_positions(T + h, hp, sp);
    earth_positions(T - h, hm, sm);
    const double scale = AU_M / (2.0 * 0.01 * DAY_S);
    for (int k = 0; k < 3; k++) {
        a[k] = (hp[k] - hm[k]) * scale;
        b[k] = (sp[k] - sm[k]) * scale;
    }
    const double ce = cos(OBLIQUITY_J2000), se = sin(OBLIQUITY_J2000);
    v_helio[0] = a[0];
    v_helio[1] = ce * a[1] - se * a[2];
    v_helio[2] = se * a[1] + ce * a[2];
    v_sun[0] = b[0];
    v_sun[1] = ce * b[1] - se * b[2];
    v_sun[2] = se * b[1] + ce * b[2];
}

/* Rotation about z by angle a applied to vector x (frame rotation R3(a)). */
static void rot3(double a, const double x[3], double y[3])
{
    const double c = cos(a), s = sin(a);
    const double x0 = x[0], x1 = x[1];
    y[0] = c * x0 + s * x1;
    y[1] = -s * x0 + c * x1;
    y[2] = x[2];
}

/* Frame rotation R2(a). */
static void rot2(double a, const double x[3], double y[3])
{
    const double c = cos(a), s = sin(a);
    const double x0 = x[0], x2 = x[2];
    y[0] = c * x0 - s * x2;
    y[1] = x[1];
    y[2] = s * x0 + c * x2;
}

void orc_lsrk_velocity(const orc_lsrk_frame *f, double mjd_utc_s, double v_rot[3],
                       double v_bary[3], double v_lsr[3])
{
    const double jd_ut = mjd_utc_s / DAY_S + 2400000.5;
    const double d_ut = jd_ut - 2451545.0;
    const double d_tt = d_ut + f->tt_minus_utc / DAY_S;
    const double T = d_tt / 36525.0;

    /* Earth rotation: GMST (Meeus 12.4), v = omega x r in the mean equator of
     * date, then precessed to J2000 (v_J2000 = R3(zeta) R2(-theta) R3(z) v). */
    double gmst = 280.46061837 + 360.98564736629 * d_ut + 0.000387933 * T * T - T * T * T / 38710000.0;
    gmst = fmod(gmst, 360.0) * D2R;
    /* the Earth's rotation rate (IERS: 1.00273781191135448 turns per UT1 day) */
    const double omega = 2.0 * M_PI * 1.00273781191135448 / DAY_S;
    const double cg = cos(gmst), sg = sin(gmst);
    const double X = f->obs_itrf[0] * cg - f->obs_itrf[1] * sg;
    const double Y = f->obs_itrf[0] * sg + f->obs_itrf[1] * cg;
    double vd[3] = {-omega * Y, omega * X, 0.0}, t1[3], t2[3];
    const double zeta = (2306.2181 * T + 0.30188 * T * T + 0.017998 * T * T * T) * AS2R;
    const double zz = (2306.2181 * T + 1.09468 * T * T + 0.018203 * T * T * T) * AS2R;
    const double th = (2004.3109 * T - 0.42665 * T * T - 0.041833 * T * T * T) * AS2R;
    rot3(zz, vd, t1);
    rot2(-th, t1, t2);
    rot3(zeta, t2, v_rot);

    /* Earth's barycentric velocity */
    double vh[3], vs[3];
    orc_earth_velocity(mjd_utc_s + f->tt_minus_utc, vh, vs);
    for (int k = 0; k < 3; k++) v_bary[k] = vh[k] + vs[k];

    /* solar motion relative to the LSR */
    const double ra = (270.0 + (3.0 * 60.0 + 50.29) * 15.0 / 3600.0) * D2R;
    const double de = (30.0 + 16.8 / 3600.0) * D2R;
    v_lsr[0] = 20000.0 * cos(de) * cos(ra);
    v_lsr[1] = 20000.0 * cos(de) * sin(ra);
    v_lsr[2] = 20000.0 * sin(de);
}

double orc_lsrk_factor(const orc_lsrk_frame *f, double mjd_utc_s)
{
    double a[3], b[3], c[3];
    orc_lsrk_velocity(f, mjd_utc_s, a, b, c);
    const double n[3] = {cos(f->dec) * cos(f->ra), cos(f->dec) * sin(f->ra), sin(f->dec)};
    double bn = 0.0, bb = 0.0;
    for (int k = 0; k < 3; k++) {
        const double beta = (a[k] + b[k] + c[k]) / C_MS;
        bn += beta * n[k];
        bb += beta * beta;
    }
    return sqrt(1.0 - bb) / (1.0 + bn);
}

void orc_lsrk_factors(const orc_lsrk_frame *f, const double *mjd_utc_s, int64_t n, double *out)
{
    /* one evaluation per run of equal times, as the reference caches the last
     * conversion (vis/FrequencyConverter.cpp:119-125) */
    double last_t = 0.0, last = 1.0;
    int have = 0;
    for (int64_t i = 0; i < n; i++) {
        if (!have || mjd_utc_s[i] != last_t) {
            last_t = mjd_utc_s[i];
            last = orc_lsrk_factor(f, last_t);
            have = 1;
        }
        out[i] = last;
    }
}
