"""Pins the oracle's TOPO -> LSRK restatement (oracle/mti_oracle_lsrk.c).

The reference converts with casacore Measures (vis/FrequencyConverter.cpp:
112-141); casacore is absent here, so parity with its numbers is unpinned.
These tests pin the restatement against facts it does not compute from:
  * the Earth's heliocentric velocity at J2000.0 from the JPL DE ephemeris
    (Horizons, Earth geocentre relative to the Sun, ICRF, au/day);
  * perihelion and aphelion speeds from the vis-viva law and their dates;
  * the Earth's orbit lying in the ecliptic (no velocity towards its pole);
  * the rotation speed omega * R of an equatorial site and its direction at the
    known Greenwich mean sidereal time of J2000.0 (18h41m50.548s);
  * the solar apex against its galactic coordinates (l, b) = (56, 23) deg via
    the Hipparcos ICRS -> galactic matrix;
  * the sign and first-order size of the shift: an observer moving towards the
    source sees TOPO above LSRK by the factor 1 + beta.
"""
import math

import numpy as np

AU = 149597870700.0
DAY = 86400.0
C = 299792458.0
J2000_MJD_S = (2451545.0 - 2400000.5) * DAY


def _radec(v):
    v = np.asarray(v) / np.linalg.norm(v)
    return math.atan2(v[1], v[0]) % (2 * math.pi), math.asin(v[2])


def test_earth_heliocentric_velocity_j2000(orc):
    vh, vs = orc.earth_velocity(J2000_MJD_S)
    jpl = np.array([-1.720762506872895e-02, -2.897070550656496e-03, -1.256077799819416e-03]) * AU / DAY
    # Keplerian elements + low-precision Moon: a few m/s (1e-8 in frequency)
    assert np.linalg.norm(vh - jpl) < 5.0, vh - jpl
    # the Sun's barycentric speed is ~13 m/s (Jupiter) plus Saturn's few m/s
    assert 8.0 < np.linalg.norm(vs) < 20.0


def test_perihelion_aphelion_speeds(orc):
    e, vmean = 0.0167086, 29784.7  # sqrt(GM_sun / 1 au) with the Earth-Moon mass
    days = np.arange(0.0, 366.0, 0.5)
    speeds = np.array([np.linalg.norm(orc.earth_velocity(J2000_MJD_S + d * DAY)[0]) for d in days])
    vp, va = vmean * math.sqrt((1 + e) / (1 - e)), vmean * math.sqrt((1 - e) / (1 + e))
    # the Moon moves the Earth by +-12.5 m/s about the Earth-Moon barycentre
    assert abs(speeds.max() - vp) < 40.0 and abs(speeds.min() - va) < 40.0
    # perihelion early January, aphelion early July (J2000.0 = Jan 1.5); the
    # speed is flat there, so the lunar term moves the extremes by days
    assert 0.0 <= days[speeds.argmax()] <= 15.0
    assert 170.0 <= days[speeds.argmin()] <= 200.0


def test_orbit_in_the_ecliptic(orc):
    eps = math.radians(23.4392911)
    pole = np.array([0.0, -math.sin(eps), math.cos(eps)])  # ecliptic north pole, J2000 equatorial
    for d in range(0, 365, 30):
        vh, _ = orc.earth_velocity(J2000_MJD_S + d * DAY)
        # only the Moon's 5-degree tilt and the EMB inclination remain
        assert abs(vh @ pole) < 3.0


def test_rotation_velocity_equatorial_site(orc):
    R = 6378137.0
    f = orc.lsrk_frame((R, 0.0, 0.0), 0.0, 0.0, tt_minus_utc=0.0)
    vrot, _, _ = orc.lsrk_velocity(f, J2000_MJD_S)
    omega = 2 * math.pi * 1.00273781191135448 / DAY
    assert abs(np.linalg.norm(vrot) - omega * R) < 1e-6
    # site meridian at RA = GMST(J2000.0 UT) = 18h41m50.548s; moving east, i.e.
    # towards RA + 90 deg
    gmst = math.radians((18 + 41 / 60 + 50.548 / 3600) * 15)
    ra, dec = _radec(vrot)
    assert abs(dec) < 1e-12
    d = (ra - (gmst + math.pi / 2)) % (2 * math.pi)
    assert min(d, 2 * math.pi - d) < math.radians(1e-4)


def test_solar_apex_galactic(orc):
    f = orc.lsrk_frame((0.0, 0.0, 0.0), 0.0, 0.0)
    _, _, vl = orc.lsrk_velocity(f, J2000_MJD_S)
    assert abs(np.linalg.norm(vl) - 20000.0) < 1e-9
    G = np.array([[-0.0548755604, -0.8734370902, -0.4838350155],
                  [0.4941094279, -0.4448296300, 0.7469822445],
                  [-0.8676661490, -0.1980763734, 0.4559837762]])
    g = G @ (vl / np.linalg.norm(vl))
    l, b = math.degrees(math.atan2(g[1], g[0])) % 360.0, math.degrees(math.asin(g[2]))
    assert abs(l - 56.2) < 0.5 and abs(b - 22.8) < 0.5, (l, b)


def test_factor_sign_and_size(orc):
    site = (5109360.0, 2006852.0, -3238948.0)
    t = J2000_MJD_S + 8000.5 * DAY
    f0 = orc.lsrk_frame(site, 0.0, 0.0)
    a, b, c = orc.lsrk_velocity(f0, t)
    v = a + b + c
    beta = np.linalg.norm(v) / C
    for sgn in (1.0, -1.0):
        ra, dec = _radec(sgn * v)
        fac = orc.lsrk_factor(orc.lsrk_frame(site, ra, dec), t)
        # moving towards the source: TOPO is blue-shifted, LSRK lies below it
        assert (fac < 1.0) if sgn > 0 else (fac > 1.0)
        assert abs(fac - (1.0 - sgn * beta)) < 2 * beta * beta
    perp = np.cross(v, [0.0, 0.0, 1.0])
    ra, dec = _radec(perp)
    fac = orc.lsrk_factor(orc.lsrk_frame(site, ra, dec), t)
    assert abs(fac - 1.0) < beta * beta


def test_factors_per_record_runs(orc):
    f = orc.lsrk_frame((5109360.0, 2006852.0, -3238948.0), 1.1, -0.6)
    t0 = J2000_MJD_S + 9000.0 * DAY
    times = np.repeat(t0 + 8.0 * np.arange(50), 7)
    out = orc.lsrk_factors(f, times)
    for k in range(0, len(times), 7):
        assert np.all(out[k:k + 7] == out[k])
        assert out[k] == orc.lsrk_factor(f, times[k])
    # the rotation term changes the factor smoothly within a day
    assert np.all(np.abs(np.diff(out[::7])) < 1e-9)
