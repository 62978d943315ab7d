"""FITS cube output on the host: byte layout (2880-byte records, 80-column
cards, big-endian float32 in [chan][stokes][m][l] order), the header values
the reference's ImageManager sets (comp/ImageManager.cpp:68-185), and the
orientation of a point source imaged by the oracle."""
import numpy as np
import pytest

import mtimager_amd as M
from mtimager_amd import fits

ARC = np.pi / 180 / 3600


def _cards(path):
    raw = open(path, "rb").read()
    cards = []
    for i in range(0, len(raw), 80):
        c = raw[i:i + 80].decode("ascii")
        cards.append(c)
        if c.startswith("END "):
            break
    return raw, cards


def test_cube_round_trip_and_header(tmp_path):
    p = M.make_params(32, 16, 2.0, 3.0, "wproj", 8, 4)
    rng = np.random.default_rng(0)
    img = rng.standard_normal((3, 4, 16, 32)).astype(np.float32)
    freqs = 1.4e9 + 1e6 * np.arange(3)
    path = tmp_path / "cube.fits"
    fits.write_cube(path, p, img, freqs, "IQUV", ra=np.radians(150.0), dec=np.radians(-30.0), telescope="MeerKAT")
    raw, cards = _cards(path)
    assert len(raw) % 2880 == 0
    assert all(len(c) == 80 for c in cards) and cards[-1].rstrip() == "END"
    for c in cards[:-1]:  # fixed format: numbers and logicals end in column 30
        if c[8:10] == "= " and not c[10:].lstrip().startswith("'"):
            assert c[29] != " " and c[10:30].strip() == c[10:].split(" / ")[0].strip()
    h, data = fits.read_cube(path)
    np.testing.assert_array_equal(data, img)
    assert (h["SIMPLE"], h["BITPIX"], h["NAXIS"]) == (True, -32, 4)
    assert (h["NAXIS1"], h["NAXIS2"], h["NAXIS3"], h["NAXIS4"]) == (32, 16, 4, 3)
    assert (h["CTYPE1"], h["CTYPE2"], h["CTYPE3"], h["CTYPE4"]) == ("RA---SIN", "DEC--SIN", "STOKES", "FREQ")
    assert h["CDELT1"] == pytest.approx(-2.0 / 3600, rel=1e-12)
    assert h["CDELT2"] == pytest.approx(3.0 / 3600, rel=1e-12)
    assert (h["CRPIX1"], h["CRPIX2"]) == (17.0, 9.0)
    assert (h["CRVAL1"], h["CRVAL2"]) == (pytest.approx(150.0), pytest.approx(-30.0))
    assert (h["CRVAL3"], h["CDELT3"], h["CRPIX3"]) == (1.0, 1.0, 1.0)
    assert (h["CRVAL4"], h["CDELT4"], h["CRPIX4"]) == (1.4e9, 1e6, 1.0)
    assert h["BUNIT"] == "JY/BEAM" and h["TELESCOP"] == "MeerKAT" and h["SPECSYS"] == "LSRK"
    assert h["EQUINOX"] == 2000.0 and h["RADESYS"] == "FK5"


def test_stokes_axis_of_two_circular_correlations(tmp_path):
    # circular RR, LL give I and V: a linear axis 1, 4
    p = M.make_params(16, 16, 2.0, 2.0, "wproj", 8, 4)
    path = tmp_path / "iv.fits"
    fits.write_cube(path, p, np.zeros((1, 2, 16, 16), np.float32), [1.4e9],
                    fits.FINALIZE_STOKES[(2, "circular")], channel_width=2e5)
    h, _ = fits.read_cube(path)
    assert (h["CRVAL3"], h["CDELT3"], h["CDELT4"]) == (1.0, 3.0, 2e5)


def test_channels_in_any_order(tmp_path):
    p = M.make_params(16, 8, 2.0, 2.0, "wproj", 8, 4)
    img = np.arange(4 * 1 * 8 * 16, dtype=np.float32).reshape(4, 1, 8, 16)
    path = tmp_path / "order.fits"
    with fits.CubeWriter(path, p, 1e9 + 1e5 * np.arange(4)) as w:
        for c in (2, 0, 3, 1):
            w.write_channel(c, img[c])
    _, data = fits.read_cube(path)
    np.testing.assert_array_equal(data, img)


@pytest.mark.parametrize("bad", [dict(freqs_hz=[1.0e9, 1.1e9, 1.3e9]), dict(stokes="QI"), dict(stokes="IQU")])
def test_bad_descriptions_are_refused(tmp_path, bad):
    a = dict(freqs_hz=[1.0e9, 1.1e9], stokes="I")
    a.update(bad)
    nst = len(a["stokes"])
    with pytest.raises(M.MTIError):
        fits.write_cube(tmp_path / "bad.fits", M.make_params(16, 16, 2.0, 2.0, "wproj", 8, 4),
                        np.zeros((len(a["freqs_hz"]), nst, 16, 16), np.float32), a["freqs_hz"], a["stokes"])


def test_point_source_pixel_and_header(orc, tmp_path):
    """A source at (l, m) = (+10, -7) cells imaged by the oracle (the
    reference's finalize: fftshift, m flip) lands at FITS pixel
    (N/2 + 1 + 10, N/2 + 7); with the reference's header (CRPIX = N/2 + 1,
    CDELT1 = -cell) that pixel reads as RA offset -10 cells and Dec offset
    +6 cells -- the reference's own orientation, reproduced for drop-in
    output."""
    n, cell = 128, 20 * ARC
    p = orc.make_params(n, n, cell, cell, "wproj", 4, 4)
    K = orc.Kernels(p)
    rng = np.random.default_rng(5)
    du = 1.0 / (cell * n)
    uvw = np.zeros((300, 3))
    uvw[:, :2] = rng.uniform(-0.8 * (n / 2 - 4) * du, 0.8 * (n / 2 - 4) * du, (300, 2))
    l0, m0 = 10 * cell, -7 * cell
    vis = np.exp(-2j * np.pi * (uvw[:, 0] * l0 + uvw[:, 1] * m0)).astype(np.complex64)
    g, nrm, _ = orc.grid(K, uvw, np.array([1.0]), vis.reshape(300, 1, 1), np.zeros((300, 1, 1), np.uint8),
                         np.ones((300, 1), np.float32), 0)
    img = orc.finalize(p, g, nrm)[0].astype(np.float32)
    path = tmp_path / "src.fits"
    fits.write_cube(path, M.make_params(n, n, 20.0, 20.0, "wproj", 4, 4), img[None, None], [1.4e9],
                    channel_width=1e6)
    h, data = fits.read_cube(path)
    r, c = np.unravel_index(np.argmax(data[0, 0]), data.shape[2:])
    x, y = c + 1, r + 1  # FITS pixels count from 1
    assert (x, y) == (n // 2 + 1 + 10, n // 2 + 7)
    assert (x - h["CRPIX1"]) * h["CDELT1"] == pytest.approx(-10 * 20.0 / 3600)
    assert (y - h["CRPIX2"]) * h["CDELT2"] == pytest.approx(6 * 20.0 / 3600)
