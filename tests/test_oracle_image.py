"""Pins orc_fft1d, orc_grid and orc_finalize: numpy's FFT, closed-form Stokes
of a delta grid, point-source position, PSF peak, and the dirty image against
a direct DFT with the w-term (and NOT against one without it)."""
import numpy as np
import pytest

ARC = np.pi / 180 / 3600


@pytest.mark.parametrize("n", [1, 2, 8, 64, 100, 127, 128, 1000])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_matches_numpy(orc, n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    np.testing.assert_allclose(orc.fft1d(x, sign), ref, rtol=0, atol=1e-11 * max(1, n))


def _corr(n, s):
    """sinc(x)/psi(|i-n/2|/(n/2)), x = (i-n/2)/(n s), evaluated on the array
    index (the reference applies it after the m flip)."""
    from oracle import gridsf
    out = np.empty(n)
    for i in range(n):
        x = (i - n // 2) / (n * s)
        sc = 1.0 if x == 0 else np.pi * x / np.sin(np.pi * x)
        out[i] = sc / gridsf(abs(i - n // 2) / (n // 2))
    return out


@pytest.mark.parametrize("npol,poltype", [(1, "linear"), (2, "linear"), (2, "circular"),
                                          (4, "linear"), (4, "circular")])
def test_stokes_of_a_delta_grid(orc, npol, poltype):
    """A delta at the uv centre transforms to a constant: every pixel is the
    Stokes combination of the normalised correlations times the corrections."""
    n, s = 16, 4
    p = orc.make_params(n, n, 1e-3, 1e-3, "wproj", 2, s)
    rng = np.random.default_rng(npol)
    a = rng.standard_normal(npol) + 1j * rng.standard_normal(npol)
    norm = rng.uniform(0.5, 2.0, npol)
    g = np.zeros((n, n, npol, 2))
    g[n // 2, n // 2, :, 0], g[n // 2, n // 2, :, 1] = a.real, a.imag
    img = orc.finalize(p, g, norm, poltype)
    b = a / norm
    if npol == 1:
        st = [b[0].real]
    elif npol == 2:
        st = [(b[0].real + b[1].real) / 2, (b[0].real - b[1].real) / 2]
    elif poltype == "linear":  # XX XY YX YY -> I Q U V
        st = [(b[0].real + b[3].real) / 2, (b[0].real - b[3].real) / 2,
              (b[1].real + b[2].real) / 2, (b[1].imag - b[2].imag) / 2]
    else:  # RR RL LR LL -> I Q U V
        st = [(b[0].real + b[3].real) / 2, (b[1].real + b[2].real) / 2,
              (b[2].imag - b[1].imag) / 2, (b[0].real - b[3].real) / 2]
    c = np.outer(_corr(n, s), _corr(n, s))
    for q in range(npol):
        np.testing.assert_allclose(img[q], st[q] * c, rtol=1e-12, atol=1e-14)


def test_ifft_sign_and_m_flip(orc):
    """Delta of value i at u = +1 cell: image(jx, jy) = Re(i e^{+2 pi i (jx-n/2)/n})
    = -sin(2 pi (jx - n/2)/n); with v = +1 cell the variation is along jy and
    appears flipped (row r = n-1-jy)."""
    n, s = 32, 4
    p = orc.make_params(n, n, 1e-3, 1e-3, "wproj", 2, s)
    cr = _corr(n, s)
    g = np.zeros((n, n, 1, 2))
    g[n // 2, n // 2 + 1, 0, 1] = 1.0
    img = orc.finalize(p, g, np.ones(1))[0]
    jx = np.arange(n)
    ref = -np.sin(2 * np.pi * (jx - n // 2) / n)[None, :] * np.outer(cr, cr)
    np.testing.assert_allclose(img, ref, atol=1e-12 * np.abs(ref).max())
    g = np.zeros((n, n, 1, 2))
    g[n // 2 + 1, n // 2, 0, 1] = 1.0
    img = orc.finalize(p, g, np.ones(1))[0]
    jy = n - 1 - np.arange(n)  # array row r holds jy = n-1-r
    ref = -np.sin(2 * np.pi * (jy - n // 2) / n)[:, None] * np.outer(cr, cr)
    np.testing.assert_allclose(img, ref, atol=1e-12 * np.abs(ref).max())


def _random_vis(nvis, umax, wmax, seed):
    rng = np.random.default_rng(seed)
    uvw = np.stack([rng.uniform(-umax, umax, nvis), rng.uniform(-umax, umax, nvis),
                    rng.uniform(-wmax, wmax, nvis)], axis=1)
    return uvw


def _image(orc, p, K, uvw, vis, npol=1, psf=False):
    nrows = len(uvw)
    flags = np.zeros((nrows, 1, npol), np.uint8)
    weight = np.ones((nrows, npol), np.float32)
    g, nrm, ng = orc.grid(K, uvw, np.array([1.0]), vis.reshape(nrows, 1, npol), flags, weight, 0,
                          psf=psf)
    return orc.finalize(p, g, nrm)[0], ng


def test_point_source_lands_on_its_pixel(orc):
    n, cell = 256, 20 * ARC
    p = orc.make_params(n, n, cell, cell, "wproj", 4, 4)
    K = orc.Kernels(p)
    du = 1.0 / (cell * n)
    uvw = _random_vis(400, 0.8 * (n / 2 - 4) * du, 0.0, 5)
    l0, m0 = 10 * cell, -7 * cell
    vis = np.exp(-2j * np.pi * (uvw[:, 0] * l0 + uvw[:, 1] * m0)).astype(np.complex64)
    img, ng = _image(orc, p, K, uvw, vis)
    assert ng == 400
    r, c = np.unravel_index(np.argmax(img), img.shape)
    assert (c, r) == (n // 2 + 10, n - 1 - (n // 2 - 7))
    assert img[r, c] == pytest.approx(1.0, abs=1e-2)


def test_psf_peak_is_one(orc):
    n, cell = 128, 20 * ARC
    p = orc.make_params(n, n, cell, cell, "wproj", 8, 4)
    K = orc.Kernels(p)
    du = 1.0 / (cell * n)
    uvw = _random_vis(300, 0.8 * (n / 2 - 4) * du, 50.0, 6)
    img, _ = _image(orc, p, K, uvw, np.zeros(300, np.complex64), psf=True)
    r, c = np.unravel_index(np.argmax(img), img.shape)
    assert (c, r) == (n // 2, n - 1 - n // 2)
    assert img[r, c] == pytest.approx(1.0, abs=1e-3)


@pytest.mark.parametrize("px,py,lo", [(-90, 70, 0.88), (-60, 40, 0.95), (30, -20, 0.95)])
def test_w_projection_refocuses_a_wide_field_source(orc, px, py, lo):
    """4.3-degree field, |w| up to w_max: at (-90, 70) pixels the w-term phase
    is ~0.4 turns, so an image made without it keeps <20% of the flux, while
    the W-projection image must recover the source on its own pixel.  The few
    per-cent shortfall is the reference's taper mismatch (screen taper
    psi(|c|/(S_t/2)) vs correction psi(|j-N/2|/(N/2)), S_t = 65 here)."""
    n, cell, s, W = 256, 60 * ARC, 4, 64
    p = orc.make_params(n, n, cell, cell, "wproj", W, s)
    K = orc.Kernels(p)
    du, wmax = 1.0 / (cell * n), 0.25 / cell
    rng = np.random.default_rng(11)
    nv = 1000
    u = rng.uniform(-0.7 * (n / 2) * du, 0.7 * (n / 2) * du, nv)
    v = rng.uniform(-0.7 * (n / 2) * du, 0.7 * (n / 2) * du, nv)
    w = rng.uniform(-wmax, wmax, nv)
    uvw = np.stack([u, v, w], 1)
    l0, m0 = px * cell, py * cell
    n0 = np.sqrt(1 - l0 * l0 - m0 * m0)
    vis = np.exp(-2j * np.pi * (u * l0 + v * m0 + w * (n0 - 1))).astype(np.complex64)
    img, _ = _image(orc, p, K, uvw, vis)
    r, c = n - 1 - (n // 2 + py), n // 2 + px
    assert np.unravel_index(np.argmax(img), img.shape) == (r, c)
    assert lo < img[r, c] < 1.02
    v64 = vis.astype(complex)
    with_w = (v64 * np.exp(2j * np.pi * (u * l0 + v * m0 + w * (n0 - 1)))).real.mean()
    without_w = (v64 * np.exp(2j * np.pi * (u * l0 + v * m0))).real.mean()
    assert with_w == pytest.approx(1.0, abs=1e-6)
    if (px, py) == (-90, 70):
        assert without_w < 0.2


def test_dirty_image_sidelobes_match_direct_dft(orc):
    """Whole inner quarter against the direct DFT at w = 0: the residual is
    the oversampled kernel's nearest-sample position error (<= 1/(2s) cell)."""
    n, cell, s = 256, 60 * ARC, 8
    p = orc.make_params(n, n, cell, cell, "wproj", 4, s)
    K = orc.Kernels(p)
    du = 1.0 / (cell * n)
    uvw = _random_vis(500, 0.7 * (n / 2) * du, 0.0, 13)
    src = [(0.6, 20 * cell, -14 * cell), (1.0, -25 * cell, 18 * cell), (0.4, 3 * cell, 30 * cell)]
    vis = np.zeros(len(uvw), complex)
    for S, l0, m0 in src:
        vis += S * np.exp(-2j * np.pi * (uvw[:, 0] * l0 + uvw[:, 1] * m0))
    img, _ = _image(orc, p, K, uvw, vis.astype(np.complex64))
    jx = np.arange(n) - n // 2
    jy = (n - 1 - np.arange(n)) - n // 2
    L, M = np.meshgrid(jx * cell, jy * cell)
    v64 = vis.astype(np.complex64).astype(complex)
    ref = np.zeros((n, n))
    for k in range(len(uvw)):
        ref += (v64[k] * np.exp(2j * np.pi * (uvw[k, 0] * L + uvw[k, 1] * M))).real
    ref /= len(uvw)
    inner = (slice(3 * n // 8, 5 * n // 8), slice(3 * n // 8, 5 * n // 8))
    err = img[inner] - ref[inner]
    assert np.sqrt(np.mean(err ** 2)) / np.abs(ref[inner]).max() < 5e-3
