"""Pins the prediction oracle (oracle/mti_oracle_degrid.c).  The reference has
no degridder, so nothing of its own fixes these results; they are pinned by
  * the direct Fourier sum: a point source of the model predicts
    S exp(-2 pi i (u l + v m + w (n - 1))) (the convention the imaging's
    point-source tests pin), with the error of nearest-sample oversampling,
    which falls as the oversampling rises and grows with the offset;
  * adjointness: the unnormalised degridder is the adjoint of orc_grid's
    scatter in the real inner product (w < 0 records, gridded conjugated,
    make the map antilinear), on one and two correlations;
  * the Stokes/correlation relations on a polarised point source: XY and YX
    of a linear feed are U +- iV and RL, LR of a circular one Q -/+ iU.
"""
import numpy as np
import pytest

ARC = np.pi / 180 / 3600


def _obs(npol=1, nant=16, ntime=30, seed=3):
    from mtimager_amd import synth
    return synth.make_observation(nant=nant, radius_m=2000.0, ntime=ntime, freqs_hz=[1.4e9], npol=npol, nsrc=1,
                                  fov_rad=512 * 2 * ARC, noise=0, flag_frac=0, seed=seed, predict=False)


def _dft(ob, l0, m0, amp=1.0):
    n = np.sqrt(1 - l0 ** 2 - m0 ** 2)
    f = ob.inv_lambda[0]
    return amp * np.exp(-2j * np.pi * f * (ob.uvw[:, 0] * l0 + ob.uvw[:, 1] * m0 + ob.uvw[:, 2] * (n - 1)))


def _point(N, off, npol=1, stokes=(1.0, 0, 0, 0)):
    img = np.zeros((npol, N, N))
    r0, jx0 = N // 2 - 1 - off, N // 2 + off  # row r holds transform row N - 1 - r
    for q in range(npol):
        img[q, r0, jx0] = stokes[q]
    return img, (off * 2 * ARC, off * 2 * ARC)


@pytest.mark.parametrize("s,off,tol", [(8, 3, 4e-3), (16, 3, 2e-3), (8, 10, 1.2e-2)])
def test_point_source_prediction_matches_direct_sum(orc, s, off, tol):
    N = 512
    p = orc.make_params(N, N, 2 * ARC, 2 * ARC, "wproj", 16, s)
    K = orc.Kernels(p)
    ob = _obs()
    img, (l0, m0) = _point(N, off)
    V, n = orc.degrid(K, ob.uvw, ob.inv_lambda[0], orc.model_grid(p, img))
    ex = _dft(ob, l0, m0)
    assert n == len(ob.uvw) and np.any(ob.uvw[:, 2] < 0)
    d = V[:, 0] - ex
    rms = np.sqrt(np.mean(np.abs(d) ** 2))
    # the mean ratio pins amplitude and phase conventions (sign, m flip,
    # normalisation); the spread is the oversampling's position error
    assert abs(np.mean(V[:, 0] / ex) - 1) < 1e-3
    assert rms < tol, rms


def test_error_falls_with_oversampling(orc):
    N, off = 512, 10
    ob = _obs()
    img, (l0, m0) = _point(N, off)
    ex = _dft(ob, l0, m0)
    errs = []
    for s in (4, 8, 16):
        p = orc.make_params(N, N, 2 * ARC, 2 * ARC, "wproj", 16, s)
        V, _ = orc.degrid(orc.Kernels(p), ob.uvw, ob.inv_lambda[0], orc.model_grid(p, img))
        errs.append(np.sqrt(np.mean(np.abs(V[:, 0] - ex) ** 2)))
    assert errs[0] > 1.7 * errs[1] > 1.7 * 1.7 * errs[2], errs


@pytest.mark.parametrize("npol", [1, 2])
def test_degrid_is_the_adjoint_of_grid(orc, npol):
    N = 128
    p = orc.make_params(N, N, 8 * ARC, 8 * ARC, "wproj", 8, 4)
    K = orc.Kernels(p)
    ob = _obs(npol=npol, nant=10, ntime=12)
    rng = np.random.default_rng(1)
    V = (rng.standard_normal((len(ob.uvw), 1, npol)) + 1j * rng.standard_normal((len(ob.uvw), 1, npol)))
    fl = np.zeros(V.shape, np.uint8)
    w = np.ones((len(ob.uvw), npol), np.float32)
    G, _, ng = orc.grid(K, ob.uvw, ob.inv_lambda, V.astype(np.complex64), fl, w, 0)
    G = G[..., 0] + 1j * G[..., 1]
    M = rng.standard_normal((N, N, npol)) + 1j * rng.standard_normal((N, N, npol))
    D, nd = orc.degrid(K, ob.uvw, ob.inv_lambda[0], M, normalise=False)
    assert ng == nd > 0 and np.any(ob.uvw[:, 2] < 0)
    lhs = np.real(np.sum(G * np.conj(M)))
    rhs = np.real(np.sum(V.astype(np.complex64).astype(np.complex128)[:, 0] * np.conj(D)))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("poltype", ["linear", "circular"])
def test_polarised_point_source(orc, poltype):
    N = 512
    p = orc.make_params(N, N, 2 * ARC, 2 * ARC, "wproj", 16, 16)
    K = orc.Kernels(p)
    ob = _obs(npol=4)
    I, Q, U, Vs = 1.0, 0.3, -0.2, 0.1
    img, (l0, m0) = _point(N, 2, npol=4, stokes=(I, Q, U, Vs))
    V, _ = orc.degrid(K, ob.uvw, ob.inv_lambda[0], orc.model_grid(p, img, poltype), )
    e = _dft(ob, l0, m0)
    if poltype == "linear":
        exp = [I + Q, U + 1j * Vs, U - 1j * Vs, I - Q]
    else:
        exp = [I + Vs, Q - 1j * U, Q + 1j * U, I - Vs]
    for q in range(4):
        assert np.sqrt(np.mean(np.abs(V[:, q] - exp[q] * e) ** 2)) < 3e-3 * abs(exp[q]), q
