"""Pins orc_weights with hand-computed textbook values (Briggs 1995) and the
R -> +inf (natural) / R -> -inf (uniform) limits.  Weighting is not in the
reference; its definitions are the textbook ones."""
import numpy as np
import pytest

import kat_binning as KAT


def _setup(orc):
    p = orc.make_params(KAT.NX, KAT.NX, KAT.CELL, KAT.CELL, "wproj", KAT.WPLANES, 4)
    K = orc.Kernels(p)
    # cells (loc): r0, r1 -> (32,32); r2, r3 -> (33,32); r4 -> (34,32); r5 off-grid
    uvw = np.array([[0, 0, 0], [0, 0, 0], [128, 0, 0], [128, 0, 0], [256, 0, 0],
                    [-3840, 0, 0]], np.float64)
    flags = np.array([0, 0, 0, 1, 0, 0], np.uint8)[:, None, None]
    weight = np.array([1, 2, 1, 5, 4, 7], np.float32)[:, None]
    return K, uvw, flags, weight


def test_uniform_by_hand(orc):
    K, uvw, flags, weight = _setup(orc)
    w = orc.weights(K, uvw, np.array([1.0]), flags, weight, 0, "uniform")[:, 0]
    # densities: 3, 1 (r3 flagged), 4
    np.testing.assert_allclose(w[[0, 1, 2, 4]], [1 / 3, 2 / 3, 1.0, 1.0], rtol=1e-15)
    assert w[5] == 7  # not gridded: weight passed through


def test_briggs_by_hand(orc):
    K, uvw, flags, weight = _setup(orc)
    w = orc.weights(K, uvw, np.array([1.0]), flags, weight, 0, "briggs", robust=0.0)[:, 0]
    f2 = 25.0 / ((9 + 1 + 16) / 8.0)
    exp = [1 / (1 + 3 * f2), 2 / (1 + 3 * f2), 1 / (1 + f2), 4 / (1 + 4 * f2)]
    np.testing.assert_allclose(w[[0, 1, 2, 4]], exp, rtol=1e-14)
    w1 = orc.weights(K, uvw, np.array([1.0]), flags, weight, 0, "briggs", robust=1.0)[:, 0]
    f2 = 0.25 / (26 / 8.0)
    np.testing.assert_allclose(w1[0], 1 / (1 + 3 * f2), rtol=1e-14)


def test_briggs_limits(orc):
    from mtimager_amd import synth
    ob = synth.make_observation(nant=8, ntime=20, freqs_hz=[1.4e9], npol=1, predict=False,
                                flag_frac=0.1, seed=4)
    p = orc.make_params(128, 128, 30 * synth.ARCSEC, 30 * synth.ARCSEC, "wproj", 8, 4)
    K = orc.Kernels(p)
    ob.weight = np.random.default_rng(1).uniform(0.5, 2.0, ob.weight.shape).astype(np.float32)
    args = (K, ob.uvw, ob.inv_lambda, ob.flags, ob.weight, 0)
    b = orc.preprocess(K, ob.uvw, ob.inv_lambda, ob.flags, 0)
    g = b["to_grid"] == 1
    nat = orc.weights(*args, "natural")[g, 0]
    uni = orc.weights(*args, "uniform")[g, 0]
    hi = orc.weights(*args, "briggs", robust=8.0)[g, 0]
    lo = orc.weights(*args, "briggs", robust=-8.0)[g, 0]
    np.testing.assert_allclose(hi, nat, rtol=1e-9)
    np.testing.assert_allclose(lo / lo.sum(), uni / uni.sum(), rtol=1e-9)
    assert not np.allclose(uni / uni.sum(), nat / nat.sum(), rtol=1e-2)


def test_mfs_density_over_channels(orc):
    """One density over several channels (MFS): a single channel reproduces
    orc_weights; two channels with the same 1/lambda double every density, so
    uniform weights halve while Briggs weights do not change (f^2 halves as
    sum W^2 / sum w doubles)."""
    K, uvw, flags, weight = _setup(orc)
    fl2 = np.repeat(flags, 2, axis=1)
    for wt in ("uniform", "briggs"):
        one = orc.weights(K, uvw, np.array([1.0]), flags, weight, 0, wt)
        m1 = orc.weights_mfs(K, uvw, np.array([1.0]), flags, weight, 0, 1, wt)
        np.testing.assert_array_equal(m1[0], one)
        m2 = orc.weights_mfs(K, uvw, np.array([1.0, 1.0]), fl2, weight, 0, 2, wt)
        np.testing.assert_array_equal(m2[0], m2[1])
        g = [0, 1, 2, 4]  # gridded, unflagged records
        if wt == "uniform":
            np.testing.assert_allclose(m2[0][g], one[g] / 2, rtol=1e-15)
        else:
            np.testing.assert_allclose(m2[0][g], one[g], rtol=1e-14)
