"""Pins orc_preprocess with hand-derived known answers and orc_sort_index
against a brute-force stable sort."""
import numpy as np

import kat_binning as KAT


def _bins(orc, s, cases):
    p = orc.make_params(KAT.NX, KAT.NX, KAT.CELL, KAT.CELL, "wproj", KAT.WPLANES, s)
    K = orc.Kernels(p)
    assert list(K.support) == [7] * KAT.WPLANES
    uvw, flags, exp = KAT.arrays(cases)
    b = orc.preprocess(K, uvw, np.array([1.0]), flags, 0)
    got = np.stack([b["to_grid"], b["plane"], b["support"], b["corner_x"], b["corner_y"],
                    b["off_x"], b["off_y"], b["conj"], b["conv_idx"]], axis=1)
    return got, exp


def test_known_answers_s4(orc):
    got, exp = _bins(orc, 4, KAT.CASES_S4)
    np.testing.assert_array_equal(got, exp)


def test_known_answers_s5(orc):
    got, exp = _bins(orc, 5, KAT.CASES_S5)
    np.testing.assert_array_equal(got, exp)


def test_sort_index_brute_force(orc):
    rng = np.random.default_rng(3)
    nant = 9
    a1 = rng.integers(0, nant, 500).astype(np.int32)
    a2 = rng.integers(0, nant, 500).astype(np.int32)
    idx = orc.sort_index(a1, a2, nant)
    keep = [i for i in range(500) if a1[i] != a2[i]]
    ref = sorted(keep, key=lambda i: (nant * nant - a1[i] * nant - a2[i] - 1, i))
    np.testing.assert_array_equal(idx, ref)


def test_sort_index_ms_layout(orc):
    """Time-major rows -> baseline-major with time order kept; autocorrelations
    dropped.  Baseline key N^2 - a1 N - a2 - 1 ascending means the last
    antenna pair comes first."""
    from mtimager_amd import synth
    ob = synth.make_observation(nant=4, ntime=3, freqs_hz=[1e9], predict=False)
    idx = orc.sort_index(ob.ant1, ob.ant2, 4)
    assert len(idx) == 6 * 3
    pairs = [(ob.ant1[i], ob.ant2[i]) for i in idx]
    assert pairs[:3] == [(2, 3)] * 3 and pairs[-3:] == [(0, 1)] * 3
    assert list(idx[:3]) == sorted(idx[:3])
