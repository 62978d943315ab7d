"""GPU parity of the TOPO -> LSRK conversion (vis/FrequencyConverter.cpp:112-141)
against the FP64 oracle: per-record factors, the reference's float 1/lambda
store, binning with per-record factors (bit-exact), and images."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

C = 299792458.0
SITE = (5109360.13, 2006852.59, -3238948.13)  # a southern-hemisphere site, ITRF metres


@pytest.fixture(scope="module")
def c1(orc):
    from mtimager_amd import synth
    return synth.small_array_obs(seed=1, ntime=120)


def _frame(mod, dec_deg=-30.0, ra=1.3):
    return mod.lsrk_frame(SITE, ra, np.deg2rad(dec_deg))


@pytest.mark.parametrize("ra,dec", [(1.3, -30.0), (4.7, 60.0), (0.0, -89.0)])
def test_lsrk_factors_match_oracle(orc, mti, ctx, c1, ra, dec):
    t = torch.tensor(c1.time, dtype=torch.float64, device="cuda")
    out = torch.empty_like(t)
    ctx.lsrk_factors(_frame(mti, dec, ra), t, out)
    got = out.cpu().numpy()
    ref = orc.lsrk_factors(_frame(orc, dec, ra), c1.time)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)
    # a real shift: tens of km/s over c
    assert 1e-6 < np.max(np.abs(got - 1.0)) < 3e-4


def test_lsrk_inv_lambda_reference_store(orc, mti, ctx, c1):
    t = torch.tensor(c1.time, dtype=torch.float64, device="cuda")
    fq = torch.tensor(c1.freq, dtype=torch.float64, device="cuda")
    store = torch.empty((len(c1.freq), len(c1.time)), dtype=torch.float32, device="cuda")
    ctx.lsrk_inv_lambda(_frame(mti), t, fq, store)
    fac = orc.lsrk_factors(_frame(orc), c1.time)
    exp = ((c1.freq[:, None] * fac[None, :]) / C).astype(np.float32)
    got = store.cpu().numpy()
    # FP64 factors agree to ~1e-15, so the float roundings agree but for ties
    ulp = np.abs(got.view(np.int32).astype(np.int64) - exp.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and np.mean(ulp == 0) > 0.999


def test_binning_with_row_doppler_bit_exact(orc, mti, ctx, c1):
    p = mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    t = torch.tensor(c1.time, dtype=torch.float64, device="cuda")
    dop = torch.empty_like(t)
    ctx.lsrk_factors(_frame(mti), t, dop)
    K = ctx.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx, row_doppler=dop)
    rows = idx.cpu().numpy()
    dop_h = dop.cpu().numpy()
    moved = 0
    for chan in (0, 7):
        bins = torch.zeros((len(rows), 8), dtype=torch.int32, device="cuda")
        ctx.preprocess(K, p, v, chan, bins)
        got = bins.cpu().numpy()
        with orc.row_doppler(dop_h[rows]):
            ref = orc.preprocess(OK, c1.uvw[rows], c1.inv_lambda, c1.flags[rows], chan)
        plain = orc.preprocess(OK, c1.uvw[rows], c1.inv_lambda, c1.flags[rows], chan)
        exp = np.stack([ref[f] for f in ("to_grid", "plane", "support", "corner_x", "corner_y",
                                          "off_x", "off_y", "conj")], axis=1)
        np.testing.assert_array_equal(got, exp)
        moved += int(np.sum((ref["off_x"] != plain["off_x"]) | (ref["off_y"] != plain["off_y"])))
    assert moved > 0  # the factors move some records to other sub-kernels


def test_lsrk_image_matches_oracle(orc, mti, ctx, c1):
    p = mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    t = torch.tensor(c1.time, dtype=torch.float64, device="cuda")
    dop = torch.empty_like(t)
    ctx.lsrk_factors(_frame(mti), t, dop)
    K = ctx.build_kernels(p)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx, row_doppler=dop)
    grids = torch.empty((1, 512, 512, 4, 2), dtype=torch.float32, device="cuda")
    norm = torch.empty((1, 4), dtype=torch.float64, device="cuda")
    ctx.grid(K, p, v, grids, norm, chan0=3, nchan=1, weighting="briggs", robust=0.0)
    img = torch.empty((1, 4, 512, 512), dtype=torch.float32, device="cuda")
    ctx.finalize(p, 4, grids, norm, img)
    got = img.cpu().numpy()[0]
    rows = idx.cpu().numpy()
    OK = orc.Kernels(H.oracle_params(orc, p))
    with orc.row_doppler(dop.cpu().numpy()[rows]):
        ref = H.oracle_image(orc, OK, H.oracle_params(orc, p), c1, 3, index=rows,
                             weighting="briggs", robust=0.0)["image"]
    e = H.image_errors(got[0], ref[0])
    print("LSRK Stokes I vs oracle:", e)
    assert e["rel_rms"] < 1e-4
