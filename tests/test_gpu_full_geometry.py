"""Parity at the full-polarisation cube's geometry: 4096^2 at 1.5 arcsec, 256
w-planes, oversampling 8 (supports 7-33), 64 antennas in an 8 km disk at dec
-45 over +-4 h, sources with 10% linear polarisation, Briggs robust 0, Stokes
IQUV.  The oracle runs on one host thread, so rows are thinned in time (24
steps instead of 1800) and the two band-edge channels (0.9 and 1.7 GHz) are
imaged; kernel tables, bins, grids and images are compared with the FP64
oracle at the full image size, plane count and oversampling."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

IMAGE_RMS_TOL = 1e-4   # relative RMS vs the FP64 oracle
GRID_RMS_TOL = 1e-5


@pytest.fixture(scope="module")
def full(orc, mti, ctx):
    from mtimager_amd import synth
    ob = synth.make_observation(nant=64, radius_m=8000.0, dec_deg=-45.0, ha_hours=(-4.0, 4.0), ntime=24,
                                freqs_hz=np.array([0.9e9, 1.7e9]), npol=4, fov_rad=4096 * 1.5 * synth.ARCSEC,
                                pol_frac=0.1, seed=2)
    p = mti.make_params(4096, 4096, 1.5, 1.5, "wproj", 256, 8)
    print("building the oracle's 256 kernel planes (about a minute)", flush=True)
    OK = orc.Kernels(H.oracle_params(orc, p))
    K = ctx.build_kernels(p)
    dev = H.Dev(ob, torch)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    return dict(ob=ob, p=p, OK=OK, K=K, dev=dev, idx=idx)


def test_full_geometry_kernel_tables(full):
    K, OK = full["K"], full["OK"]
    sup, base = K.supports()
    np.testing.assert_array_equal(sup, OK.support)
    np.testing.assert_array_equal(base, OK.base)
    assert (sup.min(), sup.max()) == (7, 33)
    tab = K.table().cpu().numpy()
    assert np.max(np.abs(tab - OK.table)) <= 2e-7 * np.max(np.abs(OK.table))


def test_full_geometry_binning_bit_exact(orc, mti, ctx, full):
    ob, p, K, OK, dev, idx = (full[k] for k in ("ob", "p", "K", "OK", "dev", "idx"))
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    rows = idx.cpu().numpy()
    for chan in (0, 1):
        bins = torch.zeros((len(rows), 8), dtype=torch.int32, device="cuda")
        ctx.preprocess(K, p, v, chan, bins)
        got = bins.cpu().numpy()
        ref = orc.preprocess(OK, ob.uvw[rows], ob.inv_lambda, ob.flags[rows], chan)
        exp = np.stack([ref[f] for f in ("to_grid", "plane", "support", "corner_x", "corner_y",
                                          "off_x", "off_y", "conj")], axis=1)
        np.testing.assert_array_equal(got, exp)
        # at 1.7 GHz the longest baselines fall off the grid (edge exclusion)
        print(f"chan {chan}: {got[:, 0].sum()} of {len(rows)} records gridded, planes up to {got[:, 1].max()}")


@pytest.fixture(scope="module")
def oracle_chans(orc, full):
    """The oracle's grids, normalisers and IQUV images of both channels."""
    ob, p, OK, idx = (full[k] for k in ("ob", "p", "OK", "idx"))
    rows = idx.cpu().numpy()
    op = H.oracle_params(orc, p)
    return [H.oracle_image(orc, OK, op, ob, chan, rows, weighting="briggs", robust=0.0) for chan in (0, 1)]


def _check_iquv(img, refs, label):
    for chan in (0, 1):
        r = refs[chan]
        ref_i = np.sqrt(np.mean(r["image"][0] ** 2))
        for q, name in enumerate("IQUV"):
            d = img[chan][q].astype(np.float64) - r["image"][q]
            e = H.image_errors(img[chan][q], r["image"][q])
            e["rel_rms_vs_I"] = float(np.sqrt(np.mean(d ** 2)) / ref_i)
            print(f"{label} chan {chan} Stokes {name}: {e}")
            assert e["rel_rms_vs_I"] < IMAGE_RMS_TOL


def test_full_geometry_briggs_iquv_images(orc, mti, ctx, full, oracle_chans):
    ob, p, K, OK, dev, idx = (full[k] for k in ("ob", "p", "K", "OK", "dev", "idx"))
    g = H.gpu_image(mti, ctx, K, p, dev, torch, [0, 1], idx, weighting="briggs", robust=0.0,
                    layout="pvu", precision="c128")
    ng = 0
    for chan in (0, 1):
        r = oracle_chans[chan]
        ng += r["gridded"]
        np.testing.assert_allclose(g["norm"][chan], r["norm"], rtol=1e-6)
        assert H.rel_rms(g["grid"][chan], r["grid"]) < GRID_RMS_TOL
    _check_iquv(g["image"], oracle_chans, "full geometry")
    assert g["stats"].gridded_records == ng


@pytest.mark.parametrize("layout", ["pvu", "vup"])
def test_full_geometry_bench_form_iquv_images(mti, ctx, full, oracle_chans, layout):
    """The form the benchmark times, against the oracle: complex64 grids written
    by the gridder (the reference's Grid element type, comp/WImager.cpp:231-232;
    FP64 sums rounded once per work unit, float atomics where a tile is split
    over units), sparse grids (tile mask) and complex64 planes between the
    transform passes."""
    p, K, dev, idx = (full[k] for k in ("p", "K", "dev", "idx"))
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    shape = (2, 4, p.ny, p.nx, 2) if layout == "pvu" else (2, p.ny, p.nx, 4, 2)
    grids = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
    norm = torch.empty((2, 4), dtype=torch.float64, device="cuda")
    tmask = torch.empty(mti.tile_mask_shape(p, 2), dtype=torch.uint8, device="cuda")
    img = torch.empty((2, 4, p.ny, p.nx), dtype=torch.float32, device="cuda")
    ctx.set_transform_storage("c64")
    try:
        ctx.grid(K, p, v, grids, norm, chan0=0, nchan=2, weighting="briggs", robust=0.0, layout=layout,
                 precision="c64", tile_mask=tmask)
        ctx.finalize(p, 4, grids, norm, img, nchan=2, layout=layout, tile_mask=tmask)
        torch.cuda.synchronize()
    finally:
        ctx.set_transform_storage("c128")
    out = img.cpu().numpy()
    assert np.isfinite(out).all()
    _check_iquv(out, oracle_chans, f"bench form ({layout})")
