"""The chunked loader on the GPU: images streamed from a container file (or
wrapped host arrays) through pinned staging buffers, in channel blocks and
`records`-row chunks (vis/MSVisibilityDataSetAsync.cpp:44-347,
vis/StorageManager.h:111-150), equal the images of the same data imaged
resident in HBM in one call (FP64 sums; only their order differs: float
atomics where a tile is split over work units add in the order units finish,
so complex128 images agree to ~1e-10 of their peak, complex64 ones to their
rounding)."""
FP64_ORDER = 1e-8  # image max difference / peak from FP64 summation order
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SITE = (5109360.13, 2006852.59, -3238948.13)


@pytest.fixture(scope="module")
def ob():
    from mtimager_amd import synth
    return synth.small_array_obs(seed=5, ntime=90, freqs_hz=np.linspace(1.40e9, 1.5e9, 7))


@pytest.fixture(scope="module")
def params(mti):
    return mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)


def _resident(mti, ctx, K, p, ob, weighting, precision, row_doppler=None):
    dev = H.Dev(ob, torch)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    nb = len(ob.freq)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx, row_doppler=row_doppler)
    gdt = torch.float32 if precision == "c64" else torch.float64
    grids = torch.empty((nb, p.ny, p.nx, 4, 2), dtype=gdt, device="cuda")
    norm = torch.empty((nb, 4), dtype=torch.float64, device="cuda")
    ctx.grid(K, p, v, grids, norm, chan0=0, nchan=nb, weighting=weighting, precision=precision)
    img = torch.empty((nb, 4, p.ny, p.nx), dtype=torch.float32, device="cuda")
    ctx.finalize(p, 4, grids, norm, img, nchan=nb, precision=precision)
    torch.cuda.synchronize()
    return img.cpu().numpy(), norm.cpu().numpy()


def _streamed(mti, ctx, K, p, ds, precision, **kw):
    from mtimager_amd import ingest
    nb = ds.nchan
    gdt = torch.float32 if precision == "c64" else torch.float64
    grids = torch.empty((nb, p.ny, p.nx, 4, 2), dtype=gdt, device="cuda")
    norm = torch.empty((nb, 4), dtype=torch.float64, device="cuda")
    ld = ingest.Loader(ctx, ds, precision=precision, **kw)
    st = ld.grid(K, p, grids, norm)
    ld.close()
    img = torch.empty((nb, 4, p.ny, p.nx), dtype=torch.float32, device="cuda")
    ctx.finalize(p, 4, grids, norm, img, nchan=nb, precision=precision)
    torch.cuda.synchronize()
    return img.cpu().numpy(), norm.cpu().numpy(), st


def _write(tmp_path, ob):
    from mtimager_amd import ingest
    path = tmp_path / "obs.mtv"
    ingest.write_dataset(path, ob.time, ob.ant1, ob.ant2, ob.uvw, ob.weight, ob.freq, ob.vis, ob.flags,
                         ob.nant, ra=1.3, dec=np.deg2rad(-30.0), obs_itrf=SITE)
    return ingest.Dataset.open(path)


def _max_rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_row_chunks_equal_resident(tmp_path, mti, ctx, ob, params, precision):
    K = ctx.build_kernels(params)
    ref, nref = _resident(mti, ctx, K, params, ob, "natural", precision)
    ds = _write(tmp_path, ob)
    print(f"[ingest] {ob.nrows} rows x {ds.nchan} ch, chunks of {ob.nrows // 4 + 1} rows, {precision}", flush=True)
    img, norm, st = _streamed(mti, ctx, K, params, ds, precision, channel_block=3,
                              chunk_rows=ob.nrows // 4 + 1, reader_threads=4)
    assert st["chunks"] == 3 * 4
    np.testing.assert_allclose(norm, nref, rtol=1e-12)
    if precision == "c128":
        assert _max_rel(img, ref) < FP64_ORDER
    else:
        # complex64 cells are rounded once per row chunk (read, add, write)
        # instead of once, and the taper correction amplifies cell roundings
        # up to 6e4-fold in the field corners: within the 1e-4 image tolerance
        rr = H.rel_rms(img, ref)
        print(f"[ingest] c64 streamed vs resident: rel RMS {rr:.2e}, max {_max_rel(img, ref):.2e}")
        assert rr < 1e-4 and _max_rel(img, ref) < 1e-3
    ds.close()


def test_briggs_channel_blocks_equal_resident(tmp_path, mti, ctx, ob, params):
    K = ctx.build_kernels(params)
    ref, _ = _resident(mti, ctx, K, params, ob, "briggs", "c128")
    ds = _write(tmp_path, ob)
    img, _, st = _streamed(mti, ctx, K, params, ds, "c128", channel_block=2, chunk_rows=0, weighting="briggs")
    assert st["chunks"] == 4
    assert _max_rel(img, ref) < FP64_ORDER
    from mtimager_amd import MTIError, ingest
    with pytest.raises(MTIError, match="every row"):
        ingest.Loader(ctx, ds, chunk_rows=100, weighting="briggs")
    ds.close()


def test_lsrk_through_the_loader(tmp_path, mti, ctx, ob, params):
    K = ctx.build_kernels(params)
    t = torch.tensor(ob.time, dtype=torch.float64, device="cuda")
    dop = torch.empty_like(t)
    ctx.lsrk_factors(mti.lsrk_frame(SITE, 1.3, np.deg2rad(-30.0), 69.184), t, dop)
    ref, _ = _resident(mti, ctx, K, params, ob, "natural", "c128", row_doppler=dop)
    plain, _ = _resident(mti, ctx, K, params, ob, "natural", "c128")
    ds = _write(tmp_path, ob)
    img, _, _ = _streamed(mti, ctx, K, params, ds, "c128", channel_block=4, chunk_rows=ob.nrows // 3 + 1,
                          lsrk=True, tt_minus_utc=69.184)
    assert _max_rel(img, ref) < FP64_ORDER
    assert _max_rel(plain, ref) > 1e-6  # the factors change the image
    ds.close()


def test_pinned_wrap_equals_resident(mti, ctx, ob, params):
    from mtimager_amd import ingest
    K = ctx.build_kernels(params)
    ref, _ = _resident(mti, ctx, K, params, ob, "natural", "c128")
    pin = lambda a: torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
    cols = [pin(ob.time), pin(ob.ant1), pin(ob.ant2), pin(ob.uvw), pin(ob.weight), pin(ob.freq),
            pin(ob.vis.transpose(1, 0, 2)), pin(ob.flags.transpose(1, 0, 2))]
    ds = ingest.Dataset.wrap(*cols, ob.nant, pinned=True)
    img, _, st = _streamed(mti, ctx, K, params, ds, "c128", channel_block=3, chunk_rows=0)
    assert st["chunks"] == 3 and st["reader_seconds"] == 0.0
    assert _max_rel(img, ref) < FP64_ORDER
    ds.close()
