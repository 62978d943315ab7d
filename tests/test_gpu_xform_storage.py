"""complex64 transform storage (mti_set_transform_storage): the packed planes
that the finalize's transform passes hand to each other stored as complex64,
the arithmetic FP64.  Images against the FP64 oracle (relative RMS 1e-4 is the
imaging path's bound; two complex64 stores cost ~1e-7) and against the
complex128 form: both column strategies (whole columns at ny <= 256, the
four-step split above), rectangular images, P = 1/2/4, complex64 and
complex128 grids, sparse grids and batched jobs."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

STORE_TOL = 1e-6     # complex64 storage vs the oracle and vs complex128 storage
IMAGE_RMS_TOL = 1e-4  # the imaging path's bound against the FP64 oracle


@pytest.fixture(scope="module")
def ctx64(mti):
    c = mti.Context(0)
    c.set_transform_storage("c64")
    yield c
    c.close()


def test_setting_round_trips_and_refuses_bad_values(mti, ctx, ctx64):
    assert ctx.get_transform_storage() == "c128" and ctx64.get_transform_storage() == "c64"
    assert mti.lib().mti_set_transform_storage(ctx64.h, 7) != 0
    assert ctx64.get_transform_storage() == "c64"


@pytest.mark.parametrize("nx,ny,npol,poltype,prec", [
    (256, 128, 1, "linear", "c128"), (512, 256, 4, "linear", "c128"), (1024, 1024, 2, "circular", "c64"),
    (2048, 512, 1, "linear", "c128"), (4096, 4096, 4, "circular", "c128"), (256, 256, 4, "circular", "c64"),
    (4096, 2048, 1, "linear", "c64")])
def test_finalize_matches_oracle(orc, mti, ctx, ctx64, nx, ny, npol, poltype, prec):
    p = mti.make_params(nx, ny, 3.0, 2.0, "wproj", 8, 8)
    rng = np.random.default_rng(nx + ny + npol)
    g = rng.standard_normal((1, p.ny, p.nx, npol, 2))
    norm = rng.uniform(1.0, 3.0, (1, npol))
    d_g = torch.tensor(g, dtype=torch.float64 if prec == "c128" else torch.float32, device="cuda")
    d_n = torch.tensor(norm, device="cuda")
    imgs = []
    for c in (ctx, ctx64):
        img = torch.empty((1, npol, p.ny, p.nx), dtype=torch.float32, device="cuda")
        c.finalize(p, npol, d_g, d_n, img, poltype=poltype, precision=prec)
        imgs.append(img.cpu().numpy()[0])
    ref = orc.finalize(H.oracle_params(orc, p), d_g.cpu().numpy()[0].astype(np.float64), norm[0], poltype)
    for q in range(npol):
        e = H.rel_rms(imgs[1][q], ref[q])
        print(f"{nx}x{ny} P={npol} {prec} grids, plane {q}: c64 storage vs oracle {e:.2e}, "
              f"c128 storage vs oracle {H.rel_rms(imgs[0][q], ref[q]):.2e}")
        assert e < STORE_TOL
        assert H.rel_rms(imgs[1][q], imgs[0][q]) < STORE_TOL


def _obs(nx, cell, nch, npol, seed, nant=16, ntime=40):
    from mtimager_amd import synth
    return synth.make_observation(nant=nant, radius_m=2000.0, ntime=ntime, freqs_hz=np.linspace(1.3e9, 1.45e9, nch),
                                  npol=npol, fov_rad=nx * cell * synth.ARCSEC, seed=seed)


def test_dirty_image_with_sparse_grids_matches_oracle(orc, mti, ctx64):
    """Gridded full-polarisation channels (Briggs weighting, sparse grids)
    finalized through complex64 planes, every Stokes image against the
    oracle."""
    P, nch, nx = 4, 2, 512
    ob = _obs(nx, 2.0, nch, P, 81)
    dev = H.Dev(ob, torch)
    p = mti.make_params(nx, nx, 2.0, 2.0, "wproj", 16, 4)
    K = ctx64.build_kernels(p)
    idx = H.gpu_sort(ctx64, dev, ob.nant, torch)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    g = torch.empty((nch, nx, nx, P, 2), dtype=torch.float64, device="cuda")
    n = torch.empty((nch, P), dtype=torch.float64, device="cuda")
    im = torch.empty((nch, P, nx, nx), dtype=torch.float32, device="cuda")
    mask = torch.empty(mti.tile_mask_shape(p, nch), dtype=torch.uint8, device="cuda")
    ctx64.grid(K, p, v, g, n, nchan=nch, weighting="briggs", robust=0.0, precision="c128", tile_mask=mask)
    ctx64.finalize(p, P, g, n, im, nchan=nch, precision="c128", tile_mask=mask)
    torch.cuda.synchronize()
    img = im.cpu().numpy()
    OK = orc.Kernels(H.oracle_params(orc, p))
    rows = idx.cpu().numpy()
    for c in range(nch):
        r = H.oracle_image(orc, OK, H.oracle_params(orc, p), ob, c, rows, weighting="briggs", robust=0.0)
        e = H.image_errors(img[c][0], r["image"][0])
        print(f"chan {c} stokes I: {e}")
        assert e["rel_rms"] < IMAGE_RMS_TOL
        ref_i = np.sqrt(np.mean(r["image"][0] ** 2))
        for q in range(1, 4):
            d = img[c][q].astype(np.float64) - r["image"][q]
            assert np.sqrt(np.mean(d ** 2)) / ref_i < IMAGE_RMS_TOL


def test_batched_jobs_match_complex128_storage(mti, ctx, ctx64):
    """Stokes-I jobs of three sizes (whole-column and split transforms),
    sparse and complete grids, batched: complex64 storage against complex128
    storage, image by image."""
    specs = [(256, 8.0, 8, 4, 3, 91, True), (512, 4.0, 16, 8, 2, 92, False), (1024, 2.0, 16, 8, 1, 93, True),
             (256, 8.0, 8, 4, 2, 94, False)]
    outs = {}
    for name, c in (("c128", ctx), ("c64", ctx64)):
        jobs, ims = [], []
        for nx, cell, W, s, nch, seed, sparse in specs:
            ob = _obs(nx, cell, nch, 1, seed, nant=12, ntime=30)
            d = H.Dev(ob, torch)
            p = mti.make_params(nx, nx, cell, cell, "wproj", W, s)
            K = c.build_kernels(p)
            g = torch.empty((nch, nx, nx, 1, 2), dtype=torch.float64, device="cuda")
            n = torch.empty((nch, 1), dtype=torch.float64, device="cuda")
            im = torch.zeros((nch, 1, nx, nx), dtype=torch.float32, device="cuda")
            m = torch.empty(mti.tile_mask_shape(p, nch), dtype=torch.uint8, device="cuda") if sparse else None
            v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
            jobs.append(c.make_job(K, p, v, g, n, im, 0, nch, ant1=d.ant1, ant2=d.ant2, nant=ob.nant, tile_mask=m))
            ims.append(im)
        c.grid_jobs(jobs, weighting="uniform")
        c.finalize_jobs(jobs)
        torch.cuda.synchronize()
        outs[name] = [x.cpu().numpy() for x in ims]
    for a, b in zip(outs["c64"], outs["c128"]):
        assert H.rel_rms(a, b) < STORE_TOL


@pytest.mark.parametrize("psf", [False, True])
def test_small_array_dirty_image_and_psf(orc, mti, ctx64, psf):
    """The small array (512^2, 32 planes, s=4, natural weighting), complex128 grids
    finalized through complex64 planes: every Stokes image and the PSF against
    the FP64 oracle.  The taper correction amplifies edge pixels by up to
    1/psi(1)^2, so this is where storage rounding shows most."""
    from mtimager_amd import synth
    c1 = synth.small_array_obs(seed=1)
    p = mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx64, dev, c1.nant, torch)
    K = ctx64.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    g = H.gpu_image(mti, ctx64, K, p, dev, torch, [0, 1], idx, psf=psf, precision="c128", layout="pvu")
    for ci in range(2):
        r = H.oracle_image(orc, OK, H.oracle_params(orc, p), c1, ci, idx.cpu().numpy(), psf=psf)
        e = H.image_errors(g["image"][ci][0], r["image"][0])
        print(f"small array chan {ci} psf={psf}, complex64 storage, stokes I: {e}")
        assert e["rel_rms"] < IMAGE_RMS_TOL
        ref_i = np.sqrt(np.mean(r["image"][0] ** 2))
        for q in range(1, 4):
            d = g["image"][ci][q].astype(np.float64) - r["image"][q]
            assert np.sqrt(np.mean(d ** 2)) / ref_i < IMAGE_RMS_TOL
