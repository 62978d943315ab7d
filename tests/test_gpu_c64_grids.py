"""complex64 grids written by the gridder itself (each cell's FP64 sum rounded
once as it is added; float atomics where a tile is split over work units):
against complex128 grids of the same call and the FP64 oracle, complete and
sparse, single calls and batched jobs mixing both grid types."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

IMAGE_RMS_TOL = 1e-4   # against the FP64 oracle
GRID_ROUND_TOL = 1e-6  # complex64 grid against complex128 grid (one rounding, 2^-24)
IMAGE_VS_C128 = 5e-5   # images of complex64 grids against complex128 grids


def _obs(nx, cell, nch, npol, seed, nant=16, ntime=40):
    from mtimager_amd import synth
    return synth.make_observation(nant=nant, radius_m=2000.0, ntime=ntime, freqs_hz=np.linspace(1.3e9, 1.45e9, nch),
                                  npol=npol, fov_rad=nx * cell * synth.ARCSEC, seed=seed)


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("layout", ["vup", "pvu"])
def test_c64_grids_match_c128_and_oracle(orc, mti, ctx, sparse, layout):
    P, nch, nx = 4, 2, 512
    ob = _obs(nx, 2.0, nch, P, 101)
    dev = H.Dev(ob, torch)
    p = mti.make_params(nx, nx, 2.0, 2.0, "wproj", 16, 4)
    K = ctx.build_kernels(p)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    shape = (nch, nx, nx, P) if layout == "vup" else (nch, P, nx, nx)
    out = {}
    for prec, dt in (("c128", torch.complex128), ("c64", torch.complex64)):
        g = torch.full(shape, complex(np.nan, np.nan), dtype=dt, device="cuda")  # unwritten tiles show
        n = torch.empty((nch, P), dtype=torch.float64, device="cuda")
        im = torch.empty((nch, P, nx, nx), dtype=torch.float32, device="cuda")
        m = torch.empty(mti.tile_mask_shape(p, nch), dtype=torch.uint8, device="cuda") if sparse else None
        ctx.grid(K, p, v, g, n, nchan=nch, weighting="briggs", layout=layout, precision=prec, tile_mask=m)
        ctx.finalize(p, P, g, n, im, nchan=nch, layout=layout, precision=prec, tile_mask=m)
        torch.cuda.synchronize()
        out[prec] = (g.cpu().numpy(), n.cpu().numpy(), im.cpu().numpy(), None if m is None else m.cpu().numpy())
    g128, n128, i128, m128 = out["c128"]
    g64, n64, i64, m64 = out["c64"]
    np.testing.assert_allclose(n64, n128, rtol=1e-12)
    if sparse:
        np.testing.assert_array_equal(m64, m128)
        cells = np.kron(m128, np.ones((32, 32), np.uint8)).astype(bool)
        sel = np.broadcast_to(cells[..., None] if layout == "vup" else cells[:, None], shape)
        g64, g128 = g64[sel], g128[sel]
    assert np.isfinite(g64).all()
    assert H.rel_rms(g64, g128) < GRID_ROUND_TOL
    OK = orc.Kernels(H.oracle_params(orc, p))
    rows = idx.cpu().numpy()
    for c in range(nch):
        r = H.oracle_image(orc, OK, H.oracle_params(orc, p), ob, c, rows, weighting="briggs", robust=0.0)
        ref_i = np.sqrt(np.mean(r["image"][0] ** 2))
        e = H.image_errors(i64[c][0], r["image"][0])
        d128 = np.sqrt(np.mean((i64[c].astype(np.float64) - i128[c]) ** 2, axis=(1, 2))) / ref_i
        print(f"{layout} sparse={sparse} chan {c}: c64 grids vs oracle {e}, vs c128 grids {d128}")
        assert e["rel_rms"] < IMAGE_RMS_TOL
        assert d128.max() < IMAGE_VS_C128
        for q in range(1, 4):
            d = i64[c][q].astype(np.float64) - r["image"][q]
            assert np.sqrt(np.mean(d ** 2)) / ref_i < IMAGE_RMS_TOL


def test_batched_jobs_mixing_grid_types_match_single_calls(mti, ctx):
    """One mti_grid_jobs / mti_finalize_jobs call over Stokes-I jobs with
    complex64 and complex128 grids of two sizes (sparse and complete): each
    job's image equals that of mti_grid + mti_finalize on the job alone (to
    the order of atomic sums in tiles split over work units: float atomics for
    complex64 grids)."""
    specs = [(256, 8.0, 8, 4, 3, 111, "c64", True), (512, 4.0, 16, 8, 2, 112, "c128", False),
             (512, 4.0, 16, 8, 2, 113, "c64", False), (256, 8.0, 8, 4, 2, 114, "c128", True),
             (1024, 2.0, 16, 8, 1, 115, "c64", True)]
    jobs, batched, single = [], [], []
    for nx, cell, W, s, nch, seed, prec, sparse in specs:
        ob = _obs(nx, cell, nch, 1, seed, nant=12, ntime=30)
        d = H.Dev(ob, torch)
        p = mti.make_params(nx, nx, cell, cell, "wproj", W, s)
        K = ctx.build_kernels(p)
        dt = torch.complex64 if prec == "c64" else torch.complex128
        mk = (lambda: torch.empty(mti.tile_mask_shape(p, nch), dtype=torch.uint8, device="cuda")) if sparse \
            else (lambda: None)
        g = torch.empty((nch, nx, nx, 1), dtype=dt, device="cuda")
        n = torch.empty((nch, 1), dtype=torch.float64, device="cuda")
        im = torch.zeros((nch, 1, nx, nx), dtype=torch.float32, device="cuda")
        v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
        jobs.append(ctx.make_job(K, p, v, g, n, im, 0, nch, ant1=d.ant1, ant2=d.ant2, nant=ob.nant, tile_mask=mk()))
        batched.append(im)
        # the same job alone
        idx = H.gpu_sort(ctx, d, ob.nant, torch)
        v1 = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight, idx)
        g1, n1 = torch.empty_like(g), torch.empty_like(n)
        i1 = torch.zeros_like(im)
        m1 = mk()
        ctx.grid(K, p, v1, g1, n1, nchan=nch, weighting="uniform", precision=prec, tile_mask=m1)
        ctx.finalize(p, 1, g1, n1, i1, nchan=nch, precision=prec, tile_mask=m1)
        single.append(i1)
    ctx.grid_jobs(jobs, weighting="uniform")
    ctx.finalize_jobs(jobs)
    torch.cuda.synchronize()
    for a, b, sp in zip(batched, single, specs):
        e = H.rel_rms(a.cpu().numpy(), b.cpu().numpy())
        print(f"{sp[0]}^2 {sp[6]} sparse={sp[7]}: batched vs single {e:.2e}")
        assert e < (5e-5 if sp[6] == "c64" else 1e-6)
