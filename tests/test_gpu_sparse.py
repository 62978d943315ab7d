"""Sparse grids on the GPU: with a tile mask the gridding call leaves the
tiles no record reaches unwritten and marks them, and the finalize reads
them as zeros.  The mask equals the tiles the records' support boxes reach
(from the binning output), and the images equal those of complete grids."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _obs(nx, cell, nch, npol, seed, nant=20, ntime=60):
    from mtimager_amd import synth
    return synth.make_observation(nant=nant, radius_m=2000.0, ntime=ntime, freqs_hz=np.linspace(1.3e9, 1.45e9, nch),
                                  npol=npol, fov_rad=nx * cell * synth.ARCSEC, seed=seed)


@pytest.mark.parametrize("layout", ["vup", "pvu"])
def test_sparse_grids_image_like_complete_grids(mti, ctx, layout):
    P, nch, nx = 4, 3, 512
    ob = _obs(nx, 2.0, nch, P, 71)
    dev = H.Dev(ob, torch)
    p = mti.make_params(nx, nx, 2.0, 2.0, "wproj", 16, 8)
    K = ctx.build_kernels(p)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    shape = (nch, nx, nx, P) if layout == "vup" else (nch, P, nx, nx)

    def run(mask):
        g = torch.full(shape, complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")  # garbage outside
        n = torch.zeros((nch, P), dtype=torch.float64, device="cuda")
        im = torch.zeros((nch, P, nx, nx), dtype=torch.float32, device="cuda")
        ctx.grid(K, p, v, g, n, nchan=nch, weighting="briggs", layout=layout, precision="c128", tile_mask=mask)
        ctx.finalize(p, P, g, n, im, nchan=nch, layout=layout, precision="c128", tile_mask=mask)
        torch.cuda.synchronize()
        return g, n.cpu().numpy(), im.cpu().numpy()

    mask = torch.full(mti.tile_mask_shape(p, nch), 7, dtype=torch.uint8, device="cuda")
    g1, n1, i1 = run(mask)
    g0 = torch.zeros(shape, dtype=torch.complex128, device="cuda")
    n0 = torch.zeros((nch, P), dtype=torch.float64, device="cuda")
    i0 = torch.zeros((nch, P, nx, nx), dtype=torch.float32, device="cuda")
    ctx.grid(K, p, v, g0, n0, nchan=nch, weighting="briggs", layout=layout, precision="c128")
    ctx.finalize(p, P, g0, n0, i0, nchan=nch, layout=layout, precision="c128")
    torch.cuda.synchronize()
    m = mask.cpu().numpy()
    assert set(np.unique(m)) <= {0, 1} and 0 < m.mean() < 1  # some tiles empty, some not
    np.testing.assert_allclose(n1, n0.cpu().numpy(), rtol=1e-12)
    # FP64 grids agree to summation order; the f32 images to its rounding
    assert H.rel_rms(i1, i0.cpu().numpy()) < 1e-9
    # marked tiles hold the complete grid's values; the complete grid is zero elsewhere
    gm = g1.cpu().numpy()
    g0h = g0.cpu().numpy()
    cells = np.kron(m, np.ones((32, 32), np.uint8)).astype(bool)  # [nch][ny][nx]
    sel = cells[..., None] if layout == "vup" else cells[:, None]
    sel = np.broadcast_to(sel, shape)
    assert H.rel_rms(gm[sel], g0h[sel]) < 1e-12
    assert np.all(g0h[~sel] == 0)


def test_mask_marks_the_tiles_the_support_boxes_reach(mti, ctx):
    P, nx = 1, 256
    ob = _obs(nx, 6.0, 2, P, 72)
    dev = H.Dev(ob, torch)
    p = mti.make_params(nx, nx, 6.0, 6.0, "wproj", 12, 4)
    K = ctx.build_kernels(p)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    g = torch.empty((2, nx, nx, 1), dtype=torch.complex128, device="cuda")
    n = torch.empty((2, 1), dtype=torch.float64, device="cuda")
    mask = torch.empty(mti.tile_mask_shape(p, 2), dtype=torch.uint8, device="cuda")
    ctx.grid(K, p, v, g, n, nchan=2, precision="c128", tile_mask=mask)
    got = mask.cpu().numpy()
    for c in range(2):
        bins = torch.zeros((len(idx), 8), dtype=torch.int32, device="cuda")
        ctx.preprocess(K, p, v, c, bins)
        b = bins.cpu().numpy()
        want = np.zeros(got.shape[1:], np.uint8)
        for to_grid, _, S, cx, cy in b[:, :5]:
            if to_grid and S > 0:
                want[cy // 32:(cy + S - 1) // 32 + 1, cx // 32:(cx + S - 1) // 32 + 1] = 1
        np.testing.assert_array_equal(got[c], want)


def test_sparse_batched_jobs(mti, ctx):
    """Jobs of two sizes and samplings (one correlation: channels share a
    packed plane across jobs of one size and sampling), masks per job."""
    specs = [(256, 8.0, 8, 4, 2, 73), (512, 2.0, 16, 8, 3, 74), (256, 8.0, 8, 4, 1, 75)]
    jobs0, jobs1, outs0, outs1 = [], [], [], []
    for nx, cell, W, s, nch, seed in specs:
        ob = _obs(nx, cell, nch, 1, seed, nant=12, ntime=30)
        d = H.Dev(ob, torch)
        p = mti.make_params(nx, nx, cell, cell, "wproj", W, s)
        K = ctx.build_kernels(p)
        for jobs, outs, masked in ((jobs0, outs0, False), (jobs1, outs1, True)):
            g = torch.zeros((nch, nx, nx, 1), dtype=torch.complex128, device="cuda")
            n = torch.zeros((nch, 1), dtype=torch.float64, device="cuda")
            im = torch.zeros((nch, 1, nx, nx), dtype=torch.float32, device="cuda")
            m = torch.empty(mti.tile_mask_shape(p, nch), dtype=torch.uint8, device="cuda") if masked else None
            v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
            jobs.append(ctx.make_job(K, p, v, g, n, im, 0, nch, ant1=d.ant1, ant2=d.ant2, nant=ob.nant,
                                     tile_mask=m))
            outs.append(im)
    for jobs in (jobs0, jobs1):
        ctx.grid_jobs(jobs, weighting="briggs")
        ctx.finalize_jobs(jobs)
    torch.cuda.synchronize()
    for a, b in zip(outs1, outs0):
        assert H.rel_rms(a.cpu().numpy(), b.cpu().numpy()) < 1e-9


def test_sparse_grids_are_refused_where_they_cannot_apply(mti, ctx):
    ob = _obs(256, 8.0, 1, 1, 76, nant=8, ntime=10)
    d = H.Dev(ob, torch)
    p = mti.make_params(256, 256, 8.0, 8.0, "wproj", 8, 4)
    K = ctx.build_kernels(p)
    v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
    mask = torch.empty(mti.tile_mask_shape(p, 1), dtype=torch.uint8, device="cuda")
    n = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    for prec, dt in (("c128", torch.complex128), ("c64", torch.complex64)):
        g = torch.zeros((1, 256, 256, 1), dtype=dt, device="cuda")
        with pytest.raises(mti.MTIError, match="accumulate"):
            ctx.grid(K, p, v, g, n, nchan=1, accumulate=True, precision=prec, tile_mask=mask)


def test_batched_sparse_jobs_refuse_accumulate(mti, ctx):
    """mti_grid_jobs refuses a job tile mask with accumulate (the unmarked
    tiles would keep whatever they held), before any work."""
    ob = _obs(256, 8.0, 1, 1, 77, nant=8, ntime=10)
    d = H.Dev(ob, torch)
    p = mti.make_params(256, 256, 8.0, 8.0, "wproj", 8, 4)
    K = ctx.build_kernels(p)
    g = torch.full((1, 256, 256, 1), 3.0 + 0j, dtype=torch.complex64, device="cuda")
    n = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    im = torch.zeros((1, 1, 256, 256), dtype=torch.float32, device="cuda")
    mask = torch.empty(mti.tile_mask_shape(p, 1), dtype=torch.uint8, device="cuda")
    v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
    job = ctx.make_job(K, p, v, g, n, im, 0, 1, ant1=d.ant1, ant2=d.ant2, nant=ob.nant, tile_mask=mask)
    with pytest.raises(mti.MTIError, match="accumulate"):
        ctx.grid_jobs([job], accumulate=True)
    torch.cuda.synchronize()
    assert float((g - 3.0).abs().max()) == 0.0
