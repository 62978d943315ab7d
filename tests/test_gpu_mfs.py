"""Multi-frequency synthesis (mti_grid_opts.mfs): a channel range gridded into
one grid and normaliser.  The reference images one plane per channel
(app/mtimager.cpp:151-170); feeding one WImager several channels' records is
its MFS.  Checked against the oracle (channels accumulated into one grid, one
Briggs density over all of them) and against the sum of per-channel grids."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c1(orc):
    from mtimager_amd import synth
    return synth.small_array_obs(seed=1, ntime=100)


@pytest.fixture(scope="module")
def p(mti):
    return mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)


def _gpu_mfs(mti, ctx, K, p, dev, idx, chan0, nch, weighting, precision="c128"):
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    gdt = torch.float32 if precision == "c64" else torch.float64
    grids = torch.empty((1, p.ny, p.nx, 4, 2), dtype=gdt, device="cuda")
    norm = torch.empty((1, 4), dtype=torch.float64, device="cuda")
    ctx.grid(K, p, v, grids, norm, chan0=chan0, nchan=nch, weighting=weighting, robust=0.0, mfs=True,
             precision=precision)
    img = torch.empty((1, 4, p.ny, p.nx), dtype=torch.float32, device="cuda")
    ctx.finalize(p, 4, grids, norm, img, precision=precision)
    torch.cuda.synchronize()
    return grids.cpu().numpy()[0], norm.cpu().numpy()[0], img.cpu().numpy()[0]


def _oracle_mfs(orc, p, c1, rows, chan0, nch, weighting):
    op = H.oracle_params(orc, p)
    OK = orc.Kernels(op)
    uvw, flags, weight, vis = c1.uvw[rows], c1.flags[rows], c1.weight[rows], c1.vis[rows]
    wts = orc.weights_mfs(OK, uvw, c1.inv_lambda, flags, weight, chan0, nch, weighting, 0.0)
    g = np.zeros((p.ny, p.nx, 4, 2))
    n = np.zeros(4)
    for c in range(nch):
        orc.grid(OK, uvw, c1.inv_lambda, vis, flags, weight, chan0 + c, wts=wts[c], grid_acc=g, norm_acc=n)
    return g, n, orc.finalize(op, g, n)


@pytest.mark.parametrize("weighting", ["natural", "briggs"])
def test_mfs_matches_oracle(orc, mti, ctx, c1, p, weighting):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(p)
    g, n, img = _gpu_mfs(mti, ctx, K, p, dev, idx, 1, 6, weighting)
    rg, rn, rimg = _oracle_mfs(orc, p, c1, idx.cpu().numpy(), 1, 6, weighting)
    # the library sums f32 sub-kernel sums and f32 effective weights in FP64
    np.testing.assert_allclose(n, rn, rtol=1e-6)
    assert H.rel_rms(g[..., 0] + 1j * g[..., 1], rg[..., 0] + 1j * rg[..., 1]) < 1e-5
    for s in range(4 if weighting == "natural" else 1):
        e = H.image_errors(img[s], rimg[s])
        print(f"MFS {weighting} Stokes {s}:", e)
        assert e["rel_rms"] < 1e-4


def test_mfs_equals_sum_of_channels(mti, ctx, c1, p):
    """Natural weighting is linear: the MFS grid is the sum of the per-channel
    grids and its normaliser the sum of theirs (FP64 grids)."""
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(p)
    g, n, _ = _gpu_mfs(mti, ctx, K, p, dev, idx, 0, 8, "natural")
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    grids = torch.empty((8, p.ny, p.nx, 4, 2), dtype=torch.float64, device="cuda")
    norm = torch.empty((8, 4), dtype=torch.float64, device="cuda")
    ctx.grid(K, p, v, grids, norm, chan0=0, nchan=8)
    np.testing.assert_allclose(n, norm.sum(0).cpu().numpy(), rtol=1e-12)
    ref = grids.sum(0).cpu().numpy()
    assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max()


def test_mfs_batched_jobs(mti, ctx, c1, p):
    """Two batched jobs, one MFS over channels 0-3 and one over 4-7, equal
    the single-call MFS grids."""
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(p)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    outs = [(torch.empty((1, p.ny, p.nx, 4, 2), dtype=torch.float64, device="cuda"),
             torch.empty((1, 4), dtype=torch.float64, device="cuda")) for _ in range(2)]
    jobs = [ctx.make_job(K, p, v, g, n, chan0=4 * i, nchan=4) for i, (g, n) in enumerate(outs)]
    ctx.grid_jobs(jobs, weighting="briggs", mfs=True)
    for i, (g, n) in enumerate(outs):
        rg, rn, _ = _gpu_mfs(mti, ctx, K, p, dev, idx, 4 * i, 4, "briggs")
        np.testing.assert_allclose(n.cpu().numpy()[0], rn, rtol=1e-12)
        gg = g.cpu().numpy()[0]
        assert np.abs(gg - rg).max() <= 1e-10 * np.abs(rg).max()
