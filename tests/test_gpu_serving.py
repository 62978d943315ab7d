"""Serving loop on the GPU: every request's image from a continuous-batching
run (mixed waves, two workers, shared kernel-table cache) equals the image of
the same request imaged alone through the single-job entry points."""
import time

import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# nx, cell arcsec, w-planes, sampling, antennas, steps, channels, seed
SPECS = [
    (256, 8.0, 8, 4, 12, 40, 3, 21),
    (512, 2.0, 16, 8, 20, 25, 2, 22),
    (256, 6.0, 12, 8, 8, 60, 1, 23),
    (256, 8.0, 8, 4, 10, 30, 2, 24),   # shares its kernel tables with the first
    (1024, 1.0, 24, 4, 16, 20, 2, 25),
    (512, 2.0, 16, 8, 14, 35, 4, 26),  # and this one with the second
]


def _request(mti, serve, rid, spec):
    from mtimager_amd import synth
    nx, cell, W, s, nant, ntime, nch, seed = spec
    ob = synth.make_observation(nant=nant, radius_m=1500.0, ntime=ntime,
                                freqs_hz=np.linspace(1.3e9, 1.45e9, nch), npol=1,
                                fov_rad=nx * cell * synth.ARCSEC, seed=seed)
    dev = H.Dev(ob, torch)
    p = mti.make_params(nx, nx, cell, cell, "wproj", W, s)
    r = serve.Request(rid, p, (dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight),
                      dev.ant1, dev.ant2, ob.nant, nch, npol=1)
    return r, dev


def _alone(mti, ctx, r, dev, weighting, precision="c128"):
    idx = H.gpu_sort(ctx, dev, r.nant, torch)
    K = ctx.build_kernels(r.params)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    nx = r.params.nx
    g = torch.zeros((r.nchan, nx, nx, 1, 2), dtype=torch.float64 if precision == "c128" else torch.float32,
                    device="cuda")
    n = torch.zeros((r.nchan, 1), dtype=torch.float64, device="cuda")
    im = torch.zeros((r.nchan, 1, nx, nx), dtype=torch.float32, device="cuda")
    ctx.grid(K, r.params, v, g, n, chan0=0, nchan=r.nchan, weighting=weighting, precision=precision)
    ctx.finalize(r.params, 1, g, n, im, nchan=r.nchan, precision=precision)
    torch.cuda.synchronize()
    K.free()
    return im.cpu().numpy(), n.cpu().numpy()


# served vs alone: the same sums in another order -- FP64 atomics for
# complex128 grids, float atomics (split tiles) for complex64 grids
SOLO_TOL = {"c128": 1e-6, "c64": 5e-5}


@pytest.mark.parametrize("weighting,precision", [("natural", "c128"), ("briggs", "c128"), ("briggs", "c64")])
def test_served_results_equal_solo_results(mti, ctx, weighting, precision):
    from mtimager_amd import serve
    rng = np.random.default_rng(7)
    reqs = [_request(mti, serve, i, s) for i, s in enumerate(SPECS * 2)]
    cache = serve.KernelCache()
    # a budget that holds about half of the requests at once, so admission waits
    budget = sum(r.nbytes for r, _ in reqs) // 2 + max(r.nbytes for r, _ in reqs)
    srv = serve.Server(budget, workers=2, max_wave=5, cache=cache, weighting=weighting, grid_precision=precision)
    try:
        for r, _ in reqs:  # Poisson arrivals at ~400 req/s
            time.sleep(float(rng.exponential(1.0 / 400.0)))
            srv.submit(r)
        assert srv.wait(timeout=120)
    finally:
        srv.close()
    assert not srv.errors, srv.errors
    distinct = len({serve.KernelCache.key(r.params) for r, _ in reqs})
    assert cache.misses == distinct and cache.hits == len(reqs) - distinct
    assert len(srv.wave_sizes) >= 2 and max(srv.wave_sizes) <= 5
    for r, dev in reqs:
        assert r.done >= r.admitted >= r.arrival
        im, n = _alone(mti, ctx, r, dev, weighting, precision)
        np.testing.assert_allclose(r.norm.cpu().numpy(), n, rtol=1e-12)
        assert H.rel_rms(r.image.cpu().numpy(), im) < SOLO_TOL[precision], r.rid
    cache.clear()


def test_plain_server_matches_batched_server(mti, ctx):
    from mtimager_amd import serve
    reqs = [_request(mti, serve, i, s) for i, s in enumerate(SPECS[:3])]
    srv = serve.Server(10 ** 10, plain=True)
    try:
        for r, _ in reqs:
            srv.submit(r)
        assert srv.wait(timeout=60)
    finally:
        srv.close()
    assert not srv.errors and srv.wave_sizes == [1, 1, 1]
    for r, dev in reqs:
        im, _ = _alone(mti, ctx, r, dev, "natural")
        assert r.gridded > 0
        assert H.rel_rms(r.image.cpu().numpy(), im) < 1e-6
