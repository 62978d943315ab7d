"""Batched imaging captured into a HIP graph: every replay, with the
program's own eager work and synchronisations between replays and new input
contents before each, equals the eager mti_grid_jobs + mti_finalize_jobs
result on the same inputs."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# nx, cell arcsec, w-planes, sampling, channels, seed
SPECS = [(256, 8.0, 8, 4, 2, 61), (512, 2.0, 16, 8, 3, 62), (256, 6.0, 12, 8, 1, 63)]


def _jobs(mti, ctx, inputs, kernels, gdt=torch.complex128):
    jobs, outs = [], []
    for (ob, d, p, nch), K in zip(inputs, kernels):
        g = torch.zeros((nch, p.ny, p.nx, 1), dtype=gdt, device="cuda")
        n = torch.zeros((nch, 1), dtype=torch.float64, device="cuda")
        im = torch.zeros((nch, 1, p.ny, p.nx), dtype=torch.float32, device="cuda")
        v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
        jobs.append(ctx.make_job(K, p, v, g, n, im, 0, nch, ant1=d.ant1, ant2=d.ant2, nant=ob.nant))
        outs.append((g, n, im))
    return jobs, outs


def _host(outs):
    return [tuple(t.cpu().numpy() for t in o) for o in outs]


# the precision forms: complex128 grids and planes; complex64 grids and planes
# (float atomics in split tiles: the last bits of those cells follow the
# order units finish, in a replay as in an eager run)
FORMS = {"c128": (torch.complex128, 1e-12, 1e-9), "c64": (torch.complex64, 1e-6, 5e-5)}


@pytest.mark.parametrize("weighting,form", [("natural", "c128"), ("briggs", "c128"), ("briggs", "c64")])
def test_graph_replays_equal_eager_runs(mti, ctx, weighting, form):
    if form == "c64":
        own = mti.Context(0)
        own.set_transform_storage("c64")
        try:
            _graph_replays(mti, own, weighting, form)
        finally:
            own.close()
    else:
        _graph_replays(mti, ctx, weighting, form)


def _graph_replays(mti, ctx, weighting, form):
    from mtimager_amd import synth
    gdt, gtol, itol = FORMS[form]
    inputs = []
    for nx, cell, W, s, nch, seed in SPECS:
        ob = synth.make_observation(nant=12, radius_m=1500.0, ntime=30, freqs_hz=np.linspace(1.3e9, 1.45e9, nch),
                                    npol=1, fov_rad=nx * cell * synth.ARCSEC, seed=seed)
        inputs.append((ob, H.Dev(ob, torch), mti.make_params(nx, nx, cell, cell, "wproj", W, s), nch))
    kernels = [ctx.build_kernels(p) for _, _, p, _ in inputs]
    # one eager run first: the context's own scratch reaches these sizes
    wjobs, _ = _jobs(mti, ctx, inputs, kernels, gdt)
    ctx.grid_jobs(wjobs, weighting=weighting)
    ctx.finalize_jobs(wjobs)
    torch.cuda.synchronize()
    gjobs, gouts = _jobs(mti, ctx, inputs, kernels, gdt)
    graph = ctx.capture_jobs(gjobs, weighting=weighting)
    before = ctx.device_bytes()
    rng = np.random.default_rng(1)
    try:
        for replay in range(3):
            # new contents of the same input buffers
            if replay == 1:
                for _, d, _, _ in inputs:
                    d.vis.mul_(2.0)
            if replay == 2:
                for _, d, _, _ in inputs:
                    rows = torch.from_numpy(rng.random(d.flags.shape[0]) < 0.2).cuda()
                    d.flags[rows] = 1
            graph.launch()
            torch.cuda.synchronize()
            got = _host(gouts)
            # the eager reference on the same inputs, into other buffers
            ejobs, eouts = _jobs(mti, ctx, inputs, kernels, gdt)
            ctx.grid_jobs(ejobs, weighting=weighting)
            ctx.finalize_jobs(ejobs)
            torch.cuda.synchronize()
            ref = _host(eouts)
            for (g1, n1, i1), (g0, n0, i0) in zip(got, ref):
                np.testing.assert_allclose(n1, n0, rtol=1e-12)
                assert H.rel_rms(g1, g0) < gtol, replay
                assert H.rel_rms(i1, i0) < itol, replay
        assert ctx.device_bytes() == before  # launches take nothing from the context
    finally:
        graph.free()
