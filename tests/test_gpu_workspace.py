"""Caller-owned workspace on the GPU: each entry point run in a workspace of
its *_workspace_size gives the results of the same call with scratch owned by
the context, the context allocates no device memory of its own, and a
workspace below the size is refused before any work is enqueued."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _obs(nx, cell, nch, npol, seed, nant=12, ntime=30):
    from mtimager_amd import synth
    return synth.make_observation(nant=nant, radius_m=1500.0, ntime=ntime,
                                  freqs_hz=np.linspace(1.3e9, 1.45e9, nch), npol=npol,
                                  fov_rad=nx * cell * synth.ARCSEC, seed=seed)


def _zeros(shape, dtype):
    return torch.zeros(shape, dtype=dtype, device="cuda")


@pytest.mark.parametrize("grid_precision", ["c128", "c64"])
def test_single_job_calls_in_a_workspace(mti, ctx, grid_precision):
    P, nch, nx = 4, 3, 256
    ob = _obs(nx, 8.0, nch, P, seed=31)
    dev = H.Dev(ob, torch)
    p = mti.make_params(nx, nx, 8.0, 8.0, "wproj", 16, 4)
    K = ctx.build_kernels(p)  # tables come from the shared context
    gdt = torch.complex128 if grid_precision == "c128" else torch.complex64
    nrows = len(ob.ant1)

    def run(c):
        idx = torch.empty(nrows, dtype=torch.int32, device="cuda")
        kept = c.sort_rows(dev.ant1, dev.ant2, ob.nant, idx)
        v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx[:kept])
        g, n, im = _zeros((nch, nx, nx, P), gdt), _zeros((nch, P), torch.float64), _zeros((nch, P, nx, nx), torch.float32)
        st = c.grid(K, p, v, g, n, nchan=nch, weighting="briggs", precision=grid_precision)
        c.finalize(p, P, g, n, im, nchan=nch, precision=grid_precision)
        torch.cuda.synchronize()
        return kept, idx[:kept].cpu().numpy(), g.cpu().numpy(), n.cpu().numpy(), im.cpu().numpy(), st, v

    k0, i0, g0, n0, im0, st0, v0 = run(ctx)
    need = max(mti.sort_rows_workspace_size(nrows),
               mti.grid_workspace_size(K.info, p, v0, nchan=nch, weighting="briggs", precision=grid_precision),
               mti.finalize_workspace_size(p, P, nch))
    w = mti.Context(0)
    try:
        buf = torch.empty(need, dtype=torch.uint8, device="cuda")
        w.set_workspace(buf)
        k1, i1, g1, n1, im1, st1, _ = run(w)
        assert w.device_bytes() == 0  # every call ran in the caller's buffer
        assert k1 == k0
        np.testing.assert_array_equal(i1, i0)
        assert st1.gridded_records == st0.gridded_records and st1.gridded_points == st0.gridded_points
        np.testing.assert_allclose(n1, n0, rtol=1e-12)
        assert H.rel_rms(g1, g0) < 1e-12
        assert H.rel_rms(im1, im0) < 1e-9
        # too small: refused with a message, nothing run
        w.set_workspace(buf[: mti.finalize_workspace_size(p, P, nch) // 2])
        im_bad = _zeros((nch, P, nx, nx), torch.float32)
        with pytest.raises(mti.MTIError, match="workspace"):
            w.finalize(p, P, torch.from_numpy(g0).cuda(), torch.from_numpy(n0).cuda(), im_bad, nchan=nch,
                       precision=grid_precision)
        torch.cuda.synchronize()
        assert float(im_bad.abs().max()) == 0.0
        # back to context scratch
        w.set_workspace(None)
        _, _, _, _, im2, _, _ = run(w)
        assert H.rel_rms(im2, im0) < 1e-9 and w.device_bytes() > 0
    finally:
        w.close()


@pytest.mark.parametrize("form", ["c128", "c64"])
def test_batched_calls_in_a_workspace(mti, ctx, form):
    """Jobs of three sizes and samplings, rows sorted inside mti_grid_jobs,
    finalized by mti_finalize_jobs: the same results in a workspace -- with
    complex128 grids and planes, and with complex64 grids and planes (whose
    requests stay within the same plans; float atomics in split tiles make
    those runs agree to the last bits of such cells)."""
    gdt = torch.complex128 if form == "c128" else torch.complex64
    gtol, itol = (1e-12, 1e-9) if form == "c128" else (1e-6, 5e-5)
    specs = [(256, 8.0, 8, 4, 2, 41), (512, 2.0, 16, 8, 3, 42), (256, 6.0, 12, 8, 1, 43)]
    jobs_in = []
    for nx, cell, W, s, nch, seed in specs:
        ob = _obs(nx, cell, nch, 1, seed)
        jobs_in.append((ob, H.Dev(ob, torch), mti.make_params(nx, nx, cell, cell, "wproj", W, s), nch))
    kernels = [ctx.build_kernels(p) for _, _, p, _ in jobs_in]

    def run(c):
        jobs, outs = [], []
        for (ob, d, p, nch), K in zip(jobs_in, kernels):
            g = _zeros((nch, p.ny, p.nx, 1), gdt)
            n = _zeros((nch, 1), torch.float64)
            im = _zeros((nch, 1, p.ny, p.nx), torch.float32)
            v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
            jobs.append(c.make_job(K, p, v, g, n, im, 0, nch, ant1=d.ant1, ant2=d.ant2, nant=ob.nant))
            outs.append((g, n, im))
        st = c.grid_jobs(jobs, weighting="briggs")
        c.finalize_jobs(jobs)
        torch.cuda.synchronize()
        return jobs, st, [tuple(t.cpu().numpy() for t in o) for o in outs]

    ref = mti.Context(0)
    w = mti.Context(0)
    try:
        for c in (ref, w):
            c.set_transform_storage(form)
        jobs0, st0, r0 = run(ref)
        need = max(mti.grid_jobs_workspace_size(jobs0, weighting="briggs"), mti.finalize_jobs_workspace_size(jobs0))
        w.set_workspace(torch.empty(need, dtype=torch.uint8, device="cuda"))
        _, st1, r1 = run(w)
        assert w.device_bytes() == 0
        assert st1.gridded_records == st0.gridded_records
        for (g0, n0, i0), (g1, n1, i1) in zip(r0, r1):
            np.testing.assert_allclose(n1, n0, rtol=1e-12)
            assert H.rel_rms(g1, g0) < gtol
            assert H.rel_rms(i1, i0) < itol
    finally:
        w.close()
        ref.close()


def test_fft2d_in_a_workspace(mti, ctx):
    rng = np.random.default_rng(5)
    x = (rng.normal(size=(3, 128, 64)) + 1j * rng.normal(size=(3, 128, 64))).astype(np.complex64)
    a = torch.from_numpy(x).cuda()
    ctx.fft2d_inverse(a, 64, 128, 3)
    w = mti.Context(0)
    try:
        w.set_workspace(torch.empty(mti.fft2d_workspace_size(64, 128, 3), dtype=torch.uint8, device="cuda"))
        b = torch.from_numpy(x).cuda()
        w.fft2d_inverse(b, 64, 128, 3)
        torch.cuda.synchronize()
        assert w.device_bytes() == 0
        np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
    finally:
        w.close()
