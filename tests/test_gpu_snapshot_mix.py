"""The snapshot-job mix and its serving stream on the GPU: 512 independent
Stokes-I jobs (10k-5M visibilities log-uniform, 256^2-2048^2 images, 8-64
w-planes, oversampling 4 or 8, seed 4).

* One job of each image size, rows thinned to an oracle-feasible count, imaged
  by one mti_grid_jobs + mti_finalize_jobs call in the benchmark's form
  (complex64 grids, complex64 transform planes, tile masks) against the FP64
  oracle.
* The whole mix at full size: the batched calls (waves cut by an HBM budget)
  against the single-job entry points, job by job.
* A Poisson stream at 800 requests/s drawn from the mix, served by the
  continuous-batching server: every request's image against the image of the
  same job imaged alone."""
import time

import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 4
NANT = 64


@pytest.fixture(scope="module")
def specs():
    from mtimager_amd import synth
    return synth.snapshot_jobs(512, seed=SEED)


def _upload(mti, sp, gen, rows=None, chans=None):
    """A job's inputs on the GPU (visibilities predicted from its sky, 5%
    random flags, weights 1), optionally restricted to the first `rows` rows
    and the channels `chans`."""
    from mtimager_amd import synth
    uvw, a1, a2, il = sp["uvw"], sp["ant1"], sp["ant2"], sp["inv_lambda"]
    if rows is not None:
        uvw, a1, a2 = uvw[:rows], a1[:rows], a2[:rows]
    if chans is not None:
        il = il[chans]
    nrows, nch = len(uvw), len(il)
    vis = synth.predict_torch(uvw, il, sp["src"], 1)
    flags = (torch.rand((nrows, nch), generator=gen, device="cuda") < 0.05).to(torch.uint8)[:, :, None].contiguous()
    nx = sp["nx"]
    return dict(p=mti.make_params(nx, nx, sp["cell"], sp["cell"], "wproj", sp["wplanes"], sp["sampling"]),
                key=(nx, sp["cell"], sp["wplanes"], sp["sampling"]), nx=nx, nch=nch, nrows=nrows,
                uvw=torch.tensor(uvw, dtype=torch.float64, device="cuda"),
                invl=torch.tensor(il, dtype=torch.float64, device="cuda"),
                vis=vis, flags=flags, weight=torch.ones((nrows, 1), dtype=torch.float32, device="cuda"),
                a1=torch.tensor(a1, device="cuda"), a2=torch.tensor(a2, device="cuda"),
                host=dict(uvw=uvw, a1=a1, a2=a2, il=il))


def _outputs(j, gdt, mask=True):
    nx, nch = j["nx"], j["nch"]
    from mtimager_amd import tile_mask_shape
    return (torch.empty((nch, nx, nx, 1, 2), dtype=gdt, device="cuda"),
            torch.empty((nch, 1), dtype=torch.float64, device="cuda"),
            torch.empty((nch, 1, nx, nx), dtype=torch.float32, device="cuda"),
            torch.empty(tile_mask_shape(j["p"], nch), dtype=torch.uint8, device="cuda") if mask else None)


def test_one_job_per_image_size_against_the_oracle(orc, mti, specs):
    """256^2, 512^2, 1024^2 and 2048^2 jobs of the mix in one batched call,
    in the benchmark's form, each channel's image within 1e-4 of the oracle."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED)
    picked = []
    for nx in (256, 512, 1024, 2048):
        sp = next(s for s in specs if s["nx"] == nx)
        nbl = int(len(np.unique(sp["ant1"] * NANT + sp["ant2"])))
        rows = min(len(sp["uvw"]), max(nbl, (6000 // nbl) * nbl))  # whole time steps
        chans = [0, len(sp["inv_lambda"]) - 1] if len(sp["inv_lambda"]) > 1 else [0]
        picked.append(_upload(mti, sp, gen, rows=rows, chans=chans))
    ctx = mti.Context(0)
    ctx.set_transform_storage("c64")
    jobs, outs = [], []
    for j in picked:
        K = ctx.build_kernels(j["p"])
        gr, nr, im, mk = _outputs(j, torch.float32)
        v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"])
        jobs.append(ctx.make_job(K, j["p"], v, gr, nr, im, 0, j["nch"], ant1=j["a1"], ant2=j["a2"], nant=NANT,
                                 tile_mask=mk))
        outs.append((K, im))
    ctx.grid_jobs(jobs)
    ctx.finalize_jobs(jobs)
    torch.cuda.synchronize()
    for j, (K, im) in zip(picked, outs):
        h = j["host"]
        op = H.oracle_params(orc, j["p"])
        OK = orc.Kernels(op)
        rows = orc.sort_index(h["a1"], h["a2"], NANT)
        vis = j["vis"].cpu().numpy()
        flags = j["flags"].cpu().numpy()
        w = np.ones((j["nrows"], 1), np.float32)
        img = im.cpu().numpy()
        for c in range(j["nch"]):
            g, n, ng = orc.grid(OK, h["uvw"][rows], h["il"], vis[rows], flags[rows], w[rows], c)
            ref = orc.finalize(op, g, n)[0]
            e = H.image_errors(img[c, 0], ref)
            print(f"\nsnapshot job {j['nx']}^2 ({j['nrows']} rows, {ng} gridded) channel {c}: {e}", flush=True)
            assert e["rel_rms"] < 1e-4
        K.free()
    ctx.close()


def test_whole_mix_batched_equals_single_job_path(mti, specs):
    """All 512 jobs at full size: waves of mti_grid_jobs + mti_finalize_jobs
    (complex128 grids: FP64 sums that differ only in order) against
    mti_sort_rows + mti_grid + mti_finalize one job at a time."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED)
    budget = 24 << 30
    ctx, one = mti.Context(0), mti.Context(0)
    cache = {}
    wave, mem, nwaves, worst, checked, offered = [], 0, 0, 0.0, 0, 0

    def run(wave):
        nonlocal worst, checked
        jobs, outs = [], []
        for j in wave:
            gr, nr, im, mk = _outputs(j, torch.float64)
            v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"])
            jobs.append(ctx.make_job(cache[j["key"]], j["p"], v, gr, nr, im, 0, j["nch"], ant1=j["a1"],
                                     ant2=j["a2"], nant=NANT, tile_mask=mk))
            outs.append((gr, nr, im, mk))
        ctx.grid_jobs(jobs, stats=False)
        ctx.finalize_jobs(jobs)
        for j, (gr, nr, im, mk) in zip(wave, outs):
            idx = torch.empty(j["nrows"], dtype=torch.int32, device="cuda")
            kept = one.sort_rows(j["a1"], j["a2"], NANT, idx)
            v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"], idx[:kept])
            g1, n1, i1, m1 = _outputs(j, torch.float64)
            one.grid(cache[j["key"]], j["p"], v, g1, n1, chan0=0, nchan=j["nch"], stats=False, tile_mask=m1)
            one.finalize(j["p"], 1, g1, n1, i1, nchan=j["nch"], tile_mask=m1)
            torch.cuda.synchronize()
            assert torch.equal(mk, m1)
            assert torch.allclose(nr, n1, rtol=1e-12, atol=0.0)
            d = float((im.double() - i1.double()).abs().max() / i1.double().abs().max().clamp_min(1e-300))
            worst = max(worst, d)
            checked += 1
            assert d < 1e-6, (j["nx"], j["nch"], j["nrows"], d)
        del jobs, outs

    for n, sp in enumerate(specs):
        j = _upload(mti, sp, gen)
        offered += j["nrows"] * j["nch"]
        if j["key"] not in cache:
            cache[j["key"]] = ctx.build_kernels(j["p"])
        # both paths' complex128 grids and f32 images, and the calls' scratch
        need = j["nch"] * j["nx"] ** 2 * 40 + 200 * j["nrows"] * j["nch"]
        if wave and mem + need > budget:
            run(wave)
            nwaves += 1
            wave, mem = [], 0
            print(f"\n{n} jobs checked, worst image difference {worst:.2e}", flush=True)
        wave.append(j)
        mem += need
    run(wave)
    nwaves += 1
    print(f"\n{checked} jobs in {nwaves} waves, {offered} visibilities, {len(cache)} kernel-table sets, "
          f"worst image difference {worst:.2e}", flush=True)
    assert checked == 512
    for K in cache.values():
        K.free()
    ctx.close()
    one.close()


def _digest(im, probe):
    """Per channel: the sum of squares and a random-sign projection of the image."""
    x = im[:, 0].double()
    return torch.stack([(x * x).sum(dim=(1, 2)), (x * probe).sum(dim=(1, 2))], dim=1).cpu().numpy()


def test_serving_stream_at_800_per_second_matches_solo_images(mti, specs):
    """A Poisson stream at 800 req/s for 2 s (about 1600 requests, the
    benchmark's serving form: complex64 grids and planes, two workers, a
    kernel-table cache) drawn from the whole mix.  Each request hands its
    image to a digest on completion (sum of squares and a random-sign
    projection per channel) compared with the digest of the same job imaged
    alone through the single-job entry points; the requests of 16 jobs (of at
    most 16 M pixels each) keep their images, which are compared pixel by
    pixel."""
    from mtimager_amd import serve
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED)
    pool = [_upload(mti, sp, gen) for sp in specs]
    probes = {}
    for nx in (256, 512, 1024, 2048):
        g = torch.Generator(device="cuda")
        g.manual_seed(nx)
        probes[nx] = (torch.randint(0, 2, (nx, nx), generator=g, device="cuda").double() * 2 - 1)
    rng = np.random.default_rng(5)
    small = [k for k, j in enumerate(pool) if j["nch"] * j["nx"] ** 2 <= (16 << 20)]
    kept = set(int(k) for k in rng.choice(small, 16, replace=False))
    # solo images: the single-job path, one job at a time
    solo_d, solo_img = {}, {}
    ctx = mti.Context(0)
    ctx.set_transform_storage("c64")
    for k, j in enumerate(pool):
        idx = torch.empty(j["nrows"], dtype=torch.int32, device="cuda")
        nk = ctx.sort_rows(j["a1"], j["a2"], NANT, idx)
        K = ctx.build_kernels(j["p"])
        v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"], idx[:nk])
        gr, nr, im, mk = _outputs(j, torch.float32)
        ctx.grid(K, j["p"], v, gr, nr, chan0=0, nchan=j["nch"], stats=False, tile_mask=mk)
        ctx.finalize(j["p"], 1, gr, nr, im, nchan=j["nch"], tile_mask=mk)
        solo_d[k] = _digest(im, probes[j["nx"]])
        if k in kept:
            solo_img[k] = im
        K.free()
    ctx.close()
    torch.cuda.synchronize()

    rate, window = 800.0, 2.0
    n = int(rng.poisson(rate * window))
    times = np.sort(rng.uniform(0.0, window, n))
    picks = rng.integers(0, len(pool), n)
    served_d, served_img = {}, []

    def hand_off(r):
        k = r.rid[1]
        served_d[r.rid] = _digest(r.image, probes[pool[k]["nx"]])
        if k in kept:
            served_img.append((k, r.image))

    cache = serve.KernelCache()
    srv = serve.Server(64 << 30, workers=2, max_wave=512, cache=cache, transform_storage="c64",
                       grid_precision="c64")
    reqs = []
    try:
        t0 = time.perf_counter()
        for i, (at, k) in enumerate(zip(times, picks)):
            j = pool[int(k)]
            r = serve.Request((i, int(k)), j["p"], (j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"]),
                              j["a1"], j["a2"], NANT, j["nch"], npol=1, keep_image=False, on_done=hand_off)
            dt = t0 + at - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            srv.submit(r)
            reqs.append(r)
        assert srv.wait(timeout=300)
    finally:
        srv.close()
        cache.clear()
    assert not srv.errors, srv.errors
    assert all(r.error is None for r in reqs)
    print(f"\n{len(reqs)} requests in {len(srv.wave_sizes)} waves (largest {max(srv.wave_sizes)}), "
          f"{len(served_img)} kept images", flush=True)
    assert len(served_d) == len(reqs) and max(srv.wave_sizes) > 1
    worst = 0.0
    bad = []
    for r in reqs:
        a, b = served_d[r.rid], solo_d[r.rid[1]]
        # complex64 grids: float atomics in split tiles add in another order
        ss = np.abs(a[:, 0] - b[:, 0]) / np.maximum(b[:, 0], 1e-300)
        pr = np.abs(a[:, 1] - b[:, 1]) / np.sqrt(np.maximum(b[:, 0], 1e-300))
        worst = max(worst, float(ss.max()), float(pr.max()))
        if not (ss.max() < 2e-4 and pr.max() < 1e-3):
            bad.append((r, a, b, ss, pr))
    for r, a, b, ss, pr in bad[:3]:
        j = pool[r.rid[1]]
        mates = sorted(q.rid[1] for q in reqs if q.wave == r.wave)
        print(f"request {r.rid}: job nx={j['nx']} nch={j['nch']} nrows={j['nrows']} "
              f"W={j['p'].wplanes} s={j['p'].sampling}; wave {r.wave} of {len(mates)} requests "
              f"(jobs {mates[:20]}...); channels off: {np.nonzero(ss >= 2e-4)[0].tolist()}", flush=True)
        print(f"  served sumsq {a[:, 0].tolist()}\n  solo sumsq {b[:, 0].tolist()}\n"
              f"  served proj {a[:, 1].tolist()}\n  solo proj {b[:, 1].tolist()}", flush=True)
    assert not bad, [(r.rid, float(ss.max()), float(pr.max())) for r, a, b, ss, pr in bad]
    for k, im in served_img:
        assert H.rel_rms(im.cpu().numpy(), solo_img[k].cpu().numpy()) < 5e-5, k
    print(f"worst digest difference {worst:.2e}", flush=True)
