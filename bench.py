#!/usr/bin/env python3
"""Imaging benchmark on one MI355X.

Workload (default): the full-polarisation cube at real size -- 64 antennas in
an 8 km disk, 2016 baselines x 1800 time steps over +-4 h at dec -45 deg, 64
channels over 0.9-1.7 GHz (232 M visibilities x 4 correlations, complex64),
20 sources with 10% linear polarisation, 4096^2 image at 1.5 arcsec,
256 w-planes, oversampling 8, Briggs robust 0, Stokes IQUV.

One step = one imaging job from HBM-resident inputs to HBM-resident images:
baseline-major row sort, W-kernel tables, binning + Briggs weights + tile
sort + gridding of all 64 channels, FP64-transform finalize of all 64
channels.  value = gridded visibilities (row x channel samples) per second.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mt-imager_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mvis/s gridded + p99 job latency, batched/served vs plain, at ≤1e-4 error"
FP64_PEAK_TFLOPS = 78.6    # MI355X published FP64 peak (vector and matrix)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(torch, cfg, seed):
    import numpy as np
    from mtimager_amd import synth
    if cfg == "c2":
        nant, radius, dec, ha, ntime = 64, 8000.0, -45.0, (-4.0, 4.0), 1800
        freqs = np.linspace(0.9e9, 1.7e9, 64)
        nx, cell, W, s, pol_frac = 4096, 1.5, 256, 8, 0.1
    else:  # c1
        nant, radius, dec, ha, ntime = 27, 3000.0, -30.0, (-2.0, 2.0), 300
        freqs = np.linspace(1.4e9, 1.5e9, 8)
        nx, cell, W, s, pol_frac = 512, 2.0, 32, 4, 0.0
    enu = synth.array_layout(nant, radius, seed)
    uvw, a1, a2 = synth.uvw_tracks(enu, -30.7, dec, ha, ntime, with_autocorr=True)
    inv_lambda = freqs / synth.C_LIGHT
    src = synth.sky(20, nx * cell * synth.ARCSEC, seed, pol_frac=pol_frac)
    nrows, nchan, P = len(uvw), len(freqs), 4
    log(f"{cfg}: {nrows} rows x {nchan} channels x {P} correlations; predicting on the GPU")
    vis = synth.predict_torch(uvw, inv_lambda, src, P)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    vis += 0.01 * torch.complex(torch.randn(vis.shape, generator=g, device="cuda"),
                                torch.randn(vis.shape, generator=g, device="cuda"))
    flags = (torch.rand((nrows, nchan), generator=g, device="cuda") < 0.05).to(torch.uint8)
    flags = flags[:, :, None].expand(nrows, nchan, P).contiguous()
    d = dict(uvw=torch.tensor(uvw, dtype=torch.float64, device="cuda"),
             invl=torch.tensor(inv_lambda, dtype=torch.float64, device="cuda"),
             vis=vis, flags=flags,
             weight=torch.ones((nrows, P), dtype=torch.float32, device="cuda"),
             a1=torch.tensor(a1, device="cuda"), a2=torch.tensor(a2, device="cuda"),
             nant=nant, nrows=nrows, nchan=nchan, P=P, nx=nx, cell=cell, W=W, s=s,
             # the cube is imaged with Briggs robust 0 weighting, the small array with natural
             weighting="briggs" if cfg == "c2" else "natural",
             host=dict(uvw=uvw, a1=a1, a2=a2, inv_lambda=inv_lambda, src=src))
    torch.cuda.synchronize()
    return d


def measured_traffic(cfg):
    """HBM bytes per gridder launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (scaled by a copy kernel of known size) for this workload, as
    recorded in profiles/; None when no measurement exists for it."""
    if cfg != "c2":
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "k_grid_tiles_counters_fullpol_cube.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def achievable_peaks():
    """Peaks re-measured on this GPU with libmti_probe: a device copy of 2 GiB
    (read + write bytes per second) and an FP64 FMA loop (8 independent chains
    per thread, 8192 x 256 threads)."""
    import ctypes
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mt-imager_amd", "libmti_probe.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.mtp_copy_gbps.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.mtp_fp64_fma_tflops.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    gbps, tflops = ctypes.c_double(0), ctypes.c_double(0)
    if L.mtp_copy_gbps(2 << 30, 5, ctypes.byref(gbps)) != 0:
        return None
    if L.mtp_fp64_fma_tflops(8192, 2048, 5, ctypes.byref(tflops)) != 0:
        return None
    return dict(copy_GBps=round(gbps.value, 1), fp64_fma_TFLOPs=round(tflops.value, 2))


def end_to_end(torch, mti, ctx, d, p, grids, norm, image, ref_image, group=8, tmask=None, pinned=True,
               reader_threads=8):
    """One imaging job from host memory through the package's chunked loader
    (mtimager_amd.ingest): the job's columns in the container's layout
    (DATA/FLAG channel-major) are wrapped as a dataset -- pinned host memory
    (copied to the device directly) or pageable (copied by the loader's reader
    threads into its pinned staging buffers first) -- and streamed in blocks
    of `group` channels on the loader's copy stream while the compute stream
    grids the previous block (Briggs densities are per channel, so blocks are
    independent); the row columns are copied, sorted and kept once.  Timed:
    kernel tables, every block's copy and gridding, the finalize.  Returns the
    host-inclusive rate and how far the images are from the HBM-resident run's."""
    from mtimager_amd import ingest
    nrows, nchan, P = d["nrows"], d["nchan"], d["P"]
    kind = "pinned" if pinned else "pageable"
    log(f"end-to-end ({kind}): staging {nchan} channels in host memory")
    host = (lambda t: t.pin_memory()) if pinned else (lambda t: t)
    vis_h = torch.empty((nchan, nrows, P), dtype=torch.complex64, pin_memory=pinned)
    flg_h = torch.empty((nchan, nrows, P), dtype=torch.uint8, pin_memory=pinned)
    for c0 in range(0, nchan, group):
        vis_h[c0:c0 + group].copy_(d["vis"][:, c0:c0 + group].permute(1, 0, 2).contiguous())
        flg_h[c0:c0 + group].copy_(d["flags"][:, c0:c0 + group].permute(1, 0, 2).contiguous())
    cols = dict(time=host(torch.zeros(nrows, dtype=torch.float64)), a1=host(d["a1"].cpu()),
                a2=host(d["a2"].cpu()), uvw=host(d["uvw"].cpu()), weight=host(d["weight"].cpu()),
                freq=host((d["invl"] * 299792458.0).cpu()))
    ds = ingest.Dataset.wrap(cols["time"], cols["a1"], cols["a2"], cols["uvw"], cols["weight"], cols["freq"],
                             vis_h, flg_h, d["nant"], pinned=pinned)
    ld = ingest.Loader(ctx, ds, channel_block=group, chunk_rows=0, weighting=d["weighting"], robust=0.0,
                       layout="pvu", precision="c64" if grids.dtype == torch.float32 else "c128",
                       reader_threads=reader_threads)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    K = ctx.build_kernels(p)
    st = ld.grid(K, p, grids, norm, tile_mask=tmask)
    ctx.finalize(p, P, grids, norm, image, nchan=nchan, layout="pvu", tile_mask=tmask)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    K.free()
    ld.close()
    ds.close()
    diff = float((image - ref_image).abs().max()) / max(float(ref_image.abs().max()), 1e-30)
    del vis_h, flg_h, cols
    return dict(host_memory=kind, seconds=round(el, 4), h2d_bytes=int(st["h2d_bytes"]),
                h2d_share_GBps=round(st["h2d_bytes"] / el / 1e9, 1), channel_block=group, chunks=st["chunks"],
                reader_seconds=round(st["reader_seconds"], 3), max_rel_diff_vs_resident=diff,
                path="mtimager_amd.ingest.Loader")


def max_rel_rms(image, ref, nchan):
    """Largest relative RMS over channels and Stokes images, each relative to
    the RMS of the channel's reference Stokes I."""
    errs = []
    for c in range(nchan):
        dd = image[c].double() - ref[c].double()
        ri = ref[c, 0].double().pow(2).mean().sqrt()
        errs.append(float((dd.pow(2).mean(dim=(1, 2)).sqrt() / ri).max()))
    return max(errs)


def precision_forms(torch, mti, ctx, d, p, grids, norm, image, ref, tmask, nchan):
    """The precision forms on the cube: kernel tables in f16 / bf16 and the
    gridder's FP32 compensated (Kahan) accumulation, each imaged once with
    stage timing; gridder milliseconds and the image error against the
    complex128 FP64 path (ref)."""
    P = d["P"]
    out = {}
    idx = torch.empty(d["nrows"], dtype=torch.int32, device="cuda")
    for name, table, acc in (("f32 tables, FP64 sums", "f32", "f64"), ("f16 tables, FP64 sums", "f16", "f64"),
                             ("bf16 tables, FP64 sums", "bf16", "f64"),
                             ("f32 tables, FP32 Kahan sums", "f32", "f32k"),
                             ("f16 tables, FP32 Kahan sums", "f16", "f32k"),
                             ("bf16 tables, FP32 Kahan sums", "bf16", "f32k")):
        kept = ctx.sort_rows(d["a1"], d["a2"], d["nant"], idx)
        K = ctx.build_kernels(p, precision=table)
        v = mti.vis_ms(d["uvw"], d["invl"], d["vis"], d["flags"], d["weight"], idx[:kept])
        torch.cuda.synchronize()
        ctx.stage_times()
        ctx.set_profiling(True)
        ctx.grid(K, p, v, grids, norm, chan0=0, nchan=nchan, weighting=d["weighting"], robust=0.0, layout="pvu",
                 stats=False, tile_mask=tmask, accumulation=acc)
        ctx.finalize(p, P, grids, norm, image, nchan=nchan, layout="pvu", tile_mask=tmask)
        torch.cuda.synchronize()
        ctx.set_profiling(False)
        stg = ctx.stage_times()
        K.free()
        err = max_rel_rms(image, ref, nchan)
        out[name] = dict(grid_ms=round(stg["grid"][0], 3), job_ms=round(sum(v[0] for v in stg.values()), 3),
                         image_max_rel_rms_vs_fp64=float(f"{err:.3e}"))
        log(f"precision form {name}: {out[name]}")
    return dict(forms=out, vs="complex128 grids and transform planes, f32 tables, FP64 sums",
                over="every channel and Stokes image, relative to the channel's Stokes I RMS")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_full(d, params_args, threads):
    """FP64 oracle on the whole job (small workloads): row sort, W-kernel
    tables, then per channel the workload's weights, gridding of every row and
    the finalize -- once on one thread, then with the channels spread over
    `threads` threads (the oracle is reentrant C; ctypes releases the
    interpreter lock).  Returns seconds and gridded records."""
    import numpy as np
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from mtimager_amd import synth
    h = d["host"]
    vis = d["vis"].cpu().numpy()
    flags = d["flags"].cpu().numpy()
    weight = d["weight"].cpu().numpy()
    nx, cell, W, s = params_args
    nchan = d["nchan"]

    def job(nthreads):
        t0 = time.perf_counter()
        rows = oracle.sort_index(h["a1"], h["a2"], d["nant"])
        p = oracle.make_params(nx, nx, cell * synth.ARCSEC, cell * synth.ARCSEC, "wproj", W, s)
        K = oracle.Kernels(p)
        uvw, vr, fr, wr = (np.ascontiguousarray(a[rows]) for a in (h["uvw"], vis, flags, weight))
        t_k = time.perf_counter() - t0

        def one(chan):
            wts = None if d["weighting"] == "natural" else oracle.weights(K, uvw, h["inv_lambda"], fr, wr, chan,
                                                                          d["weighting"], 0.0)
            g, nrm, ng = oracle.grid(K, uvw, h["inv_lambda"], vr, fr, wr, chan, wts=wts)
            oracle.finalize(p, g, nrm)
            return ng

        with ThreadPoolExecutor(max_workers=nthreads) as ex:
            ng = sum(ex.map(one, range(nchan)))
        return time.perf_counter() - t0, t_k, ng

    t1, k1, ng1 = job(1)
    tN, kN, ngN = job(threads)
    return dict(seconds_1=t1, seconds_n=tN, kernels_s=k1, gridded=ng1, threads=threads)


def cpu_baseline(d, params_args, budget_rows, threads):
    """FP64 oracle on a bounded sample: per channel, weights (the workload's) + gridding
    of `budget_rows` baseline-major rows and one finalize.  Timed on one
    thread, then on `threads` threads with one channel each (the oracle is
    reentrant C; ctypes releases the interpreter lock), and extrapolated to
    the whole cube."""
    import numpy as np
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from mtimager_amd import synth
    h = d["host"]
    rows = oracle.sort_index(h["a1"], h["a2"], d["nant"])
    sample = rows[:budget_rows]
    nx, cell, W, s = params_args
    p = oracle.make_params(nx, nx, cell * synth.ARCSEC, cell * synth.ARCSEC, "wproj", W, s)
    K = oracle.Kernels(p)
    n, P, nchan = len(sample), d["P"], d["nchan"]
    rng = np.random.default_rng(0)
    vis = (rng.standard_normal((n, 1, P)) + 1j * rng.standard_normal((n, 1, P))).astype(np.complex64)
    flags = np.zeros((n, 1, P), np.uint8)
    weight = np.ones((n, P), np.float32)
    uvw = np.ascontiguousarray(h["uvw"][sample])

    def one(chan):
        il = np.array([h["inv_lambda"][chan]])
        t0 = time.perf_counter()
        wts = None if d["weighting"] == "natural" else oracle.weights(K, uvw, il, flags, weight, 0, d["weighting"], 0.0)
        g, nrm, ng = oracle.grid(K, uvw, il, vis, flags, weight, 0, wts=wts)
        t1 = time.perf_counter()
        oracle.finalize(p, g, nrm)
        return t1 - t0, time.perf_counter() - t1, ng

    tg1, tf1, ng1 = one(nchan // 2)
    chans = [int(c) for c in np.linspace(0, nchan - 1, threads)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        rs = list(ex.map(one, chans))
    wall = time.perf_counter() - t0
    tgN = float(np.mean([r[0] / max(r[2], 1) for r in rs]))  # s per gridded record, under contention
    tfN = float(np.mean([r[1] for r in rs]))
    return dict(per_record_s=tg1 / max(ng1, 1), finalize_s=tf1, sample_rows=n, sample_gridded=ng1,
                per_record_s_threads=tgN, finalize_s_threads=tfN, threads=threads, wall_threads_s=wall)


def snapshot_pool(torch, mti, njobs, seed=4):
    """The snapshot-job mix (Stokes I, 10k-5M visibilities log-uniform,
    256^2-2048^2, 8-64 w-planes, s 4/8) with its inputs resident on the GPU:
    visibilities predicted from 20-source skies, 5% random flags, weights 1."""
    from mtimager_amd import synth
    specs = synth.snapshot_jobs(njobs, seed=seed)
    log(f"{len(specs)} snapshot jobs, predicting visibilities on the GPU")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    jobs = []
    for sp in specs:
        nrows, nch = len(sp["uvw"]), sp["nchan"]
        vis = synth.predict_torch(sp["uvw"], sp["inv_lambda"], sp["src"], 1)
        flags = (torch.rand((nrows, nch), generator=g, device="cuda") < 0.05).to(torch.uint8)[:, :, None].contiguous()
        jobs.append(dict(
            p=mti.make_params(sp["nx"], sp["nx"], sp["cell"], sp["cell"], "wproj", sp["wplanes"], sp["sampling"]),
            key=(sp["nx"], sp["cell"], sp["wplanes"], sp["sampling"]), nch=nch, nx=sp["nx"], nrows=nrows,
            uvw=torch.tensor(sp["uvw"], dtype=torch.float64, device="cuda"),
            invl=torch.tensor(sp["inv_lambda"], dtype=torch.float64, device="cuda"),
            vis=vis, flags=flags, weight=torch.ones((nrows, 1), dtype=torch.float32, device="cuda"),
            a1=torch.tensor(sp["ant1"], device="cuda"), a2=torch.tensor(sp["ant2"], device="cuda")))
    torch.cuda.synchronize()
    return jobs


def graph_replay(ctx, torch, mti, jobs, outputs, max_jobs=64, reps=7):
    """A wave of the mix's small jobs (images up to 512^2, at most 3e5
    visibilities) imaged eagerly (mti_grid_jobs + mti_finalize_jobs, with
    their host synchronisations) and as a captured HIP graph; median wall
    time per wave of each, and whether the replay equals the eager images."""
    import numpy as np
    small = [j for j in jobs if j["nx"] <= 512 and j["nrows"] * j["nch"] <= 300000][:max_jobs]
    if not small:
        return None
    kc = {}
    for j in small:
        if j["key"] not in kc:
            kc[j["key"]] = ctx.build_kernels(j["p"])

    def make():
        out, js = [], []
        for j in small:
            gr, nr, im, mk = outputs(j)
            v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"])
            js.append(ctx.make_job(kc[j["key"]], j["p"], v, gr, nr, im, 0, j["nch"],
                                   ant1=j["a1"], ant2=j["a2"], nant=64, tile_mask=mk))
            out.append(im)
        return js, out

    ejobs, eimg = make()
    gjobs, gimg = make()

    def eager():
        ctx.grid_jobs(ejobs, stats=False)
        ctx.finalize_jobs(ejobs)

    g = ctx.capture_jobs(gjobs)

    def timed(fn):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    te = timed(eager)
    tg = timed(g.launch)
    same = all(float((a - b).abs().max()) <= 1e-6 * max(float(b.abs().max()), 1e-30) for a, b in zip(gimg, eimg))
    g.free()
    for K in kc.values():
        K.free()
    nvis = sum(j["nrows"] * j["nch"] for j in small)
    log(f"c4: wave of {len(small)} small jobs: eager {te:.2f} ms, graph {tg:.2f} ms")
    return dict(jobs=len(small), visibilities=int(nvis), eager_ms=round(te, 3), graph_ms=round(tg, 3),
                speedup=round(te / tg, 3), replay_equals_eager=bool(same))


def run_c4(args, torch, mti, jobs):
    """The snapshot mix: independent jobs (Stokes I) imaged one at a time
    (plain: per job kernel tables, row sort, gridding, finalize) and batched
    (kernel-table cache, jobs admitted in waves by an HBM budget, one
    mti_grid_jobs + one mti_finalize_jobs call per wave, waves on two workers
    with a context and stream each).  Returns a dict."""
    import numpy as np
    ctx = mti.Context(0)
    ctx.set_transform_storage(args.transform_storage)
    budget = args.c4_budget_gb * (1 << 30)

    def prep(j):
        idx = torch.empty(j["nrows"], dtype=torch.int32, device="cuda")
        kept = ctx.sort_rows(j["a1"], j["a2"], 64, idx)
        return mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"], idx[:kept])

    def outputs(j):
        # grids are intermediates here: sparse (tile mask) unless asked otherwise
        nx, nch = j["nx"], j["nch"]
        mask = None if args.complete_grids else \
            torch.empty(mti.tile_mask_shape(j["p"], nch), dtype=torch.uint8, device="cuda")
        gdt = torch.float32 if args.grid_precision == "c64" else torch.float64
        return (torch.empty((nch, nx, nx, 1, 2), dtype=gdt, device="cuda"),
                torch.empty((nch, 1), dtype=torch.float64, device="cuda"),
                torch.empty((nch, 1, nx, nx), dtype=torch.float32, device="cuda"), mask)

    def plain():
        lat, gridded = [], 0
        t0 = time.perf_counter()
        for n, j in enumerate(jobs):
            if n % 64 == 0:
                log(f"c4: plain job {n}")
            v = prep(j)
            K = ctx.build_kernels(j["p"])
            gr, nr, im, mk = outputs(j)
            st = ctx.grid(K, j["p"], v, gr, nr, chan0=0, nchan=j["nch"], stats=True,
                          tile_mask=mk)
            ctx.finalize(j["p"], 1, gr, nr, im, nchan=j["nch"], tile_mask=mk)
            torch.cuda.synchronize()
            K.free()
            gridded += st.gridded_records
            lat.append(time.perf_counter() - t0)
        return time.perf_counter() - t0, gridded, lat

    # two wave workers, each with its own context and stream: one wave's host
    # work (job tables, pack descriptors, the calls' synchronisations)
    # overlaps the other wave's kernels
    wctx = [mti.Context(0) for _ in range(2)]
    for c in wctx:
        c.set_transform_storage(args.transform_storage)
    wstreams = [torch.cuda.Stream() for _ in wctx]

    def batched():
        """Waves cut by the HBM budget in arrival order (kernel tables from a
        cache built as jobs arrive), each one mti_grid_jobs + one
        mti_finalize_jobs call on whichever worker is free."""
        import queue
        import threading
        for c in wctx:
            c.set_tuning(**ctx.get_tuning())
        lat, cache, errors = [], {}, []
        gridded = [0]
        lock = threading.Lock()
        # waves dealt in turn (wave k to worker k mod 2): each worker sees the
        # same waves in every run, so its stream's cached memory fits them
        qs = [queue.Queue(maxsize=2) for _ in wctx]
        t0 = time.perf_counter()

        def worker(c, s, q):
            with torch.cuda.stream(s):
                while True:
                    wave = q.get()
                    if wave is None:
                        return
                    if errors:
                        continue
                    try:
                        items = []
                        for j in wave:
                            # the batched call puts each job's rows in baseline-major order itself
                            v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"])
                            gr, nr, im, mk = outputs(j)
                            items.append((c.make_job(cache[j["key"]], j["p"], v, gr, nr, im, 0, j["nch"],
                                                     ant1=j["a1"], ant2=j["a2"], nant=64, tile_mask=mk),
                                          gr, nr, im, mk))
                        js = [it[0] for it in items]
                        st = c.grid_jobs(js)  # on the thread's current stream, s
                        c.finalize_jobs(js)
                        s.synchronize()
                        with lock:
                            gridded[0] += st.gridded_records
                            lat.extend([time.perf_counter() - t0] * len(wave))
                        log(f"c4: wave of {len(js)} jobs done")
                        del items, js
                    except Exception as e:  # reported by the main thread
                        errors.append(e)

        ths = [threading.Thread(target=worker, args=(c, s, q)) for c, s, q in zip(wctx, wstreams, qs)]
        for t in ths:
            t.start()
        wave, mem, nw = [], 0, 0
        for j in jobs:
            if j["key"] not in cache:
                cache[j["key"]] = ctx.build_kernels(j["p"])  # complete on return
            need = j["nch"] * j["nx"] ** 2 * 20
            if wave and mem + need > budget:
                qs[nw % len(qs)].put(wave)
                nw += 1
                wave, mem = [], 0
            wave.append(j)
            mem += need
        if wave:
            qs[nw % len(qs)].put(wave)
        for q in qs:
            q.put(None)
        for t in ths:
            t.join()
        if errors:
            raise errors[0]
        el = time.perf_counter() - t0
        for K in cache.values():
            K.free()
        return el, gridded[0], lat, len(cache)

    tune_info = None
    if not args.no_autotune:
        # the project's autotuner picks the gridder's work-unit size and the
        # transform's column strategy on a sample of the mix (every 8th job,
        # up to half a wave's budget, imaged as one wave); plain and batched
        # runs below both use the choice
        from mtimager_amd import autotune
        sample, mem = [], 0
        for j in jobs[::8]:
            need = j["nch"] * j["nx"] ** 2 * 20
            if mem + need <= budget / 2:
                sample.append(j)
                mem += need
        kc = {}
        tjobs = []
        for j in sample:
            if j["key"] not in kc:
                kc[j["key"]] = ctx.build_kernels(j["p"])
            gr, nr, im, mk = outputs(j)
            v = mti.vis_ms(j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"])
            tjobs.append(ctx.make_job(kc[j["key"]], j["p"], v, gr, nr, im, 0, j["nch"],
                                      ant1=j["a1"], ant2=j["a2"], nant=64, tile_mask=mk))

        def tune_run():
            ctx.grid_jobs(tjobs, stats=False)
            ctx.finalize_jobs(tjobs)

        log(f"c4: autotuning on {len(sample)} jobs")
        best, timings = autotune.tune(ctx, tune_run, repeats=2)
        log(f"c4: tuning {best}")
        tune_info = dict(sample_jobs=len(sample), strategy="coordinate", chosen=best,
                         candidates_ms=[dict(c, ms=round(t * 1e3, 2)) for c, t in timings])
        del tjobs
        for K in kc.values():
            K.free()
        torch.cuda.synchronize()
    # memory the allocator keeps for the default stream is of no use to the
    # workers' streams
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log("c4: warm-up (batched)")
    batched()
    log("c4: plain, one job at a time")
    tp, gp, lp = plain()
    log(f"c4: plain {tp:.2f} s")
    tb, gb, lb, nk = batched()
    log(f"c4: batched {tb:.2f} s ({nk} kernel-table sets)")
    graph_info = graph_replay(ctx, torch, mti, jobs, outputs)
    for c in wctx:
        c.close()
    ctx.close()
    assert gp == gb, (gp, gb)
    lp, lb = np.array(lp) * 1e3, np.array(lb) * 1e3
    offered = sum(j["nrows"] * j["nch"] for j in jobs)
    return dict(jobs=len(jobs), offered_vis=int(offered), gridded_vis=int(gb),
                gridded_fraction=round(gb / max(offered, 1), 4),
                field=("one 0.85 deg field for every job (cell = 1.5 arcsec x 2048 / nx): on the smaller "
                       "grids the longest baselines fall outside the uv plane and are not gridded"),
                kernel_sets=nk, hbm_budget_gb=args.c4_budget_gb,
                transform_storage=args.transform_storage, grid_precision=args.grid_precision,
                autotune=tune_info, graph=graph_info,
                plain=dict(mvis_s=round(gp / tp / 1e6, 3), seconds=round(tp, 3),
                           latency_ms={"p50": round(float(np.percentile(lp, 50)), 1),
                                       "p99": round(float(np.percentile(lp, 99)), 1)}),
                batched=dict(mvis_s=round(gb / tb / 1e6, 3), seconds=round(tb, 3),
                             latency_ms={"p50": round(float(np.percentile(lb, 50)), 1),
                                         "p99": round(float(np.percentile(lb, 99)), 1)}),
                speedup=round(tp / tb, 3))


def _pct(a, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(a) * 1e3, q)), 1) if len(a) else None


def _serve_feed(srv, pool, picks, times, deadline_s):
    """Feeds requests (pool entries `picks`) into a running Server at their
    arrival times; waits until they are done or the deadline passes, then
    drops what was never admitted.  Returns the figures of this stream."""
    from mtimager_amd import serve
    reqs = []
    for i, k in enumerate(picks):
        j = pool[k]
        reqs.append(serve.Request(i, j["p"], (j["uvw"], j["invl"], j["vis"], j["flags"], j["weight"]),
                                  j["a1"], j["a2"], 64, j["nch"], npol=1, keep_image=False))
    g0, w0 = srv.gridded, len(srv.wave_sizes)
    t0 = time.perf_counter()
    for r, at in zip(reqs, times):
        dt = t0 + at - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        srv.submit(r)
    finished = srv.wait(timeout=max(1.0, t0 + deadline_s - time.perf_counter()))
    t_end = time.perf_counter()
    if not finished:
        srv.cancel_queued()
        srv.wait()
    if srv.errors:
        raise srv.errors[0]
    done = [r for r in reqs if r.done is not None and r.error is None]
    lat = [r.done - r.arrival for r in done]
    # requests dropped at the deadline count with their waiting time until then
    lat_all = lat + [t_end - r.arrival for r in reqs if r.done is None or r.error is not None]
    last = max((r.done for r in done), default=t_end)
    sizes = srv.wave_sizes[w0:]
    return dict(requests=len(reqs), completed=len(done), saturated=not finished,
                mvis_s=round((srv.gridded - g0) / max(last - t0, 1e-9) / 1e6, 3),
                latency_ms={"p50": _pct(lat_all, 50), "p99": _pct(lat_all, 99)},
                waves=len(sizes), mean_wave=round(float(sum(sizes)) / max(1, len(sizes)), 2))


def run_c5(args, torch, mti, pool):
    """Serving: Poisson streams of imaging requests drawn from the snapshot
    mix, served by continuous batching (two workers, HBM-budget admission,
    device-resident kernel-table cache) and by the plain one-job-at-a-time
    path on the same arrivals.  Both servers are long-lived (warm contexts,
    streams and memory pools) and see the rates in turn.  Inputs are resident
    on the GPU; each request images one pool job.  A stream that has not
    drained by its deadline reports saturated=true; requests never admitted
    count with their waiting time until the deadline."""
    import numpy as np
    from mtimager_amd import serve
    budget = int(args.serve_budget_gb * (1 << 30))
    cache = serve.KernelCache()
    served = serve.Server(budget, workers=2, max_wave=512, cache=cache, transform_storage=args.transform_storage,
                          grid_precision=args.grid_precision)
    plain = serve.Server(budget, plain=True, transform_storage=args.transform_storage,
                         grid_precision=args.grid_precision)
    rng = np.random.default_rng(5)
    out = []
    try:
        log("c5: warm-up (every kernel-table set into the cache, memory pools, code objects)")
        allj = np.arange(len(pool))
        _serve_feed(served, pool, allj, np.zeros(len(pool)), 600)
        _serve_feed(plain, pool, allj, np.zeros(len(pool)), 600)
        for rate in [float(x) for x in args.rates.split(",")]:
            n = int(rng.poisson(rate * args.serve_seconds))
            times = np.sort(rng.uniform(0.0, args.serve_seconds, n))
            picks = rng.integers(0, len(pool), n)
            log(f"c5: {rate:g} req/s, {n} requests, served")
            sv = _serve_feed(served, pool, picks, times, args.serve_seconds * 3)
            log(f"c5: {rate:g} req/s served {sv}")
            pl = _serve_feed(plain, pool, picks, times, args.serve_seconds * 3)
            log(f"c5: {rate:g} req/s plain {pl}")
            out.append(dict(rate=rate, served=sv, plain=pl,
                            speedup_mvis=round(sv["mvis_s"] / max(pl["mvis_s"], 1e-9), 3)))
    finally:
        served.close()
        plain.close()
    res = dict(pool=len(pool), seconds=args.serve_seconds, kernel_sets=len(cache),
               cache_hits=cache.hits, cache_misses=cache.misses, rates=out)
    cache.clear()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c2", "c1", "c4", "c5"])
    ap.add_argument("--rates", default="50,200,800", help="serving: arrival rates (req/s)")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="serving: arrival window per rate")
    ap.add_argument("--serve-budget-gb", type=float, default=160.0,
                    help="serving: HBM admission budget (the rest holds the inputs, the library's "
                         "scratch and allocator slack)")
    ap.add_argument("--jobs", type=int, default=512, help="snapshot mix: number of snapshot jobs")
    ap.add_argument("--c4-budget-gb", type=float, default=64.0, help="snapshot mix: HBM budget of a batched wave")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-c4", action="store_true",
                    help="cube runs: skip the nested snapshot-mix object (batched vs plain)")
    ap.add_argument("--no-c5", action="store_true",
                    help="cube runs: skip the nested serving object (served vs plain)")
    ap.add_argument("--c5-rate", type=float, default=800.0, help="cube runs: arrival rate of the nested serving run")
    ap.add_argument("--c5-seconds", type=float, default=10.0,
                    help="cube runs: arrival window of the nested serving run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="snapshot mix: keep the default tuning")
    ap.add_argument("--no-e2e", action="store_true", help="cube runs: skip the pinned-host end-to-end run")
    ap.add_argument("--complete-grids", action="store_true",
                    help="zero-fill and transform every grid tile (default: sparse grids with a tile mask)")
    ap.add_argument("--xcd-chunk", type=int, default=None, help="cube runs: gridder units per XCD run")
    ap.add_argument("--unit-pairs", type=int, default=None, help="cube runs: gridder work-unit size")
    ap.add_argument("--region-width", type=int, default=None,
                    help="cube runs: gridder wave regions of W x 64/W cells (4, 8 or 16)")
    ap.add_argument("--transform-storage", choices=["c128", "c64"], default="c64",
                    help="element type of the packed planes between the finalize's transform passes")
    ap.add_argument("--grid-precision", choices=["c128", "c64"], default="c64",
                    help="element type of the grids the gridder writes")
    ap.add_argument("--no-form-check", action="store_true",
                    help="cube runs: skip imaging the job again on the complex128 path")
    ap.add_argument("--no-precision-forms", action="store_true",
                    help="cube runs: skip timing the reduced-precision table / FP32 accumulation forms")
    ap.add_argument("--cpu-rows", type=int, default=60000)
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the CPU baseline (one channel each)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one elementwise negation over this many MiB before the timed region "
                         "(a kernel of known HBM traffic for calibrating FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(channels and jobs would shard one per GPU).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import mtimager_amd as mti

    torch.cuda.set_device(0)
    if args.config == "c5":
        r = run_c5(args, torch, mti, snapshot_pool(torch, mti, args.jobs))
        print(json.dumps({
            "metric": METRIC, "value": r["rates"][-1]["served"]["mvis_s"], "unit": "Mvis/s", "n_gpus": 1,
            "steps": len(r["rates"]), "warmup": 1, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"serving: Poisson arrivals at {args.rates} req/s for "
                                   f"{args.serve_seconds} s from a pool of {r['pool']} snapshot jobs, "
                                   f"HBM budget {args.serve_budget_gb} GB",
                       "path": "served vs plain"},
            "latency_ms": r["rates"][-1]["served"]["latency_ms"], "c5": r}), flush=True)
        return
    if args.config == "c4":
        r = run_c4(args, torch, mti, snapshot_pool(torch, mti, args.jobs))
        print(json.dumps({
            "metric": METRIC, "value": r["batched"]["mvis_s"], "unit": "Mvis/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": round(r["batched"]["seconds"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"batched snapshot imaging: {r['jobs']} jobs, Stokes I, "
                                   "10k-5M vis, 256^2-2048^2, 8-64 w-planes, s 4/8",
                       "path": "batched vs plain"},
            "latency_ms": r["batched"]["latency_ms"], "c4": r}), flush=True)
        return
    d = make_workload(torch, args.config, args.seed)
    p = mti.make_params(d["nx"], d["nx"], d["cell"], d["cell"], "wproj", d["W"], d["s"])
    ctx = mti.Context(0)
    ctx.set_tuning(unit_pairs=args.unit_pairs, xcd_chunk=args.xcd_chunk, region_width=args.region_width)
    ctx.set_transform_storage(args.transform_storage)
    nrows, nchan, P, nx = d["nrows"], d["nchan"], d["P"], d["nx"]
    grids = torch.empty((nchan, P, nx, nx, 2), dtype=torch.float32 if args.grid_precision == "c64" else torch.float64,
                        device="cuda")
    norm = torch.empty((nchan, P), dtype=torch.float64, device="cuda")
    image = torch.empty((nchan, P, nx, nx), dtype=torch.float32, device="cuda")
    idx = torch.empty(nrows, dtype=torch.int32, device="cuda")
    # the grids are intermediates: sparse (tiles no record reaches are neither
    # zero-filled nor read by the transform) unless --complete-grids
    tmask = None if args.complete_grids else \
        torch.empty(mti.tile_mask_shape(p, nchan), dtype=torch.uint8, device="cuda")

    def step(stats=False):
        kept = ctx.sort_rows(d["a1"], d["a2"], d["nant"], idx)
        K = ctx.build_kernels(p)
        v = mti.vis_ms(d["uvw"], d["invl"], d["vis"], d["flags"], d["weight"], idx[:kept])
        st = ctx.grid(K, p, v, grids, norm, chan0=0, nchan=nchan, weighting=d["weighting"], robust=0.0,
                      layout="pvu", stats=stats, tile_mask=tmask)
        ctx.finalize(p, P, grids, norm, image, nchan=nchan, layout="pvu", tile_mask=tmask)
        K.free()
        return st, K

    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log(f"copy probe: {n * 8} bytes read, {n * 8} bytes written (neg kernel)")
        del src, dst
    peaks = achievable_peaks()
    log(f"achievable peaks: {peaks}")
    log("warmup")
    st, _ = step(stats=True)
    torch.cuda.synchronize()
    gridded = int(st.gridded_records)
    points = int(st.gridded_points)
    Kinfo = ctx.build_kernels(p)
    sup, _ = Kinfo.supports()
    kinfo = dict(max_support=int(Kinfo.info.max_support), trial_support=int(Kinfo.info.trial_support))
    Kinfo.free()
    log(f"gridded {gridded} records, {points} points, items {st.work_items}")
    for _ in range(args.warmup - 1):
        step()
    torch.cuda.synchronize()
    ctx.stage_times()  # discard warmup events
    ctx.set_profiling(True)
    times = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"step {i}: {times[-1] * 1e3:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_all
    ctx.set_profiling(False)
    stages = ctx.stage_times()
    # the transform's row pass (k_rows_pk): its algorithmic bytes are the grid
    # cells of marked tiles it reads (every correlation plane of the channel)
    # and the packed planes it writes
    rows_pass = None
    if stages.get("fft_rows", (0, 0))[1] > 0:
        marked = 1.0 if tmask is None else float(tmask.float().mean().item())
        gsz = 8 if args.grid_precision == "c64" else 16
        xsz = 8 if args.transform_storage == "c64" else 16
        npk = nchan * P / 2.0
        rd, wr = marked * nchan * P * nx * nx * gsz, npk * nx * nx * xsz
        rms = stages["fft_rows"][0] / args.steps
        rows_pass = dict(kernel="k_rows_pk", tiles_marked=round(marked, 4), packs=npk, read_GB=round(rd / 1e9, 3),
                         write_GB=round(wr / 1e9, 3), ms_per_step=round(rms, 3),
                         algorithmic_GBps=round((rd + wr) / (rms * 1e-3) / 1e9, 1))
    # the timed job's images: the reference of the end-to-end runs (the
    # precision-form runs below reuse the image buffer)
    resident_image = None if args.no_e2e else image.clone()
    ms_per_step = elapsed / args.steps * 1e3
    value = gridded * args.steps / elapsed / 1e6

    g_ms, g_n = stages["grid"]
    grid_avg_s = g_ms / max(g_n, 1) / 1e3
    # gridder kernel: one complex multiply-add per tap per correlation, FP64
    flops = 8.0 * points
    achieved = flops / grid_avg_s / 1e12 if grid_avg_s > 0 else 0.0
    # algorithmic bytes of the gridded records (complex64 vis + u8 flags per
    # correlation, f32 weights per row-correlation + f64 uvw per row shared by
    # the channels of a row)
    rec_bytes = gridded * (9.0 * P + (4.0 * P + 24.0) / nchan)
    roofline = dict(bound="valu", kernel="k_grid_tiles", achieved=round(achieved, 3),
                    peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=round(achieved / FP64_PEAK_TFLOPS, 4),
                    traffic=measured_traffic(args.config), flops_per_launch=flops,
                    avg_launch_ms=round(grid_avg_s * 1e3, 3),
                    algorithmic_GBps=round(rec_bytes / grid_avg_s / 1e9, 1),
                    achievable=peaks,
                    frac_of_achievable=(round(achieved / peaks["fp64_fma_TFLOPs"], 4) if peaks else None))

    form_err = None
    forms = None
    if (args.grid_precision != "c128" or args.transform_storage != "c128") and not args.no_form_check:
        # the last step's job imaged again on the strict FP64 path (complex128
        # grids and transform planes, the form the parity tests hold to the
        # oracle): the difference the precision form makes, per channel and
        # Stokes image relative to the RMS of the channel's Stokes I
        log("imaging the job again with complex128 grids and planes")
        g128 = torch.empty((nchan, P, nx, nx, 2), dtype=torch.float64, device="cuda")
        n128 = torch.empty_like(norm)
        m128 = None if tmask is None else torch.empty_like(tmask)
        ref = torch.empty_like(image)
        ctx.set_transform_storage("c128")
        kept = ctx.sort_rows(d["a1"], d["a2"], d["nant"], idx)
        K = ctx.build_kernels(p)
        v = mti.vis_ms(d["uvw"], d["invl"], d["vis"], d["flags"], d["weight"], idx[:kept])
        ctx.grid(K, p, v, g128, n128, chan0=0, nchan=nchan, weighting=d["weighting"], robust=0.0, layout="pvu",
                 stats=False, tile_mask=m128)
        ctx.finalize(p, P, g128, n128, ref, nchan=nchan, layout="pvu", tile_mask=m128)
        torch.cuda.synchronize()
        K.free()
        ctx.set_transform_storage(args.transform_storage)
        del g128, n128, m128
        form_err = max_rel_rms(image, ref, nchan)
        log(f"precision form (grids {args.grid_precision}, planes {args.transform_storage}): max relative RMS "
            f"vs the complex128 path {form_err:.3e}")
        if args.config == "c2" and not args.no_precision_forms:
            forms = precision_forms(torch, mti, ctx, d, p, grids, norm, image, ref, tmask, nchan)
        del ref

    e2e = None
    if not args.no_e2e:
        ref_image = resident_image
        e2e = end_to_end(torch, mti, ctx, d, p, grids, norm, image, ref_image, tmask=tmask)
        e2e["mvis_s"] = round(gridded / e2e["seconds"] / 1e6, 3)
        log(f"end-to-end: {e2e}")
        pg = end_to_end(torch, mti, ctx, d, p, grids, norm, image, ref_image, tmask=tmask, pinned=False)
        pg["mvis_s"] = round(gridded / pg["seconds"] / 1e6, 3)
        log(f"end-to-end from pageable memory: {pg}")
        e2e["pageable"] = pg
        del ref_image, resident_image

    cpu = None
    if not args.no_cpu_baseline and args.config == "c1":
        thr = max(1, min(args.cpu_threads, os.cpu_count() or 1, nchan))
        log(f"cpu baseline (FP64 oracle, whole job, 1 and {thr} threads)")
        cb = cpu_baseline_full(d, (nx, d["cell"], d["W"], d["s"]), thr)
        cpu = dict(value=round(cb["gridded"] / cb["seconds_n"] / 1e6, 4), unit="Mvis/s", cores=thr, kind="port",
                   single_thread_value=round(cb["gridded"] / cb["seconds_1"] / 1e6, 4), cpu_model=cpu_model(),
                   host_cpus=os.cpu_count(),
                   sample=(f"FP64 oracle (gcc -O3 -march=x86-64-v3, no FP contraction), the whole job timed, "
                           f"no extrapolation: row sort, W-kernel tables ({cb['kernels_s']:.2f} s), then per "
                           f"channel {d['weighting']} weighting, gridding of every row and the {nx}^2 x {P} "
                           f"finalize; {cb['gridded']} gridded records; 1 thread {cb['seconds_1']:.2f} s, "
                           f"{thr} threads (channels spread over them) {cb['seconds_n']:.2f} s"))
    elif not args.no_cpu_baseline:
        thr = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        log(f"cpu baseline (FP64 oracle, 1 and {thr} threads)")
        cb = cpu_baseline(d, (nx, d["cell"], d["W"], d["s"]), args.cpu_rows, thr)
        per_ch_rows = gridded / nchan
        t_job1 = nchan * (per_ch_rows * cb["per_record_s"] + cb["finalize_s"])
        t_jobN = math.ceil(nchan / thr) * (per_ch_rows * cb["per_record_s_threads"] + cb["finalize_s_threads"])
        cpu = dict(value=round(gridded / t_jobN / 1e6, 4), unit="Mvis/s", cores=thr, kind="port",
                   single_thread_value=round(gridded / t_job1 / 1e6, 4), cpu_model=cpu_model(),
                   host_cpus=os.cpu_count(),
                   sample=(f"FP64 oracle: per channel, {d['weighting']} weights + gridding of {cb['sample_rows']} "
                           f"baseline-major rows ({cb['sample_gridded']} gridded) and one {nx}^2 x {P} "
                           f"finalize; 1 thread: {cb['per_record_s'] * 1e6:.2f} us/record + "
                           f"{cb['finalize_s']:.1f} s finalize; {thr} threads (one channel each, "
                           f"{cb['wall_threads_s']:.1f} s wall): {cb['per_record_s_threads'] * 1e6:.2f} "
                           f"us/record + {cb['finalize_s_threads']:.1f} s finalize per thread; "
                           f"extrapolated to {nchan} channels; W-kernel generation not counted"))

    lat = np.array(times) * 1e3
    tuning = ctx.get_tuning()
    nested = {}
    if args.config == "c2" and not (args.no_c4 and args.no_c5):
        # the metric's "batched/served vs plain": the snapshot mix (batched
        # vs one job at a time) and a serving stream (continuous batching vs
        # one request at a time), after the cube's buffers are released
        del grids, image, norm, idx, tmask, d
        torch.cuda.synchronize()
        ctx.close()  # the cube's scratch
        torch.cuda.empty_cache()
        pool = snapshot_pool(torch, mti, args.jobs)
        if not args.no_c4:
            r4 = run_c4(args, torch, mti, pool)
            if r4.get("autotune"):
                cms = r4["autotune"].pop("candidates_ms")
                # the region shapes the coordinate search timed (others at their defaults)
                regions = {c["region_width"]: c["ms"] for c in cms
                           if all(c[k] == v for k, v in cms[0].items() if k not in ("region_width", "ms"))}
                r4["autotune"] = dict(r4["autotune"], candidates=len(cms), region_width_ms=regions)
            nested["c4"] = r4
        if not args.no_c5:
            a5 = argparse.Namespace(**vars(args))
            a5.rates, a5.serve_seconds = f"{args.c5_rate:g}", args.c5_seconds
            r5 = run_c5(a5, torch, mti, pool)
            rr = r5["rates"][0]
            nested["c5"] = dict(rate_req_s=rr["rate"], seconds=r5["seconds"], hbm_budget_gb=args.serve_budget_gb,
                                pool=r5["pool"], served=rr["served"], plain=rr["plain"],
                                speedup_mvis=rr["speedup_mvis"], kernel_sets=r5["kernel_sets"],
                                cache_hits=r5["cache_hits"], cache_misses=r5["cache_misses"])
        del pool
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("full-pol cube: 64 ant / 8 km, 2016 baselines x 1800 steps, 64 ch "
                         "0.9-1.7 GHz, P=4, 4096^2 @ 1.5 arcsec, 256 w-planes, s=8, Briggs 0, IQUV")
            if args.config == "c2" else
            "small array: 27 ant / 3 km, 351 baselines x 300 steps, 8 ch, P=4, 512^2 @ 2 arcsec, 32 planes, s=4, "
            "natural weighting",
            "rows": nrows, "channels": nchan, "gridded_vis": gridded, "gridded_points": points,
            "supports": [int(sup.min()), int(sup.max())], "path": "plain single-job",
            "tuning": tuning, "grids": "complete" if args.complete_grids else "sparse (tile mask)",
            "grid_precision": args.grid_precision, "transform_storage": args.transform_storage,
        },
        "image_error_vs_fp64_path": (None if form_err is None else
                                     {"max_rel_rms": float(f"{form_err:.3e}"),
                                      "vs": "complex128 grids and complex128 transform planes",
                                      "over": "every channel and Stokes image, relative to the channel's Stokes I RMS"}),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "stages_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in stages.items()},
        # binning + record sort + plan + gridder only (the reference's gridding-rate scope)
        "gridding_stage_mvis_s": round(gridded / (sum(stages[k][0] for k in ("bin", "sort_records", "plan", "grid"))
                                                  / args.steps / 1e3) / 1e6, 3),
        "roofline": roofline,
        "transform_rows": rows_pass,
        # from pinned host memory, host-to-device copies overlapped with gridding (not `value`)
        "end_to_end": e2e,
        "cpu_baseline": cpu,
    }
    if forms is not None:
        out["c3_forms"] = forms
    out.update(nested)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
