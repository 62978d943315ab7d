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
             host=dict(uvw=uvw, a1=a1, a2=a2, inv_lambda=inv_lambda, src=src))
    torch.cuda.synchronize()
    return d


def cpu_baseline(d, params_args, budget_rows):
    """FP64 oracle on a bounded sample: one channel's gridding of
    `budget_rows` baseline-major rows (Briggs weights included) and one
    channel's finalize; extrapolated to the whole cube."""
    import numpy as np
    import oracle
    from mtimager_amd import synth
    h = d["host"]
    rows = oracle.sort_index(h["a1"], h["a2"], d["nant"])
    sample = rows[:budget_rows]
    nx, cell, W, s = params_args
    p = oracle.make_params(nx, nx, cell * synth.ARCSEC, cell * synth.ARCSEC, "wproj", W, s)
    K = oracle.Kernels(p)
    chan = d["nchan"] // 2
    rng = np.random.default_rng(0)
    n = len(sample)
    P = d["P"]
    vis = (rng.standard_normal((n, 1, P)) + 1j * rng.standard_normal((n, 1, P))).astype(np.complex64)
    flags = np.zeros((n, 1, P), np.uint8)
    weight = np.ones((n, P), np.float32)
    il = np.array([h["inv_lambda"][chan]])
    t0 = time.perf_counter()
    wts = oracle.weights(K, h["uvw"][sample], il, flags, weight, 0, "briggs", 0.0)
    g, nrm, ng = oracle.grid(K, h["uvw"][sample], il, vis, flags, weight, 0, wts=wts)
    t_grid = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.finalize(p, g, nrm)
    t_fin = time.perf_counter() - t0
    per_rec = t_grid / max(ng, 1)
    return dict(per_record_s=per_rec, finalize_s=t_fin, sample_rows=n, sample_gridded=ng)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c2", "c1"])
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=60000)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one elementwise negation over this many MiB before the timed region "
                         "(a kernel of known HBM traffic for calibrating FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(channels/jobs would shard one per GPU, see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import mtimager_amd as mti

    torch.cuda.set_device(0)
    d = make_workload(torch, args.config, args.seed)
    p = mti.make_params(d["nx"], d["nx"], d["cell"], d["cell"], "wproj", d["W"], d["s"])
    ctx = mti.Context(0)
    nrows, nchan, P, nx = d["nrows"], d["nchan"], d["P"], d["nx"]
    grids = torch.empty((nchan, P, nx, nx, 2), dtype=torch.float64, device="cuda")
    norm = torch.empty((nchan, P), dtype=torch.float64, device="cuda")
    image = torch.empty((nchan, P, nx, nx), dtype=torch.float32, device="cuda")
    idx = torch.empty(nrows, dtype=torch.int32, device="cuda")

    def step(stats=False):
        kept = ctx.sort_rows(d["a1"], d["a2"], d["nant"], idx)
        K = ctx.build_kernels(p)
        v = mti.vis_ms(d["uvw"], d["invl"], d["vis"], d["flags"], d["weight"], idx[:kept])
        st = ctx.grid(K, p, v, grids, norm, chan0=0, nchan=nchan, weighting="briggs", robust=0.0,
                      layout="pvu", precision="c128", stats=stats)
        ctx.finalize(p, P, grids, norm, image, nchan=nchan, layout="pvu", precision="c128")
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
                    traffic=None, flops_per_launch=flops, avg_launch_ms=round(grid_avg_s * 1e3, 3),
                    algorithmic_GBps=round(rec_bytes / grid_avg_s / 1e9, 1))

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP64 oracle, 1 thread)")
        cb = cpu_baseline(d, (nx, d["cell"], d["W"], d["s"]), args.cpu_rows)
        per_ch_rows = gridded / nchan
        t_job = nchan * (per_ch_rows * cb["per_record_s"] + cb["finalize_s"])
        cpu = dict(value=round(gridded / t_job / 1e6, 4), unit="Mvis/s", cores=1, kind="port",
                   sample=(f"FP64 oracle: Briggs weights + gridding of {cb['sample_rows']} baseline-major "
                           f"rows of one channel ({cb['sample_gridded']} gridded, "
                           f"{cb['per_record_s'] * 1e6:.2f} us/record) and one {nx}^2 x {P} finalize "
                           f"({cb['finalize_s']:.1f} s), extrapolated to {nchan} channels; "
                           "W-kernel generation not counted"))

    lat = np.array(times) * 1e3
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
            "workload": ("config2 full-pol cube: 64 ant / 8 km, 2016 baselines x 1800 steps, 64 ch "
                         "0.9-1.7 GHz, P=4, 4096^2 @ 1.5 arcsec, 256 w-planes, s=8, Briggs 0, IQUV")
            if args.config == "c2" else
            "config1: 27 ant / 3 km, 351 baselines x 300 steps, 8 ch, P=4, 512^2 @ 2 arcsec, 32 planes, s=4",
            "rows": nrows, "channels": nchan, "gridded_vis": gridded, "gridded_points": points,
            "supports": [int(sup.min()), int(sup.max())], "path": "plain single-job",
        },
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "stages_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in stages.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
