"""Shared helpers for the GPU parity tests: upload an observation, run the
HIP imaging path through the C-ABI, and run the oracle on the same inputs."""
import numpy as np

ARC = np.pi / 180 / 3600


def rel_rms(a, b):
    """RMS of a - b over the RMS of b; complex arrays compare both parts."""
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.sqrt(np.mean(np.abs(b) ** 2)))


def image_errors(img, ref):
    """relative RMS, peak error / peak, and dynamic range of the residual."""
    img, ref = np.asarray(img, np.float64), np.asarray(ref, np.float64)
    d = img - ref
    peak = np.abs(ref).max()
    return dict(rel_rms=rel_rms(img, ref), peak_err=float(np.abs(d).max() / peak),
                residual_dr=float(peak / max(np.abs(d).max(), 1e-300)))


class Dev:
    """Observation arrays on the GPU in MeasurementSet order."""

    def __init__(self, ob, torch):
        t = torch
        self.uvw = t.tensor(ob.uvw, dtype=t.float64, device="cuda")
        self.invl = t.tensor(ob.inv_lambda, dtype=t.float64, device="cuda")
        self.vis = t.tensor(ob.vis, dtype=t.complex64, device="cuda") if ob.vis is not None else None
        self.flags = t.tensor(ob.flags, dtype=t.uint8, device="cuda")
        self.weight = t.tensor(ob.weight, dtype=t.float32, device="cuda")
        self.ant1 = t.tensor(ob.ant1, dtype=t.int32, device="cuda")
        self.ant2 = t.tensor(ob.ant2, dtype=t.int32, device="cuda")
        self.nrows, self.nchan, self.npol = ob.flags.shape


def gpu_sort(ctx, dev, nant, torch):
    idx = torch.empty(dev.nrows, dtype=torch.int32, device="cuda")
    kept = ctx.sort_rows(dev.ant1, dev.ant2, nant, idx)
    return idx[:kept].contiguous()


def gpu_image(mti, ctx, kern, params, dev, torch, chans, index=None, psf=False,
              weighting="natural", robust=0.0, poltype="linear", layout="vup", precision="c64",
              accumulation="f64"):
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, index)
    nb, P = len(chans), dev.npol
    assert list(chans) == list(range(chans[0], chans[0] + nb))
    gdt = torch.float32 if precision == "c64" else torch.float64
    grids = torch.empty((nb, params.ny, params.nx, P, 2), dtype=gdt, device="cuda")
    norm = torch.empty((nb, P), dtype=torch.float64, device="cuda")
    st = ctx.grid(kern, params, v, grids, norm, chan0=chans[0], nchan=nb, psf=psf,
                  weighting=weighting, robust=robust, layout=layout, precision=precision,
                  accumulation=accumulation)
    g_host = grids.cpu().numpy().copy()
    if layout == "pvu":  # [nb][P][ny][nx][2] -> [nb][ny][nx][P][2]
        g_host = g_host.reshape(nb, P, params.ny, params.nx, 2).transpose(0, 2, 3, 1, 4).copy()
    img = torch.empty((nb, P, params.ny, params.nx), dtype=torch.float32, device="cuda")
    ctx.finalize(params, P, grids, norm, img, nchan=nb, poltype=poltype, layout=layout,
                 precision=precision)
    torch.cuda.synchronize()
    return dict(grid=g_host, norm=norm.cpu().numpy(), image=img.cpu().numpy(), stats=st)


def oracle_image(orc, kern, params, ob, chan, index=None, psf=False, weighting="natural",
                 robust=0.0, poltype="linear"):
    rows = index if index is not None else np.arange(ob.nrows)
    uvw, flags, weight = ob.uvw[rows], ob.flags[rows], ob.weight[rows]
    vis = ob.vis[rows] if ob.vis is not None else np.zeros(flags.shape, np.complex64)
    wts = None
    if weighting != "natural":
        wts = orc.weights(kern, uvw, ob.inv_lambda, flags, weight, chan, weighting, robust)
    g, n, ng = orc.grid(kern, uvw, ob.inv_lambda, vis, flags, weight, chan, psf=psf, wts=wts)
    img = orc.finalize(params, g, n, poltype)
    return dict(grid=g, norm=n, image=img, gridded=ng)


def oracle_params(orc, p):
    return orc.make_params(p.nx, p.ny, p.cell_l, p.cell_m,
                           "wproj" if p.mode == 1 else "normal", p.wplanes, p.sampling,
                           p.taper_support)
