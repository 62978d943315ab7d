"""Prediction (degridding) on the GPU against the FP64 oracle: model grids from
Stokes model images, degridded visibilities on the same grids (every record
binned bit-exactly as the gridder bins it), and a point-source model against
the direct Fourier sum.  No reference counterpart (the reference only images);
the oracle's own pins are in tests/test_oracle_degrid.py."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ARC = np.pi / 180 / 3600


@pytest.fixture(scope="module")
def ob():
    from mtimager_amd import synth
    return synth.make_observation(nant=20, radius_m=2500.0, ntime=40, freqs_hz=[1.40e9, 1.45e9, 1.5e9], npol=4,
                                  nsrc=1, fov_rad=512 * 2 * ARC, noise=0, flag_frac=0.05, seed=7, predict=False)


def _model(N, P, seed=2):
    rng = np.random.default_rng(seed)
    img = np.zeros((P, N, N), np.float32)
    for _ in range(12):
        r, c = rng.integers(N // 4, 3 * N // 4, 2)
        img[:, r, c] += rng.uniform(0.1, 1.0, P).astype(np.float32) * np.array([1, 0.3, -0.2, 0.1][:P], np.float32)
    return img


@pytest.mark.parametrize("P,poltype", [(1, "linear"), (4, "linear"), (4, "circular")])
def test_model_grids_and_degrid_match_oracle(orc, mti, ctx, ob, P, poltype):
    N = 512
    p = mti.make_params(N, N, 2.0, 2.0, "wproj", 16, 8)
    op = H.oracle_params(orc, p)
    K = ctx.build_kernels(p)
    OK = orc.Kernels(op)
    img = _model(N, P)
    # two channels of models: the second scaled
    models = torch.tensor(np.stack([img, 0.5 * img]), device="cuda")
    grids = torch.empty((2, P, N, N, 2), dtype=torch.float32, device="cuda")
    print(f"[degrid] model grids P={P} {poltype}", flush=True)
    ctx.model_grids(p, P, models, grids, nchan=2, poltype=poltype)
    g = grids.cpu().numpy()
    g = g[..., 0] + 1j * g[..., 1]  # [2][P][N][N]
    ref = orc.model_grid(op, img.astype(np.float64), poltype)  # [N][N][P]
    assert H.rel_rms(g[0].transpose(1, 2, 0), ref) < 1e-6
    assert H.rel_rms(g[1].transpose(1, 2, 0), 0.5 * ref) < 1e-6
    # degrid channels 1..2 of the observation from the two model channels
    uvw = torch.tensor(ob.uvw, device="cuda")
    invl = torch.tensor(ob.inv_lambda, device="cuda")
    nrows = len(ob.uvw)
    vis_in = torch.zeros((nrows, 3, P), dtype=torch.complex64, device="cuda")
    fl = torch.tensor(ob.flags[:, :, :P].copy(), device="cuda")
    w = torch.ones((nrows, P), dtype=torch.float32, device="cuda")
    a1, a2 = torch.tensor(ob.ant1, device="cuda"), torch.tensor(ob.ant2, device="cuda")
    idx = torch.empty(nrows, dtype=torch.int32, device="cuda")
    kept = ctx.sort_rows(a1, a2, ob.nant, idx)
    v = mti.vis_ms(uvw, invl, vis_in, fl, w, idx[:kept])
    out = torch.full((nrows, 3, P), complex(7.0, 7.0), dtype=torch.complex64, device="cuda")
    npred = ctx.degrid(K, p, v, grids, out, chan0=1, nchan=2)
    got = out.cpu().numpy()
    rows = idx[:kept].cpu().numpy()
    assert npred > 0.9 * 2 * kept
    for c in (1, 2):
        gm = g[c - 1].transpose(1, 2, 0).astype(np.complex128)  # the GPU's own model grid
        exp, n = orc.degrid(OK, ob.uvw[rows], ob.inv_lambda[c], gm)
        e = H.rel_rms(got[rows, c], exp)
        print(f"[degrid] P={P} {poltype} channel {c}: rel RMS vs oracle {e:.2e}", flush=True)
        assert e < 1e-6
    assert np.all(got[:, 0] == complex(7.0, 7.0))  # channel 0 not written
    auto = np.setdiff1d(np.arange(nrows), rows)
    assert np.all(got[auto] == complex(7.0, 7.0))  # autocorrelation rows not written


def test_point_source_model_matches_direct_sum(mti, ctx, ob):
    N = 512
    p = mti.make_params(N, N, 2.0, 2.0, "wproj", 16, 16)
    K = ctx.build_kernels(p)
    off = 3
    img = np.zeros((3, 1, N, N), np.float32)  # the same model for the 3 channels
    img[:, 0, N // 2 - 1 - off, N // 2 + off] = 2.0
    grids = torch.empty((3, 1, N, N, 2), dtype=torch.float32, device="cuda")
    ctx.model_grids(p, 1, torch.tensor(img, device="cuda"), grids, nchan=3)
    nrows = len(ob.uvw)
    uvw = torch.tensor(ob.uvw, device="cuda")
    invl = torch.tensor(ob.inv_lambda, device="cuda")
    v = mti.vis_ms(uvw, invl, torch.zeros((nrows, 3, 1), dtype=torch.complex64, device="cuda"),
                   torch.zeros((nrows, 3, 1), dtype=torch.uint8, device="cuda"),
                   torch.ones((nrows, 1), dtype=torch.float32, device="cuda"))
    out = torch.zeros((nrows, 3, 1), dtype=torch.complex64, device="cuda")
    npred = ctx.degrid(K, p, v, grids, out, chan0=0, nchan=3)
    g = grids.cpu().numpy()
    print(f"[degrid] point source: {npred} of {3 * nrows} predicted; model grid |G| "
          f"min {np.abs(g[0, 0, ..., 0] + 1j * g[0, 0, ..., 1]).min():.3g} max "
          f"{np.abs(g[0, 0, ..., 0] + 1j * g[0, 0, ..., 1]).max():.3g}", flush=True)
    l0 = m0 = off * 2 * ARC
    n0 = np.sqrt(1 - l0 ** 2 - m0 ** 2)
    got = out.cpu().numpy()[:, :, 0]
    for c in range(3):
        f = ob.inv_lambda[c]
        ex = 2.0 * np.exp(-2j * np.pi * f * (ob.uvw[:, 0] * l0 + ob.uvw[:, 1] * m0 + ob.uvw[:, 2] * (n0 - 1)))
        ok = np.abs(got[:, c]) > 0
        print(f"[degrid] channel {c}: {ok.mean():.3f} nonzero, first {got[:3, c]}", flush=True)
        assert ok.mean() > 0.9
        assert abs(np.mean(got[ok, c] / ex[ok]) - 1) < 1e-3
        assert np.sqrt(np.mean(np.abs(got[ok, c] - ex[ok]) ** 2)) / 2.0 < 3e-3
