"""The reference's operators with their own array contracts, through the
C-ABI: PreProcessUVW's seven outputs on the reference's chunk layout, and the
reference's ImageFinalizer chain run operator by operator
(PreIFFTUVCorrection -> IFFT2D -> ZeroToCentre2DShift -> StokesConverter ->
casagridsf correction -> SincCorrector, comp/ImageFinalizer.cpp:31-44)
against the FP64 oracle's finalize; casagridsf's generate modes against the
oracle's taper-only tables and psi."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _expected_preprocess(bins, uvw, flags, weight, sums, T, cols):
    """The reference's seven outputs from the oracle's FP64 bins, by the rules
    of op/WProjPreProcessUVW.cu:150-285 (zeros for records not gridded; a
    gridded record equal in kernel index and corner position to its
    predecessor is compressed; -1 conjugation where w < 0 and some
    correlation is unflagged)."""
    tg = bins["to_grid"].astype(bool)
    S = np.where(tg, bins["support"], 0)
    cx, cy = bins["corner_x"], bins["corner_y"]
    Ssafe = np.maximum(bins["support"], 1)
    gd = np.stack([cx, cy * cols + cx, -(cx % Ssafe), -(cy % Ssafe)], axis=1)
    gd = np.where(tg[:, None], gd, 0).astype(np.int32)
    ci = np.where(tg, bins["conv_idx"], 0).astype(np.int64)
    comp = np.zeros(len(tg), bool)
    comp[1:] = tg[1:] & (ci[1:] == ci[:-1]) & (gd[1:, 1] == gd[:-1, 1])
    anyf = (flags == 0).any(axis=1)
    conj = np.where(anyf & (uvw[:, 2] < 0), -1.0, 1.0).astype(np.float32)
    cs = np.where(tg, sums[bins["plane"], bins["off_y"] * T + bins["off_x"]], 0.0)
    csum = (cs[None, :] * weight.T.astype(np.float64) * (flags.T == 0)).astype(np.float32)
    return dict(gd=gd, ci=ci.astype(np.int32), tg=(tg & ~comp).astype(np.int32), comp=comp.astype(np.int32),
                S=S.astype(np.int32), conj=conj, csum=csum)


def test_preprocess_reference_outputs(orc, mti, ctx):
    from mtimager_amd import synth
    ob = synth.small_array_obs(seed=3, ntime=80, freqs_hz=[1.4e9, 1.5e9])
    p = mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    K = ctx.build_kernels(p)
    sup, base = K.supports()
    rows = orc.sort_index(ob.ant1, ob.ant2, ob.nant)
    # repeat some records so that compression (equal kernel and corner) occurs
    rows = np.sort(np.concatenate([rows, rows[::7]]), kind="stable")
    chan, P, T = 1, 4, K.info.T
    uvw32 = ob.uvw[rows].astype(np.float32)
    il32 = np.full(len(rows), ob.inv_lambda[chan], np.float32)
    fl = ob.flags[rows, chan, :].astype(np.int32)
    w = (ob.weight[rows] * np.random.default_rng(1).uniform(0.5, 2.0, (len(rows), P))).astype(np.float32)
    d = [torch.tensor(a, device="cuda") for a in (uvw32.T.copy(), il32, ob.vis[rows, chan, :], fl, w)]
    v = mti.vis_reference_chunk(*d)
    N = len(rows)
    outs = [torch.full((N, 4), -9, dtype=torch.int32, device="cuda")] + \
           [torch.full((N,), -9, dtype=torch.int32, device="cuda") for _ in range(4)] + \
           [torch.full((N,), 9.0, dtype=torch.float32, device="cuda"),
            torch.full((P, N), 9.0, dtype=torch.float32, device="cuda")]
    rp = mti.PreprocessRefParams(p.ny, p.nx, 1.0 / (p.cell_l * p.nx), 1.0 / (p.cell_m * p.ny), p.sampling,
                                 K.info.w_increment, K.info.nplanes)
    convinfo = torch.tensor(np.stack([sup, base], axis=1).astype(np.int32), device="cuda")
    ctx.preprocess_reference(rp, convinfo, K.sums(), v, 0, outs)
    torch.cuda.synchronize()
    got = [o.cpu().numpy() for o in outs]
    OK = orc.Kernels(H.oracle_params(orc, p))
    bins = orc.preprocess(OK, uvw32.astype(np.float64), np.array([float(il32[0])]),
                          fl.astype(np.uint8)[:, None, :], 0)
    e = _expected_preprocess(bins, uvw32.astype(np.float64), fl, w, K.sums().cpu().numpy().astype(np.float64), T,
                             p.nx)
    assert e["comp"].sum() > 0 and e["tg"].sum() > 0
    np.testing.assert_array_equal(got[0], e["gd"])
    np.testing.assert_array_equal(got[1], e["ci"])
    np.testing.assert_array_equal(got[2], e["tg"])
    np.testing.assert_array_equal(got[3], e["comp"])
    np.testing.assert_array_equal(got[4], e["S"])
    np.testing.assert_array_equal(got[5], e["conj"])
    np.testing.assert_allclose(got[6], e["csum"], rtol=1e-6, atol=0)
    print(f"\n{N} records: {int(e['tg'].sum())} gridded, {int(e['comp'].sum())} compressed")


@pytest.mark.parametrize("rows,cols,npol,poltype", [(128, 256, 4, "linear"), (60, 75, 4, "circular"),
                                                   (64, 64, 1, "linear"), (33, 48, 2, "linear")])
def test_reference_finalizer_chain_matches_oracle(orc, mti, ctx, rows, cols, npol, poltype):
    p = mti.make_params(cols, rows, 3.0, 2.0, "wproj", 8, 2)
    rng = np.random.default_rng(rows + npol)
    g = rng.standard_normal((rows, cols, npol, 2)).astype(np.float32)
    norm = rng.uniform(1.0, 3.0, npol).astype(np.float32)
    G = torch.tensor(g, device="cuda")
    X = torch.empty((npol, rows, cols, 2), dtype=torch.float32, device="cuda")
    Y = torch.empty_like(X)
    img = torch.empty((npol, rows, cols), dtype=torch.float32, device="cuda")
    mti.uv_reorder(G, X, rows, cols, npol)                       # PreIFFTUVCorrection
    ctx.fft2d_inverse(X, cols, rows, npol)                        # IFFT2D
    mti.centre_shift(X, Y, npol, rows, cols)                      # ZeroToCentre2DShift
    mti.stokes(Y, torch.tensor(norm, device="cuda"), img, npol, rows, cols, poltype)  # StokesConverter
    mti.taper_correct(img, img, npol, rows, cols)                 # casagridsf, lmPlanePostCorrection
    mti.sinc_correct(img, img, npol, rows, cols, p.sampling)      # SincCorrector
    ref = orc.finalize(H.oracle_params(orc, p), g.astype(np.float64), norm.astype(np.float64), poltype)
    out = img.cpu().numpy()
    for q in range(npol):
        assert H.rel_rms(out[q], ref[q]) < 1e-5, (q, H.rel_rms(out[q], ref[q]))


def test_normalizer_f32(mti):
    a = torch.tensor([[1.0 + 1e-12, 3.0], [2.5, 1e-3]], dtype=torch.float64, device="cuda")
    b = torch.empty((2, 2), dtype=torch.float32, device="cuda")
    mti.normalizer_f32(a, b)
    assert torch.equal(b, a.float())


@pytest.mark.parametrize("support,sampling", [(7, 4), (5, 3)])
def test_taper_kernel_is_the_taper_only_table(orc, mti, support, sampling):
    """casagridsf generate (lmPlanePostCorrection = false), reorganised to
    [oy][ox][sy][sx], equals the oracle's mode=normal table."""
    dim = support * sampling + (0 if sampling % 2 else 1)
    out = torch.empty((dim, dim), dtype=torch.float64, device="cuda")
    mti.taper_kernel(out, support, sampling)
    outc = torch.empty((dim, dim, 2), dtype=torch.float32, device="cuda")
    mti.taper_kernel(outc, support, sampling, complex_out=True)
    K = out.cpu().numpy()
    Kc = outc.cpu().numpy()
    assert np.all(Kc[..., 1] == 0) and np.allclose(Kc[..., 0], K, rtol=1e-6, atol=1e-7)
    OK = orc.Kernels(orc.make_params(256, 256, 1e-5, 1e-5, "normal", 1, sampling, support))
    T, S = OK.T, support
    tab = OK.plane(0)  # [T][T][S][S]
    for oy in range(T):
        for ox in range(T):
            sub = K[oy::sampling, ox::sampling][:S, :S]
            np.testing.assert_allclose(sub, tab[oy, ox].real, rtol=1e-12, atol=1e-15)


def test_taper_grid_is_the_psi_product(orc, mti):
    ny, nx = 50, 64
    out = torch.empty((ny, nx), dtype=torch.float64, device="cuda")
    mti.taper_correct(None, out, 1, ny, nx)
    gy = np.array([orc.gridsf(abs(j - ny // 2) / (ny // 2)) for j in range(ny)])
    gx = np.array([orc.gridsf(abs(i - nx // 2) / (nx // 2)) for i in range(nx)])
    np.testing.assert_allclose(out.cpu().numpy(), np.outer(gy, gx), rtol=1e-14, atol=0)
