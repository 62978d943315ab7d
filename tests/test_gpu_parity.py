"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle on the
same seeded inputs.  Bins, w-planes, supports and the row order are compared
bit-exactly; floating-point results against stated tolerances."""
import numpy as np
import pytest

import gpu_helpers as H
import kat_binning as KAT

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c1(orc):
    """Small-array observation (27 ant, 3 km, 300 steps, 8 ch, P=4), seed 1."""
    from mtimager_amd import synth
    return synth.small_array_obs(seed=1)


@pytest.fixture(scope="module")
def c1_params(mti):
    return mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)


# ---- a1 ------------------------------------------------------------------
def test_sort_rows_bit_exact(orc, mti, ctx, c1):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch).cpu().numpy()
    ref = orc.sort_index(c1.ant1, c1.ant2, c1.nant)
    assert len(idx) == 351 * 300
    np.testing.assert_array_equal(idx, ref)


def test_sort_rows_random_and_empty(orc, mti, ctx):
    rng = np.random.default_rng(5)
    for n, nant in [(1, 3), (5000, 7), (70000, 130)]:
        a1 = rng.integers(0, nant, n).astype(np.int32)
        a2 = rng.integers(0, nant, n).astype(np.int32)
        t1 = torch.tensor(a1, device="cuda")
        t2 = torch.tensor(a2, device="cuda")
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        kept = ctx.sort_rows(t1, t2, nant, out)
        np.testing.assert_array_equal(out[:kept].cpu().numpy(), orc.sort_index(a1, a2, nant))
    e = torch.empty(0, dtype=torch.int32, device="cuda")
    assert ctx.sort_rows(e, e, 4, e) == 0


# ---- a8-a14 ----------------------------------------------------------------
@pytest.mark.parametrize("cfg", [
    dict(nx=512, cell=2.0, W=32, s=4, mode="wproj"),      # the small-array geometry
    dict(nx=256, cell=60.0, W=64, s=4, mode="wproj"),     # wide field, growing supports
    dict(nx=128, cell=120.0, W=16, s=5, mode="wproj"),    # odd sampling
    dict(nx=256, cell=2.0, W=1, s=4, mode="normal"),      # taper only
])
def test_kernel_tables(orc, mti, ctx, cfg):
    p = mti.make_params(cfg["nx"], cfg["nx"], cfg["cell"], cfg["cell"], cfg["mode"], cfg["W"], cfg["s"])
    K = ctx.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    sup, base = K.supports()
    np.testing.assert_array_equal(sup, OK.support)
    np.testing.assert_array_equal(base, OK.base)
    assert K.info.w_increment == OK.w_increment
    tab = K.table().cpu().numpy()
    ref = OK.table
    assert np.max(np.abs(tab - ref)) <= 2e-7 * np.max(np.abs(ref))
    np.testing.assert_allclose(K.sums().cpu().numpy(), OK.sums, rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("prec,tol", [("f16", 1e-3), ("bf16", 8e-3)])
def test_kernel_tables_reduced_precision(orc, mti, ctx, prec, tol):
    p = mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    K = ctx.build_kernels(p, prec)
    OK = orc.Kernels(H.oracle_params(orc, p))
    tab = K.table().cpu().numpy()
    assert np.max(np.abs(tab - OK.table)) <= tol * np.max(np.abs(OK.table))


# ---- a4 --------------------------------------------------------------------
@pytest.mark.parametrize("s,cases", [(4, KAT.CASES_S4), (5, KAT.CASES_S5)])
def test_binning_known_answers(mti, ctx, s, cases):
    p = mti.make_params(KAT.NX, KAT.NX, KAT.CELL, KAT.CELL, "wproj", KAT.WPLANES, s, units="rad")
    K = ctx.build_kernels(p)
    uvw, flags, exp = KAT.arrays(cases)
    n = len(uvw)
    d_uvw = torch.tensor(uvw, device="cuda")
    d_fl = torch.tensor(np.repeat(flags, 1, axis=2), device="cuda")
    d_invl = torch.tensor([1.0], dtype=torch.float64, device="cuda")
    d_w = torch.ones((n, 2), dtype=torch.float32, device="cuda")
    d_vis = torch.zeros((n, 1, 2), dtype=torch.complex64, device="cuda")
    v = mti.vis_ms(d_uvw, d_invl, d_vis, d_fl, d_w)
    bins = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    ctx.preprocess(K, p, v, 0, bins)
    got = bins.cpu().numpy()
    np.testing.assert_array_equal(got, exp[:, :8])


def test_binning_bit_exact_small_array(orc, mti, ctx, c1, c1_params):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(c1_params)
    OK = orc.Kernels(H.oracle_params(orc, c1_params))
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    rows = idx.cpu().numpy()
    for chan in (0, 7):
        bins = torch.zeros((len(rows), 8), dtype=torch.int32, device="cuda")
        ctx.preprocess(K, c1_params, v, chan, bins)
        got = bins.cpu().numpy()
        ref = orc.preprocess(OK, c1.uvw[rows], c1.inv_lambda, c1.flags[rows], chan)
        exp = np.stack([ref[f] for f in ("to_grid", "plane", "support", "corner_x", "corner_y",
                                          "off_x", "off_y", "conj")], axis=1)
        np.testing.assert_array_equal(got, exp)
        assert got[:, 0].sum() > 0.9 * len(rows)


def test_binning_half_integer_ties_random(orc, mti, ctx):
    """Random records placed on exact .5 pixel and (k+1/2)^2 dw plane ties."""
    p = mti.make_params(KAT.NX, KAT.NX, KAT.CELL, KAT.CELL, "wproj", KAT.WPLANES, 4, units="rad")
    K = ctx.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    rng = np.random.default_rng(9)
    n = 20000
    u = 64.0 * rng.integers(-50, 50, n)
    v = 32.0 * rng.integers(-100, 100, n)
    w = 128.0 * (rng.integers(0, 10, n) / 2.0) ** 2 * rng.choice([-1, 1], n)
    uvw = np.stack([u, v, w], 1)
    flags = (rng.random((n, 1, 1)) < 0.1).astype(np.uint8)
    v_ = mti.vis_ms(torch.tensor(uvw, device="cuda"), torch.tensor([1.0], dtype=torch.float64, device="cuda"),
                    torch.zeros((n, 1, 1), dtype=torch.complex64, device="cuda"),
                    torch.tensor(flags, device="cuda"), torch.ones((n, 1), dtype=torch.float32, device="cuda"))
    bins = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    ctx.preprocess(K, p, v_, 0, bins)
    ref = orc.preprocess(OK, uvw, np.array([1.0]), flags, 0)
    exp = np.stack([ref[f] for f in ("to_grid", "plane", "support", "corner_x", "corner_y",
                                      "off_x", "off_y", "conj")], axis=1)
    np.testing.assert_array_equal(bins.cpu().numpy(), exp)


# ---- images ------------------------------------------------------------------
IMAGE_RMS_TOL = 1e-4   # relative RMS vs the FP64 oracle
GRID_RMS_TOL = 1e-5


@pytest.mark.parametrize("psf", [False, True])
@pytest.mark.parametrize("precision,layout", [("c64", "vup"), ("c128", "pvu")])
def test_small_array_dirty_image_and_psf(orc, mti, ctx, c1, c1_params, psf, precision, layout):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    rows = idx.cpu().numpy()
    K = ctx.build_kernels(c1_params)
    OK = orc.Kernels(H.oracle_params(orc, c1_params))
    g = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [0, 1], idx, psf=psf, precision=precision,
                    layout=layout)
    ng = 0
    for ci, chan in enumerate([0, 1]):
        r = H.oracle_image(orc, OK, H.oracle_params(orc, c1_params), c1, chan, rows, psf=psf)
        ng += r["gridded"]
        np.testing.assert_allclose(g["norm"][ci], r["norm"], rtol=1e-6)
        assert H.rel_rms(g["grid"][ci], r["grid"]) < GRID_RMS_TOL
        e = H.image_errors(g["image"][ci][0], r["image"][0])
        print(f"chan {chan} psf={psf} {precision}/{layout} stokes I: {e}")
        assert e["rel_rms"] < IMAGE_RMS_TOL
        ref_i = np.sqrt(np.mean(r["image"][0] ** 2))
        for q in range(1, 4):  # Q, U, V relative to Stokes I
            d = g["image"][ci][q].astype(np.float64) - r["image"][q]
            assert np.sqrt(np.mean(d ** 2)) / ref_i < IMAGE_RMS_TOL
    assert g["stats"].gridded_records == ng


# ---- precision forms of the kernel tables --------------------------------------------
# Kernel tables in f16 / bf16 on the small-array inputs (visibilities complex64,
# FP64 accumulation and transform, FP64 binning and w-chirp phase); every
# Stokes image and the PSF against the FP64 oracle with FP64 tables.
# Rounding a table sample (f16: 2^-11, bf16: 2^-8 relative) is noise that
# the taper correction divides by psi(l) psi(m): psi(1) = 0.004, so the gain
# is 250 at an edge and 6e4 in a corner.  A table rounded to f16 (relative
# RMS 1.9e-4 per sample) or bf16 (1.6e-3) therefore has a whole-field error
# set by the edges; the inner half of the field (psi >= 0.36 on both axes,
# gain <= 7.7) is held to the first tolerance, the whole field (relative to
# Stokes I) to the second.  f32 tables with FP32 compensated accumulation
# meet the 1e-4 image tolerance over the whole field.
PRECISION_FORM_TOL = {"f16": (2e-3, 1e-1), "bf16": (2e-2, 1.5), "f32": (1e-5, 1e-4)}


def _inner(img):
    n0, n1 = img.shape
    return img[n0 // 4: 3 * n0 // 4, n1 // 4: 3 * n1 // 4]


@pytest.mark.parametrize("table,accumulation", [("f16", "f64"), ("bf16", "f64"), ("f16", "f32k"),
                                                ("bf16", "f32k"), ("f32", "f32k")])
def test_precision_forms_small_array(orc, mti, ctx, c1, c1_params, table, accumulation):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    rows = idx.cpu().numpy()
    K = ctx.build_kernels(c1_params, precision=table)
    OK = orc.Kernels(H.oracle_params(orc, c1_params))
    op = H.oracle_params(orc, c1_params)
    tol_inner, tol_field = PRECISION_FORM_TOL[table]
    for psf in (False, True):
        g = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [0], idx, psf=psf, accumulation=accumulation)
        r = H.oracle_image(orc, OK, op, c1, 0, rows, psf=psf)
        ref_i = np.sqrt(np.mean(r["image"][0] ** 2))
        ref_i_in = np.sqrt(np.mean(_inner(r["image"][0]) ** 2))
        for q, name in enumerate("IQUV"):
            if psf and q > 0:
                break
            e = H.image_errors(g["image"][0][q], r["image"][q])
            d = g["image"][0][q].astype(np.float64) - r["image"][q]
            if q > 0:  # Q, U, V relative to Stokes I (unpolarised sky)
                e["rel_rms_vs_I"] = float(np.sqrt(np.mean(d ** 2)) / ref_i)
            e["inner_rel_rms_vs_I"] = float(np.sqrt(np.mean(_inner(d) ** 2)) / ref_i_in)
            e["field_rel_rms_vs_I"] = float(np.sqrt(np.mean(d ** 2)) / ref_i)
            print(f"precision form {table}/{accumulation} {'PSF' if psf else 'Stokes ' + name}: {e}")
            assert e["inner_rel_rms_vs_I"] < tol_inner
            assert e["field_rel_rms_vs_I"] < tol_field


def test_wide_field_wproj_image(orc, mti, ctx):
    """Wide field (4.3 deg), 64 planes, supports 7..11, P=1."""
    from mtimager_amd import synth
    ob = synth.make_observation(nant=20, radius_m=1500.0, ntime=120, freqs_hz=[1.0e9, 1.1e9],
                                npol=1, fov_rad=256 * 60 * synth.ARCSEC, seed=3)
    p = mti.make_params(256, 256, 60.0, 60.0, "wproj", 64, 4)
    dev = H.Dev(ob, torch)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    K = ctx.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    g = H.gpu_image(mti, ctx, K, p, dev, torch, [0, 1], idx)
    for ci in range(2):
        r = H.oracle_image(orc, OK, H.oracle_params(orc, p), ob, ci, idx.cpu().numpy())
        e = H.image_errors(g["image"][ci][0], r["image"][0])
        print("wide field", ci, e)
        assert e["rel_rms"] < IMAGE_RMS_TOL


@pytest.mark.parametrize("weighting,robust", [("uniform", 0.0), ("briggs", 0.0), ("briggs", -1.0)])
def test_weighting(orc, mti, ctx, c1, c1_params, weighting, robust):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    rows = idx.cpu().numpy()
    K = ctx.build_kernels(c1_params)
    OK = orc.Kernels(H.oracle_params(orc, c1_params))
    g = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [3], idx, weighting=weighting, robust=robust)
    r = H.oracle_image(orc, OK, H.oracle_params(orc, c1_params), c1, 3, rows, weighting=weighting,
                       robust=robust)
    np.testing.assert_allclose(g["norm"][0], r["norm"], rtol=1e-5)
    e = H.image_errors(g["image"][0][0], r["image"][0])
    print(weighting, robust, e)
    assert e["rel_rms"] < IMAGE_RMS_TOL


def test_batch_of_channels_equals_single_calls(mti, ctx, c1, c1_params):
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(c1_params)
    b = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [2, 3, 4, 5], idx)
    for ci, chan in enumerate([2, 3, 4, 5]):
        s = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [chan], idx)
        assert H.rel_rms(b["grid"][ci], s["grid"][0]) < 1e-6
        np.testing.assert_allclose(b["norm"][ci], s["norm"][0], rtol=1e-12)


# Two runs of one call agree to the order of atomic sums in tiles split over
# work units: FP64 atomics (complex128 grids) leave images equal to ~1e-12;
# float atomics (complex64 grids, as the reference's) change the last bits of
# those cells, which the taper correction amplifies at the field edges to
# ~2e-5 of the image RMS.
RERUN_IMAGE_TOL = {"c128": 1e-6, "c64": 5e-5}


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_planar_grid_layout_matches(orc, mti, ctx, c1, c1_params, precision):
    """[P][v][u] grids (what PreIFFTUVCorrection reorders to) give the same
    grid and image as the reference's [v][u][P] layout."""
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(c1_params)
    a = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [4], idx, layout="vup", precision=precision)
    b = H.gpu_image(mti, ctx, K, c1_params, dev, torch, [4], idx, layout="pvu", precision=precision)
    assert H.rel_rms(b["grid"][0], a["grid"][0]) < 1e-6
    assert H.rel_rms(b["image"][0], a["image"][0]) < RERUN_IMAGE_TOL[precision]
    r = H.oracle_image(orc, orc.Kernels(H.oracle_params(orc, c1_params)),
                       H.oracle_params(orc, c1_params), c1, 4, idx.cpu().numpy())
    assert H.rel_rms(b["image"][0][0], r["image"][0]) < IMAGE_RMS_TOL


def test_accumulate_over_chunks(mti, ctx, c1, c1_params):
    """Two record chunks with accumulate=1 == one call (comp/WImager.cpp:278-280)."""
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(c1_params)
    P, p = 4, c1_params
    grids = torch.empty((1, p.ny, p.nx, P, 2), dtype=torch.float32, device="cuda")
    norm = torch.empty((1, P), dtype=torch.float64, device="cuda")
    half = idx.numel() // 2
    for k, part in enumerate([idx[:half].contiguous(), idx[half:].contiguous()]):
        v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, part)
        ctx.grid(K, p, v, grids, norm, chan0=6, nchan=1, accumulate=(k > 0))
    one = H.gpu_image(mti, ctx, K, p, dev, torch, [6], idx)
    assert H.rel_rms(grids.cpu().numpy()[0], one["grid"][0]) < 1e-6
    np.testing.assert_allclose(norm.cpu().numpy()[0], one["norm"][0], rtol=1e-12)


def test_reference_chunk_layout(orc, mti, ctx, c1, c1_params):
    """The reference's per-channel chunk (uvw f32 [3][N], per-record f32
    1/lambda, int32 flags) grids to the oracle's image of the same
    (f32-rounded) inputs."""
    rows = orc.sort_index(c1.ant1, c1.ant2, c1.nant)
    chan = 5
    uvw32 = c1.uvw[rows].astype(np.float32)
    il32 = np.full(len(rows), c1.inv_lambda[chan], np.float32)
    vis = c1.vis[rows, chan, :]
    fl = c1.flags[rows, chan, :].astype(np.int32)
    w = c1.weight[rows]
    d = dict(uvw=torch.tensor(uvw32.T.copy(), device="cuda"), il=torch.tensor(il32, device="cuda"),
             vis=torch.tensor(vis, device="cuda"), fl=torch.tensor(fl, device="cuda"),
             w=torch.tensor(w, device="cuda"))
    v = mti.vis_reference_chunk(d["uvw"], d["il"], d["vis"], d["fl"], d["w"])
    K = ctx.build_kernels(c1_params)
    p = c1_params
    grids = torch.empty((1, p.ny, p.nx, 4, 2), dtype=torch.float32, device="cuda")
    norm = torch.empty((1, 4), dtype=torch.float64, device="cuda")
    ctx.grid(K, p, v, grids, norm, chan0=0, nchan=1)
    img = torch.empty((1, 4, p.ny, p.nx), dtype=torch.float32, device="cuda")
    ctx.finalize(p, 4, grids, norm, img)
    OK = orc.Kernels(H.oracle_params(orc, p))
    from mtimager_amd.synth import Observation
    ob = Observation(uvw=uvw32.astype(np.float64), inv_lambda=np.array([float(il32[0])]),
                     vis=vis[:, None, :], flags=fl.astype(np.uint8)[:, None, :], weight=w)
    r = H.oracle_image(orc, OK, H.oracle_params(orc, p), ob, 0)
    e = H.image_errors(img.cpu().numpy()[0][0], r["image"][0])
    assert e["rel_rms"] < IMAGE_RMS_TOL


@pytest.mark.parametrize("npol", [1, 2])
def test_fewer_correlations(orc, mti, ctx, npol):
    from mtimager_amd import synth
    ob = synth.small_array_obs(seed=2, npol=npol, ntime=100, freqs_hz=[1.45e9])
    p = mti.make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    dev = H.Dev(ob, torch)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    K = ctx.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    for pt in ("linear", "circular"):
        g = H.gpu_image(mti, ctx, K, p, dev, torch, [0], idx, poltype=pt)
        r = H.oracle_image(orc, OK, H.oracle_params(orc, p), ob, 0, idx.cpu().numpy(), poltype=pt)
        assert H.rel_rms(g["image"][0][0], r["image"][0]) < IMAGE_RMS_TOL


def test_normal_mode_image(orc, mti, ctx, c1):
    p = mti.make_params(512, 512, 2.0, 2.0, "normal", 1, 4, 7)
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(p)
    OK = orc.Kernels(H.oracle_params(orc, p))
    g = H.gpu_image(mti, ctx, K, p, dev, torch, [1], idx)
    r = H.oracle_image(orc, OK, H.oracle_params(orc, p), c1, 1, idx.cpu().numpy())
    assert H.rel_rms(g["image"][0][0], r["image"][0]) < IMAGE_RMS_TOL


@pytest.mark.parametrize("weighting", ["natural", "briggs"])
def test_flagged_samples_holding_nan_grid_as_zeros(mti, ctx, c1, c1_params, weighting):
    """A flagged correlation contributes nothing whatever its visibility holds
    (real data often carries NaN there): the grids and normalisers equal those
    of the same data with zeros in the flagged samples (to the order of FP64
    atomic sums)."""
    dev = H.Dev(c1, torch)
    idx = H.gpu_sort(ctx, dev, c1.nant, torch)
    K = ctx.build_kernels(c1_params)
    fl = dev.flags.bool()[:, :2]
    vz = dev.vis[:, :2].clone()
    vz[fl] = 0
    vn = vz.clone()
    vn[fl] = complex(float("nan"), float("nan"))
    out = []
    for vis in (vz, vn):
        v = mti.vis_ms(dev.uvw, dev.invl[:2].contiguous(), vis.contiguous(), dev.flags[:, :2].contiguous(),
                       dev.weight, idx)
        g = torch.empty((2, c1_params.ny, c1_params.nx, dev.npol, 2), dtype=torch.float64, device="cuda")
        n = torch.empty((2, dev.npol), dtype=torch.float64, device="cuda")
        ctx.grid(K, c1_params, v, g, n, nchan=2, weighting=weighting, precision="c128")
        torch.cuda.synchronize()
        out.append((g.cpu().numpy(), n.cpu().numpy()))
    assert fl.any() and np.isfinite(out[1][0]).all()
    # the same records and weights: equal up to the order of FP64 atomics
    # (normaliser sums across blocks, tiles split over work units)
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-12)
    assert H.rel_rms(out[1][0], out[0][0]) < 1e-12


def test_all_flagged_and_empty(mti, ctx, c1_params):
    p = c1_params
    K = ctx.build_kernels(p)
    n = 1000
    uvw = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    fl = torch.ones((n, 1, 1), dtype=torch.uint8, device="cuda")
    v = mti.vis_ms(uvw, torch.tensor([5.0], dtype=torch.float64, device="cuda"),
                   torch.zeros((n, 1, 1), dtype=torch.complex64, device="cuda"), fl,
                   torch.ones((n, 1), dtype=torch.float32, device="cuda"))
    grids = torch.full((1, p.ny, p.nx, 1, 2), 7.0, dtype=torch.float32, device="cuda")
    norm = torch.full((1, 1), 7.0, dtype=torch.float64, device="cuda")
    st = ctx.grid(K, p, v, grids, norm)
    assert st.gridded_records == 0
    assert float(grids.abs().max()) == 0.0 and float(norm.abs().max()) == 0.0
    v.nrows = 0
    grids.fill_(3.0)
    st = ctx.grid(K, p, v, grids, norm)
    assert st.gridded_records == 0 and float(grids.abs().max()) == 0.0


# ---- a17 -------------------------------------------------------------------
@pytest.mark.parametrize("nx,ny,nb", [(64, 128, 3), (512, 512, 1), (4096, 4096, 1), (2, 8, 2)])
def test_fft2d_inverse_matches_numpy(mti, ctx, nx, ny, nb):
    rng = np.random.default_rng(nx + ny)
    x = (rng.standard_normal((nb, ny, nx)) + 1j * rng.standard_normal((nb, ny, nx))).astype(np.complex64)
    d = torch.tensor(x, device="cuda")
    ctx.fft2d_inverse(d, nx, ny, nb)
    ref = np.fft.ifft2(x.astype(np.complex128)) * (nx * ny)
    got = d.cpu().numpy().astype(np.complex128)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2)) / np.sqrt(np.mean(np.abs(ref) ** 2))
    assert err < 1e-6, err


@pytest.mark.parametrize("nx,ny,npol,poltype", [
    (256, 128, 1, "linear"), (256, 128, 2, "circular"), (256, 128, 4, "linear"), (256, 128, 4, "circular"),
    # block-transform path (radix-16 passes): rectangular, P=1/2/4, the 4096^2 cube size
    (512, 256, 4, "linear"), (1024, 1024, 2, "linear"), (2048, 512, 1, "linear"), (4096, 4096, 4, "circular")])
def test_finalize_matches_oracle(orc, mti, ctx, nx, ny, npol, poltype):
    p = mti.make_params(nx, ny, 3.0, 2.0, "wproj", 8, 8)
    rng = np.random.default_rng(npol)
    g = rng.standard_normal((p.ny, p.nx, npol, 2))
    norm = rng.uniform(1.0, 3.0, npol)
    d_g = torch.tensor(g.astype(np.float32)[None], device="cuda")
    d_n = torch.tensor(norm[None], device="cuda")
    img = torch.empty((1, npol, p.ny, p.nx), dtype=torch.float32, device="cuda")
    ctx.finalize(p, npol, d_g, d_n, img, poltype=poltype)
    ref = orc.finalize(H.oracle_params(orc, p), g.astype(np.float32).astype(np.float64), norm, poltype)
    for q in range(npol):
        assert H.rel_rms(img.cpu().numpy()[0][q], ref[q]) < 1e-5


def test_invalid_parameters_raise(mti, ctx):
    bad = mti.make_params(512, 512, 2.0, 2.0, "normal", 1, 4, 6)  # even taper support
    with pytest.raises(mti.MTIError, match="odd"):
        ctx.build_kernels(bad)
    with pytest.raises(mti.MTIError):
        ctx.fft2d_inverse(torch.zeros(10, dtype=torch.complex64, device="cuda"), 0, 2)


# ---- batched jobs (the snapshot-mix path) ------------------------------------------
BATCH_JOBS = [  # nx, cell arcsec, w-planes, sampling, antennas, steps, channels, seed
    (256, 8.0, 8, 4, 12, 40, 3, 11),
    (512, 2.0, 16, 8, 20, 25, 2, 12),
    (256, 6.0, 12, 8, 8, 60, 1, 13),
    (1024, 1.0, 24, 4, 16, 20, 2, 14),
]


def _batch_job(mti, ctx, spec, P):
    from mtimager_amd import synth
    nx, cell, W, s, nant, ntime, nch, seed = spec
    ob = synth.make_observation(nant=nant, radius_m=1500.0, ntime=ntime,
                                freqs_hz=np.linspace(1.3e9, 1.45e9, nch), npol=P,
                                fov_rad=nx * cell * synth.ARCSEC, seed=seed)
    p = mti.make_params(nx, nx, cell, cell, "wproj", W, s)
    dev = H.Dev(ob, torch)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    K = ctx.build_kernels(p)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    grids = torch.zeros((nch, nx, nx, P, 2), dtype=torch.float64, device="cuda")
    norm = torch.zeros((nch, P), dtype=torch.float64, device="cuda")
    image = torch.zeros((nch, P, nx, nx), dtype=torch.float32, device="cuda")
    return dict(p=p, K=K, v=v, grids=grids, norm=norm, image=image, nch=nch, dev=dev, idx=idx, nant=ob.nant)


@pytest.mark.parametrize("weighting,order", [("natural", "index"), ("briggs", "index"), ("briggs", "in_call")])
def test_batched_jobs_match_single_jobs(mti, ctx, weighting, order):
    """Jobs of different sizes, samplings and plane counts gridded by one
    mti_grid_jobs call and finalized by one mti_finalize_jobs call give the
    per-job mti_grid + mti_finalize results (FP64 sums differ only in order).
    order "in_call": the batched call sorts each job's rows itself (antenna
    columns in the job, autocorrelation rows dropped on the device)."""
    P = 1
    js = [_batch_job(mti, ctx, s, P) for s in BATCH_JOBS]
    if order == "index":
        jobs = [ctx.make_job(j["K"], j["p"], j["v"], j["grids"], j["norm"], j["image"], 0, j["nch"]) for j in js]
    else:
        jobs = []
        for j in js:
            d = j["dev"]
            assert bool((d.ant1 == d.ant2).any())  # autocorrelations present
            v = mti.vis_ms(d.uvw, d.invl, d.vis, d.flags, d.weight)
            jobs.append(ctx.make_job(j["K"], j["p"], v, j["grids"], j["norm"], j["image"], 0, j["nch"],
                                     ant1=d.ant1, ant2=d.ant2, nant=j["nant"]))
    st = ctx.grid_jobs(jobs, weighting=weighting, robust=0.0)
    ctx.finalize_jobs(jobs)
    torch.cuda.synchronize()
    total = 0
    for j in js:
        g1 = torch.zeros_like(j["grids"])
        n1 = torch.zeros_like(j["norm"])
        i1 = torch.zeros_like(j["image"])
        s1 = ctx.grid(j["K"], j["p"], j["v"], g1, n1, chan0=0, nchan=j["nch"], weighting=weighting,
                      precision="c128")
        ctx.finalize(j["p"], P, g1, n1, i1, nchan=j["nch"], precision="c128")
        torch.cuda.synchronize()
        total += s1.gridded_records
        np.testing.assert_allclose(j["norm"].cpu().numpy(), n1.cpu().numpy(), rtol=1e-12)
        assert H.rel_rms(j["grids"].cpu().numpy(), g1.cpu().numpy()) < 1e-12
        assert H.rel_rms(j["image"].cpu().numpy(), i1.cpu().numpy()) < 1e-6
    assert st.gridded_records == total > 0
