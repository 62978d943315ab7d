"""GPU parity of the transform at image sizes that are not powers of two (and
rows above 4096): the general path (mixed-radix HBM passes, Bluestein for a
prime factor above 13) against the FP64 oracle's finalize and numpy's FFT.
The reference plans its transforms for any rows x cols
(gafw/StandardOperators/IFFT2D.cpp:86-111) and rejects only sizes below 1
(comp/WImager.cpp:262-263)."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


# 2 * 3 * 43 * 4 = 1032 has the prime factor 43 (Bluestein); 6007 is prime;
# 1000 = 8 * 125 and 768 = 256 * 3 are mixed radix; 8192 rows exceed the
# 4096-element LDS row blocks of the fast path.
@pytest.mark.parametrize("nx,ny,nb", [(100, 64, 1), (1000, 768, 2), (1032, 17, 1), (13, 7, 3),
                                      (8192, 1024, 1), (6007, 9, 1), (1, 5, 2)])
def test_fft2d_inverse_any_size_matches_numpy(mti, ctx, nx, ny, nb):
    rng = np.random.default_rng(nx * 7 + ny)
    x = (rng.standard_normal((nb, ny, nx)) + 1j * rng.standard_normal((nb, ny, nx))).astype(np.complex64)
    d = torch.tensor(x, device="cuda")
    ctx.fft2d_inverse(d, nx, ny, nb)
    ref = np.fft.ifft2(x.astype(np.complex128)) * (nx * ny)
    got = d.cpu().numpy().astype(np.complex128)
    # FP64 arithmetic, one complex64 rounding of input and output
    assert H.rel_rms(got, ref) < 1e-6


def _random_grid(rng, nx, ny, npol):
    return rng.standard_normal((ny, nx, npol, 2), dtype=np.float32)


@pytest.mark.parametrize("nx,ny,npol,poltype", [
    (17, 9, 4, "circular"), (300, 300, 1, "linear"), (1000, 768, 4, "linear"),
    (1032, 1032, 2, "circular"), (3000, 3000, 1, "linear"), (6144, 6144, 1, "linear"),
    (8192, 8192, 1, "linear")])
def test_finalize_any_size_matches_oracle(orc, mti, ctx, nx, ny, npol, poltype):
    print(f"\nfinalize {nx}x{ny} P={npol} {poltype}", flush=True)
    # oversampling 8, or 2 for the 17 x 9 image (trial support >= 3)
    p = mti.make_params(nx, ny, 1.5, 1.5, "wproj", 8, 8 if min(nx, ny) >= 64 else 2)
    rng = np.random.default_rng(nx + npol)
    g = _random_grid(rng, nx, ny, npol)
    norm = rng.uniform(1.0, 3.0, npol)
    d_g = torch.tensor(g[None], device="cuda")
    d_n = torch.tensor(norm[None], device="cuda")
    img = torch.empty((1, npol, ny, nx), dtype=torch.float32, device="cuda")
    ctx.finalize(p, npol, d_g, d_n, img, poltype=poltype)
    out = img.cpu().numpy()[0]
    del d_g
    ref = orc.finalize(H.oracle_params(orc, p), g.astype(np.float64), norm, poltype)
    for q in range(npol):
        assert H.rel_rms(out[q], ref[q]) < 1e-5, (q, H.rel_rms(out[q], ref[q]))


def test_finalize_sparse_any_size_reads_unmarked_tiles_as_zeros(mti, ctx):
    # 1000 x 768: 32 x 24 tiles, the last tile column 8 cells wide
    nx, ny, P = 1000, 768, 4
    p = mti.make_params(nx, ny, 2.0, 2.0, "wproj", 8, 4)
    rng = np.random.default_rng(3)
    ntx, nty = (nx + 31) // 32, (ny + 31) // 32
    mask = (rng.random((nty, ntx)) < 0.4).astype(np.uint8)
    g = _random_grid(rng, nx, ny, P)
    cell = np.kron(mask, np.ones((32, 32), np.uint8))[:ny, :nx].astype(bool)
    complete = np.where(cell[:, :, None, None], g, 0).astype(np.float32)
    sparse = np.where(cell[:, :, None, None], g, np.nan).astype(np.float32)
    norm = torch.tensor(rng.uniform(1, 2, (1, P)), device="cuda")
    a = torch.empty((1, P, ny, nx), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    ctx.finalize(p, P, torch.tensor(complete[None], device="cuda"), norm, a)
    ctx.finalize(p, P, torch.tensor(sparse[None], device="cuda"), norm, b,
                 tile_mask=torch.tensor(mask[None], device="cuda"))
    assert torch.equal(a, b)


def test_image_at_a_size_that_is_not_a_power_of_two(orc, mti, ctx):
    """End to end (sort, kernels, binning, gridding, finalize) at 600 x 480."""
    from mtimager_amd import synth
    ob = synth.small_array_obs(seed=4, ntime=60, freqs_hz=[1.45e9])
    p = mti.make_params(600, 480, 2.0, 2.0, "wproj", 16, 4)
    dev = H.Dev(ob, torch)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    K = ctx.build_kernels(p)
    got = H.gpu_image(mti, ctx, K, p, dev, torch, [0], index=idx, weighting="briggs", robust=0.0,
                      precision="c128")
    OK = orc.Kernels(H.oracle_params(orc, p))
    ref = H.oracle_image(orc, OK, H.oracle_params(orc, p), ob, 0, index=idx.cpu().numpy(),
                         weighting="briggs", robust=0.0)
    e = H.image_errors(got["image"][0][0], ref["image"][0])
    print(f"\n600x480 Stokes I {e}")
    assert e["rel_rms"] < 1e-4


def test_finalize_jobs_mixed_sizes_match_oracle(orc, mti, ctx):
    """One batched finalize over jobs of 300^2, 512^2 (fast path) and 1000 x 768."""
    rng = np.random.default_rng(9)
    sizes = [(300, 300), (512, 512), (1000, 768), (300, 300)]
    jobs, refs, keep = [], [], []
    for nx, ny in sizes:
        p = mti.make_params(nx, ny, 2.0, 2.0, "wproj", 8, 4)
        g = _random_grid(rng, nx, ny, 1)
        norm = rng.uniform(1, 2, 1)
        refs.append(orc.finalize(H.oracle_params(orc, p), g.astype(np.float64), norm, "linear")[0])
        dg = torch.tensor(g[None], device="cuda")
        dn = torch.tensor(norm[None], device="cuda")
        img = torch.empty((1, 1, ny, nx), dtype=torch.float32, device="cuda")
        keep += [dg, dn, img]
        j = mti.Job()  # finalize only: sizes, grids, normalisers and image
        j.params, j.nchan, j.grids_c64 = p, 1, 1
        j.vis.npol, j.vis.nchan = 1, 1
        j.grids, j.norm, j.image = dg.data_ptr(), dn.data_ptr(), img.data_ptr()
        jobs.append(j)
    ctx.finalize_jobs(jobs, poltype="linear")
    torch.cuda.synchronize()
    for j, ref in enumerate(refs):
        out = keep[3 * j + 2].cpu().numpy()[0, 0]
        assert H.rel_rms(out, ref) < 1e-5, j
