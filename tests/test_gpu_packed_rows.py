"""The packed row passes against the FP64 oracle's finalize on rectangular
power-of-two images: k_rows_pk_pipe (persistent, pipelined; complex64 grids,
rows of 256 to 4096) and k_rows_pk (complex128 grids), each with complete
and sparse grids, both grid layouts, one correlation with an odd channel count
(the last pack holds one channel), two correlations and four (linear and
circular).  The reference transforms each correlation with cuFFT and converts
to Stokes afterwards (gafw/StandardOperators/IFFT2D.cpp:86-111,
mt-imager/lib/operators/PostIFFTStokesConverter.cu:152-234); the packing is an FP64 identity, so the images match the oracle to
the rounding of the complex64 grids or transform planes and of the float32
image."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


# (nx, ny, npol, poltype, nchan, layout, sparse): rows of 4096 (one-row
# units), 2048 and 1024 (row pairs with 4-element read-ahead), 512 and 256
CASES = [
    (4096, 256, 2, "linear", 1, "vup", True),
    (2048, 512, 4, "circular", 1, "pvu", True),
    (1024, 2048, 1, "linear", 3, "vup", False),
    (512, 4096, 4, "linear", 1, "pvu", False),
    (256, 1024, 1, "linear", 3, "pvu", True),
]


def _grids(rng, nx, ny, npol, nchan, sparse):
    """Random complex grids [nchan][ny][nx][npol] (f32 values) and, if sparse,
    a tile mask with the unmarked tiles' cells set to NaN (never read)."""
    g = rng.standard_normal((nchan, ny, nx, npol, 2)).astype(np.float32)
    mask = None
    if sparse:
        mask = (rng.random((nchan, ny // 32, nx // 32)) < 0.5).astype(np.uint8)
        cell = np.kron(mask, np.ones((1, 32, 32), np.uint8)).astype(bool)
        g = np.where(cell[..., None, None], g, np.nan).astype(np.float32)
    return g, mask


@pytest.mark.parametrize("storage", ["c64", "c128"])
@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("nx,ny,npol,poltype,nchan,layout,sparse", CASES)
def test_packed_rows_match_oracle(orc, mti, ctx, nx, ny, npol, poltype, nchan, layout, sparse, precision, storage):
    print(f"\npacked rows {nx}x{ny} P={npol} {poltype} nchan={nchan} {layout} sparse={sparse} grids {precision}"
          f" planes {storage}", flush=True)
    p = mti.make_params(nx, ny, 1.5, 1.5, "wproj", 8, 8)
    rng = np.random.default_rng(nx + 3 * ny + npol)
    g, mask = _grids(rng, nx, ny, npol, nchan, sparse)
    norm = rng.uniform(1.0, 3.0, (nchan, npol))
    dev = g if layout == "vup" else np.ascontiguousarray(g.transpose(0, 3, 1, 2, 4))
    dt = torch.float32 if precision == "c64" else torch.float64
    d_g = torch.tensor(dev, dtype=dt, device="cuda")
    d_n = torch.tensor(norm, device="cuda")
    d_m = torch.tensor(mask, device="cuda") if sparse else None
    img = torch.empty((nchan, npol, ny, nx), dtype=torch.float32, device="cuda")
    before = ctx.get_transform_storage()
    ctx.set_transform_storage(storage)
    try:
        ctx.finalize(p, npol, d_g, d_n, img, nchan=nchan, poltype=poltype, layout=layout, precision=precision,
                     tile_mask=d_m)
        torch.cuda.synchronize()
    finally:
        ctx.set_transform_storage(before)
    out = img.cpu().numpy()
    op = H.oracle_params(orc, p)
    ref_in = np.nan_to_num(g.astype(np.float64), nan=0.0)  # unmarked tiles are zeros
    for c in range(nchan):
        ref = orc.finalize(op, ref_in[c], norm[c], poltype)
        for q in range(npol):
            e = H.rel_rms(out[c, q], ref[q])
            assert np.isfinite(out[c, q]).all() and e < 1e-5, (c, q, e)
