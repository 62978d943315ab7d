"""Output shapes and workspace sizes of the C-ABI (the validate() half of the
operator contract): host-only queries, no context and no device."""
import ctypes

import pytest

import mtimager_amd as M


def _params(nx=512, ny=None, wplanes=32, sampling=4):
    return M.make_params(nx, ny or nx, 2.0, 2.0, "wproj", wplanes, sampling)


def _vis(nrows, npol=4, nchan=8, row_index=False):
    v = M.Vis()
    v.nrows, v.nchan, v.npol = nrows, nchan, npol
    if row_index:  # a non-null pointer: the index length counts, the pointer is not read
        v.row_index = ctypes.c_void_p(1)
        v.nindex = nrows // 2
    return v


def _info(max_support=33, sampling=8):
    k = M.KernelInfo()
    k.nplanes, k.sampling, k.T, k.max_support = 256, sampling, sampling + 1, max_support
    return k


def test_grid_and_image_shapes_follow_the_layouts():
    p = _params(256, 128)
    assert M.grid_out_shape(p, 4, 3, "vup") == ((3, 128, 256, 4), (3, 4))
    assert M.grid_out_shape(p, 2, 1, "pvu") == ((1, 2, 128, 256), (1, 2))
    assert M.finalize_out_shape(p, 4, 5) == (5, 4, 128, 256)
    assert M.preprocess_out_shape(1000) == (1000, 8)


@pytest.mark.parametrize("bad", [dict(npol=3), dict(nchan=0)])
def test_shapes_reject_what_validation_rejects(bad):
    a = dict(npol=4, nchan=1)
    a.update(bad)
    with pytest.raises(M.MTIError):
        M.grid_out_shape(_params(), a["npol"], a["nchan"])
    with pytest.raises(M.MTIError):
        M.finalize_out_shape(_params(), a["npol"], a["nchan"])


def test_finalize_shape_takes_any_size():
    # the reference plans its transform for any rows x cols (IFFT2D.cpp:86-111)
    assert M.finalize_out_shape(M.make_params(300, 257, 2.0, 2.0, "wproj", 8, 4), 1, 1) == (1, 1, 257, 300)
    with pytest.raises(M.MTIError):
        M.finalize_out_shape(M.make_params(1, 256, 2.0, 2.0, "wproj", 8, 4), 1, 1)


def test_fft2d_workspace_holds_the_batch_planes():
    # FP64 planes X, Y of every transform of the batch plus the tables;
    # Bluestein lengths (a prime factor above 13) add two planes of 2^k >= 2n-1
    for nx, ny, nb in [(64, 128, 3), (4096, 4096, 1), (100, 64, 1)]:
        ws = M.fft2d_workspace_size(nx, ny, nb)
        assert 2 * nb * nx * ny * 16 <= ws < 2 * nb * nx * ny * 16 + (nx + ny) * 16 + 4096
    ws = M.fft2d_workspace_size(17 * 3, 8, 1)  # 51 = 3 * 17: Bluestein over 128
    assert ws >= (2 * 51 * 8 + 2 * 128 * 8) * 16
    with pytest.raises(M.MTIError):
        M.fft2d_workspace_size(0, 64, 1)


def test_sort_rows_workspace_grows_with_rows():
    a, b = M.sort_rows_workspace_size(1000), M.sort_rows_workspace_size(100000)
    assert 12 * 1000 <= a < b and b >= 12 * 100000
    assert M.sort_rows_workspace_size(0) < a


def test_finalize_workspace_holds_the_packed_transform_batch():
    p = _params(1024)
    # four correlations: two packed complex128 planes per channel
    w1, w3 = M.finalize_workspace_size(p, 4, 1), M.finalize_workspace_size(p, 4, 3)
    assert w1 >= 2 * 1024 * 1024 * 16 and w3 >= 6 * 1024 * 1024 * 16 > w1
    # one correlation: two channels share a packed plane
    assert M.finalize_workspace_size(p, 1, 2) == M.finalize_workspace_size(p, 1, 1)
    assert M.finalize_workspace_size(p, 1, 3) > M.finalize_workspace_size(p, 1, 2)


def test_grid_workspace_properties():
    p, k = _params(4096, 4096, 256, 8), _info()
    nat = M.grid_workspace_size(k, p, _vis(100000), nchan=8, weighting="natural")
    bri = M.grid_workspace_size(k, p, _vis(100000), nchan=8, weighting="briggs")
    assert bri > nat  # per-record density
    more = M.grid_workspace_size(k, p, _vis(200000), nchan=8)
    assert more > nat
    # records are (row, channel) pairs: 8 channels of 1e5 rows need at least
    # the (record, tile) pairs of 8e5 records (4 tiles each at S = 33)
    assert nat >= 8 * 100000 * 4 * 4
    # complex64 grids are written directly (no FP64 work grid): the same plan
    c64 = M.grid_workspace_size(k, p, _vis(100000), nchan=8, precision="c64")
    assert c64 == nat
    # an index selects the rows: half the rows, a smaller plan
    idx = M.grid_workspace_size(k, p, _vis(100000, row_index=True), nchan=8)
    assert idx < nat
    # no rows: statistics only
    assert M.grid_workspace_size(k, p, _vis(0), nchan=8) < 4096


def test_grid_workspace_rejects_bad_arguments():
    with pytest.raises(M.MTIError):
        M.grid_workspace_size(_info(), _params(), _vis(10, npol=3))
    with pytest.raises(M.MTIError):
        M.grid_workspace_size(_info(), _params(), _vis(10, nchan=2), chan0=1, nchan=2)
