"""Launch-shape tuning on the GPU: grids and images do not depend on the
gridder's region shape, its work-unit size or the transform's column strategy
beyond summation order, and the tuning is validated."""
import numpy as np
import pytest

import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def job(mti, ctx):
    from mtimager_amd import synth
    ob = synth.make_observation(nant=27, radius_m=3000.0, ntime=120, freqs_hz=np.array([1.40e9, 1.45e9]),
                                npol=4, fov_rad=256 * 4.0 * synth.ARCSEC, seed=51)
    dev = H.Dev(ob, torch)
    p = mti.make_params(256, 256, 4.0, 4.0, "wproj", 16, 8)
    idx = H.gpu_sort(ctx, dev, ob.nant, torch)
    v = mti.vis_ms(dev.uvw, dev.invl, dev.vis, dev.flags, dev.weight, idx)
    return dict(p=p, K=ctx.build_kernels(p), v=v)


def _image(mti, c, job, weighting="briggs"):
    p, K, v = job["p"], job["K"], job["v"]
    g = torch.zeros((2, 256, 256, 4), dtype=torch.complex128, device="cuda")
    n = torch.zeros((2, 4), dtype=torch.float64, device="cuda")
    im = torch.zeros((2, 4, 256, 256), dtype=torch.float32, device="cuda")
    st = c.grid(K, p, v, g, n, nchan=2, weighting=weighting, precision="c128")
    c.finalize(p, 4, g, n, im, nchan=2, precision="c128")
    torch.cuda.synchronize()
    return g.cpu().numpy(), n.cpu().numpy(), im.cpu().numpy(), st


def test_tuning_is_validated(mti):
    c = mti.Context(0)
    try:
        assert c.get_tuning() == {"unit_pairs": 2048, "whole_columns_ny": 256, "xcd_chunk": 0, "region_width": 8}
        for bad in (dict(unit_pairs=300), dict(unit_pairs=128), dict(unit_pairs=32768), dict(whole_columns_ny=100),
                    dict(whole_columns_ny=512), dict(xcd_chunk=3), dict(xcd_chunk=8192), dict(region_width=2),
                    dict(region_width=32), dict(region_width=12)):
            with pytest.raises(mti.MTIError):
                c.set_tuning(**bad)
        c.set_tuning(unit_pairs=512, whole_columns_ny=0, xcd_chunk=16, region_width=16)
        assert c.get_tuning() == {"unit_pairs": 512, "whole_columns_ny": 0, "xcd_chunk": 16, "region_width": 16}
        c.set_tuning(region_width=0)  # 0: the default
        assert c.get_tuning()["region_width"] == 8
    finally:
        c.close()


@pytest.mark.parametrize("tuning", [dict(unit_pairs=256), dict(unit_pairs=16384), dict(whole_columns_ny=0),
                                    dict(unit_pairs=1024, whole_columns_ny=128), dict(xcd_chunk=1),
                                    dict(xcd_chunk=32, unit_pairs=256), dict(xcd_chunk=4096), dict(region_width=4),
                                    dict(region_width=16), dict(region_width=16, unit_pairs=256, xcd_chunk=8)])
def test_results_do_not_depend_on_tuning(mti, ctx, job, tuning):
    g0, n0, i0, s0 = _image(mti, ctx, job)
    c = mti.Context(0)
    try:
        c.set_tuning(**tuning)
        g1, n1, i1, s1 = _image(mti, c, job)
    finally:
        c.close()
    assert s1.gridded_records == s0.gridded_records and s1.gridded_points == s0.gridded_points
    np.testing.assert_allclose(n1, n0, rtol=1e-12)
    assert H.rel_rms(g1, g0) < 1e-12
    assert H.rel_rms(i1, i0) < 1e-9


def test_small_units_split_crowded_tiles(mti, ctx, job):
    """With 256-pair units the crowded tiles of the inner uv plane take several
    work units each (the work-item statistic grows), with the same grids."""
    _, _, _, s0 = _image(mti, ctx, job, "natural")
    c = mti.Context(0)
    try:
        c.set_tuning(unit_pairs=256)
        _, _, _, s1 = _image(mti, c, job, "natural")
    finally:
        c.close()
    assert s1.work_items > s0.work_items


@pytest.mark.parametrize("region_width", [4, 16])
@pytest.mark.parametrize("accumulation", ["f64", "f32k"])
def test_region_shapes_with_complex64_grids(mti, ctx, job, region_width, accumulation):
    """The region variants of every accumulation form, with complex64 grids
    (cells of tiles split over several units are FP32 atomic sums in no fixed
    order): the same grids and normalisers as the 8 x 8 default."""
    p, K, v = job["p"], job["K"], job["v"]

    def run(c):
        g = torch.zeros((2, 256, 256, 4), dtype=torch.complex64, device="cuda")
        n = torch.zeros((2, 4), dtype=torch.float64, device="cuda")
        c.grid(K, p, v, g, n, nchan=2, weighting="briggs", precision="c64", accumulation=accumulation)
        torch.cuda.synchronize()
        return g.cpu().numpy(), n.cpu().numpy()

    g0, n0 = run(ctx)
    c = mti.Context(0)
    try:
        c.set_tuning(region_width=region_width, unit_pairs=256)
        g1, n1 = run(c)
    finally:
        c.close()
    np.testing.assert_allclose(n1, n0, rtol=1e-12)
    assert H.rel_rms(g1, g0) < 1e-6
