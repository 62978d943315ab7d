"""The visibility container (include/mtimager_amd_ingest.h) on the CPU: write,
map and read back every column the reference's MS loader reads
(vis/MSVisibilityDataSetAsync.cpp:44-347), both input orders, header
validation, and wrapped host arrays.  No device work."""
import numpy as np
import pytest

from mtimager_amd import MTIError, ingest, synth


@pytest.fixture(scope="module")
def ob():
    return synth.small_array_obs(seed=3, ntime=20, freqs_hz=np.linspace(1.4e9, 1.5e9, 5))


def _write(path, ob, ms_order=True):
    vis, flags = ob.vis, ob.flags
    if not ms_order:
        vis, flags = vis.transpose(1, 0, 2).copy(), flags.transpose(1, 0, 2).copy()
    ingest.write_dataset(path, ob.time, ob.ant1, ob.ant2, ob.uvw, ob.weight, ob.freq, vis, flags, ob.nant,
                         poltype="linear", ra=1.25, dec=-0.52, obs_itrf=(5109360.1, 2006852.6, -3238948.1),
                         ms_order=ms_order)


@pytest.mark.parametrize("ms_order", [True, False])
def test_round_trip(tmp_path, ob, ms_order):
    path = tmp_path / "obs.mtv"
    _write(path, ob, ms_order)
    ds = ingest.Dataset.open(path)
    assert (ds.nrows, ds.nchan, ds.npol, ds.nant) == (ob.nrows, len(ob.freq), 4, ob.nant)
    assert ds.poltype == "linear" and (ds.ra, ds.dec) == (1.25, -0.52)
    assert ds.obs_itrf == (5109360.1, 2006852.6, -3238948.1)
    np.testing.assert_array_equal(ds.column("TIME"), ob.time)
    np.testing.assert_array_equal(ds.column("ANTENNA1"), ob.ant1)
    np.testing.assert_array_equal(ds.column("ANTENNA2"), ob.ant2)
    np.testing.assert_array_equal(ds.column("UVW"), ob.uvw)
    np.testing.assert_array_equal(ds.column("WEIGHT"), ob.weight)
    np.testing.assert_array_equal(ds.column("CHAN_FREQ"), ob.freq)
    np.testing.assert_array_equal(ds.column("DATA"), ob.vis.transpose(1, 0, 2))
    np.testing.assert_array_equal(ds.column("FLAG"), ob.flags.transpose(1, 0, 2))
    ds.close()


def test_columns_are_256_aligned(tmp_path, ob):
    path = tmp_path / "obs.mtv"
    _write(path, ob)
    raw = np.fromfile(path, dtype=np.uint8)
    assert bytes(raw[:8]) == b"MTIVIS01"
    offs = raw[80:80 + 64].view(np.uint64)  # after magic, version, sizes, frame
    assert np.all(offs % 256 == 0) and np.all(np.diff(offs.astype(np.int64)) > 0)


def test_rejects_bad_files(tmp_path, ob):
    path = tmp_path / "obs.mtv"
    _write(path, ob)
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad.mtv"
    bad.write_bytes(b"NOTMTIV1" + bytes(raw[8:]))
    with pytest.raises(MTIError, match="magic"):
        ingest.Dataset.open(bad)
    bad.write_bytes(bytes(raw[: len(raw) // 2]))
    with pytest.raises(MTIError, match="outside the file"):
        ingest.Dataset.open(bad)
    bad.write_bytes(bytes(raw[:100]))
    with pytest.raises(MTIError, match="too short"):
        ingest.Dataset.open(bad)
    with pytest.raises(MTIError):
        ingest.Dataset.open(tmp_path / "missing.mtv")


def test_wrap_points_at_the_arrays(ob):
    vis_cm = ob.vis.transpose(1, 0, 2).copy()
    fl_cm = ob.flags.transpose(1, 0, 2).copy()
    ds = ingest.Dataset.wrap(ob.time, ob.ant1, ob.ant2, ob.uvw, ob.weight, ob.freq, vis_cm, fl_cm, ob.nant)
    assert ds.column("DATA").ctypes.data == vis_cm.ctypes.data
    np.testing.assert_array_equal(ds.column("UVW"), ob.uvw)
    assert ds.nrows == ob.nrows and ds.nchan == len(ob.freq)
