"""Visibility container and chunked loader (include/mtimager_amd_ingest.h).

The reference ingests a MeasurementSet through casacore
(vis/MSVisibilityDataSetAsync.cpp:44-347) and chunks it per channel into
`records`-row pieces (vis/StorageManager.h:111-150).  Here a dataset is an
"MTIVIS01" container file (write_dataset / Dataset.open) or host arrays in the
same column layout (Dataset.wrap); a Loader streams it to the device in chunks
(channel block x row run) through pinned staging buffers and grids every chunk
with mti_grid, accumulating the row chunks of one channel block.
"""
import ctypes

import numpy as np

from . import (GRID_PRECISION, LAYOUT, POLTYPE, WEIGHTING, GridStats, MTIError, _stream, lib)

COLUMNS = ["TIME", "ANTENNA1", "ANTENNA2", "UVW", "WEIGHT", "CHAN_FREQ", "DATA", "FLAG"]


class DatasetInfo(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("nchan", ctypes.c_int32), ("npol", ctypes.c_int32),
                ("nant", ctypes.c_int32), ("poltype", ctypes.c_int32), ("ra", ctypes.c_double),
                ("dec", ctypes.c_double), ("obs_itrf", ctypes.c_double * 3)]


class LoaderOpts(ctypes.Structure):
    _fields_ = [("chan0", ctypes.c_int32), ("nchan", ctypes.c_int32), ("channel_block", ctypes.c_int32),
                ("chunk_rows", ctypes.c_int64), ("lsrk", ctypes.c_int32), ("tt_minus_utc", ctypes.c_double),
                ("psf", ctypes.c_int32), ("weighting", ctypes.c_int32), ("robust", ctypes.c_double),
                ("grid_layout", ctypes.c_int32), ("grid_precision", ctypes.c_int32),
                ("reader_threads", ctypes.c_int32)]


class LoaderStats(ctypes.Structure):
    _fields_ = [("chunks", ctypes.c_int64), ("h2d_bytes", ctypes.c_int64), ("seconds", ctypes.c_double),
                ("reader_seconds", ctypes.c_double), ("grid", GridStats)]


EXPORTS = ["mti_dataset_write", "mti_dataset_open", "mti_dataset_wrap", "mti_dataset_get_info",
           "mti_dataset_column", "mti_dataset_close", "mti_loader_create", "mti_loader_grid",
           "mti_loader_last_error", "mti_loader_destroy"]


def _bind():
    L = lib()
    if getattr(L, "_ingest_bound", False):
        return L
    vp, P, cp = ctypes.c_void_p, ctypes.POINTER, ctypes.c_char_p
    L.mti_dataset_write.argtypes = [cp, P(DatasetInfo), vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int32, P(cp)]
    L.mti_dataset_open.argtypes = [cp, P(vp), P(cp)]
    L.mti_dataset_wrap.argtypes = [P(DatasetInfo), vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int32, P(vp)]
    L.mti_dataset_get_info.argtypes = [vp, P(DatasetInfo)]
    L.mti_dataset_column.argtypes = [vp, ctypes.c_int32]
    L.mti_dataset_column.restype = vp
    L.mti_dataset_close.argtypes = [vp]
    L.mti_dataset_close.restype = None
    L.mti_loader_create.argtypes = [vp, vp, P(LoaderOpts), P(vp)]
    L.mti_loader_grid.argtypes = [vp, vp, vp, vp, vp, vp, P(LoaderStats), vp]
    L.mti_loader_last_error.argtypes = [vp]
    L.mti_loader_last_error.restype = cp
    L.mti_loader_destroy.argtypes = [vp]
    L.mti_loader_destroy.restype = None
    for n in ("mti_dataset_write", "mti_dataset_open", "mti_dataset_wrap", "mti_dataset_get_info",
              "mti_loader_create", "mti_loader_grid"):
        getattr(L, n).restype = ctypes.c_int
    L._ingest_bound = True
    return L


def _info(nrows, nchan, npol, nant, poltype="linear", ra=0.0, dec=0.0, obs_itrf=(0.0, 0.0, 0.0)):
    return DatasetInfo(int(nrows), int(nchan), int(npol), int(nant), POLTYPE[poltype], float(ra), float(dec),
                       (ctypes.c_double * 3)(*[float(x) for x in obs_itrf]))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def write_dataset(path, time, ant1, ant2, uvw, weight, chan_freq, vis, flags, nant, poltype="linear",
                  ra=0.0, dec=0.0, obs_itrf=(0.0, 0.0, 0.0), ms_order=True):
    """Writes a container: vis complex64 / flags u8 in MS order [nrows][nchan][npol]
    (ms_order=True) or channel-major [nchan][nrows][npol]."""
    L = _bind()
    flags = _c(flags, np.uint8)
    nrows, nchan, npol = (flags.shape if ms_order else (flags.shape[1], flags.shape[0], flags.shape[2]))
    cols = [_c(time, np.float64), _c(ant1, np.int32), _c(ant2, np.int32), _c(uvw, np.float64),
            _c(weight, np.float32), _c(chan_freq, np.float64), _c(vis, np.complex64), flags]
    err = ctypes.c_char_p()
    info = _info(nrows, nchan, npol, nant, poltype, ra, dec, obs_itrf)
    rc = L.mti_dataset_write(str(path).encode(), ctypes.byref(info), *[c.ctypes.data for c in cols],
                             1 if ms_order else 0, ctypes.byref(err))
    if rc != 0:
        raise MTIError(f"mti_dataset_write failed (status {rc}): {(err.value or b'').decode()}")


class Dataset:
    """A mapped container file (open) or described host arrays (wrap)."""

    def __init__(self, h, keep=()):
        self.h = h
        self._keep = keep
        info = DatasetInfo()
        _bind().mti_dataset_get_info(self.h, ctypes.byref(info))
        self.nrows, self.nchan, self.npol, self.nant = info.nrows, info.nchan, info.npol, info.nant
        self.poltype = {v: k for k, v in POLTYPE.items()}[info.poltype]
        self.ra, self.dec, self.obs_itrf = info.ra, info.dec, tuple(info.obs_itrf)

    @classmethod
    def open(cls, path):
        L = _bind()
        h, err = ctypes.c_void_p(), ctypes.c_char_p()
        rc = L.mti_dataset_open(str(path).encode(), ctypes.byref(h), ctypes.byref(err))
        if rc != 0:
            raise MTIError(f"mti_dataset_open failed (status {rc}): {(err.value or b'').decode()}")
        return cls(h)

    @classmethod
    def wrap(cls, time, ant1, ant2, uvw, weight, chan_freq, vis_chan_major, flags_chan_major, nant,
             poltype="linear", ra=0.0, dec=0.0, obs_itrf=(0.0, 0.0, 0.0), pinned=False):
        """Host arrays (numpy, or CPU torch tensors -- pinned ones with
        pinned=True) in the container's column layout; kept referenced."""
        def ptr(a):
            return a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        nchan, nrows, npol = tuple(flags_chan_major.shape)
        info = _info(nrows, nchan, npol, nant, poltype, ra, dec, obs_itrf)
        arrs = (time, ant1, ant2, uvw, weight, chan_freq, vis_chan_major, flags_chan_major)
        h = ctypes.c_void_p()
        rc = _bind().mti_dataset_wrap(ctypes.byref(info), *[ptr(a) for a in arrs], 1 if pinned else 0,
                                      ctypes.byref(h))
        if rc != 0:
            raise MTIError(f"mti_dataset_wrap failed (status {rc})")
        return cls(h, arrs)

    def column(self, name):
        """numpy view of a column (valid while the dataset is open)."""
        k = COLUMNS.index(name)
        p = _bind().mti_dataset_column(self.h, k)
        n, C, P = self.nrows, self.nchan, self.npol
        shape, dt = {
            "TIME": ((n,), np.float64), "ANTENNA1": ((n,), np.int32), "ANTENNA2": ((n,), np.int32),
            "UVW": ((n, 3), np.float64), "WEIGHT": ((n, P), np.float32), "CHAN_FREQ": ((C,), np.float64),
            "DATA": ((C, n, P), np.complex64), "FLAG": ((C, n, P), np.uint8)}[name]
        count = int(np.prod(shape))
        buf = (ctypes.c_char * (count * np.dtype(dt).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)

    def close(self):
        if self.h:
            _bind().mti_dataset_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Loader:
    """Chunked streaming of a dataset into gridding calls on one context."""

    def __init__(self, ctx, dataset, chan0=0, nchan=None, channel_block=8, chunk_rows=0, lsrk=False,
                 tt_minus_utc=69.184, psf=False, weighting="natural", robust=0.0, layout="vup",
                 precision="c64", reader_threads=8):
        L = _bind()
        self.ctx, self.ds = ctx, dataset
        nchan = dataset.nchan - chan0 if nchan is None else nchan
        self.opts = LoaderOpts(int(chan0), int(nchan), int(channel_block), int(chunk_rows), 1 if lsrk else 0,
                               float(tt_minus_utc), 1 if psf else 0, WEIGHTING[weighting], float(robust),
                               LAYOUT[layout], GRID_PRECISION[precision], int(reader_threads))
        self.h = ctypes.c_void_p()
        rc = L.mti_loader_create(ctx.h, dataset.h, ctypes.byref(self.opts), ctypes.byref(self.h))
        if rc != 0:
            raise MTIError(f"mti_loader_create failed (status {rc}): {L.mti_loader_last_error(None).decode()}")

    def grid(self, kern, params, grids, norm, tile_mask=None, stream=None):
        """Streams every chunk into grids ([nchan] grids of the loader's layout
        and precision) and norm (f64 [nchan][npol]); returns the statistics."""
        L = _bind()
        st = LoaderStats()
        rc = L.mti_loader_grid(self.h, kern.h, ctypes.byref(params), grids.data_ptr(), norm.data_ptr(),
                               tile_mask.data_ptr() if tile_mask is not None else None, ctypes.byref(st),
                               _stream(stream))
        if rc != 0:
            raise MTIError(f"mti_loader_grid failed (status {rc}): {L.mti_loader_last_error(self.h).decode()}")
        return dict(chunks=st.chunks, h2d_bytes=st.h2d_bytes, seconds=st.seconds, reader_seconds=st.reader_seconds,
                    gridded_records=st.grid.gridded_records, gridded_points=st.grid.gridded_points)

    def close(self):
        if self.h:
            _bind().mti_loader_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
