"""FITS image-cube output (mti_fits_* of include/mtimager_amd_fits.h): the
image cube mti_finalize writes (f32 [nchan][npol][ny][nx]) as one primary
HDU with the header conventions of the reference's ImageManager
(comp/ImageManager.cpp:68-215): RA---SIN / DEC--SIN axes with CDELT1 = -cell,
reference pixel N/2 (+1 in FITS's count), a STOKES and a FREQ axis."""
import ctypes

from . import MTIError, lib

STOKES_CODES = {"I": 1, "Q": 2, "U": 3, "V": 4}
# the Stokes planes mti_finalize writes for each correlation count and type
FINALIZE_STOKES = {(4, "linear"): "IQUV", (4, "circular"): "IQUV", (2, "linear"): "IQ",
                   (2, "circular"): "IV", (1, "linear"): "I", (1, "circular"): "I"}


class FitsDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nstokes", ctypes.c_int32),
                ("nchan", ctypes.c_int32), ("stokes", ctypes.c_int32 * 4),
                ("cell_l", ctypes.c_double), ("cell_m", ctypes.c_double),
                ("ra", ctypes.c_double), ("dec", ctypes.c_double),
                ("freq0", ctypes.c_double), ("dfreq", ctypes.c_double), ("equinox", ctypes.c_double),
                ("specsys", ctypes.c_char_p), ("telescope", ctypes.c_char_p),
                ("observer", ctypes.c_char_p), ("date_obs", ctypes.c_char_p)]


def _bind():
    L = lib()
    if not getattr(L, "_fits_bound", False):
        vp = ctypes.c_void_p
        L.mti_fits_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(FitsDesc), ctypes.POINTER(vp),
                                      ctypes.POINTER(ctypes.c_char_p)]
        L.mti_fits_create.restype = ctypes.c_int
        L.mti_fits_write_channel.argtypes = [vp, ctypes.c_int32, vp]
        L.mti_fits_write_channel.restype = ctypes.c_int
        L.mti_fits_close.argtypes = [vp]
        L.mti_fits_close.restype = ctypes.c_int
        L._fits_bound = True
    return L


def _b(s):
    return None if s is None else str(s).encode()


class CubeWriter:
    """Writes an image cube channel by channel (channels in any order).

    params: ImageParams (nx, ny, cell_l, cell_m in radians); freqs_hz: the
    channel frequencies (equally spaced; one channel needs channel_width);
    stokes: the planes of each channel, e.g. "IQUV" (FINALIZE_STOKES gives
    what mti_finalize writes); ra/dec: phase centre in radians."""

    def __init__(self, path, params, freqs_hz, stokes="I", ra=0.0, dec=0.0, channel_width=None,
                 equinox=2000.0, specsys="LSRK", telescope=None, observer=None, date_obs=None):
        import numpy as np
        f = np.asarray(freqs_hz, dtype=np.float64).ravel()
        if len(f) > 1:
            df = np.diff(f)
            if not np.allclose(df, df[0], rtol=1e-9, atol=0.0):
                raise MTIError("channel frequencies must be equally spaced for a FITS FREQ axis")
            dfreq = float(df[0])
        else:
            dfreq = float(channel_width if channel_width is not None else 1.0)
        codes = [STOKES_CODES[s] for s in stokes]
        d = FitsDesc()
        d.nx, d.ny, d.nstokes, d.nchan = params.nx, params.ny, len(codes), len(f)
        for i, c in enumerate(codes):
            d.stokes[i] = c
        d.cell_l, d.cell_m, d.ra, d.dec = params.cell_l, params.cell_m, float(ra), float(dec)
        d.freq0, d.dfreq, d.equinox = float(f[0]), dfreq, float(equinox)
        d.specsys, d.telescope, d.observer, d.date_obs = _b(specsys), _b(telescope), _b(observer), _b(date_obs)
        self.h = ctypes.c_void_p()
        err = ctypes.c_char_p()
        rc = _bind().mti_fits_create(str(path).encode(), ctypes.byref(d), ctypes.byref(self.h), ctypes.byref(err))
        if rc != 0:
            raise MTIError(f"mti_fits_create: {err.value.decode() if err.value else 'invalid description'}")
        self.shape = (len(f), len(codes), params.ny, params.nx)

    def write_channel(self, chan, image):
        """image: f32 [npol][ny][nx], a numpy array or a torch tensor (on any device)."""
        import numpy as np
        if hasattr(image, "detach"):
            image = image.detach().to("cpu").numpy()
        a = np.ascontiguousarray(image, dtype=np.float32)
        if a.shape != self.shape[1:]:
            raise MTIError(f"channel image shape {a.shape}, expected {self.shape[1:]}")
        rc = _bind().mti_fits_write_channel(self.h, int(chan), a.ctypes.data)
        if rc != 0:
            raise MTIError(f"mti_fits_write_channel: status {rc}")

    def close(self):
        if self.h:
            rc = _bind().mti_fits_close(self.h)
            self.h = ctypes.c_void_p()
            if rc != 0:
                raise MTIError(f"mti_fits_close: status {rc}")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_cube(path, params, image, freqs_hz, stokes="I", **kw):
    """The whole cube image [nchan][npol][ny][nx] into one file."""
    with CubeWriter(path, params, freqs_hz, stokes, **kw) as w:
        for c in range(image.shape[0]):
            w.write_channel(c, image[c])


def read_cube(path):
    """(header dict, data [NAXIS4][NAXIS3][NAXIS2][NAXIS1] float32) of a file
    this module wrote -- a minimal reader for checks, not a general FITS one."""
    import numpy as np
    with open(path, "rb") as fh:
        raw = fh.read()
    hdr, pos = {}, 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            val = card[10:].split(" / ")[0].strip()
            if val.startswith("'"):
                hdr[key] = val[1:val.rindex("'")].rstrip().replace("''", "'")
            elif val in ("T", "F"):
                hdr[key] = val == "T"
            else:
                hdr[key] = float(val) if any(ch in val for ch in ".E") else int(val)
    start = (pos + 2879) // 2880 * 2880
    shape = tuple(hdr[f"NAXIS{i}"] for i in range(hdr["NAXIS"], 0, -1))
    n = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=">f4", count=n, offset=start).reshape(shape)
    return hdr, data
