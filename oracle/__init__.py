"""ctypes front-end of the FP64 oracle (oracle/mti_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as a product code path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmti_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "libmti_oracle.so"])


class Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("cell_l", ctypes.c_double), ("cell_m", ctypes.c_double),
                ("mode", ctypes.c_int), ("wplanes", ctypes.c_int),
                ("sampling", ctypes.c_int), ("taper_support", ctypes.c_int)]


class _Kernels(ctypes.Structure):
    _fields_ = [("nplanes", ctypes.c_int), ("sampling", ctypes.c_int),
                ("T", ctypes.c_int), ("trial_support", ctypes.c_int),
                ("w_increment", ctypes.c_double), ("divisor", ctypes.c_double),
                ("support", ctypes.POINTER(ctypes.c_int)),
                ("base", ctypes.POINTER(ctypes.c_int64)),
                ("total", ctypes.c_int64),
                ("table", ctypes.POINTER(ctypes.c_double)),
                ("sums", ctypes.POINTER(ctypes.c_double))]


class LsrkFrame(ctypes.Structure):
    _fields_ = [("obs_itrf", ctypes.c_double * 3), ("ra", ctypes.c_double),
                ("dec", ctypes.c_double), ("tt_minus_utc", ctypes.c_double)]


def lsrk_frame(obs_itrf, ra, dec, tt_minus_utc=69.184):
    return LsrkFrame((ctypes.c_double * 3)(*[float(x) for x in obs_itrf]), float(ra),
                     float(dec), float(tt_minus_utc))


BIN_DTYPE = np.dtype([("to_grid", np.int32), ("plane", np.int32),
                      ("support", np.int32), ("corner_x", np.int32),
                      ("corner_y", np.int32), ("off_x", np.int32),
                      ("off_y", np.int32), ("conj", np.int32),
                      ("conv_idx", np.int64)])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        L.orc_gridsf.restype = ctypes.c_double
        L.orc_gridsf.argtypes = [ctypes.c_double]
        L.orc_w_increment.restype = ctypes.c_double
        L.orc_w_increment.argtypes = [P(Params)]
        L.orc_trial_support.restype = ctypes.c_int
        L.orc_trial_support.argtypes = [P(Params)]
        L.orc_kernels_build.restype = ctypes.c_int
        L.orc_kernels_build.argtypes = [P(Params), P(_Kernels)]
        L.orc_kernels_free.argtypes = [P(_Kernels)]
        L.orc_wkernel_supports.restype = ctypes.c_int
        L.orc_wkernel_supports.argtypes = [P(Params), ctypes.c_void_p]
        L.orc_sort_index.restype = ctypes.c_int64
        L.orc_sort_index.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
        L.orc_preprocess.argtypes = [P(Params), P(_Kernels), ctypes.c_int64,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_void_p]
        L.orc_weights.restype = ctypes.c_int
        L.orc_weights.argtypes = [P(Params), P(_Kernels), ctypes.c_int64,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_void_p]
        L.orc_weights_mfs.restype = ctypes.c_int
        L.orc_weights_mfs.argtypes = [P(Params), P(_Kernels), ctypes.c_int64,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_void_p]
        L.orc_grid.restype = ctypes.c_int
        L.orc_grid.argtypes = [P(Params), P(_Kernels), ctypes.c_int64,
                               ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                               ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                               P(ctypes.c_int64)]
        L.orc_fft1d.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_finalize.restype = ctypes.c_int
        L.orc_finalize.argtypes = [P(Params), ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_lsrk_velocity.argtypes = [P(LsrkFrame), ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
        L.orc_earth_velocity.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_lsrk_factor.restype = ctypes.c_double
        L.orc_lsrk_factor.argtypes = [P(LsrkFrame), ctypes.c_double]
        L.orc_lsrk_factors.argtypes = [P(LsrkFrame), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.orc_set_row_doppler.argtypes = [ctypes.c_void_p]
        L.orc_model_grid.restype = ctypes.c_int
        L.orc_model_grid.argtypes = [P(Params), ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_model_correction.restype = ctypes.c_double
        L.orc_model_correction.argtypes = [P(Params), ctypes.c_int, ctypes.c_int]
        L.orc_degrid.restype = ctypes.c_int
        L.orc_degrid.argtypes = [P(Params), P(_Kernels), ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 P(ctypes.c_int64)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def make_params(nx, ny, cell_l, cell_m, mode="wproj", wplanes=1, sampling=4,
                taper_support=7):
    return Params(int(nx), int(ny), float(cell_l), float(cell_m),
                  1 if mode == "wproj" else 0, int(wplanes), int(sampling),
                  int(taper_support))


def gridsf(nu):
    return lib().orc_gridsf(float(nu))


class Kernels:
    """Owns an orc_kernels struct; exposes numpy views of its tables."""

    def __init__(self, params):
        self.params = params
        self._k = _Kernels()
        rc = lib().orc_kernels_build(ctypes.byref(params), ctypes.byref(self._k))
        if rc != 0:
            raise ValueError("orc_kernels_build failed (bad parameters)")
        k = self._k
        self.nplanes, self.sampling, self.T = k.nplanes, k.sampling, k.T
        self.trial_support = k.trial_support
        self.w_increment, self.divisor = k.w_increment, k.divisor
        self.support = np.ctypeslib.as_array(k.support, (k.nplanes,)).copy()
        self.base = np.ctypeslib.as_array(k.base, (k.nplanes,)).copy()
        tot = max(int(k.total), 1)
        tab = np.ctypeslib.as_array(k.table, (tot, 2)).copy()
        self.table = (tab[:, 0] + 1j * tab[:, 1])[: int(k.total)]
        self.sums = np.ctypeslib.as_array(
            k.sums, (k.nplanes, k.T * k.T)).copy()

    def plane(self, k):
        """Sub-kernels of plane k as array [T][T][S][S] (complex128)."""
        S, T = int(self.support[k]), self.T
        b = int(self.base[k])
        return self.table[b:b + T * T * S * S].reshape(T, T, S, S)

    def __del__(self):
        try:
            lib().orc_kernels_free(ctypes.byref(self._k))
        except Exception:
            pass


def supports(params):
    out = np.zeros(params.wplanes, dtype=np.int32)
    rc = lib().orc_wkernel_supports(ctypes.byref(params), _ptr(out))
    if rc != 0:
        raise ValueError("bad parameters")
    return out


def sort_index(ant1, ant2, nant):
    a1 = np.ascontiguousarray(ant1, dtype=np.int32)
    a2 = np.ascontiguousarray(ant2, dtype=np.int32)
    out = np.zeros(len(a1), dtype=np.int32)
    kept = lib().orc_sort_index(_ptr(a1), _ptr(a2), len(a1), int(nant), _ptr(out))
    return out[:kept]


def preprocess(kern, uvw, inv_lambda, flags, chan):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    inv_lambda = np.ascontiguousarray(inv_lambda, dtype=np.float64)
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    nrows, nchan, npol = flags.shape
    out = np.zeros(nrows, dtype=BIN_DTYPE)
    lib().orc_preprocess(ctypes.byref(kern.params), ctypes.byref(kern._k), nrows,
                         nchan, npol, _ptr(uvw), _ptr(inv_lambda), _ptr(flags),
                         int(chan), _ptr(out))
    return out


WEIGHTING = {"natural": 0, "uniform": 1, "briggs": 2}


def weights(kern, uvw, inv_lambda, flags, weight, chan, weighting="natural",
            robust=0.0):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    inv_lambda = np.ascontiguousarray(inv_lambda, dtype=np.float64)
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    weight = np.ascontiguousarray(weight, dtype=np.float32)
    nrows, nchan, npol = flags.shape
    out = np.zeros((nrows, npol), dtype=np.float64)
    lib().orc_weights(ctypes.byref(kern.params), ctypes.byref(kern._k), nrows,
                      nchan, npol, _ptr(uvw), _ptr(inv_lambda), _ptr(flags),
                      _ptr(weight), int(chan), WEIGHTING[weighting], float(robust),
                      _ptr(out))
    return out


def weights_mfs(kern, uvw, inv_lambda, flags, weight, chan0, nch, weighting="natural", robust=0.0):
    """Weights with one density over channels chan0 .. chan0 + nch - 1:
    [nch][nrows][npol]."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    inv_lambda = np.ascontiguousarray(inv_lambda, dtype=np.float64)
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    weight = np.ascontiguousarray(weight, dtype=np.float32)
    nrows, nchan, npol = flags.shape
    out = np.zeros((nch, nrows, npol), dtype=np.float64)
    lib().orc_weights_mfs(ctypes.byref(kern.params), ctypes.byref(kern._k), nrows, nchan, npol,
                          _ptr(uvw), _ptr(inv_lambda), _ptr(flags), _ptr(weight), int(chan0), int(nch),
                          WEIGHTING[weighting], float(robust), _ptr(out))
    return out


def grid(kern, uvw, inv_lambda, vis, flags, weight, chan, psf=False, wts=None,
         grid_acc=None, norm_acc=None):
    """Grid one channel; returns (grid [ny][nx][npol] complex128, norm[npol],
    gridded_records)."""
    p = kern.params
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    inv_lambda = np.ascontiguousarray(inv_lambda, dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex64)
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    weight = np.ascontiguousarray(weight, dtype=np.float32)
    if wts is not None:
        wts = np.ascontiguousarray(wts, dtype=np.float64)
    nrows, nchan, npol = flags.shape
    g = np.zeros((p.ny, p.nx, npol, 2)) if grid_acc is None else grid_acc
    n = np.zeros(npol) if norm_acc is None else norm_acc
    ng = ctypes.c_int64(0)
    lib().orc_grid(ctypes.byref(p), ctypes.byref(kern._k), nrows, nchan, npol,
                   _ptr(uvw), _ptr(inv_lambda), _ptr(vis), _ptr(flags),
                   _ptr(weight), _ptr(wts), int(chan), 1 if psf else 0, _ptr(g),
                   _ptr(n), ctypes.byref(ng))
    return g, n, int(ng.value)


def fft1d(x, sign):
    d = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().orc_fft1d(_ptr(d), len(d), int(sign))
    return d


POLTYPE = {"linear": 0, "circular": 1}


def finalize(params, grid_ri, norm, poltype="linear"):
    """grid_ri [ny][nx][npol][2] float64 -> image [npol][ny][nx] float64."""
    grid_ri = np.ascontiguousarray(grid_ri, dtype=np.float64)
    npol = grid_ri.shape[2]
    norm = np.ascontiguousarray(norm, dtype=np.float64)
    img = np.zeros((npol, params.ny, params.nx))
    lib().orc_finalize(ctypes.byref(params), npol, POLTYPE[poltype], _ptr(grid_ri),
                       _ptr(norm), _ptr(img))
    return img


# ---- TOPO -> LSRK (oracle/mti_oracle_lsrk.c) -------------------------------------
def lsrk_velocity(frame, mjd_utc_s):
    """(v_rot, v_bary, v_lsr) in m/s, J2000 equatorial axes."""
    a, b, c = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().orc_lsrk_velocity(ctypes.byref(frame), float(mjd_utc_s), _ptr(a), _ptr(b), _ptr(c))
    return a, b, c


def earth_velocity(mjd_tt_s):
    """(Earth relative to the Sun, Sun relative to the barycentre), m/s."""
    a, b = np.zeros(3), np.zeros(3)
    lib().orc_earth_velocity(float(mjd_tt_s), _ptr(a), _ptr(b))
    return a, b


def lsrk_factor(frame, mjd_utc_s):
    return lib().orc_lsrk_factor(ctypes.byref(frame), float(mjd_utc_s))


def lsrk_factors(frame, times):
    t = np.ascontiguousarray(times, dtype=np.float64)
    out = np.zeros(len(t))
    lib().orc_lsrk_factors(ctypes.byref(frame), _ptr(t), len(t), _ptr(out))
    return out


class row_doppler:
    """Context manager: per-record 1/lambda factors for preprocess/weights/grid
    (rows in the order those calls receive them)."""

    def __init__(self, factors):
        self.f = np.ascontiguousarray(factors, dtype=np.float64)

    def __enter__(self):
        lib().orc_set_row_doppler(_ptr(self.f))
        return self

    def __exit__(self, *exc):
        lib().orc_set_row_doppler(None)


# ---- prediction (oracle/mti_oracle_degrid.c) ---------------------------------------
def model_grid(params, image, poltype="linear"):
    """Stokes model images [npol][ny][nx] (finalize layout) -> model grids
    [ny][nx][npol] complex128."""
    image = np.ascontiguousarray(image, dtype=np.float64)
    npol = image.shape[0]
    g = np.zeros((params.ny, params.nx, npol, 2))
    lib().orc_model_grid(ctypes.byref(params), npol, POLTYPE[poltype], _ptr(image), _ptr(g))
    return g[..., 0] + 1j * g[..., 1]


def model_correction(params, jy, jx):
    return lib().orc_model_correction(ctypes.byref(params), int(jy), int(jx))


def degrid(kern, uvw, inv_lambda, grid, normalise=True):
    """Predicted visibilities [nrows][npol] complex128 of one channel from
    model grids [ny][nx][npol] complex; returns (vis, predicted count)."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    g = np.ascontiguousarray(np.stack([grid.real, grid.imag], axis=-1), dtype=np.float64)
    npol = g.shape[2]
    out = np.zeros((len(uvw), npol, 2))
    n = ctypes.c_int64(0)
    lib().orc_degrid(ctypes.byref(kern.params), ctypes.byref(kern._k), len(uvw), npol, _ptr(uvw),
                     float(inv_lambda), _ptr(g), 1 if normalise else 0, _ptr(out), ctypes.byref(n))
    return out[..., 0] + 1j * out[..., 1], int(n.value)
