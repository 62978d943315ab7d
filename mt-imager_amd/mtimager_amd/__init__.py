"""ctypes binding of the gfx950 imaging library's C-ABI (include/mtimager_amd.h).

Device memory and streams come from PyTorch; every compute call goes through
libmtimager_amd.so.  There is no CPU fallback: if the library or a GPU is
missing, the calls raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libmtimager_amd.so")

MTI_OK = 0
MODE = {"normal": 0, "wproj": 1}
PRECISION = {"f32": 0, "f16": 1, "bf16": 2}
WEIGHTING = {"natural": 0, "uniform": 1, "briggs": 2}
POLTYPE = {"linear": 0, "circular": 1}
LAYOUT = {"vup": 0, "pvu": 1}
GRID_PRECISION = {"c64": 0, "c128": 1}
ACCUMULATION = {"f64": 0, "f32k": 1}
ARCSEC = 3.141592653589793 / 180.0 / 3600.0


class MTIError(RuntimeError):
    pass


class ImageParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("cell_l", ctypes.c_double), ("cell_m", ctypes.c_double),
                ("mode", ctypes.c_int32), ("wplanes", ctypes.c_int32),
                ("sampling", ctypes.c_int32), ("taper_support", ctypes.c_int32)]


class KernelInfo(ctypes.Structure):
    _fields_ = [("nplanes", ctypes.c_int32), ("sampling", ctypes.c_int32),
                ("T", ctypes.c_int32), ("trial_support", ctypes.c_int32),
                ("max_support", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("w_increment", ctypes.c_double), ("divisor", ctypes.c_double),
                ("total", ctypes.c_int64)]


class Vis(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("nchan", ctypes.c_int32),
                ("npol", ctypes.c_int32),
                ("uvw", ctypes.c_void_p), ("uvw_f32", ctypes.c_int32),
                ("uvw_row_stride", ctypes.c_int64), ("uvw_comp_stride", ctypes.c_int64),
                ("inv_lambda", ctypes.c_void_p), ("inv_lambda_f32", ctypes.c_int32),
                ("inv_lambda_per_row", ctypes.c_int32),
                ("vis", ctypes.c_void_p), ("vis_row_stride", ctypes.c_int64),
                ("vis_chan_stride", ctypes.c_int64),
                ("flags", ctypes.c_void_p), ("flags_i32", ctypes.c_int32),
                ("flags_row_stride", ctypes.c_int64), ("flags_chan_stride", ctypes.c_int64),
                ("weight", ctypes.c_void_p), ("weight_row_stride", ctypes.c_int64),
                ("row_index", ctypes.c_void_p), ("nindex", ctypes.c_int64),
                ("row_doppler", ctypes.c_void_p)]


class LsrkFrame(ctypes.Structure):
    """mti_lsrk_frame: observatory ITRF position (m), phase centre (J2000,
    rad), TT - UTC (s)."""
    _fields_ = [("obs_itrf", ctypes.c_double * 3), ("ra", ctypes.c_double), ("dec", ctypes.c_double),
                ("tt_minus_utc", ctypes.c_double)]


def lsrk_frame(obs_itrf, ra, dec, tt_minus_utc=69.184):
    return LsrkFrame((ctypes.c_double * 3)(*[float(x) for x in obs_itrf]), float(ra), float(dec),
                     float(tt_minus_utc))


class GridOpts(ctypes.Structure):
    _fields_ = [("chan0", ctypes.c_int32), ("nchan", ctypes.c_int32),
                ("accumulate", ctypes.c_int32), ("psf", ctypes.c_int32),
                ("weighting", ctypes.c_int32), ("robust", ctypes.c_double),
                ("grid_layout", ctypes.c_int32), ("grid_precision", ctypes.c_int32),
                ("tile_mask", ctypes.c_void_p), ("mfs", ctypes.c_int32), ("accumulation", ctypes.c_int32)]


class Job(ctypes.Structure):
    """mti_job: one imaging job of a batched call."""
    _fields_ = [("kernels", ctypes.c_void_p), ("params", ImageParams), ("vis", Vis),
                ("chan0", ctypes.c_int32), ("nchan", ctypes.c_int32),
                ("grids", ctypes.c_void_p), ("norm", ctypes.c_void_p), ("image", ctypes.c_void_p),
                ("ant1", ctypes.c_void_p), ("ant2", ctypes.c_void_p), ("nant", ctypes.c_int32),
                ("tile_mask", ctypes.c_void_p), ("grids_c64", ctypes.c_int32)]


class GridStats(ctypes.Structure):
    _fields_ = [("gridded_records", ctypes.c_int64), ("gridded_points", ctypes.c_int64),
                ("work_items", ctypes.c_int64), ("flushes", ctypes.c_int64)]


class PreprocessRefParams(ctypes.Structure):
    """PreProcessUVW parameters (op/WProjPreProcessUVW.cpp:49-56)."""
    _fields_ = [("rows", ctypes.c_int32), ("columns", ctypes.c_int32),
                ("u_increment", ctypes.c_double), ("v_increment", ctypes.c_double),
                ("sampling", ctypes.c_int32), ("wsquareincrement", ctypes.c_double),
                ("wplanes", ctypes.c_int32)]


class PreprocessRefOutputs(ctypes.Structure):
    """PreProcessUVW outputs 0-6 (op/WProjPreProcessUVW.cpp:70-84)."""
    _fields_ = [("gridding_data", ctypes.c_void_p), ("conv_index", ctypes.c_void_p),
                ("to_grid", ctypes.c_void_p), ("to_compress", ctypes.c_void_p),
                ("support", ctypes.c_void_p), ("take_conjugate", ctypes.c_void_p),
                ("conv_sum", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("unit_pairs", ctypes.c_int32), ("whole_columns_ny", ctypes.c_int32), ("xcd_chunk", ctypes.c_int32),
                ("region_width", ctypes.c_int32)]


# Every symbol include/mtimager_amd.h declares.
STAGES = ["sort_rows", "kernels", "bin", "sort_records", "plan", "grid", "fft_rows",
          "fft_cols", "finalize"]

EXPORTS = ["mti_version", "mti_create", "mti_destroy", "mti_last_error",
           "mti_set_profiling", "mti_stage_times",
           "mti_kernels_build", "mti_kernels_info", "mti_kernels_supports",
           "mti_kernels_table", "mti_kernels_sums", "mti_kernels_free", "mti_kernels_position",
           "mti_sort_rows", "mti_preprocess", "mti_grid", "mti_fft2d_inverse",
           "mti_finalize", "mti_grid_jobs", "mti_finalize_jobs",
           "mti_grid_out_shape", "mti_finalize_out_shape", "mti_preprocess_out_shape",
           "mti_sort_rows_workspace_size", "mti_grid_workspace_size", "mti_grid_jobs_workspace_size",
           "mti_finalize_workspace_size", "mti_finalize_jobs_workspace_size", "mti_fft2d_workspace_size",
           "mti_set_workspace", "mti_context_device_bytes", "mti_set_tuning", "mti_get_tuning",
           "mti_graph_create", "mti_graph_launch", "mti_graph_destroy",
           "mti_fits_create", "mti_fits_write_channel", "mti_fits_close", "mti_finalize_sparse",
           "mti_set_transform_storage", "mti_get_transform_storage",
           "mti_preprocess_reference", "mti_uv_reorder", "mti_centre_shift", "mti_stokes", "mti_taper_correct",
           "mti_taper_kernel", "mti_sinc_correct", "mti_normalizer_f32",
           "mti_lsrk_factors", "mti_lsrk_inv_lambda", "mti_sort_rows_static", "mti_model_grids",
           "mti_degrid"]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MTIError(f"{LIB_PATH} not built: run __graft_entry__.build()")
    # torch first: its bundled HIP runtime (soname libamdhip64.so.7) then
    # satisfies the library's dependency, so both share one runtime.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    vp, P = ctypes.c_void_p, ctypes.POINTER
    L.mti_version.restype = ctypes.c_char_p
    L.mti_create.restype = ctypes.c_int
    L.mti_create.argtypes = [ctypes.c_int, P(vp)]
    L.mti_destroy.argtypes = [vp]
    L.mti_last_error.restype = ctypes.c_char_p
    L.mti_last_error.argtypes = [vp]
    L.mti_set_profiling.restype = ctypes.c_int
    L.mti_set_profiling.argtypes = [vp, ctypes.c_int32]
    L.mti_stage_times.restype = ctypes.c_int
    L.mti_stage_times.argtypes = [vp, vp, vp]
    L.mti_kernels_build.restype = ctypes.c_int
    L.mti_kernels_build.argtypes = [vp, P(ImageParams), ctypes.c_int32, P(vp), vp]
    L.mti_kernels_info.restype = ctypes.c_int
    L.mti_kernels_info.argtypes = [vp, P(KernelInfo)]
    L.mti_kernels_supports.restype = ctypes.c_int
    L.mti_kernels_supports.argtypes = [vp, vp, vp]
    L.mti_kernels_table.restype = vp
    L.mti_kernels_table.argtypes = [vp]
    L.mti_kernels_sums.restype = vp
    L.mti_kernels_sums.argtypes = [vp]
    L.mti_kernels_position.restype = vp
    L.mti_kernels_position.argtypes = [vp]
    L.mti_kernels_free.argtypes = [vp]
    L.mti_sort_rows.restype = ctypes.c_int
    L.mti_sort_rows.argtypes = [vp, vp, vp, ctypes.c_int64, ctypes.c_int32, vp,
                                P(ctypes.c_int64), vp]
    L.mti_sort_rows_static.restype = ctypes.c_int
    L.mti_sort_rows_static.argtypes = [vp, vp, vp, ctypes.c_int64, ctypes.c_int32, vp,
                                       P(ctypes.c_int64), vp]
    L.mti_preprocess.restype = ctypes.c_int
    L.mti_preprocess.argtypes = [vp, vp, P(ImageParams), P(Vis), ctypes.c_int32, vp, vp]
    L.mti_grid.restype = ctypes.c_int
    L.mti_grid.argtypes = [vp, vp, P(ImageParams), P(Vis), P(GridOpts), vp, vp,
                           P(GridStats), vp]
    L.mti_fft2d_inverse.restype = ctypes.c_int
    L.mti_fft2d_inverse.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp]
    L.mti_finalize.restype = ctypes.c_int
    L.mti_finalize.argtypes = [vp, P(ImageParams), ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp, vp, vp]
    L.mti_grid_jobs.restype = ctypes.c_int
    L.mti_grid_jobs.argtypes = [vp, ctypes.c_int32, vp, P(GridOpts), P(GridStats), vp]
    L.mti_finalize_jobs.restype = ctypes.c_int
    L.mti_finalize_jobs.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, ctypes.c_int32, vp]
    i32, i64, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
    for name, args in [
            ("mti_grid_out_shape", [P(ImageParams), i32, i32, i32, P(i64), P(i64)]),
            ("mti_finalize_out_shape", [P(ImageParams), i32, i32, P(i64)]),
            ("mti_preprocess_out_shape", [i64, P(i64)]),
            ("mti_sort_rows_workspace_size", [i64, P(sz)]),
            ("mti_grid_workspace_size", [P(KernelInfo), P(ImageParams), P(Vis), P(GridOpts), P(sz)]),
            ("mti_grid_jobs_workspace_size", [i32, vp, P(GridOpts), P(sz)]),
            ("mti_finalize_workspace_size", [P(ImageParams), i32, i32, P(sz)]),
            ("mti_finalize_jobs_workspace_size", [i32, vp, P(sz)]),
            ("mti_fft2d_workspace_size", [i32, i32, i32, P(sz)]),
            ("mti_set_workspace", [vp, vp, sz]),
            ("mti_context_device_bytes", [vp, P(sz)]),
            ("mti_set_tuning", [vp, P(Tuning)]),
            ("mti_get_tuning", [vp, P(Tuning)]),
            ("mti_set_transform_storage", [vp, i32]),
            ("mti_get_transform_storage", [vp, P(i32)]),
            ("mti_graph_create", [vp, i32, vp, P(GridOpts), i32, i32, P(vp)]),
            ("mti_graph_launch", [vp, vp]),
            ("mti_finalize_sparse", [vp, P(ImageParams), i32, i32, i32, i32, i32, vp, vp, vp, vp, vp]),
            ("mti_preprocess_reference", [vp, P(PreprocessRefParams), vp, vp, P(Vis), i32,
                                          P(PreprocessRefOutputs), vp]),
            ("mti_uv_reorder", [vp, vp, i32, i32, i32, i32, vp]),
            ("mti_centre_shift", [vp, vp, i64, i32, i32, i32, vp]),
            ("mti_stokes", [vp, vp, vp, i32, i32, i32, i32, vp]),
            ("mti_taper_correct", [vp, vp, i64, i32, i32, i32, vp]),
            ("mti_taper_kernel", [vp, i32, i32, i32, i32, vp]),
            ("mti_sinc_correct", [vp, vp, i64, i32, i32, i32, vp]),
            ("mti_normalizer_f32", [vp, vp, i64, vp]),
            ("mti_lsrk_factors", [vp, P(LsrkFrame), vp, i64, vp, vp]),
            ("mti_model_grids", [vp, P(ImageParams), i32, i32, i32, vp, vp, vp]),
            ("mti_degrid", [vp, vp, P(ImageParams), P(Vis), i32, i32, vp, vp, P(i64), vp]),
            ("mti_lsrk_inv_lambda", [vp, P(LsrkFrame), vp, i64, vp, i32, vp, vp])]:
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    L.mti_graph_destroy.argtypes = [vp]
    L.mti_graph_destroy.restype = None
    _lib = L
    return L


def make_params(nx, ny, cell_l, cell_m=None, mode="wproj", wplanes=1, sampling=4,
                taper_support=7, units="arcsec"):
    """Parameters from mt-imager property values: image.nx, image.ny,
    image.lintervalsize / mintervalsize (arcsec, or radians with
    units="rad"), mode, wproj.wplanes, wproj.convFunction.sampling (or
    taper.sampling), taper.support."""
    if cell_m is None:
        cell_m = cell_l
    scale = ARCSEC if units == "arcsec" else 1.0
    return ImageParams(int(nx), int(ny), float(cell_l) * scale, float(cell_m) * scale,
                       MODE[mode], int(wplanes), int(sampling), int(taper_support))


def _grid_opts(chan0=0, nchan=1, accumulate=False, psf=False, weighting="natural", robust=0.0,
               layout="vup", precision="c128", tile_mask=None, mfs=False, accumulation="f64"):
    return GridOpts(int(chan0), int(nchan), 1 if accumulate else 0, 1 if psf else 0,
                    WEIGHTING[weighting], float(robust), LAYOUT[layout], GRID_PRECISION[precision],
                    tile_mask.data_ptr() if tile_mask is not None else None, 1 if mfs else 0,
                    ACCUMULATION[accumulation])


def tile_mask_shape(params, nchan):
    """Shape of a tile mask (sparse grids): [nchan][ceil(ny/32)][ceil(nx/32)], uint8."""
    return (int(nchan), (params.ny + 31) // 32, (params.nx + 31) // 32)


def _query(fn, *args):
    out = ctypes.c_size_t(0)
    rc = fn(*args, ctypes.byref(out))
    if rc != 0:
        raise MTIError(f"{fn.__name__}: invalid arguments (status {rc})")
    return int(out.value)


# ---- output shapes and workspace sizes (no context, no device) ------------------
def grid_out_shape(params, npol, nchan, layout="vup"):
    """(grid dims, normaliser dims) of mti_grid: [nchan][ny][nx][npol] complex
    ("vup") or [nchan][npol][ny][nx] ("pvu"); [nchan][npol]."""
    g, n = (ctypes.c_int64 * 4)(), (ctypes.c_int64 * 2)()
    rc = lib().mti_grid_out_shape(ctypes.byref(params), int(npol), int(nchan), LAYOUT[layout], g, n)
    if rc != 0:
        raise MTIError(f"mti_grid_out_shape: invalid arguments (status {rc})")
    return tuple(g), tuple(n)


def finalize_out_shape(params, npol, nchan):
    d = (ctypes.c_int64 * 4)()
    rc = lib().mti_finalize_out_shape(ctypes.byref(params), int(npol), int(nchan), d)
    if rc != 0:
        raise MTIError(f"mti_finalize_out_shape: invalid arguments (status {rc})")
    return tuple(d)


def preprocess_out_shape(nrecords):
    d = (ctypes.c_int64 * 2)()
    rc = lib().mti_preprocess_out_shape(int(nrecords), d)
    if rc != 0:
        raise MTIError(f"mti_preprocess_out_shape: invalid arguments (status {rc})")
    return tuple(d)


def sort_rows_workspace_size(nrows):
    return _query(lib().mti_sort_rows_workspace_size, int(nrows))


def grid_workspace_size(kernel_info, params, vis, **opts):
    """Workspace bytes of Context.grid with the same arguments (opts as for
    grid: chan0, nchan, weighting, layout, precision, ...)."""
    return _query(lib().mti_grid_workspace_size, ctypes.byref(kernel_info), ctypes.byref(params),
                  ctypes.byref(vis), ctypes.byref(_grid_opts(**opts)))


def grid_jobs_workspace_size(jobs, **opts):
    arr = (Job * len(jobs))(*jobs)
    return _query(lib().mti_grid_jobs_workspace_size, len(jobs), ctypes.cast(arr, ctypes.c_void_p),
                  ctypes.byref(_grid_opts(**opts)))


def finalize_workspace_size(params, npol, nchan):
    return _query(lib().mti_finalize_workspace_size, ctypes.byref(params), int(npol), int(nchan))


def finalize_jobs_workspace_size(jobs):
    arr = (Job * len(jobs))(*jobs)
    return _query(lib().mti_finalize_jobs_workspace_size, len(jobs), ctypes.cast(arr, ctypes.c_void_p))


def fft2d_workspace_size(nx, ny, nbatch=1):
    return _query(lib().mti_fft2d_workspace_size, int(nx), int(ny), int(nbatch))


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream)


class Context:
    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        rc = lib().mti_create(int(device), ctypes.byref(self.h))
        if rc != MTI_OK:
            raise MTIError(f"mti_create({device}) failed with status {rc}")

    def check(self, rc, what):
        if rc != MTI_OK:
            msg = lib().mti_last_error(self.h).decode()
            raise MTIError(f"{what} failed (status {rc}): {msg}")

    def close(self):
        if self.h:
            lib().mti_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- profiling ------------------------------------------------------
    def set_profiling(self, enable=True):
        self.check(lib().mti_set_profiling(self.h, 1 if enable else 0), "mti_set_profiling")

    def stage_times(self):
        """{stage: (total ms, launches)} since the last call (waits for events)."""
        import numpy as np
        ms = np.zeros(len(STAGES), np.float64)
        n = np.zeros(len(STAGES), np.int64)
        self.check(lib().mti_stage_times(self.h, ms.ctypes.data, n.ctypes.data), "mti_stage_times")
        return {s: (float(ms[i]), int(n[i])) for i, s in enumerate(STAGES)}

    # ---- a8-a14 ---------------------------------------------------------
    def set_workspace(self, buf):
        """Caller-owned scratch: every later sort/grid/finalize/transform call
        on this context lays its scratch out in `buf` (a device tensor, kept
        referenced here) and allocates nothing; None returns to scratch owned
        by the context."""
        if buf is None:
            self.check(lib().mti_set_workspace(self.h, None, 0), "mti_set_workspace")
        else:
            nbytes = buf.numel() * buf.element_size()
            self.check(lib().mti_set_workspace(self.h, buf.data_ptr(), nbytes), "mti_set_workspace")
        self._workspace = buf

    def set_tuning(self, unit_pairs=None, whole_columns_ny=None, xcd_chunk=None, region_width=None):
        """Launch-shape tuning of this context (see mtimager_amd.autotune);
        arguments left None keep their current value."""
        t = self.get_tuning()
        for k, v in (("unit_pairs", unit_pairs), ("whole_columns_ny", whole_columns_ny), ("xcd_chunk", xcd_chunk),
                     ("region_width", region_width)):
            if v is not None:
                t[k] = int(v)
        self.check(lib().mti_set_tuning(self.h, ctypes.byref(Tuning(t["unit_pairs"], t["whole_columns_ny"],
                                                                    t["xcd_chunk"], t["region_width"]))),
                   "mti_set_tuning")

    def get_tuning(self):
        t = Tuning()
        self.check(lib().mti_get_tuning(self.h, ctypes.byref(t)), "mti_get_tuning")
        return {"unit_pairs": int(t.unit_pairs), "whole_columns_ny": int(t.whole_columns_ny),
                "xcd_chunk": int(t.xcd_chunk), "region_width": int(t.region_width)}

    def set_transform_storage(self, precision):
        """Element type of the packed planes between the finalize's transform
        passes: "c128" (default) or "c64" (half the bytes between passes; the
        arithmetic stays FP64)."""
        self.check(lib().mti_set_transform_storage(self.h, GRID_PRECISION[precision]), "mti_set_transform_storage")

    def get_transform_storage(self):
        out = ctypes.c_int32(0)
        self.check(lib().mti_get_transform_storage(self.h, ctypes.byref(out)), "mti_get_transform_storage")
        return {v: k for k, v in GRID_PRECISION.items()}[int(out.value)]

    def device_bytes(self):
        """Device bytes the context itself holds (its scratch and twiddle tables)."""
        out = ctypes.c_size_t(0)
        self.check(lib().mti_context_device_bytes(self.h, ctypes.byref(out)), "mti_context_device_bytes")
        return int(out.value)

    def build_kernels(self, params, precision="f32", stream=None):
        return Kernels(self, params, precision, stream)

    # ---- row order --------------------------------------------------------
    def sort_rows(self, ant1, ant2, nant, index_out, stream=None):
        kept = ctypes.c_int64(0)
        rc = lib().mti_sort_rows(self.h, ant1.data_ptr(), ant2.data_ptr(), ant1.numel(),
                                 int(nant), index_out.data_ptr(), ctypes.byref(kept),
                                 _stream(stream))
        self.check(rc, "mti_sort_rows")
        return int(kept.value)

    # ---- TOPO -> LSRK -------------------------------------------------------
    def lsrk_factors(self, frame, time_mjd_s, out, stream=None):
        """out[r] = f_LSRK / f_TOPO at time_mjd_s[r] (f64 device tensors;
        MJD seconds, UTC), for vis descriptors' row_doppler."""
        rc = lib().mti_lsrk_factors(self.h, ctypes.byref(frame), time_mjd_s.data_ptr(), time_mjd_s.numel(),
                                    out.data_ptr(), _stream(stream))
        self.check(rc, "mti_lsrk_factors")

    def lsrk_inv_lambda(self, frame, time_mjd_s, freq_hz, out, stream=None):
        """The reference's per-record frequency store: f32 [nchan][nrows] of
        f_LSRK / c."""
        rc = lib().mti_lsrk_inv_lambda(self.h, ctypes.byref(frame), time_mjd_s.data_ptr(), time_mjd_s.numel(),
                                       freq_hz.data_ptr(), freq_hz.numel(), out.data_ptr(), _stream(stream))
        self.check(rc, "mti_lsrk_inv_lambda")

    # ---- prediction (degridding) -----------------------------------------------
    def model_grids(self, params, npol, model, grids, nchan=1, poltype="linear", stream=None):
        """Stokes model images f32 [nchan][npol][ny][nx] (the finalize's layout)
        -> complex64 model grids [nchan][npol][ny][nx] (float32 [..][2] or
        complex64 tensor)."""
        rc = lib().mti_model_grids(self.h, ctypes.byref(params), int(npol), POLTYPE[poltype], int(nchan),
                                   model.data_ptr(), grids.data_ptr(), _stream(stream))
        self.check(rc, "mti_model_grids")

    def degrid(self, kern, params, vis, grids, vis_out, chan0=0, nchan=1, stream=None):
        """Model visibilities of vis's rows (row_index order) for channels
        chan0 .. chan0 + nchan - 1 into vis_out (complex64, vis's strides);
        returns the number of predicted records."""
        n = ctypes.c_int64(0)
        rc = lib().mti_degrid(self.h, kern.h, ctypes.byref(params), ctypes.byref(vis), int(chan0), int(nchan),
                              grids.data_ptr(), vis_out.data_ptr(), ctypes.byref(n), _stream(stream))
        self.check(rc, "mti_degrid")
        return int(n.value)

    # ---- binning ----------------------------------------------------------
    def preprocess(self, kern, params, vis, chan, bins_out, stream=None):
        rc = lib().mti_preprocess(self.h, kern.h, ctypes.byref(params), ctypes.byref(vis),
                                  int(chan), bins_out.data_ptr(), _stream(stream))
        self.check(rc, "mti_preprocess")

    def preprocess_reference(self, rp, convinfo, sums, vis, chan, outs, stream=None):
        """PreProcessUVW with the reference's seven outputs.  rp:
        PreprocessRefParams; convinfo int32 [W][2] {support, base}; sums f32
        [W][T*T]; outs: (gridding_data int32 [N][4], conv_index, to_grid,
        to_compress, support int32 [N], take_conjugate f32 [N], conv_sum f32 [P][N])."""
        o = PreprocessRefOutputs(*[t.data_ptr() for t in outs])
        rc = lib().mti_preprocess_reference(self.h, ctypes.byref(rp), convinfo.data_ptr(), sums.data_ptr(),
                                            ctypes.byref(vis), int(chan), ctypes.byref(o), _stream(stream))
        self.check(rc, "mti_preprocess_reference")

    # ---- gridding ---------------------------------------------------------
    def grid(self, kern, params, vis, grids, norm, chan0=0, nchan=1, accumulate=False,
             psf=False, weighting="natural", robust=0.0, layout="vup", precision=None,
             stats=True, stream=None, tile_mask=None, mfs=False, accumulation="f64"):
        """grids: per channel [ny][nx][npol] (layout "vup", the reference's
        Grid) or [npol][ny][nx] ("pvu"); complex64 or complex128 (precision
        "c64"/"c128", taken from the tensor dtype by default).  tile_mask
        (uint8 device tensor of tile_mask_shape): sparse grids -- tiles no
        record reaches are marked 0 and left unwritten (finalize with the
        same mask).  mfs: the nchan channels go into one grid and normaliser
        (multi-frequency synthesis).  accumulation: "f64" (default) or "f32k"
        (FP32 batch sums with Kahan-compensated totals)."""
        if precision is None:
            precision = "c128" if str(grids.dtype) in ("torch.complex128", "torch.float64") else "c64"
        o = _grid_opts(chan0, nchan, accumulate, psf, weighting, robust, layout, precision, tile_mask, mfs,
                       accumulation)
        st = GridStats()
        rc = lib().mti_grid(self.h, kern.h, ctypes.byref(params), ctypes.byref(vis),
                            ctypes.byref(o), grids.data_ptr(), norm.data_ptr(),
                            ctypes.byref(st) if stats else None, _stream(stream))
        self.check(rc, "mti_grid")
        return st

    # ---- transform --------------------------------------------------------
    def fft2d_inverse(self, data, nx, ny, nbatch=1, stream=None):
        rc = lib().mti_fft2d_inverse(self.h, data.data_ptr(), int(nx), int(ny), int(nbatch),
                                     _stream(stream))
        self.check(rc, "mti_fft2d_inverse")

    # ---- finalize -----------------------------------------------------------
    def finalize(self, params, npol, grids, norm, image, nchan=1, poltype="linear",
                 layout="vup", precision=None, stream=None, tile_mask=None):
        if precision is None:
            precision = "c128" if str(grids.dtype) in ("torch.complex128", "torch.float64") else "c64"
        rc = lib().mti_finalize_sparse(self.h, ctypes.byref(params), int(npol), POLTYPE[poltype],
                                       int(nchan), LAYOUT[layout], GRID_PRECISION[precision],
                                       grids.data_ptr(), tile_mask.data_ptr() if tile_mask is not None else None,
                                       norm.data_ptr(), image.data_ptr(), _stream(stream))
        self.check(rc, "mti_finalize")


    # ---- batched jobs ---------------------------------------------------------
    @staticmethod
    def make_job(kern, params, vis, grids, norm, image=None, chan0=0, nchan=1, ant1=None, ant2=None, nant=0,
                 tile_mask=None):
        """An mti_job; grids [nchan] grids, complex128 (complex128 or float64
        tensors) or complex64 (complex64 or float32 tensors), norm f64
        [nchan][npol], image f32 [nchan][npol][ny][nx] (for finalize_jobs).
        With ant1/ant2 (int32 device tensors of vis's rows) and nant, grid_jobs
        computes the baseline-major row order itself (vis must then have no
        row index)."""
        j = Job()
        j.kernels = kern.h
        j.params = params
        j.vis = vis
        j.chan0, j.nchan = int(chan0), int(nchan)
        j.grids, j.norm = grids.data_ptr(), norm.data_ptr()
        j.grids_c64 = 0 if str(grids.dtype) in ("torch.complex128", "torch.float64") else 1
        j.image = image.data_ptr() if image is not None else None
        if ant1 is not None:
            j.ant1, j.ant2, j.nant = ant1.data_ptr(), ant2.data_ptr(), int(nant)
        if tile_mask is not None:
            j.tile_mask = tile_mask.data_ptr()
        j._keep = (kern, vis, grids, norm, image, ant1, ant2, tile_mask)
        return j

    def grid_jobs(self, jobs, accumulate=False, psf=False, weighting="natural", robust=0.0,
                  layout="vup", stats=True, stream=None, mfs=False, accumulation="f64"):
        arr = (Job * len(jobs))(*jobs)
        o = GridOpts(0, 0, 1 if accumulate else 0, 1 if psf else 0, WEIGHTING[weighting],
                     float(robust), LAYOUT[layout], GRID_PRECISION["c128"], None, 1 if mfs else 0,
                     ACCUMULATION[accumulation])
        st = GridStats()
        rc = lib().mti_grid_jobs(self.h, len(jobs), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(o),
                                 ctypes.byref(st) if stats else None, _stream(stream))
        self.check(rc, "mti_grid_jobs")
        return st

    def capture_jobs(self, jobs, accumulate=False, psf=False, weighting="natural", robust=0.0, layout="vup",
                     finalize=True, poltype="linear"):
        """A Graph of grid_jobs (+ finalize_jobs) over these jobs: captured
        once, launched any number of times on the buffers' current contents."""
        return Graph(self, jobs, _grid_opts(0, 0, accumulate, psf, weighting, robust, layout, "c128"),
                     finalize, poltype)

    def finalize_jobs(self, jobs, poltype="linear", layout="vup", stream=None):
        arr = (Job * len(jobs))(*jobs)
        rc = lib().mti_finalize_jobs(self.h, len(jobs), ctypes.cast(arr, ctypes.c_void_p), POLTYPE[poltype],
                                     LAYOUT[layout], _stream(stream))
        self.check(rc, "mti_finalize_jobs")


# ---- the reference's operators with their own array contracts ---------------
def _op(rc, what):
    if rc != MTI_OK:
        raise MTIError(f"{what} failed (status {rc})")


def _c128(t):
    return 1 if str(t.dtype) in ("torch.complex128", "torch.float64") else 0


def uv_reorder(inp, out, rows, cols, npol, stream=None):
    """PreIFFTUVCorrection: complex [rows][cols][npol] -> [npol][rows][cols]."""
    _op(lib().mti_uv_reorder(inp.data_ptr(), out.data_ptr(), int(rows), int(cols), int(npol), _c128(inp),
                             _stream(stream)), "mti_uv_reorder")


def centre_shift(inp, out, planes, rows, cols, stream=None):
    """ZeroToCentre2DShift of every [rows][cols] complex plane."""
    _op(lib().mti_centre_shift(inp.data_ptr(), out.data_ptr(), int(planes), int(rows), int(cols), _c128(inp),
                               _stream(stream)), "mti_centre_shift")


def stokes(inp, norm_f32, out, npol, rows, cols, poltype="linear", stream=None):
    """StokesConverter: complex64 [npol][rows][cols] + f32 normaliser -> f32 Stokes, rows flipped."""
    _op(lib().mti_stokes(inp.data_ptr(), norm_f32.data_ptr(), out.data_ptr(), int(npol), int(rows), int(cols),
                         POLTYPE[poltype], _stream(stream)), "mti_stokes")


def taper_correct(inp, out, planes, rows, cols, stream=None):
    """casagridsf lmPlanePostCorrection: divide by the taper (inp None: write the taper)."""
    f64 = 1 if str(out.dtype) == "torch.float64" else 0
    _op(lib().mti_taper_correct(inp.data_ptr() if inp is not None else None, out.data_ptr(), int(planes),
                                int(rows), int(cols), f64, _stream(stream)), "mti_taper_correct")


def taper_kernel(out, support, sampling, complex_out=False, stream=None):
    """casagridsf generate: the oversampled taper-only gridding function."""
    f64 = 1 if str(out.dtype) in ("torch.float64", "torch.complex128") else 0
    _op(lib().mti_taper_kernel(out.data_ptr(), int(support), int(sampling), 1 if complex_out else 0, f64,
                               _stream(stream)), "mti_taper_kernel")


def sinc_correct(inp, out, planes, rows, cols, sampling, stream=None):
    """SincCorrector of f32 planes."""
    _op(lib().mti_sinc_correct(inp.data_ptr(), out.data_ptr(), int(planes), int(rows), int(cols), int(sampling),
                               _stream(stream)), "mti_sinc_correct")


def normalizer_f32(norm, out, stream=None):
    """The reference's f32 NormalizerValue from FP64 normalisers."""
    _op(lib().mti_normalizer_f32(norm.data_ptr(), out.data_ptr(), int(norm.numel()), _stream(stream)),
        "mti_normalizer_f32")


class Graph:
    """mti_graph: batched imaging of a fixed job list captured into a HIP
    graph (the jobs' tensors, kernel tables and the context must outlive it)."""

    def __init__(self, ctx, jobs, opts, finalize=True, poltype="linear"):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        arr = (Job * len(jobs))(*jobs)
        rc = lib().mti_graph_create(ctx.h, len(jobs), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(opts),
                                    1 if finalize else 0, POLTYPE[poltype], ctypes.byref(self.h))
        ctx.check(rc, "mti_graph_create")
        self._keep = list(jobs)

    def launch(self, stream=None):
        self.ctx.check(lib().mti_graph_launch(self.h, _stream(stream)), "mti_graph_launch")

    def free(self):
        if self.h:
            lib().mti_graph_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Kernels:
    def __init__(self, ctx, params, precision="f32", stream=None):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        rc = lib().mti_kernels_build(ctx.h, ctypes.byref(params), PRECISION[precision],
                                     ctypes.byref(self.h), _stream(stream))
        ctx.check(rc, "mti_kernels_build")
        self.info = KernelInfo()
        lib().mti_kernels_info(self.h, ctypes.byref(self.info))

    def supports(self):
        import numpy as np
        n = self.info.nplanes
        s = np.zeros(n, dtype=np.int32)
        b = np.zeros(n, dtype=np.int64)
        lib().mti_kernels_supports(self.h, s.ctypes.data, b.ctypes.data)
        return s, b

    def table(self):
        """Device table copied to a torch tensor (complex64 [total])."""
        import torch
        n = int(self.info.total)
        prec = self.info.precision
        ptr = lib().mti_kernels_table(self.h)
        esz = 8 if prec == 0 else 4
        raw = torch.empty(n * esz, dtype=torch.uint8, device="cuda")
        if n:
            _memcpy_d2d(raw.data_ptr(), ptr, n * esz)
        if prec == 0:
            return raw.view(torch.complex64)
        if prec == 1:
            h = raw.view(torch.float16).float().view(-1, 2)
        else:
            h = raw.view(torch.bfloat16).float().view(-1, 2)
        return torch.complex(h[:, 0], h[:, 1])

    def sums(self):
        import torch
        n = self.info.nplanes * self.info.T * self.info.T
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        _memcpy_d2d(out.data_ptr(), lib().mti_kernels_sums(self.h), n * 4)
        return out.view(self.info.nplanes, self.info.T * self.info.T)

    def free(self):
        if self.h:
            lib().mti_kernels_free(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _memcpy_d2d(dst, src, nbytes):
    """Device-to-device copy through the HIP runtime the library is bound to."""
    import torch
    hip = lib()
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes), 3)
    if rc != 0:
        raise MTIError(f"hipMemcpy failed ({rc})")


def vis_ms(uvw, inv_lambda, vis, flags, weight, row_index=None, row_doppler=None):
    """mti_vis for MeasurementSet-ordered torch tensors on the GPU:
    uvw f64 [nrows][3], inv_lambda f64 [nchan], vis complex64 [nrows][nchan][npol],
    flags u8 [nrows][nchan][npol], weight f32 [nrows][npol], row_index int32 or None,
    row_doppler f64 [nrows] (TOPO -> LSRK factors) or None.
    The descriptor keeps references to the tensors it points into."""
    nrows, nchan, npol = flags.shape
    for t in (uvw, inv_lambda, vis, flags, weight):
        assert t.is_contiguous() and t.is_cuda
    v = Vis()
    v.nrows, v.nchan, v.npol = nrows, nchan, npol
    v.uvw, v.uvw_f32, v.uvw_row_stride, v.uvw_comp_stride = uvw.data_ptr(), 0, 3, 1
    v.inv_lambda, v.inv_lambda_f32, v.inv_lambda_per_row = inv_lambda.data_ptr(), 0, 0
    v.vis, v.vis_row_stride, v.vis_chan_stride = vis.data_ptr(), nchan * npol, npol
    v.flags, v.flags_i32 = flags.data_ptr(), 0
    v.flags_row_stride, v.flags_chan_stride = nchan * npol, npol
    v.weight, v.weight_row_stride = weight.data_ptr(), npol
    if row_index is not None:
        v.row_index, v.nindex = row_index.data_ptr(), row_index.numel()
    else:
        v.row_index, v.nindex = None, 0
    if row_doppler is not None:
        assert row_doppler.is_contiguous() and row_doppler.is_cuda and row_doppler.numel() == nrows
        v.row_doppler = row_doppler.data_ptr()
    v._keep = (uvw, inv_lambda, vis, flags, weight, row_index, row_doppler)
    return v


def vis_chan_major(uvw, inv_lambda, vis, flags, weight, row_index=None, row_doppler=None):
    """mti_vis for channel-major torch tensors on the GPU (a block of channels
    uploaded as one contiguous piece): uvw f64 [nrows][3], inv_lambda f64
    [nchan], vis complex64 [nchan][nrows][npol], flags u8 [nchan][nrows][npol],
    weight f32 [nrows][npol]."""
    nchan, nrows, npol = flags.shape
    for t in (uvw, inv_lambda, vis, flags, weight):
        assert t.is_contiguous() and t.is_cuda
    v = Vis()
    v.nrows, v.nchan, v.npol = nrows, nchan, npol
    v.uvw, v.uvw_f32, v.uvw_row_stride, v.uvw_comp_stride = uvw.data_ptr(), 0, 3, 1
    v.inv_lambda, v.inv_lambda_f32, v.inv_lambda_per_row = inv_lambda.data_ptr(), 0, 0
    v.vis, v.vis_row_stride, v.vis_chan_stride = vis.data_ptr(), npol, nrows * npol
    v.flags, v.flags_i32 = flags.data_ptr(), 0
    v.flags_row_stride, v.flags_chan_stride = npol, nrows * npol
    v.weight, v.weight_row_stride = weight.data_ptr(), npol
    if row_index is not None:
        v.row_index, v.nindex = row_index.data_ptr(), row_index.numel()
    else:
        v.row_index, v.nindex = None, 0
    if row_doppler is not None:
        assert row_doppler.is_contiguous() and row_doppler.is_cuda and row_doppler.numel() == nrows
        v.row_doppler = row_doppler.data_ptr()
    v._keep = (uvw, inv_lambda, vis, flags, weight, row_index, row_doppler)
    return v


def vis_reference_chunk(uvw_t, inv_lambda_rows, vis, flags_i32, weight):
    """mti_vis for the reference's per-channel chunk layout
    (vis/StorageManager.h:147-150): uvw f32 [3][N], 1/lambda f32 [N],
    vis f32 [N][2P], flags int32 [N][P], weights f32 [N][P]."""
    n, npol = flags_i32.shape
    v = Vis()
    v.nrows, v.nchan, v.npol = n, 1, npol
    v.uvw, v.uvw_f32, v.uvw_row_stride, v.uvw_comp_stride = uvw_t.data_ptr(), 1, 1, n
    v.inv_lambda, v.inv_lambda_f32, v.inv_lambda_per_row = inv_lambda_rows.data_ptr(), 1, 1
    v.vis, v.vis_row_stride, v.vis_chan_stride = vis.data_ptr(), npol, 0
    v.flags, v.flags_i32 = flags_i32.data_ptr(), 1
    v.flags_row_stride, v.flags_chan_stride = npol, 0
    v.weight, v.weight_row_stride = weight.data_ptr(), npol
    v.row_index, v.nindex = None, 0
    v._keep = (uvw_t, inv_lambda_rows, vis, flags_i32, weight)
    return v
