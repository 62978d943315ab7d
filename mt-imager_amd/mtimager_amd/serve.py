"""Serving loop for imaging requests on one GPU: continuous batching over the
batched C-ABI entry points (mti_grid_jobs + mti_finalize_jobs).

The reference images one job at a time (one WImager DAG per channel,
comp/WImager.cpp:167-234, driven by app/mtimager.cpp:151-170); it has no
request loop.  This module adds one above the same operators:

* a request carries its visibilities (already on the device), its imaging
  parameters and its channel range;
* admission is by an HBM budget: a request is admitted when the bytes of its
  grids, image and per-record work space fit beside those of the requests
  already admitted, so memory, not a queue length, bounds what is in flight;
* W-kernel tables are built once per (nx, ny, cell, mode, w-planes,
  oversampling, taper support) and kept on the device in a cache shared by
  all workers;
* each worker owns a library context and a HIP stream; whenever it is free it
  takes every admitted request waiting (first come, first served) and images
  them as one wave, so requests that arrive while a wave runs join the next
  one and the waves of different workers overlap on the device.

Results do not depend on which wave a request lands in: each job of a wave
has its own grids, normaliser and density, and the FP64 accumulation differs
from a solo run only in summation order.
"""
import collections
import threading
import time

from . import Context, MTIError, vis_ms

def work_bytes_per_vis(npol):
    """Scratch of the batched gridding call per (row, channel) record: the
    16-byte descriptor and 4-byte tag, the weighted visibility (8 bytes per
    correlation), flag-masked weights and density (4 + 4 per correlation),
    and the (record, tile) pairs with their sort keys (8 + 8 bytes each, up
    to four tiles for the supports of these tables)."""
    return 20 + 16 * int(npol) + 4 * 16


def request_bytes(nx, ny, npol, nchan, nvis, grid_bytes=16):
    """Device bytes a request holds while admitted: grids (complex128: 16
    bytes a cell and correlation, complex64: 8), f64 normaliser, f32 image and
    the per-record work space of its wave."""
    pix = int(nx) * int(ny) * int(npol) * int(nchan)
    return pix * (int(grid_bytes) + 4) + 8 * npol * nchan + work_bytes_per_vis(npol) * int(nvis)


class Request:
    """One imaging request.  vis_args: (uvw, inv_lambda, vis, flags, weight)
    device tensors in MeasurementSet order (see mtimager_amd.vis_ms);
    ant1/ant2: the antenna columns of those rows (int32 device tensors), for
    the baseline-major row order; params: ImageParams; nchan channels from
    chan0; npol correlations.

    After the request is done: image f32 [nchan][npol][ny][nx] and norm f64
    [nchan][npol] (if keep_image), arrival/admitted/done perf_counter stamps,
    on_done(request) called by the worker once the request's wave has
    finished on the device (image and norm set for the call, then dropped
    unless keep_image) -- a result hand-off that keeps no image alive,
    error (an exception, or None), wave (the wave number it ran in) and
    gridded (its gridded (row, channel) count when it ran alone in its wave;
    a batched wave reports one count for all of its requests, summed in
    Server.gridded)."""

    __slots__ = ("rid", "params", "vis_args", "ant1", "ant2", "nant", "chan0", "nchan", "npol",
                 "nvis", "keep_image", "arrival", "admitted", "done", "image", "norm",
                 "gridded", "error", "nbytes", "wave", "on_done")

    def __init__(self, rid, params, vis_args, ant1, ant2, nant, nchan, npol=1, chan0=0,
                 nvis=None, keep_image=True, on_done=None):
        self.rid = rid
        self.params = params
        self.vis_args = vis_args
        self.ant1, self.ant2, self.nant = ant1, ant2, int(nant)
        self.chan0, self.nchan, self.npol = int(chan0), int(nchan), int(npol)
        self.nvis = int(nvis) if nvis is not None else int(ant1.numel()) * int(nchan)
        self.keep_image = keep_image
        self.on_done = on_done
        self.arrival = self.admitted = self.done = None
        self.image = self.norm = None
        self.gridded = 0
        self.error = None
        self.nbytes = request_bytes(params.nx, params.ny, npol, nchan, self.nvis)
        self.wave = -1


class KernelCache:
    """Device-resident W-kernel tables keyed by the parameters they depend on.
    Thread-safe; a miss builds on the caller's context and stream."""

    def __init__(self, precision="f32"):
        self.precision = precision
        self._lock = threading.Lock()
        self._tab = {}
        self.hits = self.misses = 0

    @staticmethod
    def key(p):
        return (p.nx, p.ny, p.cell_l, p.cell_m, p.mode, p.wplanes, p.sampling, p.taper_support)

    def get(self, ctx, params, stream=None):
        k = self.key(params)
        with self._lock:
            K = self._tab.get(k)
            if K is not None:
                self.hits += 1
                return K
            self.misses += 1
            K = ctx.build_kernels(params, self.precision, stream)
            self._tab[k] = K
            return K

    def __len__(self):
        return len(self._tab)

    def clear(self):
        with self._lock:
            for K in self._tab.values():
                K.free()
            self._tab.clear()


class Server:
    """Continuous-batching server.  submit() hands in a request (time-stamped
    on arrival); workers image admitted requests in waves; wait() blocks until
    every submitted request is done.

    hbm_budget: bytes of grids/images/work space admitted at once;
    workers: contexts (and streams) working concurrently;
    max_wave: most requests in one wave;
    cache: a KernelCache, or None to build the tables of every request afresh;
    plain: the one-job-at-a-time baseline (one worker, waves of one request,
    the single-job entry points mti_grid + mti_finalize, tables built per
    request);
    weighting/robust/poltype: imaging options shared by all requests;
    transform_storage: "c128" or "c64", the element type of the packed
    planes between the finalize's transform passes (Context.set_transform_storage);
    grid_precision: "c128" or "c64", the element type of the requests' grids
    (complex64 grids take half the device memory, so more requests fit the
    budget)."""

    def __init__(self, hbm_budget, workers=2, max_wave=256, cache=None, device=0,
                 weighting="natural", robust=0.0, poltype="linear", plain=False, transform_storage="c128",
                 grid_precision="c128"):
        import torch
        self.torch = torch
        self.budget = int(hbm_budget)
        self.plain = bool(plain)
        if self.plain:
            workers, max_wave, cache = min(int(workers), 1), 1, None
        self.max_wave = int(max_wave)
        self.cache = cache
        self.device = device
        self.weighting, self.robust, self.poltype = weighting, robust, poltype
        self.transform_storage = transform_storage
        if grid_precision not in ("c128", "c64"):
            raise ValueError(f"grid_precision {grid_precision!r}")
        self.grid_precision = grid_precision
        self.gridded = 0
        self.errors = []
        self.wave_sizes = []
        # per wave: requests, visibilities, pixels, host preparation and device seconds
        self.wave_log = []
        self._cv = threading.Condition()
        self._queue = collections.deque()
        self._in_use = 0
        self._pending = 0
        self._running = 0
        # device bytes each worker's context keeps after its last wave (its
        # scratch grows to the largest wave it ran and is not returned)
        self._ctx_bytes = [0] * int(workers)
        self._stop = False
        self._threads = []
        for w in range(int(workers)):
            t = threading.Thread(target=self._worker, args=(w,), daemon=True)
            t.start()
            self._threads.append(t)

    # ---- client side --------------------------------------------------------
    def submit(self, req, now=None):
        p = req.params
        req.nbytes = request_bytes(p.nx, p.ny, req.npol, req.nchan, req.nvis,
                                   8 if self.grid_precision == "c64" else 16)
        if req.nbytes > self.budget:
            raise MTIError(f"request {req.rid} needs {req.nbytes} bytes, above the HBM budget {self.budget}")
        req.arrival = time.perf_counter() if now is None else now
        with self._cv:
            self._queue.append(req)
            self._pending += 1
            self._cv.notify_all()

    def wait(self, timeout=None):
        """Blocks until every submitted request is done; False on timeout."""
        end = None if timeout is None else time.perf_counter() + timeout
        with self._cv:
            while self._pending:
                left = None if end is None else end - time.perf_counter()
                if left is not None and left <= 0:
                    return False
                self._cv.wait(left)
        return True

    def cancel_queued(self):
        """Drops the requests not yet admitted (their error is set); returns
        how many.  Waves already running finish normally."""
        with self._cv:
            dropped = list(self._queue)
            self._queue.clear()
            self._pending -= len(dropped)
            self._cv.notify_all()
        for r in dropped:
            r.error = MTIError("cancelled before admission")
        return len(dropped)

    def close(self):
        """Stops the workers after their current wave; requests still queued
        are dropped with an error (and no longer count as pending)."""
        self.cancel_queued()
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    # ---- admission (called with the lock held) --------------------------------
    def _head_fits(self):
        """The head request fits beside the admitted requests and the scratch
        the workers' contexts keep; with nothing in flight it is admitted
        against the budget alone (it was checked against it on submit)."""
        if not self._queue:
            return False
        if self._running == 0 and self._in_use == 0:
            return True
        return self._in_use + sum(self._ctx_bytes) + self._queue[0].nbytes <= self.budget

    def _take_wave(self):
        """Requests of the next wave: first come first served while they fit
        the budget.  A request that does not fit waits at the head of the
        queue, so later small requests cannot starve it."""
        wave = []
        while len(wave) < self.max_wave and self._head_fits():
            r = self._queue.popleft()
            self._in_use += r.nbytes
            wave.append(r)
        return wave

    # ---- workers ------------------------------------------------------------
    def _worker(self, wid_slot):
        torch = self.torch
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        ctx = Context(self.device)
        ctx.set_transform_storage(self.transform_storage)
        try:
            while True:
                with self._cv:
                    while not self._stop and not self._head_fits():
                        self._cv.wait()
                    if self._stop:
                        return
                    wave = self._take_wave()
                    self.wave_sizes.append(len(wave))
                    wid = len(self.wave_sizes)
                    self._running += 1
                t_adm = time.perf_counter()
                for r in wave:
                    r.admitted, r.wave = t_adm, wid
                try:
                    self._run_wave(ctx, stream, wave)
                except Exception as e:  # the wave fails as a whole; its requests carry the error
                    for r in wave:
                        r.error = e
                    self.errors.append(e)
                t = time.perf_counter()
                for r in wave:
                    r.done = t
                    if r.on_done is not None and r.error is None:
                        try:
                            r.on_done(r)
                        except Exception as e:  # a failing hand-off fails its request only
                            r.error = e
                    if not r.keep_image:
                        r.image = r.norm = None
                try:
                    held = ctx.device_bytes()
                except Exception:
                    held = 0
                with self._cv:
                    for r in wave:
                        self._in_use -= r.nbytes
                        self._pending -= 1
                    self._running -= 1
                    self._ctx_bytes[wid_slot] = held
                    self._cv.notify_all()
        finally:
            ctx.close()

    def _outputs(self, r, stream):
        """A request's grids, normaliser and image.  They come from the
        allocator's pool of the device's default stream and are recorded on
        the worker's stream, so the workers share one pool of freed blocks
        (separate per-stream pools would each keep their peak).  A block of
        that pool may have been freed by default-stream work that is still
        queued (a temporary of another thread's result hand-off, say): the
        worker's stream waits for the default stream before it writes them."""
        torch = self.torch
        p = r.params
        from . import tile_mask_shape
        default = torch.cuda.default_stream(self.device)
        with torch.cuda.stream(default):
            gdt = torch.float32 if self.grid_precision == "c64" else torch.float64
            out = (torch.empty((r.nchan, p.ny, p.nx, r.npol, 2), dtype=gdt, device="cuda"),
                   torch.empty((r.nchan, r.npol), dtype=torch.float64, device="cuda"),
                   torch.empty((r.nchan, r.npol, p.ny, p.nx), dtype=torch.float32, device="cuda"),
                   # the grids are internal: sparse, with a tile mask
                   torch.empty(tile_mask_shape(p, r.nchan), dtype=torch.uint8, device="cuda"))
        for t in out:
            t.record_stream(stream)
        stream.wait_stream(default)
        return out

    def _run_wave(self, ctx, stream, wave):
        torch = self.torch
        sid = stream.cuda_stream
        t0 = time.perf_counter()
        with torch.cuda.stream(stream):
            jobs, owned = [], []
            for r in wave:
                if self.plain:
                    idx = torch.empty(r.ant1.numel(), dtype=torch.int32, device="cuda")
                    kept = ctx.sort_rows(r.ant1, r.ant2, r.nant, idx, stream=sid)
                    v = vis_ms(*r.vis_args, row_index=idx[:kept])
                else:
                    # the batched call sorts the rows itself, with no host synchronisation
                    v = vis_ms(*r.vis_args)
                p = r.params
                if self.cache is not None:
                    K = self.cache.get(ctx, p, sid)
                else:
                    K = ctx.build_kernels(p, "f32", sid)
                grids, norm, image, mask = self._outputs(r, stream)
                jobs.append(ctx.make_job(K, p, v, grids, norm, image, r.chan0, r.nchan,
                                         ant1=r.ant1, ant2=r.ant2, nant=r.nant, tile_mask=mask))
                owned.append((r, K, v, grids, norm, image, mask))
            t1 = time.perf_counter()
            if self.plain:
                r, K, v, grids, norm, image, mask = owned[0]
                st = ctx.grid(K, r.params, v, grids, norm, chan0=r.chan0, nchan=r.nchan,
                              weighting=self.weighting, robust=self.robust, precision=self.grid_precision, stream=sid,
                              tile_mask=mask)
                ctx.finalize(r.params, r.npol, grids, norm, image, nchan=r.nchan, poltype=self.poltype,
                             precision=self.grid_precision, stream=sid, tile_mask=mask)
            else:
                st = ctx.grid_jobs(jobs, weighting=self.weighting, robust=self.robust, stream=sid)
                ctx.finalize_jobs(jobs, poltype=self.poltype, stream=sid)
            stream.synchronize()
        t2 = time.perf_counter()
        with self._cv:
            self.gridded += st.gridded_records
            self.wave_log.append((len(wave), sum(r.nvis for r in wave),
                                  sum(r.params.nx * r.params.ny * r.nchan * r.npol for r in wave), t1 - t0, t2 - t1))
        if len(owned) == 1:
            owned[0][0].gridded = st.gridded_records
        for r, K, v, grids, norm, image, mask in owned:
            if r.keep_image or r.on_done is not None:
                r.image, r.norm = image, norm
            if self.cache is None:
                K.free()
