"""Host logic of the serving loop (no GPU): request sizing, HBM-budget
admission order and the kernel-table cache."""
import threading
import types

import pytest

from mtimager_amd import MTIError, make_params
from mtimager_amd import serve


class _Ant:
    def __init__(self, n):
        self.n = n

    def numel(self):
        return self.n


def _req(rid, nx, nchan, nrows, npol=1):
    p = make_params(nx, nx, 2.0, 2.0, "wproj", 8, 4)
    return serve.Request(rid, p, None, _Ant(nrows), _Ant(nrows), 8, nchan, npol=npol)


def _server(budget, max_wave=256):
    """A Server with no worker threads, to drive admission by hand."""
    s = serve.Server(budget, workers=0, max_wave=max_wave)
    return s


def test_request_bytes():
    # grids c128 + image f32 per pixel and correlation, norm f64, work per
    # visibility: descriptor + tag 20, 16 per correlation, pairs and keys 64
    assert serve.work_bytes_per_vis(1) == 100 and serve.work_bytes_per_vis(4) == 148
    assert serve.request_bytes(256, 128, 2, 3, 1000) == 256 * 128 * 2 * 3 * 20 + 8 * 2 * 3 + 116 * 1000
    r = _req(0, 512, 4, 100)
    assert r.nvis == 400
    assert r.nbytes == serve.request_bytes(512, 512, 1, 4, 400)


def test_complex64_grids_halve_the_grid_bytes():
    assert serve.request_bytes(256, 128, 2, 3, 1000, grid_bytes=8) == 256 * 128 * 2 * 3 * 12 + 8 * 2 * 3 + 116 * 1000
    r = _req(0, 512, 4, 100)
    s = serve.Server(1 << 40, workers=0, grid_precision="c64")
    s.submit(r)
    assert r.nbytes == serve.request_bytes(512, 512, 1, 4, 400, grid_bytes=8)
    with pytest.raises(ValueError):
        serve.Server(1 << 40, workers=0, grid_precision="c32")


def test_admission_fifo_within_budget():
    a, b, c = _req(0, 256, 1, 10), _req(1, 256, 1, 10), _req(2, 256, 1, 10)
    s = _server(2 * a.nbytes + 10)
    for r in (a, b, c):
        s.submit(r)
    with s._cv:
        w1 = s._take_wave()
    assert [r.rid for r in w1] == [0, 1]
    assert s._in_use == a.nbytes + b.nbytes
    with s._cv:
        assert s._take_wave() == []          # c waits for memory
        s._in_use -= a.nbytes                 # a finished
        w2 = s._take_wave()
    assert [r.rid for r in w2] == [2]


def test_head_of_line_request_is_not_starved():
    small, big, small2 = _req(0, 256, 1, 10), _req(1, 1024, 8, 10), _req(2, 256, 1, 10)
    s = _server(big.nbytes + 10)
    for r in (small, big, small2):
        s.submit(r)
    with s._cv:
        assert [r.rid for r in s._take_wave()] == [0]
        # big does not fit beside small; small2 must not overtake it
        assert s._take_wave() == []
        s._in_use -= small.nbytes
        assert [r.rid for r in s._take_wave()] == [1]


def test_max_wave():
    reqs = [_req(i, 256, 1, 10) for i in range(7)]
    s = _server(10 ** 12, max_wave=3)
    for r in reqs:
        s.submit(r)
    with s._cv:
        sizes = [len(s._take_wave()) for _ in range(3)]
    assert sizes == [3, 3, 1]


def test_request_above_budget_is_refused():
    r = _req(0, 2048, 64, 10)
    s = _server(r.nbytes - 1)
    with pytest.raises(MTIError, match="above the HBM budget"):
        s.submit(r)


def test_plain_mode_is_one_worker_one_job_no_cache():
    s = serve.Server(10 ** 9, workers=0, max_wave=64, cache=serve.KernelCache(), plain=True)
    assert s.max_wave == 1 and s.cache is None
    s.close()


def test_wait_times_out_with_pending_requests():
    s = _server(10 ** 12)
    s.submit(_req(0, 256, 1, 10))
    assert s.wait(timeout=0.05) is False


class _FakeCtx:
    def __init__(self):
        self.built = []

    def build_kernels(self, params, precision="f32", stream=None):
        k = types.SimpleNamespace(freed=False, key=serve.KernelCache.key(params))
        k.free = lambda: setattr(k, "freed", True)
        self.built.append(k)
        return k


def test_kernel_cache_keys_and_hits():
    c = serve.KernelCache()
    ctx = _FakeCtx()
    p1 = make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)
    p2 = make_params(512, 512, 2.0, 2.0, "wproj", 32, 8)   # another oversampling
    p3 = make_params(512, 512, 2.0, 2.0, "wproj", 32, 4)   # same as p1
    k1, k2, k3 = c.get(ctx, p1), c.get(ctx, p2), c.get(ctx, p3)
    assert k1 is k3 and k1 is not k2
    assert (c.hits, c.misses, len(c)) == (1, 2, 2)
    c.clear()
    assert all(k.freed for k in ctx.built) and len(c) == 0


def test_kernel_cache_builds_once_under_concurrency():
    c = serve.KernelCache()
    ctx = _FakeCtx()
    p = make_params(256, 256, 8.0, 8.0, "wproj", 16, 8)
    out = []
    ts = [threading.Thread(target=lambda: out.append(c.get(ctx, p))) for _ in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len(ctx.built) == 1 and all(k is out[0] for k in out)


def test_cancel_queued_drops_unadmitted_requests():
    a, b = _req(0, 256, 1, 10), _req(1, 256, 1, 10)
    s = _server(a.nbytes + 10)
    s.submit(a)
    s.submit(b)
    with s._cv:
        assert [r.rid for r in s._take_wave()] == [0]
    assert s.cancel_queued() == 1
    assert b.error is not None and s._pending == 1 and not s._queue


def test_close_drops_queued_requests_and_releases_wait():
    s = _server(1 << 40)
    a, b = _req(0, 256, 1, 10), _req(1, 256, 1, 10)
    s.submit(a)
    s.submit(b)
    s.close()
    assert a.error is not None and b.error is not None
    assert s.wait(timeout=0.1)  # nothing pending any more


def test_retained_context_scratch_counts_against_the_budget():
    a, b = _req(0, 256, 1, 10), _req(1, 256, 1, 10)
    s = _server(a.nbytes + b.nbytes + 10)
    s.submit(a)
    s.submit(b)
    with s._cv:
        s._ctx_bytes = [b.nbytes]          # a worker keeps scratch of this size
        s._running = 1                      # and another wave is in flight
        w = s._take_wave()
        assert [r.rid for r in w] == [0]    # b no longer fits beside a and the scratch
        s._in_use -= a.nbytes
        s._running = 0                      # nothing in flight: the head is admitted
        assert [r.rid for r in s._take_wave()] == [1]
