"""The allocation-free row sort (mti_sort_rows_static) through the C-ABI: its
order and kept count equal mti_sort_rows's bit for bit on two calls in a row
(the second smaller), the context holds no device memory afterwards, and a
row count above its static scratch is refused before any work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _rows(n, nant, seed):
    rng = np.random.default_rng(seed)
    a1 = rng.integers(0, nant, n).astype(np.int32)
    a2 = rng.integers(0, nant, n).astype(np.int32)
    return torch.tensor(a1, device="cuda"), torch.tensor(a2, device="cuda")


def test_static_sort_matches_sort_rows(mti):
    L = mti.lib()
    fresh = mti.Context(0)  # no scratch of its own yet
    ref_ctx = mti.Context(0)
    try:
        for n, nant, seed in ((3_000_000, 64, 11), (70_001, 27, 12)):
            print(f"[no-alloc] nrows={n} nant={nant} seed={seed}", flush=True)
            a1, a2 = _rows(n, nant, seed)
            exp = torch.empty(n, dtype=torch.int32, device="cuda")
            kept_exp = ref_ctx.sort_rows(a1, a2, nant, exp)
            got = torch.full((n,), -7, dtype=torch.int32, device="cuda")
            kept = ctypes.c_int64(-1)
            st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            rc = L.mti_sort_rows_static(fresh.h, a1.data_ptr(), a2.data_ptr(), n, nant, got.data_ptr(),
                                        ctypes.byref(kept), st)
            fresh.check(rc, "mti_sort_rows_static")
            torch.cuda.synchronize()
            assert kept.value == kept_exp
            assert torch.equal(got[:kept_exp], exp[:kept_exp])
            assert fresh.device_bytes() == 0
    finally:
        fresh.close()
        ref_ctx.close()


def test_static_sort_refuses_rows_above_its_scratch(mti):
    L = mti.lib()
    ctx = mti.Context(0)
    try:
        dummy = torch.zeros(1, dtype=torch.int32, device="cuda")
        kept = ctypes.c_int64(0)
        rc = L.mti_sort_rows_static(ctx.h, dummy.data_ptr(), dummy.data_ptr(), (1 << 22) + 1, 4,
                                    dummy.data_ptr(), ctypes.byref(kept), None)
        assert rc == 1  # MTI_ERR_INVALID, before any device work
        assert "MTI_SORT_STATIC_MAX_ROWS" in L.mti_last_error(ctx.h).decode()
    finally:
        ctx.close()
