"""b64_ring_switch_fold_1b_rows_noalloc (partial sums in the library's static
device array, no allocation) against b64_ring_switch_fold_1b_rows (partial sums
from the stream-ordered pool) and the oracle, through the C ABI: two calls in a
row, the second on a smaller input, so stale partial sums of the first call
would show up in the second."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


def _run(b64, fn, mat, vec, log_len):
    out = b64.DeviceBuffer(128 * 16)
    b64.check(getattr(b64.lib(), fn)(mat.ptr, vec.ptr, log_len, out.ptr, None), fn)
    return out.to_host(128)


def test_noalloc_matches_allocating_entry_point(b64, orc):
    L = b64.lib()
    # 2^20 rows fill all 2048 partial-sum blocks; 2^9 rows use 2 of them
    for call, log_len in enumerate((20, 9)):
        print(f"noalloc call {call}: log_len={log_len}", flush=True)
        mat = orc.random_elems(300 + log_len, 1 << log_len)
        vec = orc.random_elems(400 + log_len, 1 << log_len)
        d_mat, d_vec = b64.DeviceBuffer.from_host(mat), b64.DeviceBuffer.from_host(vec)
        got = _run(b64, "b64_ring_switch_fold_1b_rows_noalloc", d_mat, d_vec, log_len)
        b64.check(L.b64_stream_synchronize(None), "sync")
        want = _run(b64, "b64_ring_switch_fold_1b_rows", d_mat, d_vec, log_len)
        assert np.array_equal(got, want), f"call {call}: noalloc != allocating entry point"
        if log_len <= 12:
            assert np.array_equal(got, orc.rs_fold_1b_rows(mat, vec)), f"call {call}: noalloc != oracle"


def test_noalloc_rejects_oversize(b64):
    buf = b64.DeviceBuffer(16)
    rc = b64.lib().b64_ring_switch_fold_1b_rows_noalloc(buf.ptr, buf.ptr, 41, buf.ptr, None)
    assert rc == 1  # B64_ERR_INVALID_ARGS
    rc = b64.lib().b64_ring_switch_fold_1b_rows_noalloc(None, buf.ptr, 4, buf.ptr, None)
    assert rc == 1
