"""GPU ring-switch folds (prover/prover/src/ring_switch.rs) through the C ABI
against the oracle, bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


@pytest.mark.parametrize("log_len", [0, 1, 2, 3, 6, 7, 10, 14, 17, 22])
def test_fold_1b_rows(b64, orc, log_len):
    mat, vec = orc.random_elems(110 + log_len, 1 << log_len), orc.random_elems(120 + log_len, 1 << log_len)
    out = b64.ring_switch_fold_1b_rows(b64.DeviceBuffer.from_host(mat), b64.DeviceBuffer.from_host(vec), log_len)
    assert np.array_equal(out.to_host(128), orc.rs_fold_1b_rows(mat, vec))


def test_fold_1b_rows_special_rows(b64, orc):
    """All-ones and all-zero bit rows: out[b] = XOR of vec over the all-ones rows."""
    log_len = 9
    n = 1 << log_len
    mat = np.zeros((n, 2), dtype=np.uint64)
    mat[::3] = np.uint64(0xFFFFFFFFFFFFFFFF)
    vec = orc.random_elems(5, n)
    out = b64.ring_switch_fold_1b_rows(b64.DeviceBuffer.from_host(mat), b64.DeviceBuffer.from_host(vec),
                                       log_len).to_host(128)
    want = np.bitwise_xor.reduce(vec[::3], axis=0)
    assert all(np.array_equal(out[b], want) for b in range(128))


@pytest.mark.parametrize("log_len", [0, 4, 10, 16, 20])
def test_fold_elems(b64, orc, log_len):
    elems, vec = orc.random_elems(130 + log_len, 1 << log_len), orc.random_elems(140 + log_len, 128)
    buf = b64.DeviceBuffer.from_host(elems)
    b64.ring_switch_fold_elems(buf, log_len, vec)
    assert np.array_equal(buf.to_host(1 << log_len), orc.rs_fold_elems(elems, vec))


def test_fold_elems_monomial_basis_is_identity(b64, orc):
    elems = orc.random_elems(6, 1 << 12)
    buf = b64.DeviceBuffer.from_host(elems)
    b64.ring_switch_fold_elems(buf, 12, orc.to_elems([1 << b for b in range(128)]))
    assert np.array_equal(buf.to_host(1 << 12), elems)


def test_ring_switch_errors(b64):
    buf = b64.DeviceBuffer(16)
    with pytest.raises(b64.B64Error) as e:
        b64.ring_switch_fold_1b_rows(buf, buf, 41)
    assert e.value.code == 1
