"""Full-size configurations (log_n = 24: a 2^24-element message, a 2^25-element
codeword of 512 MiB) on the device.

* The fused encode + Merkle commit + interleaved fold equals the oracle's
  multithreaded restatement bit for bit (codeword, every tree node, folded oracle).
* Size-independent properties of the transform at 2^25: forward then inverse is
  the identity, and the encoder is linear.
* The complete one-bit PCS evaluation proof at 2^24 B128 elements is
  byte-identical to the oracle's.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 24
THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


@pytest.fixture(scope="module")
def setup(b64, orc):
    params, dom = b64.setup_commit_params(N, 1, 96)
    assert (params.log_dim, params.log_batch_size, params.fold_arities) == (20, 4, [4, 4, 4, 1])
    ref = orc.Domain(dom.subspace(params.rs_log_len()))
    msg = orc.random_elems(2424, 1 << N)
    return params, dom, ref, msg


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_commit_fold_log_n24_matches_oracle(b64, orc, setup):
    params, dom, ref, msg = setup
    say("commit_fold log_n=24")
    chal = orc.random_elems(7, params.log_batch_size)
    d_msg = b64.DeviceBuffer.from_host(msg)
    cw = b64.DeviceBuffer(16 << params.log_len())
    tree = b64.DeviceBuffer(32 * ((2 << params.rs_log_len()) - 1))
    folded = b64.DeviceBuffer(16 << params.rs_log_len())
    root = np.zeros(32, dtype=np.uint8)
    pc = params.to_c()
    b64.check(b64.lib().b64_fri_commit_fold(C.byref(pc), dom.ptr, d_msg.ptr, cw.ptr, tree.ptr,
                                            b64._p64(np.ascontiguousarray(chal)), folded.ptr, b64._p8(root), None))
    cw_ref = orc.encode_batch(ref, msg, params.log_dim, 1, params.log_batch_size, threads=THREADS)
    assert np.array_equal(cw.to_host(1 << params.log_len()), cw_ref)
    tree_ref = orc.merkle_build(cw_ref, 1 << params.log_batch_size, threads=THREADS)
    n_nodes = (2 << params.rs_log_len()) - 1
    got_tree = np.frombuffer(tree.to_bytes(32 * n_nodes), dtype=np.uint8).reshape(n_nodes, 32)
    assert np.array_equal(got_tree, tree_ref)
    assert root.tobytes() == tree_ref[-1].tobytes()
    f_ref = orc.fold_interleaved(cw_ref, params.rs_log_len(), params.log_batch_size, chal, threads=THREADS)
    assert np.array_equal(folded.to_host(1 << params.rs_log_len()), f_ref)


def test_ntt_roundtrip_and_linearity_log_d25(b64, orc, setup):
    params, dom, ref, msg = setup
    say("ntt roundtrip / linearity log_d=25")
    log_d = params.log_len()
    ntt = b64.AdditiveNTT(dom)
    code = b64.ReedSolomonCode(dom, params.log_dim, 1)
    other = orc.random_elems(2525, 1 << N)
    enc_a = code.encode_batch(b64.DeviceBuffer.from_host(msg), params.log_batch_size)
    enc_b = code.encode_batch(b64.DeviceBuffer.from_host(other), params.log_batch_size)
    enc_ab = code.encode_batch(b64.DeviceBuffer.from_host(msg ^ other), params.log_batch_size)
    a, b, ab = (x.to_host(1 << log_d) for x in (enc_a, enc_b, enc_ab))
    assert np.array_equal(a ^ b, ab)
    # forward then inverse over the same layer window is the identity
    data = orc.random_elems(2526, 1 << log_d)
    buf = b64.DeviceBuffer.from_host(data)
    ntt.forward_transform(buf, log_d, 1, params.log_batch_size)
    mid = buf.to_host(1 << log_d)
    assert not np.array_equal(mid, data)
    ntt.inverse_transform(buf, log_d, 1, params.log_batch_size)
    assert np.array_equal(buf.to_host(1 << log_d), data)


def test_pcs_log_n24_matches_oracle(b64, orc, setup):
    params, dom, ref, msg = setup
    say("pcs log_n=24")
    point = orc.random_elems(2427, N + 7)
    co = b64.commit_interleaved(params, dom, b64.DeviceBuffer.from_host(msg))
    t = b64.ProverTranscript()
    t.message(co.commitment)
    b64.OneBitPCSProver(dom, params).prove(co.codeword, co.tree, b64.DeviceBuffer.from_host(msg), point, t)
    got = t.finalize()
    cw = co.codeword.to_host(1 << params.log_len())
    n_nodes = (2 << params.rs_log_len()) - 1
    tree = np.frombuffer(co.tree.to_bytes(32 * n_nodes), dtype=np.uint8).reshape(n_nodes, 32).copy()
    ot = orc.Transcript()
    ot.message(co.commitment)
    orc.pcs_prove(ref, msg, point, cw, tree, params.log_dim, 1, params.log_batch_size, params.fold_arities,
                  params.n_test_queries, ot, threads=THREADS)
    assert got == ot.tape()
