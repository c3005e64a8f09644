"""GPU parity: every kernel path through the C ABI against the oracle, bit-exact.

Sizes are ones the oracle finishes in seconds; the full-size configurations are
covered in test_gpu_large.py through the oracle's multithreaded restatement and
size-independent properties.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_device_mul_kats_and_random(b64, orc):
    kats = GOLD["ghash_mul_kats"]
    a = orc.to_elems(int(x[0], 16) for x in kats)
    b = orc.to_elems(int(x[1], 16) for x in kats)
    assert orc.to_ints(b64.gf_mul_device(a, b)) == [int(x[2], 16) for x in kats]
    a = orc.random_elems(1, 1 << 16)
    b = orc.random_elems(2, 1 << 16)
    assert np.array_equal(b64.gf_mul_device(a, b), orc.gf_mul(a, b))


def test_device_mul_aes_embedding(b64, orc):
    phi = orc.to_elems(int(x, 16) for x in GOLD["aes_to_ghash"])
    A, B = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    got = b64.gf_mul_device(phi[A.ravel()], phi[B.ravel()])
    assert np.array_equal(got, orc.gf_mul(phi[A.ravel()], phi[B.ravel()]))


@pytest.mark.parametrize("c", [0, 1, 2, 0x87, (1 << 127), 0x494EF99794D5244F9152DF59D87A9186])
def test_device_mul_const_table(b64, orc, c):
    a = orc.random_elems(3, 5000)
    got = b64.gf_mul_const_device(a, c)
    assert np.array_equal(got, orc.gf_mul(a, orc.to_elems([c] * a.shape[0])))


def ntt_dev(b64, dom, data, se, sl, inverse=False):
    buf = b64.DeviceBuffer.from_host(data)
    ntt = b64.AdditiveNTT(dom)
    n = data.shape[0]
    log_d = n.bit_length() - 1
    if inverse:
        ntt.inverse_transform(buf, log_d, se, sl)
    else:
        ntt.forward_transform(buf, log_d, se, sl)
    return buf.to_host(n)


@pytest.mark.parametrize("log_d", [0, 1, 2, 5, 8, 11, 12, 14, 17])
def test_ntt_forward_windows(b64, orc, log_d):
    dom, ref = b64.DomainContext.standard(log_d + 1), orc.Domain(dim=log_d + 1)
    data = orc.random_elems(100 + log_d, 1 << log_d)
    h = log_d // 2
    for se, sl in sorted({(0, 0), (1, 0), (0, 1), (h, 0), (0, log_d - h), (1, 4), (3, 2)}):
        if se + sl > log_d:
            continue
        say("ntt", log_d, se, sl)
        assert np.array_equal(ntt_dev(b64, dom, data, se, sl), orc.ntt_forward(ref, data, se, sl)), (se, sl)


@pytest.mark.parametrize("variant,max_k,tile_log,tab_min,tab_group,waves", [
    (0, 7, 11, 1, 2, 0), (1, 7, 11, 1, 2, 0), (1, 7, 11, 128, 2, 0), (1, 7, 11, 1 << 20, 2, 0), (1, 3, 6, 4, 2, 0),
    (1, 1, 1, 1, 2, 0), (1, 5, 9, 64, 2, 0), (0, 2, 4, 1, 2, 0), (1, 4, 11, 32, 2, 0), (1, 7, 11, 1, 4, 0),
    (1, 6, 11, 16, 8, 0), (1, 5, 10, 8, 8, 0), (1, 3, 11, 1, 4, 0), (1, 5, 10, 16, 2, 4), (1, 4, 9, 1, 4, 4),
    (1, 7, 11, 1, 2, 4), (1, 3, 10, 1, 8, 4), (1, 4, 11, 64, 1, 0), (1, 7, 11, 1, 1, 0),
    (1, 3, 6, 4, 1, 4), (1, 5, 10, 16, 1, 0), (2, 4, 11, 128, 2, 0), (2, 4, 6, 8, 2, 0),
    (3, 4, 9, 16, 2, 0), (3, 4, 11, 8, 2, 0), (4, 4, 10, 16, 2, 0), (3, 7, 11, 1, 2, 0), (4, 3, 9, 1 << 20, 2, 0),
    (4, 5, 10, 2, 2, 0), (3, 4, 9, 16, 1, 0), (4, 4, 10, 16, 1, 0), (3, 3, 8, 1, 1, 0)])
def test_ntt_tuning_variants(b64, orc, variant, max_k, tile_log, tab_min, tab_group, waves):
    """Every multiply variant and pass/tile shape gives the same transform."""
    old = {k: b64.config_get(k) for k in ("ntt.variant", "ntt.max_k", "ntt.tile_log", "ntt.tab_min", "ntt.tab_group",
                                          "ntt.waves")}
    b64.config_set("ntt.clear_sizes", 0)  # per-size schedules would override the defaults set here
    try:
        b64.config_set("ntt.variant", variant)
        b64.config_set("ntt.max_k", max_k)
        b64.config_set("ntt.tile_log", tile_log)
        b64.config_set("ntt.tab_min", tab_min)
        b64.config_set("ntt.tab_group", tab_group)
        b64.config_set("ntt.waves", waves)
        for log_d, se, sl in ((13, 1, 4), (12, 0, 0), (9, 2, 1), (3, 0, 0)):
            dom, ref = b64.DomainContext.standard(log_d), orc.Domain(dim=log_d)
            data = orc.random_elems(7 * log_d, 1 << log_d)
            say("ntt-variant", variant, max_k, tile_log, log_d, se, sl)
            assert np.array_equal(ntt_dev(b64, dom, data, se, sl), orc.ntt_forward(ref, data, se, sl))
            assert np.array_equal(ntt_dev(b64, dom, data, se, sl, inverse=True), orc.ntt_inverse(ref, data, se, sl))
    finally:
        for k, v in old.items():
            b64.config_set(k, v)
        from binius64_amd._abi import _apply_tuning
        _apply_tuning(b64.lib())


@pytest.mark.parametrize("variant,max_k,tile_log,tab_min,tab_group", [(1, 4, 6, 8, 2), (1, 7, 11, 128, 2),
                                                                    (1, 4, 11, 64, 1), (1, 5, 10, 64, 2),
                                                                    (2, 4, 11, 128, 2), (2, 4, 10, 64, 2),
                                                                    (3, 4, 9, 16, 2), (4, 4, 9, 16, 2),
                                                                    (3, 4, 11, 8, 2), (4, 4, 10, 32, 2),
                                                                    (3, 4, 9, 16, 1), (3, 4, 10, 16, 1)])
def test_ntt_schedules_log_d20(b64, orc, variant, max_k, tile_log, tab_min, tab_group):
    """Schedules at a size where the last passes take extra index bits into
    their tiles, and where variant 2 runs the upper 4-layer passes in registers
    (including the message-gathering one): forward, inverse and the encode."""
    log_d = 20
    cand = (variant, max_k, tile_log, tab_min, tab_group)
    dom, ref = b64.DomainContext.standard(log_d), orc.Domain(dim=log_d)
    data = orc.random_elems(2024, 1 << log_d)
    try:
        for key, v in zip(("ntt.variant", "ntt.max_k", "ntt.tile_log", "ntt.tab_min", "ntt.tab_group"), cand):
            b64.config_set(f"{key}:{log_d}", v)
        say("ntt-many-tiles", cand)
        assert np.array_equal(ntt_dev(b64, dom, data, 1, 4), orc.ntt_forward(ref, data, 1, 4))
        assert np.array_equal(ntt_dev(b64, dom, data, 0, 0, inverse=True), orc.ntt_inverse(ref, data, 0, 0))
        msg = data[: 1 << (log_d - 1)]
        out = b64.ReedSolomonCode(dom, log_d - 5, 1).encode_batch(b64.DeviceBuffer.from_host(msg), 4)
        assert np.array_equal(out.to_host(1 << log_d), orc.encode_batch(ref, msg, log_d - 5, 1, 4, threads=8))
    finally:
        from binius64_amd._abi import _apply_tuning
        b64.config_set("ntt.clear_sizes", 0)
        _apply_tuning(b64.lib())


@pytest.mark.parametrize("variant", [2, 3, 4])
@pytest.mark.parametrize("log_dim,lir,lbs", [(16, 2, 0), (15, 1, 2), (14, 3, 1)])
def test_encode_register_passes(b64, orc, log_dim, lir, lbs, variant):
    """Variants 2..4 on encodes whose message-gathering pass runs in registers, with
    rates 1/2..1/8 (the replicas are the hi groups of that pass) and batches
    of 1, 2 and 4 polynomials; a nonstandard basis for the first case."""
    log_d = log_dim + lir + lbs
    if log_dim == 16:
        basis = [(0x9E3779B97F4A7C15 * (i + 1)) ^ (i << 70) for i in range(log_dim + lir)]
        dom, ref = b64.DomainContext(orc.to_elems(basis)), orc.Domain(orc.to_elems(basis))
    else:
        dom, ref = b64.DomainContext.standard(log_dim + lir), orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(4242 + log_d, 1 << (log_dim + lbs))
    try:
        for key, v in zip(("ntt.variant", "ntt.max_k", "ntt.tile_log", "ntt.tab_min", "ntt.tab_group"),
                          (variant, 4, 11, 64, 2)):
            b64.config_set(f"{key}:{log_d}", v)
        say("encode-r16", variant, log_dim, lir, lbs)
        out = b64.ReedSolomonCode(dom, log_dim, lir).encode_batch(b64.DeviceBuffer.from_host(msg), lbs)
        assert np.array_equal(out.to_host(1 << log_d), orc.encode_batch(ref, msg, log_dim, lir, lbs, threads=8))
    finally:
        from binius64_amd._abi import _apply_tuning
        b64.config_set("ntt.clear_sizes", 0)
        _apply_tuning(b64.lib())


def test_per_size_tuning_override(b64, orc):
    """A per-size schedule applies to its transform size only, and any schedule
    gives the oracle's transform."""
    log_d = 14
    dom, ref = b64.DomainContext.standard(log_d), orc.Domain(dim=log_d)
    data = orc.random_elems(77, 1 << log_d)
    want = orc.ntt_forward(ref, data, 1, 4)
    try:
        for cand in ((1, 3, 8, 16, 4), (0, 6, 11, 1, 2), (1, 7, 11, 1, 8)):
            for key, v in zip(("ntt.variant", "ntt.max_k", "ntt.tile_log", "ntt.tab_min", "ntt.tab_group"), cand):
                b64.config_set(f"{key}:{log_d}", v)
            assert b64.config_get(f"ntt.max_k:{log_d}") == cand[1]
            assert np.array_equal(ntt_dev(b64, dom, data, 1, 4), want), cand
    finally:
        from binius64_amd._abi import _apply_tuning
        b64.config_set("ntt.clear_sizes", 0)
        _apply_tuning(b64.lib())


@pytest.mark.parametrize("basis", [[5, 7, 22, 95], [0x1234, 0x98765, 0xABCDEF, 0x1111111111, 0x3 << 100, 7, 9, 1 << 77]])
def test_ntt_nonstandard_basis(b64, orc, basis):
    dom, ref = b64.DomainContext(orc.to_elems(basis)), orc.Domain(orc.to_elems(basis))
    d = len(basis)
    data = orc.random_elems(9, 1 << d)
    assert np.array_equal(ntt_dev(b64, dom, data, 0, 0), orc.ntt_forward(ref, data, 0, 0))


@pytest.mark.parametrize("log_d", [1, 6, 13])
def test_ntt_inverse(b64, orc, log_d):
    dom, ref = b64.DomainContext.standard(log_d), orc.Domain(dim=log_d)
    data = orc.random_elems(4, 1 << log_d)
    for se, sl in ((0, 0), (1, 0), (0, 1), (1, 1)):
        if se + sl > log_d:
            continue
        got = ntt_dev(b64, dom, data, se, sl, inverse=True)
        assert np.array_equal(got, orc.ntt_inverse(ref, data, se, sl))
        assert np.array_equal(ntt_dev(b64, dom, got, se, sl), data)


def test_ntt_precondition_errors(b64):
    dom = b64.DomainContext.standard(4)
    buf = b64.DeviceBuffer(16 << 6)
    with pytest.raises(b64.B64Error):
        b64.AdditiveNTT(dom).forward_transform(buf, 6, 0, 0)
    with pytest.raises(b64.B64Error):
        b64.AdditiveNTT(dom).forward_transform(buf, 3, 2, 2)


@pytest.mark.parametrize("log_dim,lir,lbs", [(4, 2, 0), (6, 2, 1), (8, 3, 2), (1, 2, 0), (16, 1, 0), (12, 1, 4)])
def test_encode_batch(b64, orc, log_dim, lir, lbs):
    """Config 1 is (16, 1, 0): 2^16 message elements, standard 17-dim subspace."""
    say("encode", log_dim, lir, lbs)
    dom, ref = b64.DomainContext.standard(log_dim + lir), orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(50, 1 << (log_dim + lbs))
    code = b64.ReedSolomonCode(dom, log_dim, lir)
    out = code.encode_batch(b64.DeviceBuffer.from_host(msg), lbs)
    got = out.to_host(1 << (log_dim + lir + lbs))
    assert np.array_equal(got, orc.encode_batch(ref, msg, log_dim, lir, lbs, threads=8))


def test_encode_dimension_mismatch(b64):
    dom = b64.DomainContext.standard(5)
    with pytest.raises(b64.B64Error) as e:
        b64.ReedSolomonCode(dom, 5, 1).encode_batch(b64.DeviceBuffer(16 << 5), 0)
    assert e.value.code == 3


@pytest.mark.parametrize("n_elems,batch", [(1, 1), (2, 1), (16, 16), (64, 4), (1 << 12, 16), (256, 2),
                                           (1 << 10, 128), (8, 8), (12, 3),
                                           # leaves of 32..128 elements: one block per leaf up to
                                           # 256 leaves, one thread per leaf above
                                           (1 << 11, 32), (1 << 14, 64), (1 << 15, 64), (1 << 8, 128),
                                           # 2^18 and 2^19 leaves: the widest levels two at a time
                                           # per thread, then the subtree steps
                                           (1 << 18, 1), (1 << 20, 2)])
def test_merkle_commit(b64, orc, n_elems, batch):
    elems = orc.random_elems(60 + n_elems, n_elems)
    root, tree = b64.BinaryMerkleTreeProver.commit(b64.DeviceBuffer.from_host(elems), n_elems, batch)
    ref = orc.merkle_build(elems, batch)
    n_leaves = n_elems // batch
    got = np.frombuffer(tree.to_bytes(32 * (2 * n_leaves - 1)), dtype=np.uint8).reshape(-1, 32)
    assert np.array_equal(got, ref)
    assert root == ref[-1].tobytes()
    log_len = n_leaves.bit_length() - 1
    for depth in range(log_len + 1):
        layer = b64.BinaryMerkleTreeProver.layer(tree, log_len, depth)
        start = (2 * n_leaves - 1) + 1 - (2 << depth)
        assert layer == [ref[start + i].tobytes() for i in range(1 << depth)]
    for idx in (0, n_leaves - 1, n_leaves // 3):
        br = b64.BinaryMerkleTreeProver.branch(tree, log_len, idx, 0)
        assert br == [x.tobytes() for x in orc.merkle_branch(ref, log_len, idx, 0)]


def test_merkle_errors(b64):
    buf = b64.DeviceBuffer(16 * 12)
    with pytest.raises(b64.B64Error) as e:
        b64.BinaryMerkleTreeProver.commit(buf, 12, 5)
    assert e.value.code == 4
    with pytest.raises(b64.B64Error) as e:
        b64.BinaryMerkleTreeProver.commit(buf, 12, 4)
    assert e.value.code == 5


@pytest.mark.parametrize("log_len,lbs", [(0, 0), (3, 1), (10, 4), (14, 4), (9, 6)])
def test_fold_interleaved(b64, orc, log_len, lbs):
    cw = orc.random_elems(70, 1 << (log_len + lbs))
    chal = orc.random_elems(71, lbs)
    out = b64.DeviceBuffer(16 << log_len)
    d_cw = b64.DeviceBuffer.from_host(cw)
    ch = np.ascontiguousarray(chal, dtype=np.uint64)
    b64.check(b64.lib().b64_fri_fold_interleaved(d_cw.ptr, log_len, lbs, b64._p64(ch) if lbs else None, out.ptr, None))
    assert np.array_equal(out.to_host(1 << log_len), orc.fold_interleaved(cw, log_len, lbs, chal))


@pytest.mark.parametrize("log_len,n_chal", [(1, 1), (2, 2), (4, 4), (6, 2), (5, 3), (13, 4), (17, 4), (21, 4),
                                              (9, 1), (12, 6)])
def test_fold_codeword(b64, orc, log_len, n_chal):
    dom, ref = b64.DomainContext.standard(log_len + 2), orc.Domain(dim=log_len + 2)
    cw = orc.random_elems(80, 1 << log_len)
    chal = orc.random_elems(81, n_chal)
    d_cw = b64.DeviceBuffer.from_host(cw)
    out = b64.DeviceBuffer(16 << (log_len - n_chal))
    ch = np.ascontiguousarray(chal, dtype=np.uint64)
    b64.check(b64.lib().b64_fri_fold_codeword(dom.ptr, d_cw.ptr, log_len, b64._p64(ch), n_chal, out.ptr, None))
    assert np.array_equal(out.to_host(1 << (log_len - n_chal)), orc.fold_codeword(ref, cw, log_len, chal))


@pytest.mark.parametrize("n", [8, 12, 16])
def test_commit_fold_fused(b64, orc, n):
    """b64_fri_commit_fold (encode + leaf hash + first interleaved fold in one
    pass over the codeword) = oracle encode, Merkle root and fold_interleaved."""
    import ctypes as C
    params, dom = b64.setup_commit_params(n, 1, 96)
    lbs, rs = params.log_batch_size, params.rs_log_len()
    msg = orc.random_elems(300 + n, 1 << n)
    chal = np.ascontiguousarray(orc.random_elems(301, lbs))
    d_msg = b64.DeviceBuffer.from_host(msg)
    cw = b64.DeviceBuffer(16 << params.log_len())
    tree = b64.DeviceBuffer(32 * ((2 << rs) - 1))
    folded = b64.DeviceBuffer(16 << rs)
    root = np.zeros(32, dtype=np.uint8)
    pc = params.to_c()
    b64.check(b64.lib().b64_fri_commit_fold(C.byref(pc), dom.ptr, d_msg.ptr, cw.ptr, tree.ptr, b64._p64(chal),
                                            folded.ptr, b64._p8(root), None))
    ref_dom = orc.Domain(dom.subspace(rs))
    cw_ref = orc.encode_batch(ref_dom, msg, params.log_dim, 1, lbs, threads=8)
    assert np.array_equal(cw.to_host(1 << params.log_len()), cw_ref)
    tree_ref = orc.merkle_build(cw_ref, 1 << lbs, threads=8)
    assert root.tobytes() == tree_ref[-1].tobytes()
    got_tree = np.frombuffer(tree.to_bytes(32 * ((2 << rs) - 1)), dtype=np.uint8).reshape(-1, 32)
    assert np.array_equal(got_tree, tree_ref)
    assert np.array_equal(folded.to_host(1 << rs), orc.fold_interleaved(cw_ref, rs, lbs, chal, threads=8))


def test_pack_witness(b64, orc):
    words = orc.random_words(3, 1000)
    d = b64.DeviceBuffer.from_host(words)
    out = b64.pack_witness(d, 1000, 10)
    got = out.to_host(1024)
    want = np.zeros((1024, 2), dtype=np.uint64)
    want[:500] = orc.pack_witness(words)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("log_dim,lir,lbs,arities,nq", [
    (8, 2, 0, [1] * 7, 3), (8, 2, 0, [3, 2, 1], 3), (6, 2, 2, [3, 2, 1], 3), (4, 2, 2, [], 3),
    (8, 1, 4, [4, 1], 232), (10, 1, 4, [4, 4, 1], 232),
])
def test_fri_proof_tape_matches_oracle(b64, orc, log_dim, lir, lbs, arities, nq):
    say("fri", log_dim, lir, lbs, arities)
    dom, ref = b64.DomainContext.standard(log_dim + lir), orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(90, 1 << (log_dim + lbs))
    params = b64.FRIParams.new(log_dim, lir, lbs, arities, nq)
    tape, roots, chal = b64.fri_prove(params, dom, msg)
    want = orc.fri_prove(ref, msg, log_dim, lir, lbs, arities, nq, threads=8)
    assert roots == want["roots"]
    assert orc.to_ints(want["challenges"]) == chal
    assert tape == want["tape"]


def test_fri_commit_config2_setup_chain(b64, orc):
    """Config 2: 2^21 random u64 witness words -> pack_witness -> 2^20 B128,
    Verifier::setup parameters (rs (16,1), batch 2^4, arities [4,4,1], 232 queries);
    every root, every folded oracle and the proof tape are byte-identical."""
    n = 20
    params, dom = b64.setup_commit_params(n, 1, 96)
    assert (params.log_dim, params.log_batch_size, params.fold_arities) == (16, 4, [4, 4, 1])
    words = orc.random_words(1, 1 << (n + 1))
    d_words = b64.DeviceBuffer.from_host(words)
    msg_dev = b64.pack_witness(d_words, 1 << (n + 1), n)
    msg = orc.pack_witness(words)
    assert np.array_equal(msg_dev.to_host(1 << n), msg)
    tape, roots, chal = b64.fri_prove(params, dom, msg)
    ref = orc.Domain(dom.subspace(params.rs_log_len()))
    want = orc.fri_prove(ref, msg, params.log_dim, 1, params.log_batch_size, params.fold_arities, 232, threads=8)
    assert roots == want["roots"]
    assert tape == want["tape"]
