"""The CPU baseline (oracle/libb64cpu.so: the reference's x86-64 commit path with
4-wide VPCLMULQDQ products, SHA extensions and a persistent thread pool) gives
the scalar oracle's outputs byte for byte, on every path the host supports:
4-wide and 1-wide products, SHA extensions and scalar rounds, 1..8 threads."""
import hashlib
import itertools

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cb():
    import subprocess
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    subprocess.run(["make", "-s", "-C", str(root / "oracle")], check=True)
    from oracle import cpu_baseline
    yield cpu_baseline
    cpu_baseline.restrict(True, True)


PATHS = list(itertools.product([True, False], [True, False]))


@pytest.mark.parametrize("wide,sha_ni", PATHS)
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_commit_step_matches_oracle(cb, orc, wide, sha_ni, threads):
    cb.restrict(wide, sha_ni)
    n = 12
    p = orc.fri_params_setup(n, 1, 96)
    basis = orc.Domain(dim=n + 1).subspace(p["rs_log_len"])
    dom = orc.Domain(basis)
    msg = orc.random_elems(77 + threads, 1 << n)
    chal = orc.random_elems(78, 4)
    base = cb.Baseline(basis, threads)
    cw, tree, folded = base.commit_step(msg, p["log_dim"], 1, p["log_batch_size"], chal)
    want_cw = orc.encode_batch(dom, msg, p["log_dim"], 1, p["log_batch_size"])
    assert np.array_equal(cw, want_cw)
    assert np.array_equal(tree, orc.merkle_build(want_cw, 16))
    assert np.array_equal(folded, orc.fold_interleaved(want_cw, p["rs_log_len"], 4, chal))


@pytest.mark.parametrize("wide", [True, False])
@pytest.mark.parametrize("log_d,se,sl,threads", [(2, 0, 0, 4), (3, 1, 0, 2), (6, 0, 1, 4), (11, 0, 0, 8),
                                                 (13, 2, 3, 5), (14, 1, 4, 16), (12, 0, 0, 1)])
def test_ntt_matches_oracle(cb, orc, wide, log_d, se, sl, threads):
    """NeighborsLastMultiThread with 4-element packing, including the layers that
    pair lanes of one packed element (skip_late < 2) and more shares than the
    transform allows."""
    cb.restrict(wide, True)
    basis = orc.to_elems([(0x9E3779B97F4A7C15 * (i + 3)) ^ (i << 90) for i in range(log_d)])
    data = orc.random_elems(5 * log_d + se, 1 << log_d)
    got = cb.Baseline(basis, threads).ntt_forward(data, se, sl)
    assert np.array_equal(got, orc.ntt_forward(orc.Domain(basis), data, se, sl))


@pytest.mark.parametrize("sha_ni", [True, False])
def test_leaf_and_node_hashes(cb, orc, sha_ni):
    """Leaves are SHA-256 of the LE16 bytes (= hashlib); nodes Sha256Compression."""
    cb.restrict(True, sha_ni)
    elems = orc.random_elems(9, 64)
    tree = cb.Baseline(orc.to_elems([1, 2]), 4).merkle_build(elems, 4)
    for i in range(16):
        assert tree[i].tobytes() == hashlib.sha256(elems[4 * i:4 * i + 4].tobytes()).digest()
    assert np.array_equal(tree, orc.merkle_build(elems, 4))


def test_features_reported(cb):
    f = cb.features()
    assert set(f) == {"vpclmulqdq_4x128", "sha_ni"}
