"""Pins the oracle's FRI parameters, folds and FRI-only prover flow.

* calculate_n_test_queries / estimate_optimal_arity KATs
  (verifier/verifier/src/fri/common.rs:225-247).
* the parameter table derived from Verifier::setup (verify.rs:61-106) for the
  configurations this repository measures.
* fold(encode(m)) = encode(fold_cols(m, eq(r))) (prover/prover/src/fri/fold.rs:334-368).
* the final FRI value equals the multilinear evaluation of the message at the
  reversed challenges (prover/prover/src/fri/tests.rs:149-156).
* every opening in the proof tape verifies against the committed roots with the
  verifier's branch rule (verifier/verifier/src/merkle_tree/scheme.rs:132-160).
"""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def test_fri_kats(orc):
    for sec, lir, want in GOLD["fri_kats"]["n_test_queries"]:
        assert orc.n_test_queries(sec, lir) == want
    lo, hi = GOLD["fri_kats"]["optimal_arity_digest256"]["log_block_length"]
    for L in range(lo, hi):
        assert orc.estimate_optimal_arity(L, 256, 128) == 4
    lo, hi = GOLD["fri_kats"]["optimal_arity_digest1024"]["log_block_length"]
    for L in range(lo, hi):
        assert orc.estimate_optimal_arity(L, 1024, 128) == 6


@pytest.mark.parametrize("n,want", [
    (12, dict(log_batch_size=4, log_dim=8, rs_log_len=9, fold_arities=[1])),
    (16, dict(log_batch_size=4, log_dim=12, rs_log_len=13, fold_arities=[4, 1])),
    (20, dict(log_batch_size=4, log_dim=16, rs_log_len=17, fold_arities=[4, 4, 1])),
    (24, dict(log_batch_size=4, log_dim=20, rs_log_len=21, fold_arities=[4, 4, 4, 1])),
])
def test_setup_parameter_table(orc, n, want):
    p = orc.fri_params_setup(n, 1, 96)
    for k, v in want.items():
        assert p[k] == v, k
    assert p["n_test_queries"] == 232


def fold_cols(orc, mat, vec):
    m = vec.shape[0]
    n = mat.shape[0] // m
    out = np.zeros((n, 2), dtype=np.uint64)
    for i in range(n):
        acc = np.zeros(2, dtype=np.uint64)
        prods = orc.gf_mul(mat[i * m:(i + 1) * m], vec)
        for p in prods:
            acc ^= p
        out[i] = acc
    return out


@pytest.mark.parametrize("log_dim,arity", [(0, 0), (3, 1), (5, 3), (7, 2), (6, 0)])
def test_fold_commutes_with_encoding(orc, log_dim, arity):
    dom = orc.Domain(dim=32)
    msg = orc.random_elems(40 + log_dim, 1 << (log_dim + arity))
    chal = orc.random_elems(41, arity)
    query = orc.eq_ind(chal)
    folded_msg = fold_cols(orc, msg, query)
    cw = orc.ntt_forward(dom, msg, 0, 0)
    folded_cw = orc.fold_codeword(dom, cw, log_dim + arity, chal)
    assert np.array_equal(folded_cw, orc.ntt_forward(dom, folded_msg, 0, 0))


def test_eq_ind_tensor(orc):
    r = orc.random_elems(3, 3)
    t = orc.eq_ind(r)
    ri = orc.to_ints(r)
    for idx in range(8):
        want = orc.to_elems([1])
        for i in range(3):
            f = ri[i] if (idx >> i) & 1 else ri[i] ^ 1
            want = orc.gf_mul(want, orc.to_elems([f]))
        assert orc.to_ints(t[idx:idx + 1]) == orc.to_ints(want)


def test_fold_interleaved_matches_rows(orc):
    log_len, lbs = 6, 3
    cw = orc.random_elems(4, 1 << (log_len + lbs))
    chal = orc.random_elems(5, lbs)
    t = orc.eq_ind(chal)
    got = orc.fold_interleaved(cw, log_len, lbs, chal)
    assert np.array_equal(got, fold_cols(orc, cw, t))
    big = orc.random_elems(6, 1 << (14 + 2))
    ch2 = orc.random_elems(7, 2)
    assert np.array_equal(orc.fold_interleaved(big, 14, 2, ch2, threads=4), orc.fold_interleaved(big, 14, 2, ch2))


def reverse_bits(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def binius_compress_py(orc, left, right):
    iv = list(struct.unpack("<8I", hashlib.sha256(b"BINIUS SHA-256 COMPRESS").digest()))
    return struct.pack("<8I", *orc.sha256_compress_block(iv, left + right))


class TapeReader:
    def __init__(self, data):
        self.data, self.pos = data, 0

    def take(self, n):
        out = self.data[self.pos:self.pos + n]
        assert len(out) == n
        self.pos += n
        return out


@pytest.mark.parametrize("log_dim,lir,lbs,arities", [
    (8, 2, 0, [1] * 7), (8, 2, 0, [3, 2, 1]), (6, 2, 2, [3, 2, 1]), (4, 2, 2, []), (8, 1, 4, [4, 1]),
])
def test_fri_flow_final_value_and_openings(orc, log_dim, lir, lbs, arities):
    n_queries = 3
    dom = orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(77, 1 << (log_dim + lbs))
    res = orc.fri_prove(dom, msg, log_dim, lir, lbs, arities, n_queries)
    chal = res["challenges"]
    # final value: terminal codeword is constant-on-cosets encoding of the
    # remaining message; fold it fully and compare with the MLE evaluation
    # (prover/prover/src/fri/tests.rs:149-156).
    n_final = log_dim - sum(arities)
    term = res["terminal"]
    remaining = chal[len(chal) - n_final:] if n_final else chal[:0]
    log_term = n_final + lir
    if n_final:
        folded = orc.fold_codeword(dom, term, log_term, remaining)
    else:
        folded = term
    point = chal[::-1].copy()
    want = orc.evaluate_mle(msg, point)
    assert all(v == want for v in orc.to_ints(folded))
    # tape structure: terminal scalars, optimal layers, then query openings.
    tape = TapeReader(res["tape"])
    # message() writes: the codeword root, then one root per commit round
    for root in res["roots"]:
        assert tape.take(32) == root
    assert tape.take(16 << log_term) == term.tobytes()
    cap = (n_queries - 1).bit_length()
    depths, log_n_cosets = [], log_dim + lir + lbs
    for a in [lbs] + list(arities):
        log_n_cosets -= a
        depths.append(min(cap, log_n_cosets))
    layers = [[tape.take(32) for _ in range(1 << d)] for d in depths]
    # each layer hashes up to the committed root
    for layer, root in zip(layers, res["roots"]):
        nodes = list(layer)
        while len(nodes) > 1:
            nodes = [binius_compress_py(orc, nodes[2 * i], nodes[2 * i + 1]) for i in range(len(nodes) // 2)]
        assert nodes[0] == root
    # query openings verify against the layers
    arity_seq = [lbs] + list(arities)
    log_lens = [log_dim + lir]
    for a in arities:
        log_lens.append(log_lens[-1] - a)
    for _ in range(n_queries):
        index = None
        for k, a in enumerate(arity_seq):
            if k == 0:
                # index is re-derived from the tape position only through the branch check
                pass
            vals = tape.take(16 << a)
            d = hashlib.sha256(vals).digest()
            tree_depth = log_lens[k]
            br = [tape.take(32) for _ in range(tree_depth - depths[k])]
            # find the coset index consistent with this opening
            matches = []
            for idx in range(1 << tree_depth):
                if index is not None and idx >> 0 != index:
                    continue
                h, j = d, idx
                for node in br:
                    h = binius_compress_py(orc, h, node) if j & 1 == 0 else binius_compress_py(orc, node, h)
                    j >>= 1
                if h == layers[k][j]:
                    matches.append(idx)
                    break
            assert matches, "opening does not verify"
            index = matches[0] >> arity_seq[k + 1] if k + 1 < len(arity_seq) else None
    assert tape.pos == len(res["tape"])


def test_fri_flow_is_deterministic_and_threads_agree(orc):
    dom = orc.Domain(dim=11)
    msg = orc.random_elems(5, 1 << 14)
    a = orc.fri_prove(dom, msg, 10, 1, 4, [4, 1], 8)
    b = orc.fri_prove(dom, msg, 10, 1, 4, [4, 1], 8, threads=4)
    assert a["tape"] == b["tape"] and a["roots"] == b["roots"]
