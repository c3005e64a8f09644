"""Verifier-side restatements used to check proof tapes (test infrastructure).

Each function restates a reference verifier routine from its source, in Python,
over the oracle's field primitives (pinned separately) and hashlib:

* VerifierTape          -- VerifierTranscript over a prover tape: message reads
                           are observed by the challenger, decommitment reads
                           are not (verifier/transcript/src/transcript.rs).
* recover_round         -- RoundProof::recover (verifier protocols/sumcheck/common.rs:82-109).
* commit_rounds         -- calculate_fri_commit_rounds (verifier fri/fold.rs:257-278).
* verify_fri_queries    -- FRIQueryVerifier::verify (verifier fri/verify.rs:97-289)
                           with MerkleTreeScheme::verify_{vector,layer,opening}
                           (merkle_tree/scheme.rs:88-160).
* basefold_verify       -- protocols/basefold.rs verify (verifier side).
"""
from __future__ import annotations

import hashlib
import struct

import numpy as np


def as_elems(orc, x) -> np.ndarray:
    if isinstance(x, np.ndarray):
        return np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
    return orc.to_elems(x)


def fmul(orc, a: int, b: int) -> int:
    return orc.to_ints(orc.gf_mul(orc.to_elems([a]), orc.to_elems([b])))[0]


def poly_eval(orc, coeffs, x: int) -> int:
    acc = 0
    for c in reversed(coeffs):
        acc = fmul(orc, acc, x) ^ c
    return acc


def elem_bytes(v: int) -> bytes:
    return v.to_bytes(16, "little")


def bytes_elems(b: bytes):
    return [int.from_bytes(b[16 * i:16 * i + 16], "little") for i in range(len(b) // 16)]


class VerifierTape:
    def __init__(self, orc, tape: bytes):
        self.t = orc.Transcript()
        self.tape, self.pos = tape, 0

    def _take(self, n: int) -> bytes:
        out = self.tape[self.pos:self.pos + n]
        assert len(out) == n, "proof tape too short"
        self.pos += n
        return out

    def message(self, n: int) -> bytes:
        b = self._take(n)
        self.t.observe(b)
        return b

    def decommit(self, n: int) -> bytes:
        return self._take(n)

    def sample_elem(self) -> int:
        return int.from_bytes(self.t.sample(16), "little")

    def sample_bits(self, bits: int) -> int:
        return self.t.sample_bits(bits)

    def done(self) -> bool:
        return self.pos == len(self.tape)


def recover_round(c0: int, c1: int, claim: int):
    """Degree-2 round polynomial from its truncated coefficients: r(0) + r(1) = claim."""
    return [c0, c1, claim ^ c1]


def commit_rounds(lbs: int, arities, n_rounds: int):
    out = [False] * n_rounds
    idx = 0
    for a in [lbs] + list(arities):
        idx += a
        if idx < n_rounds:
            out[idx] = True
        elif idx == n_rounds:
            break
    return out


_IV = list(struct.unpack("<8I", hashlib.sha256(b"BINIUS SHA-256 COMPRESS").digest()))


def compress(orc, left: bytes, right: bytes) -> bytes:
    return struct.pack("<8I", *orc.sha256_compress_block(_IV, left + right))


class Sha256Scheme:
    """SHA-256 leaves, Sha256Compression nodes (verifier/verifier/src/hash)."""

    def __init__(self, orc):
        self.orc = orc

    def leaf(self, data: bytes) -> bytes:
        return hashlib.sha256(data).digest()

    def compress(self, left: bytes, right: bytes) -> bytes:
        return compress(self.orc, left, right)


class VisionScheme:
    """VisionHasherDigest leaves (Vision-4 or -6), VisionCompression nodes."""

    def __init__(self, leaf_vision, node_vision):
        self.leaf_v, self.node_v = leaf_vision, node_vision

    def leaf(self, data: bytes) -> bytes:
        return self.leaf_v.digest(data)

    def compress(self, left: bytes, right: bytes) -> bytes:
        return self.node_v.compress(left, right)


def fold_digests(orc, digests, scheme=None):
    scheme = scheme or Sha256Scheme(orc)
    nodes = list(digests)
    while len(nodes) > 1:
        nodes = [scheme.compress(nodes[2 * i], nodes[2 * i + 1]) for i in range(len(nodes) // 2)]
    return nodes[0]


def verify_fri_queries(orc, dom, vt: VerifierTape, log_dim: int, lir: int, lbs: int, arities, n_queries: int,
                       codeword_root: bytes, round_roots, challenges, scheme=None) -> int:
    """Returns the final FRI value; raises AssertionError on a rejected proof."""
    scheme = scheme or Sha256Scheme(orc)
    n_final = log_dim - sum(arities)
    rs_log_len = log_dim + lir
    assert len(round_roots) == len(arities) + 1
    assert len(challenges) == log_dim + lbs
    interleave = challenges[:lbs]
    fold_ch = challenges[lbs:]
    tensor = orc.to_ints(orc.eq_ind(orc.to_elems(interleave))) if lbs else [1]

    # verify_last_oracle: the terminal codeword hashes to the last root, and
    # folding each coset with the final challenges gives one repeated value.
    term = bytes_elems(vt.decommit(16 << (n_final + lir)))
    leaves = [scheme.leaf(b"".join(elem_bytes(v) for v in term[i:i + (1 << n_final)]))
              for i in range(0, len(term), 1 << n_final)]
    assert fold_digests(orc, leaves, scheme) == round_roots[-1], "terminal codeword root"
    final_ch = orc.to_elems(fold_ch[len(fold_ch) - n_final:]) if n_final else None
    rep = []
    for i in range(len(term) >> n_final):
        chunk = orc.to_elems(term[i << n_final:(i + 1) << n_final])
        rep.append(orc.fold_chunk(dom, n_final + lir, i, chunk, final_ch) if n_final else term[i])
    assert all(v == rep[0] for v in rep), "IncorrectDegree"
    final_value = rep[0]

    # optimal layers (vcs_optimal_layers_depths_iter, verifier fri/common.rs:154-168)
    cap = (n_queries - 1).bit_length()
    depths, log_n_cosets = [], log_dim + lir + lbs
    for a in [lbs] + list(arities):
        log_n_cosets -= a
        depths.append(min(cap, log_n_cosets))
    layers = [[vt.decommit(32) for _ in range(1 << d)] for d in depths]
    for root, layer in zip([codeword_root] + list(round_roots), layers):
        assert fold_digests(orc, layer, scheme) == root, "layer does not hash to its root"

    def opening(index, log_coset, depth, tree_depth, layer):
        vals = bytes_elems(vt.decommit(16 << log_coset))
        h = scheme.leaf(b"".join(elem_bytes(v) for v in vals))
        j = index
        for _ in range(tree_depth - depth):
            node = vt.decommit(32)
            h = scheme.compress(h, node) if j & 1 == 0 else scheme.compress(node, h)
            j >>= 1
        assert h == layer[j], "opening does not verify"
        return vals

    for _ in range(n_queries):
        index = vt.sample_bits(rs_log_len)
        vals = opening(index, lbs, depths[0], rs_log_len, layers[0])
        nxt = 0
        for v, t in zip(vals, tensor):
            nxt ^= fmul(orc, v, t)
        fold_round, log_nc = 0, rs_log_len
        for k, a in enumerate(arities):
            coset = index >> a
            log_nc -= a
            vals = opening(coset, a, depths[k + 1], log_nc, layers[k + 1])
            assert nxt == vals[index % (1 << a)], "IncorrectFold"
            nxt = orc.fold_chunk(dom, rs_log_len - fold_round, coset, orc.to_elems(vals),
                                 orc.to_elems(fold_ch[fold_round:fold_round + a]))
            index = coset
            fold_round += a
        assert nxt == term[index], "IncorrectFold (terminal)"
    return final_value


def basefold_verify(orc, dom, vt: VerifierTape, log_dim: int, lir: int, lbs: int, arities, n_queries: int,
                    codeword_root: bytes, claim: int, scheme=None):
    """basefold::verify (verifier protocols/basefold.rs): returns
    (final_fri_value, final_sumcheck_value, challenges)."""
    n_vars = log_dim + lbs
    commits = commit_rounds(lbs, arities, n_vars + 1)
    roots, challenges, s = [], [], claim
    for rnd in range(n_vars):
        c0, c1 = bytes_elems(vt.message(32))
        if commits[rnd]:
            roots.append(vt.message(32))
        coeffs = recover_round(c0, c1, s)
        r = vt.sample_elem()
        s = poly_eval(orc, coeffs, r)
        challenges.append(r)
    if commits[n_vars]:
        roots.append(vt.message(32))
    final_fri = verify_fri_queries(orc, dom, vt, log_dim, lir, lbs, arities, n_queries, codeword_root, roots,
                                   challenges, scheme)
    return final_fri, s, challenges


def rs_eq_ind_definition(orc, suffix_point, batching) -> np.ndarray:
    """The ring-switching equality indicator from its definition: entry x is the
    sum over bits b of eq(suffix_point)[x] of eq(batching)[b]
    (prover/prover/src/ring_switch.rs:130-176 applied to the eq tensor)."""
    eq_s = np.asarray(orc.eq_ind(as_elems(orc, suffix_point)), dtype=np.uint64)
    eq_b = np.asarray(orc.eq_ind(as_elems(orc, batching)), dtype=np.uint64)
    out = np.zeros_like(eq_s)
    for x in range(eq_s.shape[0]):
        v = int(eq_s[x, 0]) | (int(eq_s[x, 1]) << 64)
        bits = [b for b in range(128) if (v >> b) & 1]
        if bits:
            out[x] = np.bitwise_xor.reduce(eq_b[bits], axis=0)
    return out
