"""Pins the oracle's bivariate-product sumcheck, BaseFold prover and one-bit PCS
opening with restated verifiers (tests/fri_verify.py):

* sumcheck: every round polynomial recovered from the tape satisfies the
  running claim, and the final claim equals a(r) * b(r) computed by direct
  multilinear evaluation at the reversed challenges
  (prover protocols/sumcheck/bivariate_product.rs:133-214 test flow).
* BaseFold: basefold::verify accepts (round polynomials + FRI queries), and
  sumcheck_fri_consistency holds: final sumcheck value = final FRI value *
  eq(eval point, reversed challenges) (prover protocols/basefold.rs:186-254).
* PCS: pcs::verify (verifier/verifier/src/pcs.rs:41-113) accepts, with the
  evaluation claim computed by brute force over the B1 multilinear and the
  ring-switching equality indicator from its definition.
Each accepting check is paired with a tampered claim that the same check
rejects, so a check that could not fail pins nothing.
"""
import numpy as np
import pytest

import fri_verify as fv


def _claim_of(orc, a, b) -> int:
    acc = 0
    for x, y in zip(orc.to_ints(a), orc.to_ints(b)):
        acc ^= fv.fmul(orc, x, y)
    return acc


def _sumcheck_verify(orc, tape, n_vars, claim):
    vt = fv.VerifierTape(orc, tape)
    s, ch = claim, []
    for _ in range(n_vars):
        c0, c1 = fv.bytes_elems(vt.message(32))
        coeffs = fv.recover_round(c0, c1, s)
        r = vt.sample_elem()
        s = fv.poly_eval(orc, coeffs, r)
        ch.append(r)
    return s, ch


@pytest.mark.parametrize("n_vars", [0, 1, 2, 5, 8])
def test_sumcheck_verifies(orc, n_vars):
    a = orc.random_elems(100 + n_vars, 1 << n_vars)
    b = orc.random_elems(200 + n_vars, 1 << n_vars)
    claim = _claim_of(orc, a, b)
    t = orc.Transcript()
    out = orc.sumcheck_prove(a, b, claim, t)
    tape = t.tape()
    assert len(tape) == 32 * n_vars
    s, ch = _sumcheck_verify(orc, tape, n_vars, claim)
    assert orc.to_ints(out["challenges"]) == ch
    point = orc.to_elems(ch[::-1])
    ea, eb = (orc.evaluate_mle(a, point), orc.evaluate_mle(b, point)) if n_vars else (
        orc.to_ints(a)[0], orc.to_ints(b)[0])
    assert out["evals"] == [ea, eb]
    assert s == fv.fmul(orc, ea, eb)


def test_sumcheck_wrong_claim_fails_final_check(orc):
    n_vars = 6
    a = orc.random_elems(7, 1 << n_vars)
    b = orc.random_elems(8, 1 << n_vars)
    claim = _claim_of(orc, a, b) ^ 1
    t = orc.Transcript()
    out = orc.sumcheck_prove(a, b, claim, t)
    s, _ = _sumcheck_verify(orc, t.tape(), n_vars, claim)
    ea, eb = out["evals"]
    assert s != fv.fmul(orc, ea, eb)


def _commit(orc, dom, msg, log_dim, lir, lbs):
    cw = orc.encode_batch(dom, msg, log_dim, lir, lbs)
    tree = orc.merkle_build(cw, 1 << lbs)
    return cw, tree


# (log_dim, lir, lbs, arities, n_queries): the reference test's shape
# (lbs 0, arity-1 folds), mixed arities, and the Verifier::setup chain at n = 12, 16.
BASEFOLD_CASES = [
    (7, 1, 1, [1] * 6, 3),
    (6, 2, 2, [3, 2, 1], 4),
    (8, 1, 4, [1], 232),
    (12, 1, 4, [4, 1], 232),
]


@pytest.mark.parametrize("log_dim,lir,lbs,arities,nq", BASEFOLD_CASES)
def test_basefold_verifies_and_is_consistent(orc, log_dim, lir, lbs, arities, nq):
    n_vars = log_dim + lbs
    dom = orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(31 + n_vars, 1 << n_vars)
    point = orc.random_elems(41 + n_vars, n_vars)
    eq = orc.eq_ind(point)
    claim = orc.evaluate_mle(msg, point)
    cw, tree = _commit(orc, dom, msg, log_dim, lir, lbs)
    root = tree[-1].tobytes()
    t = orc.Transcript()
    t.message(root)
    ch = orc.basefold_prove(dom, msg, eq, claim, cw, tree, log_dim, lir, lbs, arities, nq, t, threads=4)
    tape = t.tape()

    vt = fv.VerifierTape(orc, tape)
    assert vt.message(32) == root
    final_fri, final_sum, chals = fv.basefold_verify(orc, dom, vt, log_dim, lir, lbs, arities, nq, root, claim)
    assert vt.done()
    assert orc.to_ints(ch) == chals
    rev = orc.to_elems(chals[::-1])
    # sumcheck_fri_consistency
    assert final_sum == fv.fmul(orc, final_fri, orc.evaluate_mle(eq, rev))
    # the FRI final value is the message's multilinear evaluation (fri/tests.rs:149-156)
    assert final_fri == orc.evaluate_mle(msg, rev)


def test_basefold_wrong_claim_is_inconsistent(orc):
    log_dim, lir, lbs, arities, nq = 6, 1, 2, [2, 2], 3
    n_vars = log_dim + lbs
    dom = orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(5, 1 << n_vars)
    point = orc.random_elems(6, n_vars)
    eq = orc.eq_ind(point)
    claim = orc.evaluate_mle(msg, point) ^ 1
    cw, tree = _commit(orc, dom, msg, log_dim, lir, lbs)
    root = tree[-1].tobytes()
    t = orc.Transcript()
    t.message(root)
    orc.basefold_prove(dom, msg, eq, claim, cw, tree, log_dim, lir, lbs, arities, nq, t)
    vt = fv.VerifierTape(orc, t.tape())
    vt.message(32)
    final_fri, final_sum, chals = fv.basefold_verify(orc, dom, vt, log_dim, lir, lbs, arities, nq, root, claim)
    assert final_sum != fv.fmul(orc, final_fri, orc.evaluate_mle(eq, orc.to_elems(chals[::-1])))


def test_tensor_transpose_is_bit_transpose(orc):
    e = orc.random_elems(9, 128)
    t = orc.to_ints(orc.tensor_transpose(e))
    v = orc.to_ints(e)
    for i in range(128):
        for j in (0, 1, 63, 64, 127, (i * 37) % 128):
            assert (t[i] >> j) & 1 == (v[j] >> i) & 1
    assert orc.to_ints(orc.tensor_transpose(orc.tensor_transpose(e))) == v


def _b1_evaluation(orc, packed, point) -> int:
    """The B1 multilinear (bit b of element i is entry 128 i + b) evaluated at point."""
    eq_lo = np.asarray(orc.eq_ind(point[:7]), dtype=np.uint64)
    eq_hi = orc.to_ints(orc.eq_ind(point[7:]))
    acc = 0
    for i, v in enumerate(orc.to_ints(packed)):
        bits = [b for b in range(128) if (v >> b) & 1]
        if bits:
            w = np.bitwise_xor.reduce(eq_lo[bits], axis=0)
            acc ^= fv.fmul(orc, int(w[0]) | (int(w[1]) << 64), eq_hi[i])
    return acc


def pcs_verify(orc, dom, tape, root, claim, point, log_dim, lir, lbs, arities, nq):
    """pcs::verify (verifier/verifier/src/pcs.rs:41-113); True iff accepted."""
    vt = fv.VerifierTape(orc, tape)
    assert vt.message(32) == root
    s_hat_v = orc.to_elems(fv.bytes_elems(vt.message(16 * 128)))
    if orc.evaluate_mle(s_hat_v, point[:7]) != claim:
        return False  # EvaluationClaimMismatch
    s_hat_u = orc.tensor_transpose(s_hat_v)
    batching = [vt.sample_elem() for _ in range(7)]
    sc_claim = orc.evaluate_mle(s_hat_u, orc.to_elems(batching))
    final_fri, final_sum, chals = fv.basefold_verify(orc, dom, vt, log_dim, lir, lbs, arities, nq, root, sc_claim)
    assert vt.done()
    rs_eq = fv.rs_eq_ind_definition(orc, point[7:], batching)
    rs_eq_at = orc.evaluate_mle(rs_eq, orc.to_elems(chals[::-1]))
    return final_sum == fv.fmul(orc, final_fri, rs_eq_at)


PCS_CASES = [
    (4, 1, 1, [1] * 3, 3),
    (6, 1, 2, [2, 2], 3),
    (4, 1, 4, [], 232),
]


@pytest.mark.parametrize("log_dim,lir,lbs,arities,nq", PCS_CASES)
def test_pcs_prove_verifies(orc, log_dim, lir, lbs, arities, nq):
    n = log_dim + lbs
    dom = orc.Domain(dim=log_dim + lir)
    packed = orc.random_elems(61 + n, 1 << n)
    point = orc.random_elems(71 + n, n + 7)
    claim = _b1_evaluation(orc, packed, point)
    cw, tree = _commit(orc, dom, packed, log_dim, lir, lbs)
    root = tree[-1].tobytes()
    t = orc.Transcript()
    t.message(root)
    orc.pcs_prove(dom, packed, point, cw, tree, log_dim, lir, lbs, arities, nq, t, threads=2)
    tape = t.tape()
    assert pcs_verify(orc, dom, tape, root, claim, point, log_dim, lir, lbs, arities, nq)
    assert not pcs_verify(orc, dom, tape, root, claim ^ 1, point, log_dim, lir, lbs, arities, nq)
    # a wrong point also fails (the claim no longer matches s_hat_v)
    bad = point.copy()
    bad[0, 0] ^= 1
    assert not pcs_verify(orc, dom, tape, root, claim, bad, log_dim, lir, lbs, arities, nq)


@pytest.fixture(scope="module")
def visions(orc):
    import json
    from pathlib import Path
    gold = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
    return {m: orc.Vision(m, gold[f"vision{m}"]) for m in (4, 6)}


@pytest.mark.parametrize("leaf_m", [4, 6])
def test_fri_and_basefold_with_vision_merkle(orc, visions, leaf_m):
    """The Vision Merkle scheme (setup_vision4; Vision-6 leaves as in the Merkle
    bench) through the FRI-only flow and BaseFold: the restated verifiers accept
    with the same scheme and reject with SHA-256."""
    scheme = (visions[leaf_m], visions[4])
    vs = fv.VisionScheme(*scheme)
    log_dim, lir, lbs, arities, nq = 6, 1, 2, [2, 2], 3
    n_vars = log_dim + lbs
    dom = orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(900 + leaf_m, 1 << n_vars)
    res = orc.fri_prove(dom, msg, log_dim, lir, lbs, arities, nq, scheme=scheme)
    # replay the FRI-only flow (fri/tests.rs:58-89): root, then per round a
    # challenge and, on commit rounds, that round's root
    vt = fv.VerifierTape(orc, res["tape"])
    roots, chal = [vt.message(32)], []
    commits = fv.commit_rounds(lbs, arities, n_vars + 1)
    for rnd in range(n_vars + 1):
        if rnd:
            chal.append(vt.sample_elem())
        if commits[rnd]:
            roots.append(vt.message(32))
    assert chal == orc.to_ints(res["challenges"]) and roots == res["roots"]
    final = fv.verify_fri_queries(orc, dom, vt, log_dim, lir, lbs, arities, nq, roots[0], roots[1:], chal, vs)
    assert vt.done()
    assert final == orc.evaluate_mle(msg, orc.to_elems(chal[::-1]))
    # BaseFold over a Vision-committed codeword
    point = orc.random_elems(901 + leaf_m, n_vars)
    eq = orc.eq_ind(point)
    claim = orc.evaluate_mle(msg, point)
    cw = orc.encode_batch(dom, msg, log_dim, lir, lbs)
    tree = orc.merkle_build_vision(scheme[0], scheme[1], cw, 1 << lbs)
    root = tree[-1].tobytes()
    t = orc.Transcript()
    t.message(root)
    orc.basefold_prove(dom, msg, eq, claim, cw, tree, log_dim, lir, lbs, arities, nq, t, scheme=scheme)
    tape = t.tape()
    vt = fv.VerifierTape(orc, tape)
    vt.message(32)
    final_fri, final_sum, chals = fv.basefold_verify(orc, dom, vt, log_dim, lir, lbs, arities, nq, root, claim, vs)
    assert vt.done()
    assert final_sum == fv.fmul(orc, final_fri, orc.evaluate_mle(eq, orc.to_elems(chals[::-1])))
    vt = fv.VerifierTape(orc, tape)
    vt.message(32)
    with pytest.raises(AssertionError):
        fv.basefold_verify(orc, dom, vt, log_dim, lir, lbs, arities, nq, root, claim)  # SHA-256 scheme
