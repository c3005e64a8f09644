"""GPU sumcheck, BaseFold and one-bit PCS opening through the C ABI against the
oracle: eq tensors, round polynomials, evaluations and whole proof tapes are
byte-identical."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [0, 1, 3, 8, 9, 10, 11, 12, 13, 16, 17, 20, 21, 24])
def test_eq_ind_partial_eval(b64, orc, n):
    say("eq", n)
    point = orc.random_elems(300 + n, n) if n else np.zeros((0, 2), dtype=np.uint64)
    got = b64.eq_ind_partial_eval(point).to_host(1 << n)
    assert np.array_equal(got, orc.eq_ind(point))


@pytest.mark.parametrize("n_vars", [0, 1, 2, 5, 8, 12, 16, 19])
def test_sumcheck_matches_oracle(b64, orc, n_vars):
    say("sumcheck", n_vars)
    a = orc.random_elems(400 + n_vars, 1 << n_vars)
    b = orc.random_elems(500 + n_vars, 1 << n_vars)
    # the claimed sum over the hypercube: XOR of the products a[x] b[x]
    claim = orc.to_ints(np.bitwise_xor.reduce(orc.gf_mul(a, b), axis=0, keepdims=True))[0]
    t = b64.ProverTranscript()
    prover = b64.BivariateProductSumcheckProver(b64.DeviceBuffer.from_host(a), b64.DeviceBuffer.from_host(b),
                                                n_vars, claim)
    out = b64.prove_single(prover, t)
    ot = orc.Transcript()
    want = orc.sumcheck_prove(a, b, claim, ot, threads=8)
    assert out["challenges"] == orc.to_ints(want["challenges"])
    assert out["multilinear_evals"] == want["evals"]
    assert t.finalize() == ot.tape()


def test_sumcheck_call_order_errors(b64, orc):
    a, b = orc.random_elems(1, 8), orc.random_elems(2, 8)
    p = b64.BivariateProductSumcheckProver(b64.DeviceBuffer.from_host(a), b64.DeviceBuffer.from_host(b), 3, 0)
    with pytest.raises(b64.B64Error) as e:
        p.fold(5)
    assert e.value.code == 13  # ExpectedExecute
    p.execute()
    with pytest.raises(b64.B64Error) as e:
        p.execute()
    assert e.value.code == 12  # ExpectedFold
    with pytest.raises(b64.B64Error) as e:
        p.finish()
    assert e.value.code == 12
    p.fold(5)
    with pytest.raises(b64.B64Error) as e:
        p.finish()
    assert e.value.code == 13


def _commit_both(b64, orc, log_dim, lir, lbs, arities, nq, msg):
    dom, ref = b64.DomainContext.standard(log_dim + lir), orc.Domain(dim=log_dim + lir)
    params = b64.FRIParams.new(log_dim, lir, lbs, arities, nq)
    co = b64.commit_interleaved(params, dom, b64.DeviceBuffer.from_host(msg))
    cw = orc.encode_batch(ref, msg, log_dim, lir, lbs)
    tree = orc.merkle_build(cw, 1 << lbs)
    assert co.commitment == tree[-1].tobytes()
    return dom, ref, params, co, cw, tree


BASEFOLD_CASES = [
    (7, 1, 1, [1] * 6, 3),
    (6, 2, 2, [3, 2, 1], 4),
    (8, 1, 4, [1], 232),
    (12, 1, 4, [4, 1], 232),
    (16, 1, 4, [4, 4, 1], 232),
]


@pytest.mark.parametrize("log_dim,lir,lbs,arities,nq", BASEFOLD_CASES)
def test_basefold_tape_matches_oracle(b64, orc, log_dim, lir, lbs, arities, nq):
    say("basefold", log_dim, lir, lbs, arities)
    n_vars = log_dim + lbs
    msg = orc.random_elems(31 + n_vars, 1 << n_vars)
    point = orc.random_elems(41 + n_vars, n_vars)
    eq = orc.eq_ind(point)
    claim = orc.evaluate_mle(msg, point)
    dom, ref, params, co, cw, tree = _commit_both(b64, orc, log_dim, lir, lbs, arities, nq, msg)
    t = b64.ProverTranscript()
    t.message(co.commitment)
    b64.basefold_prove(params, dom, b64.DeviceBuffer.from_host(msg), b64.DeviceBuffer.from_host(eq), claim,
                       co.codeword, co.tree, t)
    ot = orc.Transcript()
    ot.message(co.commitment)
    orc.basefold_prove(ref, msg, eq, claim, cw, tree, log_dim, lir, lbs, arities, nq, ot, threads=8)
    assert t.finalize() == ot.tape()


PCS_CASES = [
    (4, 1, 1, [1] * 3, 3),
    (6, 1, 2, [2, 2], 3),
    (4, 1, 4, [], 232),
    (10, 1, 4, [4, 1], 232),
]


@pytest.mark.parametrize("log_dim,lir,lbs,arities,nq", PCS_CASES)
def test_pcs_tape_matches_oracle(b64, orc, log_dim, lir, lbs, arities, nq):
    say("pcs", log_dim, lir, lbs, arities)
    n = log_dim + lbs
    packed = orc.random_elems(61 + n, 1 << n)
    point = orc.random_elems(71 + n, n + 7)
    dom, ref, params, co, cw, tree = _commit_both(b64, orc, log_dim, lir, lbs, arities, nq, packed)
    t = b64.ProverTranscript()
    t.message(co.commitment)
    b64.OneBitPCSProver(dom, params).prove(co.codeword, co.tree, b64.DeviceBuffer.from_host(packed), point, t)
    ot = orc.Transcript()
    ot.message(co.commitment)
    orc.pcs_prove(ref, packed, point, cw, tree, log_dim, lir, lbs, arities, nq, ot, threads=8)
    assert t.finalize() == ot.tape()


def test_pcs_config2_setup_chain(b64, orc):
    """The Verifier::setup chain at 2^20 B128 elements (rs (16, 1), batch 2^4,
    arities [4, 4, 1], 232 queries): the full evaluation proof is byte-identical."""
    n = 20
    params, dom = b64.setup_commit_params(n, 1, 96)
    packed = orc.random_elems(2020, 1 << n)
    point = orc.random_elems(2021, n + 7)
    co = b64.commit_interleaved(params, dom, b64.DeviceBuffer.from_host(packed))
    ref = orc.Domain(dom.subspace(params.rs_log_len()))
    cw = orc.encode_batch(ref, packed, params.log_dim, 1, params.log_batch_size, threads=8)
    tree = orc.merkle_build(cw, 1 << params.log_batch_size, threads=8)
    assert co.commitment == tree[-1].tobytes()
    t = b64.ProverTranscript()
    t.message(co.commitment)
    b64.OneBitPCSProver(dom, params).prove(co.codeword, co.tree, b64.DeviceBuffer.from_host(packed), point, t)
    ot = orc.Transcript()
    ot.message(co.commitment)
    orc.pcs_prove(ref, packed, point, cw, tree, params.log_dim, 1, params.log_batch_size, params.fold_arities, 232,
                  ot, threads=8)
    tape = t.finalize()
    assert tape == ot.tape()
    # the proof as the reference's prover CLI writes it (Proof::owned(transcript, StdChallenger::NAME))
    wire = b64.Proof.owned(tape, b64.STD_CHALLENGER_NAME).serialize()
    assert wire[:4] == (1).to_bytes(4, "little") and len(wire) == 12 + 24 + len(tape)
    assert b64.Proof.deserialize(wire).as_slice() == ot.tape()
