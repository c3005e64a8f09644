"""GPU Vision-4 / Vision-6 hash through the C ABI against the oracle, bit-exact:
the permutation (including the reference's test vectors), sponge digests of
every message length around the rate, and Vision Merkle trees."""
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


@pytest.fixture(scope="module")
def oracles(orc):
    return {m: orc.Vision(m, GOLD[f"vision{m}"]) for m in (4, 6)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("m", [4, 6])
def test_permutation_vectors_and_random(b64, orc, oracles, m):
    say("vision permute", m)
    kats = GOLD[f"vision{m}"]["permutation_kats"]
    states = [int(x, 16) for inp, _ in kats for x in inp]
    rnd = orc.random_elems(600 + m, 300 * m)
    # zeros inside states exercise the zero branch of the batch inversion
    rnd[5] = 0
    rnd[7 * m:8 * m] = 0
    all_states = np.concatenate([orc.to_elems(states), rnd])
    buf = b64.DeviceBuffer.from_host(all_states)
    n = all_states.shape[0] // m
    b64.VisionHasher(m).permute(buf, n)
    got = buf.to_host(n * m)
    assert orc.to_ints(got[:2 * m]) == [int(x, 16) for _, exp in kats for x in exp]
    assert np.array_equal(got, oracles[m].permutation(all_states))


@pytest.mark.parametrize("m", [4, 6])
def test_digest_lengths(b64, orc, oracles, m):
    say("vision digest", m)
    rate = m - 2
    for msg_elems in list(range(0, 2 * rate + 2)) + [16]:
        n = 37
        msgs = orc.random_elems(700 + msg_elems, max(n * msg_elems, 1))[:n * msg_elems]
        d_msgs = b64.DeviceBuffer.from_host(msgs if msgs.size else np.zeros((1, 2), dtype=np.uint64))
        out = b64.VisionHasher(m).digest_many(d_msgs, n, msg_elems)
        got = out.to_bytes(32 * n)
        want = b"".join(oracles[m].digest(msgs[k * msg_elems:(k + 1) * msg_elems].tobytes()) for k in range(n))
        assert got == want, msg_elems


# (4, 17, 1): lane-form digests and widest level, quad-form upper levels;
# (6, 15, 2): quad-form digests at the largest count that takes them
@pytest.mark.parametrize("leaf_m,log_leaves,batch", [(4, 0, 16), (4, 6, 16), (6, 6, 16), (4, 5, 2), (6, 9, 16),
                                                     (4, 17, 1), (6, 15, 2)])
def test_merkle_vision(b64, orc, oracles, leaf_m, log_leaves, batch):
    say("vision merkle", leaf_m, log_leaves, batch)
    n = batch << log_leaves
    elems = orc.random_elems(800 + leaf_m + log_leaves, n)
    root, tree = b64.merkle_commit_vision(b64.DeviceBuffer.from_host(elems), n, batch, leaf_m)
    want = orc.merkle_build_vision(oracles[leaf_m], oracles[4], elems, batch, threads=8)
    n_nodes = 2 * (n // batch) - 1
    got = np.frombuffer(tree.to_bytes(32 * n_nodes), dtype=np.uint8).reshape(n_nodes, 32)
    assert np.array_equal(got, want)
    assert root == want[-1].tobytes()


def test_merkle_vision_errors(b64):
    buf = b64.DeviceBuffer(16 * 48)
    with pytest.raises(b64.B64Error) as e:
        b64.merkle_commit_vision(buf, 48, 16, 4)  # 3 leaves
    assert e.value.code == 5
    with pytest.raises(b64.B64Error) as e:
        b64.merkle_commit_vision(buf, 48, 5, 4)
    assert e.value.code == 4
    with pytest.raises(b64.B64Error) as e:
        b64.merkle_commit_vision(buf, 32, 16, 5)
    assert e.value.code == 1


@pytest.mark.parametrize("leaf_m,log_dim,lir,lbs,arities,nq", [
    (4, 6, 1, 2, [2, 2], 3), (6, 8, 1, 4, [4, 1], 232), (4, 10, 1, 4, [4, 4, 1], 232)])
def test_fri_proof_tape_with_vision_merkle(b64, orc, oracles, leaf_m, log_dim, lir, lbs, arities, nq):
    """FRI-only flow with the Vision Merkle scheme: roots, challenges and the proof
    tape equal the oracle's."""
    say("fri vision", leaf_m, log_dim, lbs, arities)
    dom, ref = b64.DomainContext.standard(log_dim + lir), orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(950 + log_dim, 1 << (log_dim + lbs))
    params = b64.FRIParams.new(log_dim, lir, lbs, arities, nq)
    tape, roots, chal = b64.fri_prove(params, dom, msg, merkle_hash=leaf_m)
    want = orc.fri_prove(ref, msg, log_dim, lir, lbs, arities, nq, threads=8, scheme=(oracles[leaf_m], oracles[4]))
    assert roots == want["roots"]
    assert orc.to_ints(want["challenges"]) == chal
    assert tape == want["tape"]


def test_pcs_with_vision_merkle(b64, orc, oracles):
    """The evaluation proof over a Vision-4-committed codeword (setup_vision4)
    with the Verifier::setup shape at 2^12 elements equals the oracle's."""
    n = 12
    params, dom = b64.setup_commit_params(n, 1, 96)
    packed = orc.random_elems(1212, 1 << n)
    point = orc.random_elems(1213, n + 7)
    pcs = b64.OneBitPCSProver(dom, params, merkle_hash=b64.MERKLE_VISION4)
    co = pcs.commit(b64.DeviceBuffer.from_host(packed))
    ref = orc.Domain(dom.subspace(params.rs_log_len()))
    cw = orc.encode_batch(ref, packed, params.log_dim, 1, params.log_batch_size)
    tree = orc.merkle_build_vision(oracles[4], oracles[4], cw, 1 << params.log_batch_size, threads=8)
    assert co.commitment == tree[-1].tobytes()
    t = b64.ProverTranscript()
    t.message(co.commitment)
    pcs.prove(co.codeword, co.tree, b64.DeviceBuffer.from_host(packed), point, t)
    ot = orc.Transcript()
    ot.message(co.commitment)
    orc.pcs_prove(ref, packed, point, cw, tree, params.log_dim, 1, params.log_batch_size, params.fold_arities,
                  params.n_test_queries, ot, threads=8, scheme=(oracles[4], oracles[4]))
    assert t.finalize() == ot.tape()
