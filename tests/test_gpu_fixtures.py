"""GPU path against the committed oracle fixtures (tests/golden/oracle_fixtures.json):
the same inputs (oracle SplitMix64 seeds) through the C ABI give the stored
outputs byte for byte."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FX = json.loads((Path(__file__).parent / "golden" / "oracle_fixtures.json").read_text())


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


def sha(arr) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def test_mul_fixture(b64, orc):
    f = FX["mul"]
    got = b64.gf_mul_device(orc.random_elems(f["seed_a"], f["n"]), orc.random_elems(f["seed_b"], f["n"]))
    assert [f"{v:032x}" for v in orc.to_ints(got)] == f["ab"]


@pytest.mark.parametrize("i", range(len(FX["ntt"])))
def test_ntt_fixture(b64, orc, i):
    f = FX["ntt"][i]
    log_d = f["log_d"]
    if f["domain"] == "standard":
        dom = b64.DomainContext.standard(log_d)
    else:
        dom = b64.DomainContext(b64.DomainContext.standard(11).subspace(log_d))
    buf = b64.DeviceBuffer.from_host(orc.random_elems(f["seed"], 1 << log_d))
    b64.AdditiveNTT(dom).forward_transform(buf, log_d, f["skip_early"], f["skip_late"])
    assert sha(buf.to_host(1 << log_d)) == f["sha256"]


@pytest.mark.parametrize("i", range(len(FX["encode"])))
def test_encode_fixture(b64, orc, i):
    f = FX["encode"][i]
    log_dim, lir, lbs = f["log_dim"], f["lir"], f["lbs"]
    dom = b64.DomainContext.standard(log_dim + lir)
    msg = b64.DeviceBuffer.from_host(orc.random_elems(f["seed"], 1 << (log_dim + lbs)))
    out = b64.ReedSolomonCode(dom, log_dim, lir).encode_batch(msg, lbs)
    assert sha(out.to_host(1 << (log_dim + lir + lbs))) == f["sha256"]


def test_hash_fixtures(b64, orc):
    f = FX["sha256_leaf"]
    root, _ = b64.BinaryMerkleTreeProver.commit(b64.DeviceBuffer.from_host(orc.random_elems(f["seed"], 16)), 16, 16)
    assert root.hex() == f["digest"]
    f = FX["compress"]
    root, tree = b64.BinaryMerkleTreeProver.commit(b64.DeviceBuffer.from_host(orc.random_elems(f["seed"], 32)), 32, 16)
    nodes = tree.to_bytes(96)
    assert (nodes[:32].hex(), nodes[32:64].hex(), root.hex()) == (f["left"], f["right"], f["node"])
    f = FX["merkle_2p10_batch16"]
    root, _ = b64.BinaryMerkleTreeProver.commit(
        b64.DeviceBuffer.from_host(orc.random_elems(f["seed"], 1 << 10)), 1 << 10, 16)
    assert root.hex() == f["root"]


@pytest.mark.parametrize("i", range(len(FX["fri_transcript"])))
def test_fri_transcript_fixture(b64, orc, i):
    f = FX["fri_transcript"][i]
    n = f["n"]
    params, dom = b64.setup_commit_params(n, 1, 96)
    assert (params.log_dim, params.log_batch_size, params.fold_arities, params.n_test_queries) == \
        (f["log_dim"], f["log_batch_size"], f["fold_arities"], f["n_test_queries"])
    tape, roots, _ = b64.fri_prove(params, dom, orc.random_elems(f["seed"], 1 << n))
    assert [r.hex() for r in roots] == f["roots"]
    assert len(tape) == f["tape_len"] and hashlib.sha256(tape).hexdigest() == f["tape_sha256"]
