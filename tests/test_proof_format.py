"""Proof wire format through the C ABI (host code, no GPU): the reference's
fixture verifier/core/test_data/proof_v1.bin (copied to tests/golden/) and the
cases of its Proof tests (constraint_system.rs:1601-1760), plus truncation,
UTF-8 and buffer-size errors."""
from pathlib import Path

import numpy as np
import pytest

import binius64_amd as b64
from binius64_amd import B64Error, Proof

FIXTURE = (Path(__file__).parent / "golden" / "proof_v1.bin").read_bytes()
FIXTURE_DATA = bytes([0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF, 0xFE, 0xDC, 0xBA, 0x98, 0x76, 0x54,
                      0x32, 0x10, 0xDE, 0xAD, 0xBE, 0xEF, 0xCA, 0xFE, 0xBA, 0xBE])
NOT_ENOUGH_BYTES, INVALID_CONSTRUCTION, FROM_UTF8, BUFFER_TOO_SMALL = 14, 15, 16, 10


def code_of(fn):
    with pytest.raises(B64Error) as e:
        fn()
    return e.value.code


def test_reference_fixture_deserializes():
    p = Proof.deserialize(FIXTURE)
    assert p.len() == 24
    assert p.challenger_type() == "HasherChallenger<Sha256>" == b64.STD_CHALLENGER_NAME
    assert p.as_slice() == FIXTURE_DATA
    assert FIXTURE[:4] == (1).to_bytes(4, "little")


def test_serialize_reproduces_fixture():
    assert Proof.owned(FIXTURE_DATA, b64.STD_CHALLENGER_NAME).serialize() == FIXTURE


@pytest.mark.parametrize("data,challenger", [
    (bytes([0x12, 0x34, 0x56, 0x78, 0x9A, 0xBC, 0xDE, 0xF0]), "HasherChallenger<Sha256>"),
    (bytes([0xAA, 0xBB, 0xCC, 0xDD]), "TestChallenger"),
    (b"", "EmptyProof"),
    (b"\x42", "HasherChallenger<Blake2b>"),
    (b"\x42", "CustomChallenger"),
    (b"\x42", ""),
    (b"\x00\x01", "Chällenger ✓ \U0001F600"),
])
def test_round_trip(data, challenger):
    p = Proof.owned(data, challenger)
    q = Proof.deserialize(p.serialize())
    assert q == p and q.as_slice() == data and q.challenger_type() == challenger
    assert q.is_empty() == (len(data) == 0)
    assert q.into_owned() == (data, challenger) and q.to_owned() == q


def test_large_transcript():
    data = np.random.default_rng(12345).integers(0, 256, 10000, dtype=np.uint8).tobytes()
    q = Proof.deserialize(Proof.owned(data, "LargeProofChallenger").serialize())
    assert q.len() == 10000 and q.as_slice() == data


def test_version_mismatch():
    buf = (999).to_bytes(4, "little") + Proof.owned(b"\xaa", "TestChallenger").serialize()[4:]
    assert code_of(lambda: Proof.deserialize(buf)) == INVALID_CONSTRUCTION


def test_truncated_buffers():
    for n in range(len(FIXTURE)):
        assert code_of(lambda: Proof.deserialize(FIXTURE[:n])) == NOT_ENOUGH_BYTES, n


def test_trailing_bytes_ignored():
    assert Proof.deserialize(FIXTURE + b"\x00\x11").as_slice() == FIXTURE_DATA


def test_challenger_utf8_matches_strict_decoder():
    """The challenger type must be UTF-8 as String::from_utf8 requires; Python's
    strict decoder applies the same rules (no overlongs, surrogates or code
    points above U+10FFFF)."""
    rng = np.random.default_rng(7)
    cases = [b"\xff", b"\xc0\xaf", b"\xed\xa0\x80", b"\xf4\x90\x80\x80", b"\xe0\x80\x80", b"\xc2",
             "߿￿\U0010ffff".encode()]
    cases += [rng.integers(0, 256, int(rng.integers(1, 7)), dtype=np.uint8).tobytes() for _ in range(3000)]
    for ct in cases:
        buf = (1).to_bytes(4, "little") + len(ct).to_bytes(4, "little") + ct + (0).to_bytes(4, "little")
        try:
            ct.decode("utf-8")
            ok = True
        except UnicodeDecodeError:
            ok = False
        if ok:
            assert Proof.deserialize(buf).challenger_type() == ct.decode("utf-8"), ct
        else:
            assert code_of(lambda: Proof.deserialize(buf)) == FROM_UTF8, ct


def test_serialize_buffer_sizes():
    import ctypes as C
    L = b64.lib()
    data = np.frombuffer(FIXTURE_DATA, dtype=np.uint8).copy()
    n = C.c_uint64()
    p8 = data.ctypes.data_as(C.POINTER(C.c_uint8))
    assert L.b64_proof_serialize(p8, 24, b"HasherChallenger<Sha256>", None, 0, C.byref(n)) == BUFFER_TOO_SMALL
    assert n.value == len(FIXTURE)
    out = np.zeros(len(FIXTURE), dtype=np.uint8)
    po = out.ctypes.data_as(C.POINTER(C.c_uint8))
    assert L.b64_proof_serialize(p8, 24, b"HasherChallenger<Sha256>", po, len(FIXTURE) - 1, C.byref(n)) \
        == BUFFER_TOO_SMALL
    assert L.b64_proof_serialize(p8, 24, b"HasherChallenger<Sha256>", po, len(FIXTURE), C.byref(n)) == 0
    assert out.tobytes() == FIXTURE
