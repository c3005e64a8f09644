"""Pins the oracle's GF(2^128) arithmetic against reference-held vectors.

* test_ghash_mul KATs (verifier/field/src/ghash.rs:798-865)
* the AES -> GHASH embedding table (ghash.rs:520-783) together with the
  homomorphism the reference proptests (ghash.rs:928-935): phi(a)*phi(b) =
  phi(a x b) in Rijndael's GF(2^8) (aes_field.rs:29, x^8+x^4+x^3+x+1).  This is
  65,536 independent product checks against a textbook field.
* mul_x / mul_inv_x consistency (ghash.rs:873-925)
* MULTIPLICATIVE_GENERATOR has order 2^128 - 1 (ghash.rs:502, :868-870)
* three independent restatements (PCLMUL, BearSSL portable, bit-serial) agree.
"""
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())

# Prime factorisation of 2^128 - 1 (Fermat numbers F0..F6 = 3*5*17*257*65537*(641*6700417)*
# (274177*67280421310721)); a wrong multiplier fails the order test.
FACTORS_2_128_M1 = [3, 5, 17, 257, 641, 65537, 274177, 6700417, 67280421310721]


def aes_mul(a: int, b: int) -> int:
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        if a & 0x100:
            a ^= 0x11B
        b >>= 1
    return r


def test_factorisation_is_complete():
    prod = 1
    for p in FACTORS_2_128_M1:
        prod *= p
    assert prod == (1 << 128) - 1


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_ghash_mul_kats(orc, variant):
    a = orc.to_elems(int(x[0], 16) for x in GOLD["ghash_mul_kats"])
    b = orc.to_elems(int(x[1], 16) for x in GOLD["ghash_mul_kats"])
    want = [int(x[2], 16) for x in GOLD["ghash_mul_kats"]]
    assert orc.to_ints(orc.gf_mul(a, b, variant)) == want


@pytest.mark.parametrize("variant", [0, 1])
def test_aes_embedding_homomorphism(orc, variant):
    phi = [int(x, 16) for x in GOLD["aes_to_ghash"]]
    A, B = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    A, B = A.ravel(), B.ravel()
    phi_arr = orc.to_elems(phi)
    prod = orc.gf_mul(phi_arr[A], phi_arr[B], variant)
    aes = np.array([[aes_mul(a, b) for b in range(256)] for a in range(256)]).ravel()
    want = phi_arr[aes]
    assert np.array_equal(prod, want)


def test_aes_embedding_additive(orc):
    phi = [int(x, 16) for x in GOLD["aes_to_ghash"]]
    for a in range(256):
        for b in (1, 2, 4, 8, 16, 32, 64, 128):
            assert phi[a ^ b] == phi[a] ^ phi[b]


def test_mul_x_matches_mul_by_two(orc):
    for key in ("mul_x_cases", "mul_inv_x_cases"):
        vals = orc.to_elems(int(x, 16) for x in GOLD[key])
        two = orc.to_elems([2] * vals.shape[0])
        assert np.array_equal(orc.gf_mul_x(vals), orc.gf_mul(vals, two))
        inv2 = orc.gf_inv(orc.to_elems([2]))
        assert np.array_equal(orc.gf_mul_inv_x(vals), orc.gf_mul(vals, np.repeat(inv2, vals.shape[0], 0)))


def test_multiplicative_generator_order(orc):
    g = int(GOLD["multiplicative_generator"], 16)
    order = (1 << 128) - 1
    assert orc.gf_pow(g, order) == 1
    for p in FACTORS_2_128_M1:
        assert orc.gf_pow(g, order // p) != 1, p


def test_restatements_agree_random(orc):
    a = orc.random_elems(11, 4000)
    b = orc.random_elems(12, 4000)
    p0 = orc.gf_mul(a, b, 0)
    assert np.array_equal(p0, orc.gf_mul(a, b, 1))
    assert np.array_equal(p0[:500], orc.gf_mul(a[:500], b[:500], 2))


def test_field_axioms(orc):
    a = orc.random_elems(21, 512)
    b = orc.random_elems(22, 512)
    c = orc.random_elems(23, 512)
    # distributivity and commutativity
    assert np.array_equal(orc.gf_mul(a, b ^ c), orc.gf_mul(a, b) ^ orc.gf_mul(a, c))
    assert np.array_equal(orc.gf_mul(a, b), orc.gf_mul(b, a))
    # associativity
    assert np.array_equal(orc.gf_mul(orc.gf_mul(a, b), c), orc.gf_mul(a, orc.gf_mul(b, c)))
    # inverses
    one = orc.to_elems([1] * 512)
    assert np.array_equal(orc.gf_mul(a, orc.gf_inv(a)), one)
    assert orc.to_ints(orc.gf_inv(orc.to_elems([0]))) == [0]
    assert np.array_equal(orc.gf_square(a), orc.gf_mul(a, a))
