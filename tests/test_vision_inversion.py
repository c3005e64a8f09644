"""The facts the device's S-box inversion rests on, checked on the CPU with the
oracle's field arithmetic (oracle/b64_oracle.c gf_mul / gf_inv, the GHASH field of
verifier/field/src/ghash.rs):

* for d = 64, 32, 16, 8 the elements fixed by x -> x^(2^d) form a d-dimensional
  subspace of GF(2^128) (the subfield GF(2^d));
* in the GHASH representation an element of GF(2^64) is determined by its 32-bit
  words 0 and 2, one of GF(2^32) (so also of GF(2^16)) by word 0, and one of
  GF(2^8) by nibbles 0 and 4 of word 0, the bits the device's subfield tables
  read;
* the norm chain n1 = a a^(2^64), n2 = n1 n1^(2^32), n3 = n2 n2^(2^16),
  n4 = n3 n3^(2^8) stays in those subfields and
  a^-1 = a^(2^64) n1^(2^32) n2^(2^16) n3^(2^8) n4^-1, equal to invert_or_zero
  (verifier/math/src/batch_invert.rs) on every nonzero element tried.
"""
import numpy as np
import pytest


def _frob(orc, arr, k):
    for _ in range(k):
        arr = orc.gf_square(arr)
    return arr


def _rank(vs):
    piv, r = {}, 0
    for v in vs:
        while v:
            h = v.bit_length() - 1
            if h in piv:
                v ^= piv[h]
            else:
                piv[h] = v
                r += 1
                break
    return r


def _kernel(cols):
    """Basis of {c : XOR of cols[j] over the bits j of c = 0}."""
    piv, null = {}, []
    for j, col in enumerate(cols):
        v, c = col, 1 << j
        while v:
            h = v.bit_length() - 1
            if h in piv:
                pv, pc = piv[h]
                v ^= pv
                c ^= pc
            else:
                piv[h] = (v, c)
                break
        if v == 0:
            null.append(c)
    return null


@pytest.fixture(scope="module")
def subfields(orc):
    unit = orc.to_elems([1 << j for j in range(128)])
    out = {}
    img = unit
    done = 0
    for d in (8, 16, 32, 64):
        img = _frob(orc, img, d - done)
        done = d
        cols = [a ^ (1 << j) for j, a in enumerate(orc.to_ints(img))]
        out[d] = _kernel(cols)
    return out


def test_subfield_dimensions(subfields):
    assert {d: len(b) for d, b in subfields.items()} == {8: 8, 16: 16, 32: 32, 64: 64}


def test_device_projections_are_one_to_one(subfields):
    w0 = lambda x: x & 0xFFFFFFFF
    w02 = lambda x: (x & 0xFFFFFFFF) | (((x >> 64) & 0xFFFFFFFF) << 32)
    n04 = lambda x: (x & 0xF) | (((x >> 16) & 0xF) << 4)
    assert _rank([w02(b) for b in subfields[64]]) == 64
    assert _rank([w0(b) for b in subfields[32]]) == 32
    assert _rank([w0(b) for b in subfields[16]]) == 16
    assert _rank([n04(b) for b in subfields[8]]) == 8
    # and the next cheaper choices are not: word 0 alone for GF(2^64), one byte for GF(2^8)
    assert _rank([w0(b) for b in subfields[64]]) < 64
    assert all(_rank([(b >> (8 * k)) & 0xFF for b in subfields[8]]) < 8 for k in range(16))


def test_norm_chain_inverts(orc, subfields):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 2 ** 63, size=(512, 2), dtype=np.uint64) ^ (
        rng.integers(0, 2, size=(512, 2), dtype=np.uint64) << np.uint64(63))
    # ones and elements of the subfields themselves
    a[0] = (1, 0)
    for i, d in enumerate((8, 16, 32, 64)):
        v = 0
        for b in subfields[d][1::2]:
            v ^= b
        a[1 + i] = (v & (2 ** 64 - 1), v >> 64)
    a[5] = (0, 1)
    t64 = _frob(orc, a, 64)
    n1 = orc.gf_mul(a, t64)
    t32 = _frob(orc, n1, 32)
    n2 = orc.gf_mul(n1, t32)
    t16 = _frob(orc, n2, 16)
    n3 = orc.gf_mul(n2, t16)
    t8 = _frob(orc, n3, 8)
    n4 = orc.gf_mul(n3, t8)
    # each norm lies in its subfield
    assert np.array_equal(_frob(orc, n1, 64), n1)
    assert np.array_equal(_frob(orc, n2, 32), n2)
    assert np.array_equal(_frob(orc, n3, 16), n3)
    assert np.array_equal(_frob(orc, n4, 8), n4)
    # the GF(2^8) inverse the device reads from its table
    i8 = orc.gf_inv(n4)
    assert np.array_equal(orc.gf_mul(i8, n4), np.tile(np.array([1, 0], dtype=np.uint64), (len(a), 1)))
    inv = orc.gf_mul(orc.gf_mul(orc.gf_mul(orc.gf_mul(t64, t32), t16), t8), i8)
    assert np.array_equal(inv, orc.gf_inv(a))
