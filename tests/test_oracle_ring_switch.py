"""Pins of the oracle's ring-switch folds (prover/prover/src/ring_switch.rs)."""
import numpy as np
import pytest


def xor_reduce(elems: np.ndarray) -> np.ndarray:
    return np.bitwise_xor.reduce(elems.reshape(-1, 2), axis=0) if elems.size else np.zeros(2, np.uint64)


@pytest.mark.parametrize("log_len", [0, 1, 2, 3, 5, 10, 13])
def test_fold_1b_rows_four_russians_equals_definition(orc, log_len):
    """The reference's chunked method of four Russians with the 4x4 nibble
    transpose (ring_switch.rs:216-257) equals the definition bit by bit."""
    mat, vec = orc.random_elems(10 + log_len, 1 << log_len), orc.random_elems(20 + log_len, 1 << log_len)
    assert np.array_equal(orc.rs_fold_1b_rows(mat, vec, 0), orc.rs_fold_1b_rows(mat, vec, 1))


def test_fold_1b_rows_unit_rows(orc):
    """A single row r with vec = e and mat = 2^b gives e at output b only."""
    for b in (0, 1, 63, 64, 127):
        mat = orc.to_elems([1 << b])
        vec = orc.random_elems(b, 1)
        out = orc.rs_fold_1b_rows(mat, vec, 0)
        want = np.zeros((128, 2), dtype=np.uint64)
        want[b] = vec[0]
        assert np.array_equal(out, want)


@pytest.mark.parametrize("n", [1, 5, 1000])
def test_fold_elems_tables_equal_inner_product(orc, n):
    """fold_b128_elems_inplace equals fold_elems_inplace (ring_switch.rs:439-465)."""
    elems, vec = orc.random_elems(30 + n, n), orc.random_elems(40 + n, 128)
    assert np.array_equal(orc.rs_fold_elems(elems, vec, 0), orc.rs_fold_elems(elems, vec, 1))


def test_fold_elems_monomial_basis_is_identity(orc):
    """With vec[b] = x^b the fold is x -> sum_b bit_b(x) x^b = x."""
    elems = orc.random_elems(5, 300)
    basis = orc.to_elems([1 << b for b in range(128)])
    assert np.array_equal(orc.rs_fold_elems(elems, basis, 0), elems)


def test_fold_1b_rows_and_fold_elems_are_one_bilinear_form(orc):
    """sum_b fold_1b_rows(mat, vec)[b] w[b] = sum_r vec[r] fold_elems(mat, w)[r]:
    both are sum_{r,b} vec[r] w[b] bit_b(mat[r])."""
    mat, vec, w = orc.random_elems(1, 1 << 9), orc.random_elems(2, 1 << 9), orc.random_elems(3, 128)
    lhs = xor_reduce(orc.gf_mul(orc.rs_fold_1b_rows(mat, vec), w))
    rhs = xor_reduce(orc.gf_mul(vec, orc.rs_fold_elems(mat, w)))
    assert np.array_equal(lhs, rhs)


def test_fold_tensor_product(orc):
    """ring_switch.rs:384-431: the inner product of a bit matrix with the tensor of
    a full point equals folding with the tensor of the first 7 coordinates and
    evaluating the result at the rest."""
    n = 6
    mat = orc.random_elems(77, 1 << n)
    point = orc.random_elems(78, n + 7)
    full = orc.eq_ind(point)
    bits = np.unpackbits(mat.view(np.uint8), bitorder="little").astype(bool)  # bit i of the bit matrix
    want = xor_reduce(full[bits])
    folded = orc.rs_fold_elems(mat, orc.eq_ind(point[:7]))
    got = orc.evaluate_mle(folded, point[7:])
    assert orc.to_ints(want.reshape(1, 2))[0] == got
