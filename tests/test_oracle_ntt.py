"""Pins the oracle's additive NTT, domain and RS encoder.

* NTT output = evaluation of the novel-basis polynomial on the subspace, brute
  force (verifier/math/src/ntt/tests_evaluation.rs:187-295), for the standard
  basis and the reference's non-standard basis [5, 7, 22, 95].  The check builds
  the subspace polynomials W_i from their roots, independently of the twiddle
  tables.
* on-the-fly twiddles (subset sums of the normalized evals, domain_context.rs:91-94)
  equal the pre-expanded tables (:114-132, :149-151).
* the multithreaded NeighborsLast restatement equals NeighborsLastReference for
  every skip window and share count the reference tests use
  (ntt/tests_reference.rs:24-112), and forward o inverse = id (:229-258).
* encode_batch = NTT of the zero-padded bit-reversed message
  (reed_solomon.rs:229-292), for the reference's (log_dim, lir, lbs) triples.
"""
import numpy as np
import pytest


def span(orc, basis_ints):
    pts = [0]
    for b in basis_ints:
        pts = pts + [p ^ b for p in pts]
    return pts


def eval_subspace_poly(orc, roots, xs):
    """prod_{a in roots} (x + a) for every x in xs (vectorised over xs)."""
    acc = orc.to_elems([1] * len(xs))
    X = orc.to_elems(xs)
    for a in roots:
        acc = orc.gf_mul(acc, X ^ orc.to_elems([a])[0])
    return acc


def novel_basis_evals(orc, basis_ints):
    """M[k, j] = X_j(s_k) with X_j = prod_{i in bits(j)} W_i / W_i(beta_i)."""
    d = len(basis_ints)
    pts = span(orc, basis_ints)  # s_k, bit i of k <-> beta_i
    what = []
    for i in range(d):
        roots = span(orc, basis_ints[:i])
        w_pts = eval_subspace_poly(orc, roots, pts)
        w_beta = eval_subspace_poly(orc, roots, [basis_ints[i]])
        inv = orc.gf_inv(w_beta)
        what.append(orc.gf_mul(w_pts, np.repeat(inv, len(pts), 0)))
    n = 1 << d
    M = np.zeros((n, n, 2), dtype=np.uint64)
    for j in range(n):
        col = orc.to_elems([1] * n)
        for i in range(d):
            if (j >> i) & 1:
                col = orc.gf_mul(col, what[i])
        M[:, j] = col
    return M


@pytest.mark.parametrize("basis", [None, [5, 7, 22, 95]])
def test_ntt_equals_novel_basis_evaluation(orc, basis):
    if basis is None:
        basis = [1 << i for i in range(6)]
    d = len(basis)
    dom = orc.Domain(orc.to_elems(basis))
    coeffs = orc.random_elems(100 + d, 1 << d)
    out = orc.ntt_forward(dom, coeffs, 0, 0)
    M = novel_basis_evals(orc, basis)
    want = np.zeros_like(coeffs)
    for j in range(1 << d):
        want ^= orc.gf_mul(M[:, j], np.repeat(coeffs[j:j + 1], 1 << d, 0))
    assert np.array_equal(out, want)


def test_standard_domain_matches_explicit_basis(orc):
    a = orc.Domain(dim=9)
    b = orc.Domain(orc.to_elems([1 << i for i in range(9)]))
    assert np.array_equal(a.evals(), b.evals())


@pytest.mark.parametrize("basis", [None, [5, 7, 22, 95, 0x1234, 0xFFFF01]])
def test_on_the_fly_twiddles_equal_pre_expanded(orc, basis):
    dom = orc.Domain(orc.to_elems(basis)) if basis else orc.Domain(dim=8)
    ev = dom.evals()
    d = dom.dim
    for layer in range(d):
        row = orc.to_ints(ev[d - 1 - layer])
        for block in range(1 << layer):
            want = 0
            for b in range(layer):
                if (block >> b) & 1:
                    want ^= row[1 + b]
            assert dom.twiddle(layer, block) == want


def test_subspace_accessor(orc):
    dom = orc.Domain(dim=7)
    ev = dom.evals()
    for i in range(1, 8):
        assert np.array_equal(dom.subspace(i), ev[7 - i, :i])
    assert orc.to_ints(dom.subspace(7)) == [1 << i for i in range(7)]


@pytest.mark.parametrize("log_d", [1, 8, 12])
def test_multithread_equals_reference(orc, log_d):
    dom = orc.Domain(dim=log_d)
    data = orc.random_elems(7 + log_d, 1 << log_d)
    h = log_d // 2
    windows = {(0, 0), (1, 0), (0, 1), (h, 0), (0, log_d - h), (1, 1)}
    for se, sl in sorted(windows):
        if se + sl > log_d:
            continue
        ref = orc.ntt_forward(dom, data, se, sl)
        for shares in (0, 1, 2, 3, 1000):
            got = orc.ntt_forward(dom, data, se, sl, threads=4, log_num_shares=shares)
            assert np.array_equal(got, ref), (se, sl, shares)


@pytest.mark.parametrize("log_d", [1, 8])
def test_forward_inverse_roundtrip(orc, log_d):
    dom = orc.Domain(dim=log_d + 2)
    data = orc.random_elems(31, 1 << log_d)
    for se in range(0, log_d + 1):
        for sl in range(0, log_d + 1 - se):
            f = orc.ntt_forward(dom, data, se, sl)
            assert np.array_equal(orc.ntt_inverse(dom, f, se, sl), data)


def test_ntt_precondition(orc):
    dom = orc.Domain(dim=4)
    with pytest.raises(ValueError):
        orc.ntt_forward(dom, orc.random_elems(1, 1 << 6), 0, 0)  # log_d - skip_late > domain
    with pytest.raises(ValueError):
        orc.ntt_forward(dom, orc.random_elems(1, 1 << 3), 2, 2)  # double skip


def reverse_bits(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("log_dim,lir,lbs", [(4, 2, 0), (6, 2, 1), (8, 3, 2), (1, 2, 0), (10, 1, 4)])
def test_encode_batch_closed_form(orc, log_dim, lir, lbs):
    dom = orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(5, 1 << (log_dim + lbs))
    got = orc.encode_batch(dom, msg, log_dim, lir, lbs)
    m = log_dim + lbs
    ref = np.zeros(((1 << (log_dim + lir + lbs)), 2), dtype=np.uint64)
    for i in range(1 << m):
        ref[reverse_bits(i, m)] = msg[i]
    ref = orc.ntt_forward(dom, ref, 0, lbs)
    assert np.array_equal(got, ref)
    assert np.array_equal(orc.encode_batch(dom, msg, log_dim, lir, lbs, threads=4), ref)


def test_rs_code_is_evaluation_of_low_degree_poly(orc):
    """Rate-1/4 codeword positions are evaluations: re-encoding the NTT inverse
    of the codeword reproduces it and the high coefficients vanish."""
    log_dim, lir = 5, 2
    dom = orc.Domain(dim=log_dim + lir)
    msg = orc.random_elems(8, 1 << log_dim)
    cw = orc.encode_batch(dom, msg, log_dim, lir, 0)
    coeffs = orc.ntt_inverse(dom, cw, 0, 0)
    n = 1 << log_dim
    assert not coeffs[n:].any()
    assert np.array_equal(coeffs[:n], msg[[reverse_bits(i, log_dim) for i in range(n)]])
