/*
 * CPU oracle for the Binius64 B128 RS-encode + Merkle-commit + FRI-fold path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry and bench.py's CPU-baseline leg may load it.  The product
 * (binius64_amd / libb64hip.so) never links, calls or falls back to it.
 *
 * Every function restates the algorithm of the reference file:line it cites
 * (paths relative to the reference repository root).  Field elements cross the
 * ABI as pairs of little-endian u64 words (lo, hi), i.e. the u128 LE byte order
 * of verifier/utils/src/serialization.rs:66-72.
 */
#ifndef B64_ORACLE_H
#define B64_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bo_domain bo_domain;

/* ---- field ------------------------------------------------------------- */
/* variant: 0 = PCLMUL (arch/shared/ghash.rs:19-45), 1 = portable BearSSL
 * (arch/portable/packed_ghash_128.rs:33-97), 2 = bit-serial textbook. */
void bo_gf_mul_many(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n, int variant);
void bo_gf_square_many(const uint64_t *a, uint64_t *out, size_t n);
void bo_gf_inv_many(const uint64_t *a, uint64_t *out, size_t n);
void bo_gf_mul_x_many(const uint64_t *a, uint64_t *out, size_t n);
void bo_gf_mul_inv_x_many(const uint64_t *a, uint64_t *out, size_t n);
void bo_gf_pow(const uint64_t a[2], const uint64_t e[2], uint64_t out[2]);

/* ---- domain / twiddles (verifier/math/src/ntt/domain_context.rs) -------- */
bo_domain *bo_domain_new(const uint64_t *basis, int dim);
bo_domain *bo_domain_new_standard(int dim);
void bo_domain_free(bo_domain *d);
int bo_domain_dim(const bo_domain *d);
/* evals[i][k] (k < dim - i) written at out[2*(i*dim + k)] */
void bo_domain_evals(const bo_domain *d, uint64_t *out);
void bo_domain_twiddle(const bo_domain *d, int layer, uint64_t block, uint64_t out[2]);
/* basis of subspace(i): i elements */
void bo_domain_subspace(const bo_domain *d, int i, uint64_t *out);

/* ---- NTT ----------------------------------------------------------------- */
int bo_ntt_forward_ref(const bo_domain *d, uint64_t *data, int log_d, int skip_early, int skip_late);
int bo_ntt_inverse_ref(const bo_domain *d, uint64_t *data, int log_d, int skip_early, int skip_late);
int bo_ntt_forward_mt(const bo_domain *d, uint64_t *data, int log_d, int skip_early, int skip_late,
                      int log_num_shares, int nthreads);

/* ---- RS encode (verifier/math/src/reed_solomon.rs:128-201) -------------- */
int bo_encode_batch(const bo_domain *d, const uint64_t *msg, int log_dim, int log_inv_rate,
                    int log_batch, uint64_t *out, int nthreads);

/* ---- hashing ------------------------------------------------------------- */
void bo_sha256(const uint8_t *data, size_t len, uint8_t out[32]);
void bo_sha256_compress_block(const uint32_t iv[8], const uint8_t block[64], uint32_t state_out[8]);
void bo_binius_compress(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);
/* tree_out holds (2*n_leaves - 1) digests; returns 0 or an error code */
int bo_merkle_build(const uint64_t *elems, size_t n_elems, size_t batch, uint8_t *tree_out, int nthreads);
int bo_merkle_branch(const uint8_t *tree, int log_len, size_t index, int layer_depth, uint8_t *out);

/* ---- FRI ----------------------------------------------------------------- */
size_t bo_n_test_queries(int security_bits, int log_inv_rate);
int bo_estimate_optimal_arity(int log_block_length, int digest_size, int field_size);
/* Verifier::setup -> FRIParams::choose_with_constant_fold_arity.  Writes
 * [log_batch, log_dim, rs_log_len, n_queries, n_arities, arities...] into out
 * (cap >= 8 + 64). */
int bo_fri_params_setup(int log_msg_len, int log_inv_rate, int security_bits, int *out);
void bo_eq_ind(const uint64_t *point, int n, uint64_t *out);
int bo_fold_interleaved(const uint64_t *cw, int log_len, int log_batch, const uint64_t *chal,
                        uint64_t *out, int nthreads);
int bo_fold_codeword(const bo_domain *d, const uint64_t *cw, int log_len, const uint64_t *chal,
                     int n_chal, uint64_t *out, int nthreads);
/* fold_chunk (verifier fri/fold.rs:61-99): one coset of 2^n_chal values of a
 * codeword of length 2^log_len folded to a single value. */
void bo_fold_chunk(const bo_domain *d, int log_len, uint64_t chunk_index, const uint64_t *vals,
                   const uint64_t *chal, int n_chal, uint64_t out[2]);
void bo_evaluate_mle(const uint64_t *evals, int log_n, const uint64_t *point, uint64_t out[2]);
void bo_pack_witness(const uint64_t *words, size_t n_elems, uint64_t *out);

/* ---- ring switch (prover/prover/src/ring_switch.rs) ------------------------ */
/* fold_1b_rows_for_b128 (:197-265): out[b] = sum over rows r of vec[r] * bit_b(mat[r]),
 * b = 0..127; mat and vec hold 2^log_len elements, out 128.  variant 0 = the
 * reference's method of four Russians over 4-row chunks with a 4x4 bit transpose,
 * variant 1 = the definition, bit by bit. */
void bo_rs_fold_1b_rows(const uint64_t *mat, const uint64_t *vec, int log_len, uint64_t *out, int variant,
                        int nthreads);
/* fold_b128_elems_inplace (:130-176) / fold_elems_inplace (:53-113) for B128:
 * elems[x] <- sum over b of bit_b(elems[x]) * vec[b] (vec: 128 elements).
 * variant 0 = 8-bit subset-sum tables (fold_b128_elems_inplace), variant 1 = the
 * inner product over B1 components (fold_elems_inplace's fallback). */
void bo_rs_fold_elems(uint64_t *elems, size_t n, const uint64_t *vec, int variant, int nthreads);

/* Full FRI-only prover flow (prover/prover/src/fri/tests.rs:58-89 with
 * FRIFoldProver, fri/fold.rs:117-244, and FRIQueryProver, fri/query.rs:51-134).
 * challenges_in == NULL: challenges are sampled from the transcript; otherwise
 * the given log_msg_len challenges are used (they are still not observed).
 * Outputs: the proof tape, every committed root, every challenge used, the
 * terminal codeword. Returns 0 on success. */
int bo_fri_prove(const bo_domain *d, const uint64_t *msg, int log_dim, int log_inv_rate,
                 int log_batch, const int *arities, int n_arities, int n_queries,
                 const uint64_t *challenges_in, uint8_t *tape, size_t tape_cap,
                 size_t *tape_len, uint8_t *roots_out, uint64_t *challenges_out,
                 uint64_t *terminal_out, int nthreads);

/* ---- transcript primitives (verifier/transcript/src) --------------------- */
typedef struct bo_transcript bo_transcript;
typedef struct bo_vision bo_vision;

/* The same provers with the Vision Merkle scheme (leaf: Vision-4 or -6 digest,
 * node: Vision-4 compression, prover/examples/src/lib.rs setup_vision4); NULL
 * for both selects SHA-256 leaves and Sha256Compression nodes. */
int bo_fri_prove_ex(const bo_domain *d, const uint64_t *msg, int log_dim, int log_inv_rate, int log_batch,
                    const int *arities, int n_arities, int n_queries, const uint64_t *challenges_in, uint8_t *tape,
                    size_t tape_cap, size_t *tape_len, uint8_t *roots_out, uint64_t *challenges_out,
                    uint64_t *terminal_out, int nthreads, const bo_vision *leaf, const bo_vision *node);

/* ---- sumcheck, BaseFold, one-bit PCS opening ------------------------------ */
/* prove_single (prover protocols/sumcheck/prove.rs:17-60) over
 * BivariateProductSumcheckProver (sumcheck/bivariate_product.rs): a and b
 * (2^n_vars elements each) are folded in place; writes each round's c0, c1 as a
 * message, samples the challenge; evals_out = (a(r), b(r)). */
int bo_sumcheck_prove(uint64_t *a, uint64_t *b, int n_vars, const uint64_t sum[2], bo_transcript *t,
                      uint64_t *challenges_out, uint64_t evals_out[4], int nthreads);
/* BaseFoldProver::prove (prover protocols/basefold.rs:131-166).  The caller has
 * written the codeword root.  multilinear and transparent (2^(log_dim+lbs)) are
 * folded in place; codeword / tree are the committed interleaved codeword and
 * its Merkle tree (bo_encode_batch / bo_merkle_build). */
int bo_basefold_prove(const bo_domain *d, uint64_t *multilinear, uint64_t *transparent, const uint64_t claim[2],
                      const uint64_t *codeword, const uint8_t *tree, int log_dim, int lir, int lbs,
                      const int *arities, int n_ar, int n_queries, bo_transcript *t, uint64_t *challenges_out,
                      int nthreads);
int bo_basefold_prove_ex(const bo_domain *d, uint64_t *multilinear, uint64_t *transparent, const uint64_t claim[2],
                         const uint64_t *codeword, const uint8_t *tree, int log_dim, int lir, int lbs,
                         const int *arities, int n_ar, int n_queries, bo_transcript *t, uint64_t *challenges_out,
                         int nthreads, const bo_vision *leaf, const bo_vision *node);
/* TensorAlgebra<B1, B128>::transpose (verifier/math/src/tensor_algebra.rs:137-141):
 * out[i] bit j = in[j] bit i over 128 elements. */
void bo_tensor_transpose(const uint64_t *in, uint64_t *out);
/* OneBitPCSProver::prove (prover/prover/src/pcs.rs:107-170); packed
 * (2^(log_dim+lbs)) is consumed; eval_point has log_dim + lbs + 7 coordinates. */
int bo_pcs_prove(const bo_domain *d, uint64_t *packed, const uint64_t *eval_point, const uint64_t *codeword,
                 const uint8_t *tree, int log_dim, int lir, int lbs, const int *arities, int n_ar, int n_queries,
                 bo_transcript *t, int nthreads);
int bo_pcs_prove_ex(const bo_domain *d, uint64_t *packed, const uint64_t *eval_point, const uint64_t *codeword,
                    const uint8_t *tree, int log_dim, int lir, int lbs, const int *arities, int n_ar, int n_queries,
                    bo_transcript *t, int nthreads, const bo_vision *leaf, const bo_vision *node);
bo_transcript *bo_transcript_new(void);
void bo_transcript_free(bo_transcript *t);
void bo_transcript_message(bo_transcript *t, const uint8_t *bytes, size_t n);
void bo_transcript_observe(bo_transcript *t, const uint8_t *bytes, size_t n);
void bo_transcript_decommit(bo_transcript *t, const uint8_t *bytes, size_t n);
void bo_transcript_sample(bo_transcript *t, uint8_t *out, size_t n);
uint32_t bo_transcript_sample_bits(bo_transcript *t, int bits);
size_t bo_transcript_tape(const bo_transcript *t, uint8_t *out, size_t cap);

/* ---- Vision-M hash (verifier/verifier/src/hash/vision_{4,6}) --------------- */
/* m = 4 or 6; b_fwd_coeffs: 4 elements (B_FWD_COEFFS); initial / matrix /
 * constant: the round-key schedule's INITIAL_CONSTANT (m), CONSTANTS_MATRIX
 * (m*m, row-major), CONSTANTS_CONSTANT (m).  Round constants and B^-1 are
 * derived from them.  NULL for a bad m or a singular B. */
bo_vision *bo_vision_new(int m, const uint64_t *b_fwd_coeffs, const uint64_t *initial, const uint64_t *matrix,
                         const uint64_t *constant);
void bo_vision_free(bo_vision *v);
void bo_vision_round_constants(const bo_vision *v, uint64_t *out); /* (1 + 2*8) * m elements */
void bo_vision_b_maps(const bo_vision *v, const uint64_t *x, uint64_t *fwd, uint64_t *inv, size_t n);
void bo_vision_batch_invert(uint64_t *x, int n); /* batch_invert: invert_or_zero of n <= 6 elements */
void bo_vision_permutation(const bo_vision *v, uint64_t *states, size_t n_states);
void bo_vision_digest(const bo_vision *v, const uint8_t *data, size_t len, uint8_t out[32]);
void bo_vision_compress(const bo_vision *v4, const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);
int bo_merkle_build_vision(const bo_vision *leaf, const bo_vision *node4, const uint64_t *elems, size_t n_elems,
                           size_t batch, uint8_t *tree, int nthreads);

/* ---- seeded input generator (SplitMix64; fixed for the whole repository) -- */
void bo_splitmix_fill(uint64_t seed, uint64_t *out, size_t n_words);

#ifdef __cplusplus
}
#endif
#endif
