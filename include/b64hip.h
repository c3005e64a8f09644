/*
 * libb64hip: MI355X (gfx950) implementation of Binius64's polynomial-commitment
 * hot path -- B128 additive NTT / Reed-Solomon encode, SHA-256 Merkle commit,
 * FRI interleaved and coset folds, and the FRI fold/query prover -- behind a
 * C ABI.
 *
 * Conventions
 *  - Field elements are BinaryField128bGhash values, 16 bytes each, stored as the
 *    little-endian u128 (verifier/utils/src/serialization.rs:66-72); as uint64_t
 *    pairs they are (lo, hi).
 *  - Pointers named d_* are device (HBM) pointers; every other pointer is host.
 *  - `stream` is a hipStream_t (NULL = the null stream).  Calls are stream-ordered;
 *    calls that return host data (roots, layers, branches) synchronize the stream.
 *  - Every function returns a b64 status code; no exception crosses the ABI.
 *    Error codes mirror the reference's error variants where one exists.
 *  - Digests are the raw 32 bytes of Output<Sha256>.
 *
 * Each declaration cites the reference interface it replaces (paths relative to
 * the reference repository root).
 */
#ifndef B64HIP_H
#define B64HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define B64_OK 0
#define B64_ERR_INVALID_ARGS 1                /* fri::Error::InvalidArgs, assert! preconditions */
#define B64_ERR_ENCODER_SUBSPACE_MISMATCH 2   /* reed_solomon.rs:206 */
#define B64_ERR_SUBSPACE_DIMENSION_MISMATCH 3 /* reed_solomon.rs:208 */
#define B64_ERR_INCORRECT_BATCH_SIZE 4        /* merkle_tree::Error::IncorrectBatchSize */
#define B64_ERR_POWER_OF_TWO_LENGTH_REQUIRED 5 /* merkle_tree::Error::PowerOfTwoLengthRequired */
#define B64_ERR_EARLY_PROVER_FINISH 6         /* fri::Error::EarlyProverFinish */
#define B64_ERR_INCORRECT_LAYER_DEPTH 7       /* merkle_tree::Error::IncorrectLayerDepth */
#define B64_ERR_INDEX_OUT_OF_RANGE 8          /* merkle_tree::Error::IndexOutOfRange */
#define B64_ERR_INVALID_FOLD_ARITY_SEQUENCE 9 /* fri::Error::InvalidFoldAritySequence */
#define B64_ERR_BUFFER_TOO_SMALL 10
#define B64_ERR_NOT_COMMIT_ROUND 11
#define B64_ERR_EXPECTED_FOLD 12              /* sumcheck::Error::ExpectedFold */
#define B64_ERR_EXPECTED_EXECUTE 13           /* sumcheck::Error::ExpectedExecute */
#define B64_ERR_NOT_ENOUGH_BYTES 14           /* SerializationError::NotEnoughBytes */
#define B64_ERR_INVALID_CONSTRUCTION 15       /* SerializationError::InvalidConstruction ("Proof::version") */
#define B64_ERR_FROM_UTF8 16                  /* SerializationError::FromUtf8Error */
#define B64_ERR_USIZE_TOO_LARGE 17            /* SerializationError::UsizeTooLarge */
#define B64_ERR_NO_DEVICE 20
#define B64_ERR_HIP 21
#define B64_ERR_OUT_OF_MEMORY 22
#define B64_ERR_ADMISSION_REJECTED 23

/* Merkle schemes of the commitments (the MerkleTreeProver type parameter of the
 * reference's provers): SHA-256 leaves with Sha256Compression nodes (the
 * default, verifier/verifier/src/config.rs), or VisionHasherDigest leaves of
 * width 4 or 6 with VisionCompression nodes (prover/examples/src/lib.rs
 * setup_vision4; Vision-6 leaves as in prover/prover/benches/binary_merkle_tree.rs). */
#define B64_MERKLE_SHA256 0u
#define B64_MERKLE_VISION4 4u
#define B64_MERKLE_VISION6 6u

typedef struct b64_domain b64_domain;
typedef struct b64_transcript b64_transcript;
typedef struct b64_fri_prover b64_fri_prover;
typedef struct b64_sumcheck b64_sumcheck;

/* FRIParams (verifier/verifier/src/fri/common.rs:15-31). */
#define B64_MAX_FOLD_ARITIES 32
typedef struct b64_fri_params {
    uint32_t log_dim;        /* rs_code.log_dim() */
    uint32_t log_inv_rate;   /* rs_code.log_inv_rate() */
    uint32_t log_batch_size;
    uint32_t n_test_queries;
    uint32_t n_fold_arities;
    uint32_t fold_arities[B64_MAX_FOLD_ARITIES];
    uint32_t log_terminal_dim; /* n_final_challenges() */
} b64_fri_params;

/* ---- library / device ----------------------------------------------------- */
const char *b64_version(void);
const char *b64_status_string(int status);
/* Number of visible HIP devices (0 when none). */
int b64_device_count(void);
int b64_device_alloc(size_t bytes, void **d_ptr);
int b64_device_free(void *d_ptr);
int b64_memcpy_htod(void *d_dst, const void *src, size_t bytes, void *stream);
int b64_memcpy_dtoh(void *dst, const void *d_src, size_t bytes, void *stream);
int b64_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream);
int b64_stream_synchronize(void *stream);

/* Kernel tuning knobs (no counterpart in the reference; the autotuner's
 * interface).  Keys: "ntt.variant" (0 = general multiply, 1 = per layer, LDS
 * window tables where a twiddle covers at least "ntt.tab_min" butterflies of a
 * tile, 2 = as 1 with the forward 4-layer passes whose flipped bits start at
 * index bit 10 or above run in registers, 3 = as 2 with the forward 4-layer
 * passes at index bits 8..9 in registers through 2 KiB tables and the other
 * forward passes on LDS tiles through 2 KiB tables -- or, with "ntt.tab_group"
 * 1 and at most 30 twiddles per tile, 1 KiB tables all built in one round --,
 * 4 = as 3 with every register pass through 2 KiB tables), "ntt.max_k" (layers per NTT pass, 1..7), "ntt.tile_log" (log2 elements
 * per LDS tile), "ntt.tab_group" (window tables resident in LDS at once: 1, 2,
 * 4 or 8), "ntt.waves" (4 = tables-only passes use a kernel limited to 128 VGPRs,
 * 0 = default).  "ntt.<knob>:<log_d>" sets the knob for transforms of size
 * 2^log_d only (the autotuner's per-size table); "ntt.clear_sizes" drops every
 * per-size setting.  "batch.<knob>[:<log_d>]" and "batch.clear_sizes" do the same
 * for the NTT passes of batch plans: variants 0-1 run every pass in the multi-job
 * LDS-tile kernel (max_k, tile_log, tab_min, tab_group), variants 3-4 run each
 * pass in the multi-job form of the kernel a single-job launch would pick (register
 * passes, the one-round small-table tile), the rest in the table kernel.
 * "host.pclmul" (default 1) lets the host's field products use PCLMULQDQ where
 * the CPU has it; 0 selects the portable shift-and-XOR loop (same results);
 * reading it reports whether PCLMULQDQ is in use.  Returns INVALID_ARGS for an
 * unknown key or value.  Settings are process-wide; change them only while no
 * call is running. */
int b64_config_set(const char *key, int64_t value);
int b64_config_get(const char *key, int64_t *value);

/* ---- field ------------------------------------------------------------------ */
/* BinaryField128bGhash Mul (verifier/field/src/ghash.rs:188-195), host side. */
void b64_gf_mul(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
/* invert_or_zero (arch/portable/packed_ghash_128.rs:117-128), host side. */
void b64_gf_inv(const uint64_t a[2], uint64_t out[2]);
/* Element-wise d_out[i] = d_a[i] * d_b[i] on the device (the device multiply
 * every kernel uses; exposed for parity testing). */
int b64_gf_mul_device(const void *d_a, const void *d_b, void *d_out, uint64_t n, void *stream);
/* d_out[i] = d_a[i] * c for a constant c through the LDS window-table path. */
int b64_gf_mul_const_device(const void *d_a, const uint64_t c[2], void *d_out, uint64_t n, void *stream);

/* ---- DomainContext (verifier/math/src/ntt/mod.rs:118-179) --------------------- */
/* GenericPreExpanded::generate_from_subspace (ntt/domain_context.rs:99-132) over
 * an explicit basis of `dim` elements. */
int b64_domain_create(const uint64_t *basis, uint32_t dim, b64_domain **out);
/* ... over BinarySubspace::with_dim(dim) (verifier/math/src/binary_subspace.rs:81-86). */
int b64_domain_create_standard(uint32_t dim, b64_domain **out);
void b64_domain_destroy(b64_domain *d);
/* DomainContext::log_domain_size (ntt/mod.rs:130) */
uint32_t b64_domain_log_size(const b64_domain *d);
/* DomainContext::subspace(i) basis, i elements (ntt/mod.rs:143) */
int b64_domain_subspace(const b64_domain *d, uint32_t i, uint64_t *basis_out);
/* DomainContext::twiddle(layer, block) (ntt/mod.rs:162) */
int b64_domain_twiddle(const b64_domain *d, uint32_t layer, uint64_t block, uint64_t out[2]);

/* ---- AdditiveNTT (verifier/math/src/ntt/mod.rs:58-113) ------------------------ */
/* forward_transform(data, skip_early, skip_late) in place on 2^log_len elements;
 * preconditions of input_check (ntt/reference.rs:89-100) return INVALID_ARGS. */
int b64_ntt_forward(const b64_domain *d, void *d_data, uint32_t log_len, uint32_t skip_early,
                    uint32_t skip_late, void *stream);
/* inverse_transform (ntt/reference.rs:49-76) */
int b64_ntt_inverse(const b64_domain *d, void *d_data, uint32_t log_len, uint32_t skip_early,
                    uint32_t skip_late, void *stream);

/* No counterpart in the reference: the pass schedule the library runs for a
 * transform of 2^log_len elements over layers skip_early .. log_len - skip_late
 * (encode != 0: the first pass gathers from the message, as in
 * b64_rs_encode_batch), for the autotuner and the benches.  Per pass: its first
 * layer, layer count, the low index bits held in a tile, the index bits above
 * the flipped ones held in a tile, the kernel, and the work items of its grid on
 * a device of n_cus compute units (0 = the current device).  A forward or
 * inverse transform whose plan has a pass of 2^32 or more work items is refused
 * with B64_ERR_INVALID_ARGS before anything is launched. */
#define B64_NTT_KERNEL_GENERAL 0    /* LDS tile, general product in every layer */
#define B64_NTT_KERNEL_TAB 1        /* LDS tile, window tables in every layer */
#define B64_NTT_KERNEL_TAB_HYBRID 2 /* LDS tile, tables where a twiddle covers >= ntt.tab_min butterflies */
#define B64_NTT_KERNEL_TAB_W4 3     /* as TAB, 2^10-element tiles, at most 128 VGPRs */
#define B64_NTT_KERNEL_R16 4        /* 4 layers in registers, 15 resident tables */
#define B64_NTT_KERNEL_T8 5         /* LDS tile, 2 KiB tables per twiddle where one covers >= ntt.tab_min butterflies */
#define B64_NTT_KERNEL_R16W8 6      /* 4 layers in registers, 15 resident 2 KiB tables, 256-thread blocks */
#define B64_NTT_KERNEL_T4U 7        /* LDS tile, every twiddle's 1 KiB table built in one round */
typedef struct {
    uint32_t layer0, n_layers, low_bits, extra_bits, kernel;
    uint64_t workitems;
} b64_ntt_pass_info;
int b64_ntt_plan(uint32_t log_len, uint32_t skip_early, uint32_t skip_late, int inverse, int encode, uint32_t n_cus,
                 b64_ntt_pass_info *passes, uint32_t cap, uint32_t *n_passes);

/* ---- ReedSolomonCode::encode_batch (verifier/math/src/reed_solomon.rs:128-201) -- */
/* The code is ReedSolomonCode::with_domain_context_subspace(d, log_dim, log_inv_rate)
 * (reed_solomon.rs:54-65): SUBSPACE_DIMENSION_MISMATCH if log_dim+log_inv_rate
 * exceeds the domain.  d_msg holds 2^(log_dim+log_batch) elements, d_out
 * 2^(log_dim+log_batch+log_inv_rate). */
int b64_rs_encode_batch(const b64_domain *d, uint32_t log_dim, uint32_t log_inv_rate, const void *d_msg,
                        uint32_t log_batch, void *d_out, void *stream);

/* ---- MerkleTreeProver (prover/prover/src/merkle_tree/mod.rs:16-70) ----------- */
/* BinaryMerkleTreeProver::commit (mod.rs:26-38, prover.rs:93-118) with
 * Sha256 leaves and Sha256Compression nodes.  d_tree receives the flattened
 * inner_nodes [leaves | ... | root], (2*n_leaves - 1) * 32 bytes. */
int b64_merkle_commit(const void *d_elems, uint64_t n_elems, uint64_t batch_size, void *d_tree,
                      uint8_t root_out[32], void *stream);
/* BinaryMerkleTree::layer (binary_merkle_tree.rs:142-149): 2^depth digests. */
int b64_merkle_layer(const void *d_tree, uint32_t log_len, uint32_t depth, uint8_t *out, void *stream);
/* BinaryMerkleTree::branch (binary_merkle_tree.rs:154-169): log_len-depth digests. */
int b64_merkle_branch(const void *d_tree, uint32_t log_len, uint64_t index, uint32_t depth, uint8_t *out,
                      void *stream);

/* ---- FRI folds (prover/prover/src/fri/fold.rs:260-318) ------------------------ */
/* fold_interleaved: d_cw has 2^(log_len+log_batch) elements, challenges log_batch
 * host elements, d_out 2^log_len elements. */
int b64_fri_fold_interleaved(const void *d_cw, uint32_t log_len, uint32_t log_batch,
                             const uint64_t *challenges, void *d_out, void *stream);
/* fold_codeword: d_cw has 2^log_len elements, n_chal host challenges,
 * d_out 2^(log_len-n_chal) elements. */
int b64_fri_fold_codeword(const b64_domain *d, const void *d_cw, uint32_t log_len,
                          const uint64_t *challenges, uint32_t n_chal, void *d_out, void *stream);

/* ---- FRIParams (verifier/verifier/src/fri/common.rs) -------------------------- */
uint32_t b64_calculate_n_test_queries(uint32_t security_bits, uint32_t log_inv_rate); /* :182-186 */
uint32_t b64_estimate_optimal_arity(uint32_t log_block_length, uint32_t digest_size,
                                    uint32_t field_size); /* :191-219 */
/* FRIParams::new (:35-54) */
int b64_fri_params_new(uint32_t log_dim, uint32_t log_inv_rate, uint32_t log_batch_size,
                       const uint32_t *fold_arities, uint32_t n_fold_arities, uint32_t n_test_queries,
                       b64_fri_params *out);
/* FRIParams::choose_with_constant_fold_arity (:64-115) for an NTT over `d`.
 * INVALID_ARGS when the arity list would exceed B64_MAX_FOLD_ARITIES entries. */
int b64_fri_params_choose(const b64_domain *d, uint32_t log_msg_len, uint32_t security_bits,
                          uint32_t log_inv_rate, uint32_t arity, b64_fri_params *out);
/* Verifier::setup + Prover::setup parameter chain (verifier/verifier/src/verify.rs:61-106,
 * prover/prover/src/prove.rs:93-115): FRI params for 2^log_witness_elems B128
 * elements, and the prover's NTT domain (GenericPreExpanded over the RS code's
 * subspace).  The caller destroys *domain_out. */
int b64_setup_commit_params(uint32_t log_witness_elems, uint32_t log_inv_rate, uint32_t security_bits,
                            b64_fri_params *params_out, b64_domain **domain_out);

/* ---- commit_interleaved (prover/prover/src/fri/commit.rs:25-75) --------------- */
/* d_codeword: 2^(rs_log_len+log_batch) elements; d_tree: (2^(rs_log_len+1)-1)*32 bytes. */
int b64_fri_commit_interleaved(const b64_fri_params *p, const b64_domain *d, const void *d_msg,
                               void *d_codeword, void *d_tree, uint8_t root_out[32], void *stream);

/* The same with the Merkle scheme merkle_hash (B64_MERKLE_*). */
int b64_fri_commit_interleaved_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash,
                                  const void *d_msg, void *d_codeword, void *d_tree, uint8_t root_out[32],
                                  void *stream);

/* commit_interleaved followed by the first fold_interleaved (prover fri/fold.rs:260-282)
 * with log_batch_size challenges supplied up front, fused so the codeword is read
 * once for leaf hashing and folding.  In the transcript-driven protocol the
 * challenges are sampled after the root is written (protocols/basefold.rs:139-147),
 * so this form serves benchmark / pre-seeded-challenge callers.  d_folded receives
 * 2^rs_log_len elements. */
int b64_fri_commit_fold(const b64_fri_params *p, const b64_domain *d, const void *d_msg, void *d_codeword,
                        void *d_tree, const uint64_t *challenges, void *d_folded, uint8_t root_out[32],
                        void *stream);

/* ---- ProverTranscript<HasherChallenger<Sha256>> (verifier/transcript/src) ----- */
int b64_transcript_create(b64_transcript **out);
void b64_transcript_destroy(b64_transcript *t);
void b64_transcript_message(b64_transcript *t, const uint8_t *bytes, size_t n);   /* transcript.rs:318 */
void b64_transcript_observe(b64_transcript *t, const uint8_t *bytes, size_t n);   /* transcript.rs:283 */
void b64_transcript_decommit(b64_transcript *t, const uint8_t *bytes, size_t n);  /* transcript.rs:308 */
void b64_transcript_sample(b64_transcript *t, uint8_t *out, size_t n);            /* CanSample, :403 */
uint32_t b64_transcript_sample_bits(b64_transcript *t, uint32_t bits);            /* :413 */
size_t b64_transcript_tape(const b64_transcript *t, uint8_t *out, size_t cap);    /* finalize, :245 */

/* ---- Proof wire format (verifier/core/src/constraint_system.rs:1014-1100) ------ */
/* Proof::SERIALIZATION_VERSION and the challenger name the reference's prover
 * CLI writes (StdChallenger::NAME, verifier/verifier/src/config.rs:19-21). */
#define B64_PROOF_SERIALIZATION_VERSION 1u
#define B64_STD_CHALLENGER_NAME "HasherChallenger<Sha256>"
/* Proof::owned(data, challenger_type).serialize: u32 LE version, then the
 * challenger type and the transcript bytes, each as a u32 LE length and the
 * bytes (usize serialises as u32, utils/serialization.rs:47-54).  *out_len
 * always receives the encoded size; out may be NULL to ask for it.
 * USIZE_TOO_LARGE if a length exceeds u32, BUFFER_TOO_SMALL if out_cap is less
 * than the encoded size. */
int b64_proof_serialize(const uint8_t *data, uint64_t data_len, const char *challenger_type, uint8_t *out,
                        uint64_t out_cap, uint64_t *out_len);
/* Proof::deserialize, zero-copy: the challenger type and the transcript data
 * are returned as (offset, length) ranges of buf.  INVALID_CONSTRUCTION for a
 * version other than 1, NOT_ENOUGH_BYTES for a truncated buffer, FROM_UTF8 if
 * the challenger type is not UTF-8.  Bytes after the proof are ignored, as a
 * read buffer's remainder is. */
int b64_proof_deserialize(const uint8_t *buf, uint64_t buf_len, uint64_t *challenger_off, uint64_t *challenger_len,
                          uint64_t *data_off, uint64_t *data_len);

/* ---- FRIFoldProver / FRIQueryProver (prover/prover/src/fri/fold.rs, query.rs) - */
/* FRIFoldProver::new (fold.rs:57-82).  The prover keeps pointers to the committed
 * codeword and tree (device) and owns the folded oracles it commits. */
int b64_fri_prover_create(const b64_fri_params *p, const b64_domain *d, const void *d_codeword,
                          const void *d_tree, void *stream, b64_fri_prover **out);
/* The same with the Merkle scheme merkle_hash (B64_MERKLE_*) for the folded oracles. */
int b64_fri_prover_create_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash, const void *d_codeword,
                             const void *d_tree, void *stream, b64_fri_prover **out);
void b64_fri_prover_destroy(b64_fri_prover *fp);
uint32_t b64_fri_prover_n_rounds(const b64_fri_prover *fp);              /* fold.rs:85-87 */
uint32_t b64_fri_prover_curr_round(const b64_fri_prover *fp);            /* fold.rs:90-92 */
int b64_fri_prover_receive_challenge(b64_fri_prover *fp, const uint64_t challenge[2]); /* :107-110 */
/* execute_fold_round (fold.rs:117-175): *committed = 1 and root_out set when the
 * round commits a folded oracle. */
int b64_fri_prover_execute_fold_round(b64_fri_prover *fp, uint8_t root_out[32], int *committed);
/* finish_proof (fold.rs:221-244): terminal codeword, optimal layers and the
 * n_test_queries coset openings written to the transcript's tape. */
int b64_fri_prover_finish_proof(b64_fri_prover *fp, b64_transcript *t);
/* Committed folded oracle k (0-based): device pointer and log length (for tests). */
int b64_fri_prover_oracle(const b64_fri_prover *fp, uint32_t k, const void **d_codeword,
                          uint32_t *log_len, const void **d_tree);

/* ---- commit jobs: plain and batched (no single counterpart in the reference) -- */
/* A job is the FRI commit phase of one message of 2^log_msg_len B128 elements
 * with the Verifier::setup parameters (verify.rs:61-106: rate 1/2, batch 2^4,
 * arity 4) and pre-seeded challenges: commit_interleaved (fri/commit.rs:25-75),
 * then every commit round of FRIFoldProver::execute_fold_round (fri/fold.rs:117-175).
 * roots_out receives the codeword root and each folded oracle's root, the same
 * digests the transcript-driven prover writes with these challenges. */
typedef struct b64_job_desc {
    uint32_t log_msg_len;       /* 8 .. 40 (batch 2^4) */
    const void *d_msg;          /* device, 2^log_msg_len elements */
    const uint64_t *challenges; /* host, log_msg_len elements (read when the job is run / planned) */
    void *d_workspace;          /* device, b64_job_workspace_size(log_msg_len) bytes */
    uint8_t *roots_out;         /* host, b64_job_n_roots(log_msg_len) * 32 bytes */
} b64_job_desc;
typedef struct b64_batch_plan b64_batch_plan;

size_t b64_job_workspace_size(uint32_t log_msg_len);
uint32_t b64_job_n_roots(uint32_t log_msg_len);
/* Oracle k of a job's workspace: k = 0 the interleaved codeword, k >= 1 folded oracle k-1. */
int b64_job_oracle(const b64_job_desc *job, uint32_t k, const void **d_codeword, uint32_t *log_len);
/* Plain path: the job's kernels launched one stage at a time, then its roots read back. */
int b64_job_run(const b64_job_desc *job, void *stream);
/* Batched path: each launch runs one stage for every job that has it; with
 * use_graph the launch sequence is captured once in a hipGraph and replayed. */
int b64_batch_plan_create(const b64_job_desc *jobs, uint32_t n_jobs, int use_graph, void *stream,
                          b64_batch_plan **out);
/* run = enqueue (kernels + roots copied to pinned memory, asynchronous) +
 * stream synchronize + collect (roots delivered to each job's roots_out). */
int b64_batch_plan_run(b64_batch_plan *plan, void *stream);
int b64_batch_plan_enqueue(b64_batch_plan *plan, void *stream);
int b64_batch_plan_collect(b64_batch_plan *plan);
uint32_t b64_batch_plan_n_launches(const b64_batch_plan *plan);
void b64_batch_plan_destroy(b64_batch_plan *plan);

/* ---- batched proofs: FRI driven by each job's own transcript ------------------ */
/* No single counterpart in the reference: for every job, the sequence of
 * prover/prover/src/fri/tests.rs:57-88 (commit_interleaved, fri/commit.rs:25-75;
 * the root written to the transcript; FRIFoldProver::execute_fold_round with
 * each challenge sampled from the transcript, fri/fold.rs:117-175; finish_proof,
 * :221-244), with the Verifier::setup parameters of the job's size (as
 * b64_job_desc).  Each round's device work runs as one launch per stage across
 * the jobs that have it; the host samples every job's challenges between
 * rounds (one device round trip per round), and a job's query phase starts as
 * soon as its own rounds are over.  Each transcript ends byte-identical to the
 * one-job prover's. */
typedef struct b64_prove_job {
    uint32_t log_msg_len;        /* 8 .. 40 (batch 2^4) */
    const void *d_msg;           /* device, 2^log_msg_len elements */
    void *d_workspace;           /* device, b64_job_workspace_size(log_msg_len) bytes */
    b64_transcript *transcript;  /* the job's prover transcript, continued */
    uint8_t *roots_out;          /* host, b64_job_n_roots(log_msg_len) * 32 bytes, or NULL */
} b64_prove_job;
int b64_batch_prove(const b64_prove_job *jobs, uint32_t n_jobs, void *stream);

/* ---- serving: continuous batching of commit jobs (no counterpart in the reference) */
/* A worker thread admits submitted jobs in arrival order whenever their HBM
 * footprint (message + job workspace) fits in the server's budget, groups the
 * admitted jobs into batches of up to max_batch, and keeps up to n_streams
 * batches in flight on separate streams, so jobs that arrive while others run
 * are admitted as memory and stream slots free up.  Results equal b64_job_run's. */
typedef struct b64_server b64_server;
typedef struct b64_server_config {
    size_t hbm_budget_bytes; /* device memory pool for admitted jobs */
    uint32_t max_batch;      /* jobs per batch launch */
    uint32_t n_streams;      /* batches in flight */
} b64_server_config;
int b64_server_create(const b64_server_config *cfg, b64_server **out);
/* Queues a job and returns a ticket.  The host message (2^log_msg_len B128
 * elements; pinned memory makes its upload asynchronous) is read when the job is
 * admitted and must stay valid until poll/wait reports the job done; the
 * challenges are copied.  ADMISSION_REJECTED if the job alone exceeds the budget. */
int b64_server_submit(b64_server *srv, uint32_t log_msg_len, const uint64_t *msg, const uint64_t *challenges,
                      uint64_t *ticket);
/* Queues a proof job: the job of b64_batch_prove (commitment and FRI proof under
 * `transcript`, which must stay valid and untouched until poll/wait reports the
 * job done; the proof is then on its tape).  Proof jobs are batched with other
 * proof jobs; each of their rounds waits for the host to sample challenges, and
 * the slot's other work and the other slots keep running meanwhile. */
int b64_server_submit_proof(b64_server *srv, uint32_t log_msg_len, const uint64_t *msg, b64_transcript *transcript,
                            uint64_t *ticket);
/* *done = 1 and roots (b64_job_n_roots * 32 bytes) written once the job has
 * finished; latency_s = submit-to-completion time measured by the server. */
int b64_server_poll(b64_server *srv, uint64_t ticket, uint8_t *roots_out, int *done, double *latency_s);
/* Blocks until the job has finished. */
int b64_server_wait(b64_server *srv, uint64_t ticket, uint8_t *roots_out, double *latency_s);
/* jobs completed, batches launched, peak bytes of the pool in use */
int b64_server_stats(b64_server *srv, uint64_t *jobs_done, uint64_t *batches, size_t *peak_bytes);
void b64_server_destroy(b64_server *srv);

/* ---- witness packing (prover/prover/src/prove.rs:286-327) --------------------- */
/* B128 element i = word[2i+1] << 64 | word[2i]; zero-padded to 2^log_elems. */
int b64_pack_witness(const void *d_words, uint64_t n_words, uint32_t log_elems, void *d_out, void *stream);

/* ---- ring switch (prover/prover/src/ring_switch.rs) --------------------------- */
/* fold_1b_rows_for_b128 (ring_switch.rs:197-265): d_out[b] (128 elements) = sum over
 * rows r < 2^log_len of d_vec[r] * bit b of d_mat[r], bit b of the LE u128. */
int b64_ring_switch_fold_1b_rows(const void *d_mat, const void *d_vec, uint32_t log_len, void *d_out,
                                 void *stream);
/* No counterpart in the reference: b64_ring_switch_fold_1b_rows for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  Same result; its per-block partial sums live in a file-scope
 * device array of B64_RS_SCRATCH_BLOCKS * 128 elements (4 MiB), held by every GPU
 * the library runs on until the process ends, and the call allocates nothing.
 * Calls on one GPU share that array: two calls that overlap in time on one GPU
 * (different streams, different host threads) race and must be ordered by the
 * caller.  Every block writes all 128 of its partial sums before they are read,
 * so no clearing is needed.  log_len > 40 is rejected (B64_ERR_INVALID_ARGS); a
 * launch needing more blocks than the array holds is rejected with
 * B64_ERR_BUFFER_TOO_SMALL (no log_len <= 40 does: blocks are capped at 2048). */
#define B64_RS_SCRATCH_BLOCKS 2048
int b64_ring_switch_fold_1b_rows_noalloc(const void *d_mat, const void *d_vec, uint32_t log_len, void *d_out,
                                         void *stream);
/* fold_b128_elems_inplace (ring_switch.rs:130-176), in place on 2^log_len elements:
 * x <- sum over b of bit b of x times vec[b]; vec is 128 host elements. */
int b64_ring_switch_fold_elems(void *d_elems, uint32_t log_len, const uint64_t *vec, void *stream);

/* ---- eq_ind_partial_eval (verifier/math/src/multilinear/eq.rs:123-131) ------ */
/* d_out[x] = prod over k < n of (bit k of x ? point[k] : 1 + point[k]),
 * 2^n elements; n <= 40. */
int b64_eq_ind_partial_eval(const uint64_t *point, uint32_t n, void *d_out, void *stream);

/* ---- BivariateProductSumcheckProver (prover/prover/src/protocols/sumcheck/
 *      bivariate_product.rs) through the SumcheckProver round interface
 *      (sumcheck/common.rs:28-53).  Binds the highest variable first. ------- */
/* new (:25-35): the prover takes d_a and d_b (2^n_vars elements each) as its
 * multilinears and folds them in place, as the reference takes the buffers by
 * value; their contents are consumed. */
int b64_sumcheck_create(void *d_a, void *d_b, uint32_t n_vars, const uint64_t sum[2], void *stream,
                        b64_sumcheck **out);
void b64_sumcheck_destroy(b64_sumcheck *sc);
uint32_t b64_sumcheck_n_vars(const b64_sumcheck *sc);                  /* :39-41 */
/* execute (:43-82): the round polynomial's coefficients c0, c1, c2.
 * EXPECTED_FOLD if called twice without a fold. */
int b64_sumcheck_execute(b64_sumcheck *sc, uint64_t coeffs_out[6]);
/* fold (:84-98).  EXPECTED_EXECUTE unless execute came first. */
int b64_sumcheck_fold(b64_sumcheck *sc, const uint64_t challenge[2]);
/* finish (:100-121): the two multilinears' evaluations at the challenge point. */
int b64_sumcheck_finish(b64_sumcheck *sc, uint64_t evals_out[4]);

/* ---- BaseFoldProver::prove (prover/prover/src/protocols/basefold.rs:131-166) -- */
/* The caller has committed the multilinear (b64_fri_commit_interleaved) and
 * written the root to the transcript.  d_multilinear and d_transparent
 * (2^(log_dim+log_batch_size) elements each) are consumed; claim is the sum of
 * their product.  Writes every round polynomial, round commitment and the FRI
 * query proof to t. */
int b64_basefold_prove(const b64_fri_params *p, const b64_domain *d, void *d_multilinear, void *d_transparent,
                       const uint64_t claim[2], const void *d_codeword, const void *d_tree, b64_transcript *t,
                       void *stream);

int b64_basefold_prove_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash, void *d_multilinear,
                          void *d_transparent, const uint64_t claim[2], const void *d_codeword, const void *d_tree,
                          b64_transcript *t, void *stream);

/* ---- OneBitPCSProver::prove (prover/prover/src/pcs.rs:107-170) ---------------- */
/* Proves the evaluation of the committed B1 multilinear (d_packed: 2^(log_dim +
 * log_batch_size) B128 elements, 128 bits each; consumed) at eval_point
 * (n_point = log_dim + log_batch_size + 7 coordinates): ring-switch partial
 * evaluations, row-batching challenges, then BaseFold. */
int b64_pcs_prove(const b64_fri_params *p, const b64_domain *d, const void *d_codeword, const void *d_tree,
                  void *d_packed, const uint64_t *eval_point, uint32_t n_point, b64_transcript *t, void *stream);
/* The same over a codeword committed with the Merkle scheme merkle_hash (B64_MERKLE_*). */
int b64_pcs_prove_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash, const void *d_codeword,
                     const void *d_tree, void *d_packed, const uint64_t *eval_point, uint32_t n_point,
                     b64_transcript *t, void *stream);

/* ---- Vision-4 / Vision-6 hash (verifier/verifier/src/hash/vision_{4,6},
 *      prover/prover/src/hash/vision_{4,6}) ------------------------------------ */
/* The permutation (vision_{m}/permutation.rs) of n states of m = 4 or 6
 * elements each, in place (d_states: n * m elements). */
int b64_vision_permute(uint32_t m, void *d_states, uint64_t n, void *stream);
/* VisionHasherDigest (vision_{m}/digest.rs; the prover's multi-digest gives the
 * same values) of n messages of msg_elems B128 elements each, message k at
 * d_msgs + 16 * k * msg_elems bytes; d_out receives n 32-byte digests. */
int b64_vision_digest(uint32_t m, const void *d_msgs, uint64_t n, uint64_t msg_elems, void *d_out, void *stream);
/* BinaryMerkleTreeProver::commit (merkle_tree/prover.rs) with VisionHasherDigest
 * leaves of width leaf_m (4: prover/examples/src/lib.rs setup_vision4; 6: the
 * leaf hasher of prover/prover/benches/binary_merkle_tree.rs) and
 * VisionCompression (vision_4/compression.rs) nodes.  Layout and error codes as
 * b64_merkle_commit. */
int b64_merkle_commit_vision(uint32_t leaf_m, const void *d_elems, uint64_t n_elems, uint64_t batch, void *d_tree,
                             uint8_t root_out[32], void *stream);

#ifdef __cplusplus
}
#endif
#endif
