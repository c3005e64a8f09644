// Internal interfaces shared by the kernel launchers (kernels.hip) and the C ABI
// (api.hip).  Nothing here crosses the library boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>
#include <stddef.h>

#include "../../include/b64hip.h"

namespace b64 {

constexpr int MAX_DIM = 64;          // largest supported domain dimension
constexpr int W_STRIDE = MAX_DIM;    // W table row stride (elements)

struct u128h {
    uint64_t lo, hi;
};

// Per-domain device data: W[layer][i] = evals[dim-1-layer][1+i] for i < layer,
// the F2-basis of the layer's twiddles (ntt/domain_context.rs:117-128,149-151):
// twiddle(layer, block) = XOR over set bits i of block of W[layer][i].
// X (when not null) holds the same twiddles pre-expanded for layers below
// x_layers (the domain's GenericPreExpanded form, domain_context.rs:117-151):
// X[2^layer + block] = twiddle(layer, block).
constexpr int X_MAX_LAYERS = 21;  // at most 2^21 entries (32 MiB) per domain
struct DeviceDomain {
    int dim;
    const uint4 *W;  // dim * W_STRIDE entries
    const uint4 *X = nullptr;
    int x_layers = 0;
};

// Sha256Compression IV (LE words of SHA-256("BINIUS SHA-256 COMPRESS")).
struct CompressIV {
    uint32_t w[8];
};

constexpr int MAX_FOLD_CHAL = 12;
constexpr int MERKLE_RED_LOG = 9;  // a subtree block reduces a chunk of up to 2^9 nodes
// Levels wider than 2^MERKLE_WIDE_LOG nodes are reduced MERKLE_WIDE_LEVELS at a
// time, so most compressions run in full wavefronts.
constexpr int MERKLE_WIDE_LOG = 13;
constexpr int MERKLE_WIDE_LEVELS = 4;
// Levels of at least 2^MERKLE_QUAD_LOG nodes are reduced two at a time by one
// thread per 4 nodes (k_merkle_quad: no LDS, no barrier, every lane busy).
constexpr int MERKLE_QUAD_LOG = 18;
// Nodes per block of a subtree launch over a level of n_in nodes.
inline uint64_t merkle_chunk(uint64_t n_in) {
    return n_in < (1ull << MERKLE_RED_LOG) ? n_in : (1ull << MERKLE_RED_LOG);
}

// Per-stage kernel arguments.  Single-job launches pass them by value; the
// batched path keeps one table per launch in device memory.
struct NttPassArgs {
    uint4 *data;
    const uint4 *W;
    // First encode pass only: read the tile straight from the message,
    // out[r*2^m + i] = msg[rev_m(i)] (reed_solomon.rs:164-187), instead of data.
    const uint4 *msg;
    int log_d, layer0, k, c, m;
    // index bits lb+k .. lb+k+e-1, just above the flipped bits, also held in the
    // tile (independent butterfly groups), when the c low bits cannot fill it
    int e;
    int tab_min;  // layers whose twiddle covers >= tab_min butterflies use window tables
    // pre-expanded twiddles of the layers below x_layers (DeviceDomain::X), or null
    const uint4 *X = nullptr;
    int x_layers = 0;
};
struct LeafArgs {
    const uint4 *elems;
    uint64_t n_leaves, batch;
    uint4 *digests;
};
struct Leaf16Args {
    const uint4 *cw;
    uint64_t n_leaves;
    uint4 *digests;
    uint4 *folded;  // nullptr: hash only
    uint4 chal[4];
};
struct SubtreeArgs {
    uint4 *tree;
    uint64_t in_off, n_in;
    int levels;
    CompressIV iv;
};
struct FoldCwArgs {
    const uint4 *cw;
    uint4 *out;
    uint64_t n_out;
    int log_len, n_chal;
    const uint4 *W;
    uint4 chal[MAX_FOLD_CHAL];
};
// One job's fold_interleaved of a batch-2^4 codeword (16 elements per output).
struct FoldIlArgs {
    const uint4 *cw;
    uint4 *out;
    uint64_t n_out;
    uint4 chal[4];
};
// A pass of one job in a multi-job register or small-table pass launch.
struct NttJobArgs {
    NttPassArgs a;
    uint64_t per_block;  // positions per block (register passes), else 0
};
enum class MultiKind {
    NttPass,  // multi-job LDS-tile kernel with 32-row tables (NttPassArgs)
    NttR16, NttR16Msg, NttR16W8, NttR16W8Msg, NttT4U9, NttT4U10,  // (NttJobArgs)
    Leaf16Fold, LeafHash, Subtree, FoldCodeword,
    LeafHashLong,  // one block per leaf, for few leaves of 32..128 elements (LeafArgs)
    MerkleQuad,    // two tree levels, one thread per 4 input nodes (SubtreeArgs, levels = 2)
    FoldInterleaved  // fold_interleaved of batch 2^4 (FoldIlArgs)
};

struct NttTuning {
    int variant;   // 0 = general product, 1 = per-layer choice (window tables where amortised),
                   // 2 = as 1, with forward 4-layer passes above index bit 10 in registers,
                   // 3 = as 2, forward passes at bits 8..9 in registers and the remaining
                   //     forward passes on LDS tiles, both through 2 KiB tables,
                   // 4 = as 3, every register pass through 2 KiB tables
    int max_k;     // layers per pass
    int tile_log;  // log2 elements per LDS tile
    int tab_min;   // variants >= 1: butterflies per twiddle (in a tile) needed to build a table
    int tab_group; // variant 1: window tables resident in LDS at once (1, 2, 4 or 8; 8 KiB each)
    int waves;     // tables-only passes: 4 = kernel limited to 128 VGPRs (4 waves per SIMD), 0 = default
};
extern NttTuning g_ntt_tuning;  // defaults
// Per transform size (log_d) overrides; variant < 0 marks an unset entry.
extern NttTuning g_ntt_tuning_size[MAX_DIM + 1];
const NttTuning &ntt_tuning_for(int log_d);
// Schedules of the NTT passes in batch plans (the multi-job table kernel; only
// max_k, tile_log, tab_min, tab_group and variant 0 = no tables are read), kept
// apart from the single-job schedules ("batch.<knob>[:<log_d>]").
extern NttTuning g_batch_tuning;
extern NttTuning g_batch_tuning_size[MAX_DIM + 1];
const NttTuning &batch_tuning_for(int log_d);

// Kernel launchers.  All return hipError_t of the launch.
hipError_t launch_gf_mul(const uint4 *a, const uint4 *b, uint4 *out, uint64_t n, hipStream_t s);
hipError_t launch_gf_mul_const(const uint4 *a, uint4 c, uint4 *out, uint64_t n, hipStream_t s);
hipError_t launch_bitrev_replicate(const uint4 *msg, uint4 *out, int m, int lir, hipStream_t s);
// X[2^layer + block] = twiddle(layer, block) for layer < x_layers (X[0] = 0).
hipError_t launch_expand_twiddles(const uint4 *W, uint4 *X, int x_layers, hipStream_t s);
hipError_t launch_ntt(const DeviceDomain &dd, uint4 *data, int log_d, int skip_early, int skip_late,
                      bool inverse, hipStream_t s, const uint4 *msg = nullptr, int msg_log = 0);
hipError_t launch_leaf_hash(const uint4 *elems, uint64_t n_leaves, int log_batch, uint64_t batch,
                            uint4 *digests, hipStream_t s);
// Whether leaves of `batch` elements, n_leaves of them, go to the one-block-per-leaf
// kernel (MultiKind::LeafHashLong, one block per leaf).
bool leaf_hash_long_fits(uint64_t batch, uint64_t n_leaves);
hipError_t launch_merkle_levels(uint4 *tree, uint64_t n_leaves, const CompressIV &iv, hipStream_t s);
// 16-element leaves; with folded != nullptr also fold_interleaved with 4 challenges.
hipError_t launch_leaf16_fold(const uint4 *cw, uint64_t n_leaves, uint4 *digests, uint4 *folded, const uint4 *chal,
                              hipStream_t s);
hipError_t launch_fold_interleaved(const uint4 *cw, uint4 *out, uint64_t n_out, int log_batch,
                                   const uint4 *eq_host, hipStream_t s);
hipError_t launch_fold_codeword(const DeviceDomain &dd, const uint4 *cw, uint4 *out, int log_len,
                                const uint4 *chal_host, int n_chal, hipStream_t s);
// Subtree-reduction schedule of a tree with n_leaves leaves: per launch the
// input level offset (in digests), its node count and the levels reduced.
// quad[i] != 0 marks a two-level step for k_merkle_quad.
int merkle_reduce_steps(uint64_t n_leaves, uint64_t *offs, uint64_t *ns, int *levels, int *quad);
// Blocks of one step's launch.
uint32_t merkle_step_blocks(uint64_t n_in, int quad);
// FRIFoldProver::execute_fold_round without the final wait: the fold, the
// Merkle commit and, on a committed round, the copy of the root into pinned
// host memory (*root_host, valid once the stream is synchronised).
int fri_fold_round_enqueue(b64_fri_prover *fp, int *committed, const uint8_t **root_host);
// Pass schedule of a forward NTT (layer0, k, c, e per pass); returns the count.
// A pass that gathers from the message (the first forward pass of an encode,
// msg_first) keeps e = 0.
int ntt_pass_plan(const NttTuning &tu, int log_d, int skip_early, int skip_late, bool msg_first, int *layer0, int *k,
                  int *c, int *e);
// Kernel (B64_NTT_KERNEL_*) of one pass of that schedule, and the work items
// (blocks x threads) of its grid on a device of `cus` compute units.
int ntt_pass_kernel(const NttTuning &tu, int log_d, int layer0, int k, int c, int e, bool inverse);
uint64_t ntt_pass_workitems(int kernel, int log_d, int layer0, int k, int c, int e, uint32_t cus);
uint32_t ntt_device_cus();
// Batch-plan kernel of one job's pass (see kernels.hip); returns the B64_NTT_KERNEL_* it runs.
int ntt_multi_pass(const NttTuning &tu, int log_d, int layer0, int k, int c, int e, bool msg, MultiKind *kind,
                   uint64_t *per_block, uint32_t *blocks);
// One launch of `kind` over n_jobs jobs: d_jobs is a device array of the
// kind's args, d_start[j] the first block of job j, `blocks` the total.
// fold_nc: for FoldCodeword, the n_chal (1..4) every job shares, else 0.
hipError_t launch_multi(MultiKind kind, const void *d_jobs, const uint32_t *d_start, int n_jobs, uint32_t blocks,
                        hipStream_t s, int fold_nc = 0);
// 128 B128 elements passed by value to a kernel (2 KiB of arguments).
struct RsVec {
    uint4 v[128];
};
hipError_t launch_rs_fold_elems(uint4 *elems, uint64_t n, const RsVec &v, hipStream_t s);
// Blocks (and 128-element partial sums) of a fold_1b_rows launch over n rows.
int rs_fold_1b_rows_blocks(uint64_t n);
hipError_t launch_rs_fold_1b_rows(const uint4 *mat, const uint4 *vec, uint64_t n, uint4 *partial, uint4 *out,
                                  hipStream_t s);
hipError_t launch_pack_witness(const uint64_t *words, uint64_t n_words, uint4 *out, uint64_t n_elems,
                               hipStream_t s);

// A committed FRI oracle on the device: codeword 2^log_len elements, Merkle tree
// over 2^tree_log_len cosets.
struct FriOracleView {
    const uint4 *cw;
    const uint4 *tree;
    int log_len, tree_log_len;
};

// One opened oracle of a query phase on the device: a codeword of
// 2^(tree_log_len + log_coset) elements, its Merkle tree over 2^tree_log_len
// cosets, the depth of the layer sent whole (optimal_verify_layer), and the
// granules of one query's opening: 2^log_coset values, then two per branch digest.
struct QueryOracle {
    const uint4 *cw;
    const uint4 *tree;
    int tree_log_len, log_coset, depth;
    uint32_t n_gran;
};
constexpr int MAX_OPENED = B64_MAX_FOLD_ARITIES + 1;
// A job's query phase as the 16-byte granules of the tape it writes, in tape
// order: the terminal codeword (term_gran), the layer of every opened oracle
// (layers_gran), then per query index the opening of each opened oracle
// (per_query).  idx_at and out_at: offsets of the job's indices and of its
// granules in the arrays a launch shares.
struct QueryArgs {
    const uint4 *term;
    uint64_t term_gran, layers_gran, n_gran;
    uint64_t idx_at, out_at;
    uint32_t nq, per_query;
    int n_opened;
    QueryOracle o[MAX_OPENED];
};

// The query phase of FRIFoldProver::finish_proof (prover fri/fold.rs:221-244,
// query.rs:51-134), in steps so that many proofs can share device round trips:
//   prepare -- the oracles (the interleaved codeword, then every committed one);
//   sample  -- the query indices.  Decommitments are not observed, so every
//              index can be sampled before any of the phase's tape is written;
//   args    -- the device arguments of the tape (idx_at = out_at = 0).
// One launch_fri_query then writes the phase's whole tape, which is
// decommitted as one piece.
struct FriQuery {
    b64_fri_params p;
    FriOracleView term;
    std::vector<FriOracleView> opened;
    std::vector<int> depths;
    std::vector<uint64_t> idx;
    int prepare(const b64_fri_params &params, const FriOracleView &first, const FriOracleView *committed,
                size_t n_committed);
    int sample(b64_transcript *t);
    int args(QueryArgs *a) const;
};
// One launch writing the query-phase tapes of n_jobs jobs: d_jobs a device
// array of their arguments, d_start[j] the first block of job j (BLOCK granules
// per block), `blocks` the total; job j reads d_idx + idx_at and writes
// d_out + out_at.
hipError_t launch_fri_query(const QueryArgs *d_jobs, const uint32_t *d_start, int n_jobs, uint32_t blocks,
                            const uint64_t *d_idx, uint4 *d_out, hipStream_t s);
// Blocks of one job in that launch.
uint32_t fri_query_blocks(const QueryArgs &a);

// Batched proofs (b64_batch_prove) as a resumable sequence of phases: enqueue a
// phase (nothing waits), wait for the stream by any means, then step() does the
// host part (roots into transcripts, challenges, the next phase's tables) and
// reports whether a next phase exists; enqueue that one, then drain() decommits
// the completed phase's query tapes while the device runs it, and appends the
// jobs it completed to `finished` (indices into the batch, may be null).  Every
// proof is complete once step() reports no next phase and drain() has run.
// Buffers persist across batches.
struct ProveBuffers;
struct ProveBatch;
// reserve: bytes of stage tables (twice that of results and query tapes) to
// allocate up front, in stream order on s for the device part (0: none).
ProveBuffers *prove_buffers_create(size_t reserve, hipStream_t s);
void prove_buffers_destroy(ProveBuffers *b);
int prove_batch_create(const b64_prove_job *jobs, uint32_t n_jobs, ProveBuffers *buf, ProveBatch **out);
int prove_batch_enqueue(ProveBatch *pb, hipStream_t s);
int prove_batch_step(ProveBatch *pb, bool *more);
void prove_batch_drain(ProveBatch *pb, std::vector<uint32_t> *finished);
void prove_batch_destroy(ProveBatch *pb);

// Caller-owned buffers for a batch plan (the serving loop reuses one per stream
// slot, so building a plan neither allocates nor synchronizes the device).
struct BatchStorage {
    uint8_t *d_tables;  // device, tables_cap bytes
    uint8_t *h_stage;   // coherent pinned host staging for the table upload (a kernel reads it), tables_cap bytes
    size_t tables_cap;
    uint8_t *h_roots;   // coherent pinned host (hipHostMallocCoherent), roots_cap digests: kernels write the roots there
    size_t roots_cap;
};

}  // namespace b64

// Batch plan over caller storage (st may be null: the plan allocates its own);
// tables are uploaded on upload_stream.
int b64_batch_plan_build(const b64_job_desc *jobs, uint32_t n_jobs, int use_graph, b64::BatchStorage *st,
                         hipStream_t upload_stream, b64_batch_plan **out);

// Vision Merkle tree (leaf_m = 4 or 6 leaves, Vision-4 compression nodes) into
// tree; a b64 status code.
namespace b64 {
int vision_merkle_commit(int leaf_m, const uint4 *elems, uint64_t n_elems, uint64_t batch, uint4 *tree,
                         hipStream_t s);
}

// Device view of a domain (uploads its twiddle basis on first use).
int b64_domain_device_view_internal(b64_domain *d, b64::DeviceDomain *dd);
