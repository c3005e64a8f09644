// Internal interfaces shared by the kernel launchers (kernels.hip) and the C ABI
// (api.hip).  Nothing here crosses the library boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
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
struct DeviceDomain {
    int dim;
    const uint4 *W;  // dim * W_STRIDE entries
};

// Sha256Compression IV (LE words of SHA-256("BINIUS SHA-256 COMPRESS")).
struct CompressIV {
    uint32_t w[8];
};

struct NttTuning {
    int variant;   // 0 = general product, 1 = LDS window tables
    int max_k;     // layers per pass
    int tile_log;  // log2 elements per LDS tile
};
extern NttTuning g_ntt_tuning;

// Kernel launchers.  All return hipError_t of the launch.
hipError_t launch_gf_mul(const uint4 *a, const uint4 *b, uint4 *out, uint64_t n, hipStream_t s);
hipError_t launch_gf_mul_const(const uint4 *a, uint4 c, uint4 *out, uint64_t n, hipStream_t s);
hipError_t launch_bitrev_replicate(const uint4 *msg, uint4 *out, int m, int lir, hipStream_t s);
hipError_t launch_ntt(const DeviceDomain &dd, uint4 *data, int log_d, int skip_early, int skip_late,
                      bool inverse, hipStream_t s, const uint4 *msg = nullptr, int msg_log = 0);
hipError_t launch_leaf_hash(const uint4 *elems, uint64_t n_leaves, int log_batch, uint64_t batch,
                            uint4 *digests, hipStream_t s);
hipError_t launch_merkle_levels(uint4 *tree, uint64_t n_leaves, const CompressIV &iv, hipStream_t s);
// 16-element leaves; with folded != nullptr also fold_interleaved with 4 challenges.
hipError_t launch_leaf16_fold(const uint4 *cw, uint64_t n_leaves, uint4 *digests, uint4 *folded, const uint4 *chal,
                              hipStream_t s);
hipError_t launch_fold_interleaved(const uint4 *cw, uint4 *out, uint64_t n_out, int log_batch,
                                   const uint4 *eq_host, hipStream_t s);
hipError_t launch_fold_codeword(const DeviceDomain &dd, const uint4 *cw, uint4 *out, int log_len,
                                const uint4 *chal_host, int n_chal, hipStream_t s);
hipError_t launch_pack_witness(const uint64_t *words, uint64_t n_words, uint4 *out, uint64_t n_elems,
                               hipStream_t s);

}  // namespace b64
