// Evaluation-proof side of the polynomial commitment on the device:
//  * eq_ind_partial_eval (verifier/math/src/multilinear/eq.rs:123-131)
//  * BivariateProductSumcheckProver (prover/prover/src/protocols/sumcheck/bivariate_product.rs)
//  * BaseFoldProver (prover/prover/src/protocols/basefold.rs), driving the
//    sumcheck and the FRI fold prover round by round on one transcript
//  * OneBitPCSProver::prove (prover/prover/src/pcs.rs:107-170): the ring-switch
//    reduction to a B128 sumcheck claim, then BaseFold.
// The multilinears stay in HBM from the first round to the last; per round the
// host reads back two field elements (the round evaluations), as the
// Fiat-Shamir transcript needs them before the next challenge exists.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <string.h>

#include <new>
#include <vector>

#include "abi_util.hpp"
#include "gf128_dev.hpp"
#include "internal.hpp"

using b64::abi::as_stream;
using b64::abi::from_u4;
using b64::abi::ld128;
using b64::abi::st128;
using b64::abi::to_u4;
using b64::host::u128;

namespace b64 {
namespace {

constexpr int SC_BLOCK = 256;
constexpr int SC_WAVES = SC_BLOCK / 64;
constexpr int SC_MAX_BLOCKS = 2048;

__device__ __forceinline__ uint4 wave_xor(uint4 v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        v.x ^= __shfl_xor(v.x, m);
        v.y ^= __shfl_xor(v.y, m);
        v.z ^= __shfl_xor(v.z, m);
        v.w ^= __shfl_xor(v.w, m);
    }
    return v;
}

// XOR of (p, q) over the block; thread 0 gets the result.
__device__ __forceinline__ void block_xor2(uint4 &p, uint4 &q) {
    __shared__ uint4 red[2][SC_WAVES];
    p = wave_xor(p);
    q = wave_xor(q);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = p;
        red[1][w] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < SC_WAVES; k++) {
            p = gxor(p, red[0][k]);
            q = gxor(q, red[1][k]);
        }
    }
}

// One sumcheck round of the bivariate product A*B over `half` pairs
// (bivariate_product.rs:43-82): y1 = sum A(v||1) B(v||1), yinf = sum
// (A(v||0)+A(v||1)) (B(v||0)+B(v||1)), v over the low variables.  With fold set,
// the previous round's fold_highest_var_inplace (multilinear/fold.rs:14-29,
// zero += r (one - zero)) is applied first, in place, to arrays of 4*half
// elements: every lane reads the four elements whose folds it needs and writes
// the two it owns, so the in-place update is race-free.  The multiplies by r go
// through a window table of r; the products A*B are general products.
// Each block writes its two partial sums; k_sc_reduce adds them up.  (Letting
// the last block to finish do the reduction costs more than the launch it
// saves: each block's device-scope release then writes back the L2 the round's
// folded arrays are dirty in.)
// Announces a round's totals, already stored in host-visible memory: the
// stores are made visible system-wide before the round's sequence number is.
__device__ __forceinline__ void publish_seq(uint32_t *flag, uint32_t seq) {
    __threadfence_system();
    *(volatile uint32_t *)flag = seq;
}

__global__ __launch_bounds__(SC_BLOCK) void k_sc_round(uint4 *__restrict__ a, uint4 *__restrict__ b, uint64_t half,
                                                       int fold, uint4 r, uint4 *__restrict__ partial,
                                                       uint32_t *flag, uint32_t seq) {
    __shared__ __attribute__((aligned(256))) uint4 tab[WIN_ENTRIES];
    if (fold) {
        if (threadIdx.x < WIN_ROWS) win_build_row(tab, r, threadIdx.x);
        __syncthreads();
    }
    const uint32_t tb = lds_addr(tab);
    const uint64_t len = 2 * half;
    uint4 y1 = make_uint4(0, 0, 0, 0), yi = make_uint4(0, 0, 0, 0);
    for (uint64_t i = (uint64_t)blockIdx.x * SC_BLOCK + threadIdx.x; i < half; i += (uint64_t)gridDim.x * SC_BLOCK) {
        uint4 a0, a1, b0, b1;
        if (fold) {
            const uint4 x0 = a[i], x1 = a[i + half], x2 = a[i + len], x3 = a[i + len + half];
            const uint4 z0 = b[i], z1 = b[i + half], z2 = b[i + len], z3 = b[i + len + half];
            a0 = win_mul_perm(tb, gxor(x0, x2), x0);
            a1 = win_mul_perm(tb, gxor(x1, x3), x1);
            b0 = win_mul_perm(tb, gxor(z0, z2), z0);
            b1 = win_mul_perm(tb, gxor(z1, z3), z1);
            a[i] = a0;
            a[i + half] = a1;
            b[i] = b0;
            b[i + half] = b1;
        } else {
            a0 = a[i];
            a1 = a[i + half];
            b0 = b[i];
            b1 = b[i + half];
        }
        y1 = gxor(y1, gf_mul(a1, b1));
        yi = gxor(yi, gf_mul(gxor(a0, a1), gxor(b0, b1)));
    }
    block_xor2(y1, yi);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = y1;
        partial[2 * blockIdx.x + 1] = yi;
        if (flag) publish_seq(flag, seq);  // a one-block round writes the totals itself
    }
}

// The round's totals: XOR of the n_blocks partial pairs, written to `result`,
// host-visible pinned memory read after the stream synchronises.
__global__ __launch_bounds__(SC_BLOCK) void k_sc_reduce(const uint4 *__restrict__ partial, int n_blocks,
                                                        uint4 *__restrict__ result, uint32_t *flag, uint32_t seq) {
    uint4 y1 = make_uint4(0, 0, 0, 0), yi = make_uint4(0, 0, 0, 0);
    for (int k = threadIdx.x; k < n_blocks; k += SC_BLOCK) {
        y1 = gxor(y1, partial[2 * k]);
        yi = gxor(yi, partial[2 * k + 1]);
    }
    block_xor2(y1, yi);
    if (threadIdx.x == 0) {
        result[0] = y1;
        result[1] = yi;
        publish_seq(flag, seq);
    }
}

// finish (bivariate_product.rs:100-121): the pending fold of both length-2
// multilinears, if any, and their single values written to host-visible memory.
__global__ void k_sc_final(const uint4 *__restrict__ a, const uint4 *__restrict__ b, int fold, uint4 r,
                           uint4 *__restrict__ result) {
    if (threadIdx.x != 0) return;
    uint4 x = a[0], z = b[0];
    if (fold) {
        x = gxor(x, gf_mul(r, gxor(x, a[1])));
        z = gxor(z, gf_mul(r, gxor(z, b[1])));
    }
    result[0] = x;
    result[1] = z;
}

// ---- eq_ind_partial_eval -----------------------------------------------------
// eq(r)[x] = prod_k (bit k of x ? r_k : 1 + r_k).  The coordinates are cut into
// groups of 8; group g's table holds the eq values of its coordinates (256
// entries, built by doubling as tensor_prod_eq_ind does, eq.rs:39-71).  Then
// eq[x] = G0[x & 255] * c(x >> 8) with c(y) = prod_{g>=1} G_g[byte g-1 of y]:
// each run of 256 outputs shares c(y), so the run is one window table of c(y)
// applied to G0.  One small launch builds the group tables (k_eq_groups); then
// every wave of k_eq_runs takes a range of runs with a table of its own, so no
// wave waits for another.
constexpr int EQ_MAX_VARS = 40;
constexpr int EQ_GROUPS = (EQ_MAX_VARS + 7) / 8;
constexpr int EQ_MAX_RUNS_PER_WAVE = 32;
constexpr int EQ_TARGET_WAVES = 2048;  // 8 waves on each of 256 CUs
constexpr int EQ_MAX_BLOCKS = 16384;
struct EqPoint {
    uint4 r[EQ_MAX_VARS];
};

// Orders one wave's LDS accesses across its lanes: a wave's LDS instructions
// execute in issue order, so keeping the compiler from moving accesses across
// this point is enough.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint4 readlane4(uint4 v, int lane) {
    return make_uint4(__builtin_amdgcn_readlane(v.x, lane), __builtin_amdgcn_readlane(v.y, lane),
                      __builtin_amdgcn_readlane(v.z, lane), __builtin_amdgcn_readlane(v.w, lane));
}

// One block: wave w doubles groups w, w + SC_WAVES, ... in LDS and writes them
// to gtab[g * 256 ..].  With n <= 8 the single group is the whole tensor.
__global__ __launch_bounds__(SC_BLOCK) void k_eq_groups(EqPoint p, int n, uint4 *__restrict__ gtab) {
    __shared__ uint4 g[EQ_GROUPS][256];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int n_groups = n ? (n + 7) / 8 : 1;  // n = 0: the single entry 1
    for (int gi = w; gi < n_groups; gi += SC_WAVES) {
        const int len = (n - 8 * gi) < 8 ? (n - 8 * gi) : 8;
        uint4 *G = g[gi];
        if (lane == 0) G[0] = make_uint4(1, 0, 0, 0);
        wave_lds_sync();
        for (int k = 0; k < len; k++) {
            const uint4 r = p.r[8 * gi + k];
            for (int i = lane; i < (1 << k); i += 64) {
                const uint4 prod = gf_mul(G[i], r);
                G[i + (1 << k)] = prod;
                G[i] = gxor(G[i], prod);
            }
            wave_lds_sync();
        }
        const int size = 1 << len;
        for (int i = lane; i < size; i += 64) gtab[256 * gi + i] = G[i];
    }
}

// Runs [gw * rpw, (gw + 1) * rpw) for each wave gw; one wave per block.
__global__ __launch_bounds__(64) void k_eq_runs(const uint4 *__restrict__ gtab, int n, int rpw,
                                                uint4 *__restrict__ out) {
    __shared__ __attribute__((aligned(256))) uint4 wtab[WIN_ENTRIES];
    const int lane = threadIdx.x;
    const int n_groups = (n + 7) / 8;
    uint4 lo[4];
#pragma unroll
    for (int k = 0; k < 4; k++) lo[k] = gtab[lane + 64 * k];
    const uint64_t n_runs = 1ull << (n - 8);
    const uint64_t n_waves = n_runs / (uint64_t)rpw;
    const uint32_t tb = lds_addr(wtab);
    for (uint64_t gw = blockIdx.x; gw < n_waves; gw += gridDim.x) {
        const uint64_t y0 = gw * (uint64_t)rpw;
        // lane j (mod rpw) holds c(y0 + j)
        const uint64_t y = y0 + (uint64_t)(lane & (rpw - 1));
        uint4 c = gtab[256 + (y & 255)];
        for (int gi = 2; gi < n_groups; gi++) c = gf_mul(c, gtab[256 * gi + ((y >> (8 * (gi - 1))) & 255)]);
        for (int j = 0; j < rpw; j++) {
            if (lane < WIN_ROWS) win_build_row(wtab, readlane4(c, j), lane);
            wave_lds_sync();
            uint4 *o = out + ((y0 + j) << 8);
#pragma unroll
            for (int k = 0; k < 4; k++) o[lane + 64 * k] = win_mul_perm(tb, lo[k]);
            wave_lds_sync();  // every lane's lookups precede the next run's table
        }
    }
}

// gtab: EQ_GROUPS * 256 elements of device scratch (unused for n <= 8).
hipError_t launch_eq_expand(const EqPoint &p, int n, uint4 *gtab, uint4 *out, hipStream_t s) {
    if (n <= 8) {
        k_eq_groups<<<1, SC_BLOCK, 0, s>>>(p, n, out);
        return hipGetLastError();
    }
    k_eq_groups<<<1, SC_BLOCK, 0, s>>>(p, n, gtab);
    const uint64_t n_runs = 1ull << (n - 8);
    int rpw = 1;
    while (rpw < EQ_MAX_RUNS_PER_WAVE && n_runs / (uint64_t)(2 * rpw) >= (uint64_t)EQ_TARGET_WAVES) rpw *= 2;
    uint64_t blocks = n_runs / (uint64_t)rpw;
    if (blocks > EQ_MAX_BLOCKS) blocks = EQ_MAX_BLOCKS;
    k_eq_runs<<<(unsigned)blocks, 64, 0, s>>>(gtab, n, rpw, out);
    return hipGetLastError();
}

int sc_blocks(uint64_t half) {
    const uint64_t b = (half + SC_BLOCK - 1) / SC_BLOCK;
    return (int)(b < SC_MAX_BLOCKS ? (b ? b : 1) : SC_MAX_BLOCKS);
}

// 128 x 128 bit transpose: out[i] bit j = in[j] bit i
// (TensorAlgebra<B1, B128>::transpose, verifier/math/src/tensor_algebra.rs:137-141).
void transpose128(const u128 *in, u128 *out) {
    for (int i = 0; i < 128; i++) {
        u128 v = 0;
        for (int j = 0; j < 128; j++) v |= ((in[j] >> i) & 1) << j;
        out[i] = v;
    }
}

// eq_ind_partial_eval on the host for a handful of coordinates.
void eq_host(const u128 *r, int n, u128 *out) {
    out[0] = 1;
    for (int k = 0; k < n; k++)
        for (size_t j = 0; j < ((size_t)1 << k); j++) {
            const u128 prod = host::gf_mul(out[j], r[k]);
            out[j + ((size_t)1 << k)] = prod;
            out[j] ^= prod;
        }
}

}  // namespace
}  // namespace b64

namespace b64 {

// Waits for a round's totals: polls the sequence number the last kernel
// publishes in host memory after them (visible a few microseconds after the
// kernel ends, where a blocking stream wait takes tens).  A launch that failed
// never publishes; past the polling bound the stream wait takes over and
// reports the error.
int wait_published(volatile uint32_t *flag, uint32_t seq, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; *flag != seq; i++) {
        if ((i & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(100)) {
            HIP_TRY(hipStreamSynchronize(s));
            if (*flag != seq) return B64_ERR_HIP;
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);  // the totals were stored before the number
    return B64_OK;
}

}  // namespace b64

// ===========================================================================
// BivariateProductSumcheckProver
// ===========================================================================
struct b64_sumcheck {
    uint4 *a = nullptr, *b = nullptr;  // caller's buffers, folded in place
    uint32_t n_vars = 0;
    bool has_coeffs = false;  // RoundCoeffsOrSum::Coeffs
    bool pending = false;     // a fold by r_pending not yet applied on the device
    u128 sum = 0, r_pending = 0;
    u128 coeffs[3] = {0, 0, 0};
    uint4 *d_partial = nullptr;  // 2 * SC_MAX_BLOCKS partial sums
    uint4 *h_res = nullptr;      // pinned, 2 elements, written by the kernels
    uint4 *h_res_dev = nullptr;  // h_res as the device addresses it
    uint32_t seq = 0;            // round sequence number, published after the totals
    hipStream_t s = nullptr;

    volatile uint32_t *flag() const { return (volatile uint32_t *)(h_res + 2); }
    uint32_t *flag_dev() const { return (uint32_t *)(h_res_dev + 2); }

    ~b64_sumcheck() {
        if (d_partial) (void)hipFreeAsync(d_partial, s);
        if (h_res) {
            (void)hipStreamSynchronize(s);  // the kernels write h_res
            b64::abi::pinned_small_put(h_res);
        }
    }
};

extern "C" {

int b64_sumcheck_create(void *d_a, void *d_b, uint32_t n_vars, const uint64_t sum[2], void *stream,
                        b64_sumcheck **out) {
    if (!d_a || !d_b || !sum || !out || n_vars > 40) return B64_ERR_INVALID_ARGS;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return B64_ERR_NO_DEVICE;
    b64_sumcheck *sc = new (std::nothrow) b64_sumcheck();
    if (!sc) return B64_ERR_OUT_OF_MEMORY;
    sc->a = (uint4 *)d_a;
    sc->b = (uint4 *)d_b;
    sc->n_vars = n_vars;
    sc->sum = ld128(sum);
    sc->s = as_stream(stream);
    b64::abi::keep_pool_memory();
    if (hipMallocAsync((void **)&sc->d_partial, 2 * b64::SC_MAX_BLOCKS * sizeof(uint4), sc->s) != hipSuccess ||
        !(sc->h_res = (uint4 *)b64::abi::pinned_small_get()) ||
        hipHostGetDevicePointer((void **)&sc->h_res_dev, sc->h_res, 0) != hipSuccess) {
        delete sc;
        return B64_ERR_OUT_OF_MEMORY;
    }
    *sc->flag() = 0;
    *out = sc;
    return B64_OK;
}

void b64_sumcheck_destroy(b64_sumcheck *sc) { delete sc; }

uint32_t b64_sumcheck_n_vars(const b64_sumcheck *sc) { return sc ? sc->n_vars : 0; }

int b64_sumcheck_execute(b64_sumcheck *sc, uint64_t coeffs_out[6]) {
    if (!sc || !coeffs_out) return B64_ERR_INVALID_ARGS;
    if (sc->has_coeffs) return B64_ERR_EXPECTED_FOLD;
    if (sc->n_vars == 0) return B64_ERR_INVALID_ARGS;  // assert!(n_vars_remaining > 0)
    const uint64_t half = 1ull << (sc->n_vars - 1);
    const int blocks = b64::sc_blocks(half);
    // a one-block round writes its totals straight to the pinned result
    const uint32_t seq = ++sc->seq;
    if (blocks == 1) {
        b64::k_sc_round<<<1, b64::SC_BLOCK, 0, sc->s>>>(sc->a, sc->b, half, sc->pending ? 1 : 0,
                                                          to_u4(sc->r_pending), sc->h_res_dev, sc->flag_dev(), seq);
    } else {
        b64::k_sc_round<<<blocks, b64::SC_BLOCK, 0, sc->s>>>(sc->a, sc->b, half, sc->pending ? 1 : 0,
                                                               to_u4(sc->r_pending), sc->d_partial, nullptr, 0);
        b64::k_sc_reduce<<<1, b64::SC_BLOCK, 0, sc->s>>>(sc->d_partial, blocks, sc->h_res_dev, sc->flag_dev(), seq);
    }
    HIP_TRY(hipGetLastError());
    B64_TRY(b64::wait_published(sc->flag(), seq, sc->s));
    sc->pending = false;
    // RoundEvals2::interpolate (round_evals.rs:73-78, :126-140)
    const u128 y1 = from_u4(sc->h_res[0]), yinf = from_u4(sc->h_res[1]);
    const u128 y0 = sc->sum ^ y1;
    sc->coeffs[0] = y0;
    sc->coeffs[2] = yinf;
    sc->coeffs[1] = y1 ^ y0 ^ yinf;
    sc->has_coeffs = true;
    for (int k = 0; k < 3; k++) st128(coeffs_out + 2 * k, sc->coeffs[k]);
    return B64_OK;
}

int b64_sumcheck_fold(b64_sumcheck *sc, const uint64_t challenge[2]) {
    if (!sc || !challenge) return B64_ERR_INVALID_ARGS;
    if (!sc->has_coeffs) return B64_ERR_EXPECTED_EXECUTE;
    const u128 r = ld128(challenge);
    // RoundCoeffs::evaluate (Horner)
    u128 acc = 0;
    for (int k = 2; k >= 0; k--) acc = b64::host::gf_mul(acc, r) ^ sc->coeffs[k];
    sc->sum = acc;
    sc->n_vars--;
    sc->pending = true;  // applied by the next execute, fused with its reduction, or by finish
    sc->r_pending = r;
    sc->has_coeffs = false;
    return B64_OK;
}

int b64_sumcheck_finish(b64_sumcheck *sc, uint64_t evals_out[4]) {
    if (!sc || !evals_out) return B64_ERR_INVALID_ARGS;
    if (sc->n_vars > 0) return sc->has_coeffs ? B64_ERR_EXPECTED_FOLD : B64_ERR_EXPECTED_EXECUTE;
    if (sc->has_coeffs) return B64_ERR_EXPECTED_FOLD;
    b64::k_sc_final<<<1, 64, 0, sc->s>>>(sc->a, sc->b, sc->pending ? 1 : 0, to_u4(sc->r_pending), sc->h_res_dev);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(sc->s));
    sc->pending = false;
    st128(evals_out, from_u4(sc->h_res[0]));
    st128(evals_out + 2, from_u4(sc->h_res[1]));
    return B64_OK;
}

// ===========================================================================
// eq_ind_partial_eval
// ===========================================================================
int b64_eq_ind_partial_eval(const uint64_t *point, uint32_t n, void *d_out, void *stream) {
    if ((!point && n) || !d_out || n > (uint32_t)b64::EQ_MAX_VARS) return B64_ERR_INVALID_ARGS;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return B64_ERR_NO_DEVICE;
    b64::EqPoint p;
    memset(&p, 0, sizeof(p));
    for (uint32_t k = 0; k < n; k++) p.r[k] = to_u4(ld128(point + 2 * k));
    const hipStream_t s = as_stream(stream);
    uint4 *gtab = nullptr;
    if (n > 8) {
        b64::abi::keep_pool_memory();
        HIP_TRY(hipMallocAsync((void **)&gtab, b64::EQ_GROUPS * 256 * sizeof(uint4), s));
    }
    const hipError_t e = b64::launch_eq_expand(p, (int)n, gtab, (uint4 *)d_out, s);
    if (gtab) (void)hipFreeAsync(gtab, s);
    HIP_TRY(e);
    return B64_OK;
}

// ===========================================================================
// BaseFoldProver::prove (protocols/basefold.rs:131-166)
// ===========================================================================
int b64_basefold_prove(const b64_fri_params *p, const b64_domain *d, void *d_multilinear, void *d_transparent,
                       const uint64_t claim[2], const void *d_codeword, const void *d_tree, b64_transcript *t,
                       void *stream) {
    return b64_basefold_prove_ex(p, d, B64_MERKLE_SHA256, d_multilinear, d_transparent, claim, d_codeword, d_tree, t,
                                 stream);
}

int b64_basefold_prove_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash, void *d_multilinear,
                          void *d_transparent, const uint64_t claim[2], const void *d_codeword, const void *d_tree,
                          b64_transcript *t, void *stream) {
    if (!p || !d || !d_multilinear || !d_transparent || !claim || !d_codeword || !d_tree || !t)
        return B64_ERR_INVALID_ARGS;
    const uint32_t n_vars = p->log_dim + p->log_batch_size;
    b64_sumcheck *sc = nullptr;
    B64_TRY(b64_sumcheck_create(d_multilinear, d_transparent, n_vars, claim, stream, &sc));
    b64_fri_prover *fp = nullptr;
    int rc = b64_fri_prover_create_ex(p, d, merkle_hash, d_codeword, d_tree, stream, &fp);
    if (rc != B64_OK) {
        b64_sumcheck_destroy(sc);
        return rc;
    }
    if (b64_fri_prover_n_rounds(fp) != n_vars) rc = B64_ERR_INVALID_ARGS;
    for (uint32_t i = 0; rc == B64_OK && i < n_vars; i++) {
        // execute (:99-113): the sumcheck round and the FRI fold round.  The two
        // are independent device work on one stream: the fold round is
        // enqueued first and the sumcheck's wait for its coefficients also
        // covers the fold round's root.
        int committed = 0;
        const uint8_t *root = nullptr;
        rc = b64::fri_fold_round_enqueue(fp, &committed, &root);
        if (rc != B64_OK) break;
        uint64_t co[6];
        rc = b64_sumcheck_execute(sc, co);
        if (rc != B64_OK) break;
        b64_transcript_message(t, (const uint8_t *)co, 32);  // RoundCoeffs::truncate: c0, c1
        if (committed) b64_transcript_message(t, root, 32);
        uint64_t r[2];
        b64_transcript_sample(t, (uint8_t *)r, 16);
        // fold (:119-123)
        rc = b64_sumcheck_fold(sc, r);
        if (rc == B64_OK) rc = b64_fri_prover_receive_challenge(fp, r);
    }
    if (rc == B64_OK) {
        // finish (:158-166)
        uint8_t root[32];
        int committed = 0;
        rc = b64_fri_prover_execute_fold_round(fp, root, &committed);
        if (rc == B64_OK && committed) b64_transcript_message(t, root, 32);
        if (rc == B64_OK) rc = b64_fri_prover_finish_proof(fp, t);
    }
    b64_fri_prover_destroy(fp);
    b64_sumcheck_destroy(sc);
    return rc;
}

// ===========================================================================
// OneBitPCSProver::prove (prover/prover/src/pcs.rs:107-170)
// ===========================================================================
int b64_pcs_prove(const b64_fri_params *p, const b64_domain *d, const void *d_codeword, const void *d_tree,
                  void *d_packed, const uint64_t *eval_point, uint32_t n_point, b64_transcript *t, void *stream) {
    return b64_pcs_prove_ex(p, d, B64_MERKLE_SHA256, d_codeword, d_tree, d_packed, eval_point, n_point, t, stream);
}

int b64_pcs_prove_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash, const void *d_codeword,
                     const void *d_tree, void *d_packed, const uint64_t *eval_point, uint32_t n_point,
                     b64_transcript *t, void *stream) {
    if (!p || !d || !d_codeword || !d_tree || !d_packed || !eval_point || !t) return B64_ERR_INVALID_ARGS;
    const uint32_t log_len = p->log_dim + p->log_batch_size;
    constexpr uint32_t KAPPA = 7;  // log2 of B128's degree over B1
    if (n_point != log_len + KAPPA || log_len > (uint32_t)b64::EQ_MAX_VARS) return B64_ERR_INVALID_ARGS;
    const hipStream_t s = as_stream(stream);
    void *d_suffix = nullptr, *d_shv = nullptr;
    b64::abi::keep_pool_memory();
    HIP_TRY(hipMallocAsync(&d_suffix, (16ull << log_len) + 128 * 16, s));
    d_shv = (uint8_t *)d_suffix + (16ull << log_len);
    int rc = B64_OK;
    do {
        // suffix tensor eq(eval_point[7..]) and the partial evaluations s_hat_v
        rc = b64_eq_ind_partial_eval(eval_point + 2 * KAPPA, log_len, d_suffix, stream);
        if (rc != B64_OK) break;
        rc = b64_ring_switch_fold_1b_rows(d_packed, d_suffix, log_len, d_shv, stream);
        if (rc != B64_OK) break;
        u128 s_hat_v[128], s_hat_u[128], rpp[KAPPA], eq_rpp[128];
        if (hipMemcpyAsync(s_hat_v, d_shv, sizeof(s_hat_v), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = B64_ERR_HIP;
            break;
        }
        b64_transcript_message(t, (const uint8_t *)s_hat_v, sizeof(s_hat_v));
        b64::transpose128(s_hat_v, s_hat_u);
        // row-batching challenges
        b64_transcript_sample(t, (uint8_t *)rpp, sizeof(rpp));
        b64::eq_host(rpp, KAPPA, eq_rpp);
        u128 claim = 0;
        for (int i = 0; i < 128; i++) claim ^= b64::host::gf_mul(s_hat_u[i], eq_rpp[i]);
        // rs_eq_ind = fold_b128_elems_inplace(suffix tensor, eq(r''))
        uint64_t vec[256];
        for (int i = 0; i < 128; i++) st128(vec + 2 * i, eq_rpp[i]);
        rc = b64_ring_switch_fold_elems(d_suffix, log_len, vec, stream);
        if (rc != B64_OK) break;
        uint64_t cl[2];
        st128(cl, claim);
        rc = b64_basefold_prove_ex(p, d, merkle_hash, d_packed, d_suffix, cl, d_codeword, d_tree, t, stream);
    } while (false);
    (void)hipFreeAsync(d_suffix, s);
    return rc;
}

}  // extern "C"
