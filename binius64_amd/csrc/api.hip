// C ABI of libb64hip (include/b64hip.h): domain setup, NTT / RS encode, Merkle
// commit, FRI folds, FRI parameters, the Fiat-Shamir transcript and the FRI
// fold / query prover.  Host orchestration is C++; all bulk work runs in the
// gfx950 kernels of kernels.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "abi_util.hpp"
#include "host_util.hpp"
#include "internal.hpp"

using b64::host::u128;

#define B64_VERSION "0.1.0"

using b64::abi::as_stream;
using b64::abi::ld128;
using b64::abi::log2_ceil_u64;
using b64::abi::st128;
using b64::abi::to_u4;

namespace {

int device_ok() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    return 1;
}

b64::CompressIV compress_iv() {
    static b64::CompressIV iv;
    static std::once_flag once;
    std::call_once(once, [] {
        static const char tag[] = "BINIUS SHA-256 COMPRESS";
        uint8_t d[32];
        b64::host::Sha256::digest((const uint8_t *)tag, sizeof(tag) - 1, d);
        // compress.rs:37-45: the digest bytes reinterpreted as native (LE) u32 words
        for (int k = 0; k < 8; k++)
            iv.w[k] = (uint32_t)d[4 * k] | ((uint32_t)d[4 * k + 1] << 8) | ((uint32_t)d[4 * k + 2] << 16) |
                      ((uint32_t)d[4 * k + 3] << 24);
    });
    return iv;
}

}  // namespace

// ===========================================================================
// Domain
// ===========================================================================
struct b64_domain {
    int dim = 0;
    std::vector<u128> evals;  // evals[i*dim + k], k < dim - i
    std::vector<uint4> W;     // dim * W_STRIDE
    uint4 *d_W = nullptr;
    uint4 *d_X = nullptr;     // pre-expanded twiddles of layers < x_layers (2^x_layers entries)
    int x_layers = 0;
    std::mutex mu;

    u128 eval(int i, int k) const { return evals[(size_t)i * dim + k]; }
    // twiddle(layer, block) = subset sum of evals[dim-1-layer][1..] (domain_context.rs:91-94,149-151)
    u128 twiddle(int layer, uint64_t block) const {
        u128 t = 0;
        for (int b = 0; b < layer; b++)
            if ((block >> b) & 1) t ^= eval(dim - 1 - layer, 1 + b);
        return t;
    }
    int device_view(b64::DeviceDomain &dd) {
        std::lock_guard<std::mutex> g(mu);
        if (!d_W) {
            if (!device_ok()) return B64_ERR_NO_DEVICE;
            // the basis, and the pre-expanded low layers read by the LDS-tile
            // passes over the last layers (a twiddle per 16..128 butterflies there)
            const int xl = dim < b64::X_MAX_LAYERS ? dim : b64::X_MAX_LAYERS;
            uint4 *w = nullptr, *x = nullptr;
            hipError_t e = hipMalloc(&w, W.size() * sizeof(uint4));
            if (e == hipSuccess) e = hipMalloc(&x, ((size_t)1 << xl) * sizeof(uint4));
            if (e == hipSuccess) e = hipMemcpy(w, W.data(), W.size() * sizeof(uint4), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = b64::launch_expand_twiddles(w, x, xl, nullptr);
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
            if (e != hipSuccess) {
                if (w) (void)hipFree(w);
                if (x) (void)hipFree(x);
                return b64::abi::hip_status(e);
            }
            d_W = w;
            d_X = x;
            x_layers = xl;
        }
        dd.dim = dim;
        dd.W = d_W;
        dd.X = d_X;
        dd.x_layers = x_layers;
        return B64_OK;
    }
};

int b64_domain_device_view_internal(b64_domain *d, b64::DeviceDomain *dd) {
    if (!d || !dd) return B64_ERR_INVALID_ARGS;
    return d->device_view(*dd);
}

extern "C" {

const char *b64_version(void) { return B64_VERSION; }

const char *b64_status_string(int s) {
    switch (s) {
        case B64_OK: return "ok";
        case B64_ERR_INVALID_ARGS: return "invalid arguments";
        case B64_ERR_ENCODER_SUBSPACE_MISMATCH: return "the evaluation domain of the code does not match the subspace of the NTT encoder";
        case B64_ERR_SUBSPACE_DIMENSION_MISMATCH: return "the dimension of the evaluation domain of the code does not match the parameters";
        case B64_ERR_INCORRECT_BATCH_SIZE: return "incorrect batch size";
        case B64_ERR_POWER_OF_TWO_LENGTH_REQUIRED: return "power of two length required";
        case B64_ERR_EARLY_PROVER_FINISH: return "prover finished early";
        case B64_ERR_INCORRECT_LAYER_DEPTH: return "incorrect layer depth";
        case B64_ERR_INDEX_OUT_OF_RANGE: return "index out of range";
        case B64_ERR_INVALID_FOLD_ARITY_SEQUENCE: return "invalid fold arity sequence";
        case B64_ERR_BUFFER_TOO_SMALL: return "buffer too small";
        case B64_ERR_NOT_COMMIT_ROUND: return "not a commit round";
        case B64_ERR_EXPECTED_FOLD: return "expected a call to fold";
        case B64_ERR_EXPECTED_EXECUTE: return "expected a call to execute";
        case B64_ERR_NOT_ENOUGH_BYTES: return "not enough data in read buffer to deserialize";
        case B64_ERR_INVALID_CONSTRUCTION: return "invalid construction of Proof::version";
        case B64_ERR_FROM_UTF8: return "string is not valid UTF-8";
        case B64_ERR_USIZE_TOO_LARGE: return "length too large to serialize (max is u32::MAX)";
        case B64_ERR_NO_DEVICE: return "no HIP device";
        case B64_ERR_HIP: return "HIP runtime error";
        case B64_ERR_OUT_OF_MEMORY: return "out of device memory";
        case B64_ERR_ADMISSION_REJECTED: return "admission rejected";
        default: return "unknown status";
    }
}

int b64_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int b64_device_alloc(size_t bytes, void **d_ptr) {
    if (!d_ptr) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    HIP_TRY(hipMalloc(d_ptr, bytes ? bytes : 16));
    return B64_OK;
}

int b64_device_free(void *d_ptr) {
    if (!d_ptr) return B64_OK;
    HIP_TRY(hipFree(d_ptr));
    return B64_OK;
}

int b64_memcpy_htod(void *d_dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return B64_OK;
    HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return B64_OK;
}

int b64_memcpy_dtoh(void *dst, const void *d_src, size_t bytes, void *stream) {
    if (!bytes) return B64_OK;
    HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return B64_OK;
}

int b64_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream) {
    if (!bytes) return B64_OK;
    HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return B64_OK;
}

int b64_stream_synchronize(void *stream) {
    HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return B64_OK;
}

// "ntt.<field>" addresses the defaults, "ntt.<field>:<log_d>" the tuning of
// transforms of size 2^log_d (first set copies the defaults).
static bool parse_tuning_key(const char *key, char *field, size_t cap, int *log_d) {
    const char *colon = strchr(key, ':');
    const size_t n = colon ? (size_t)(colon - key) : strlen(key);
    if (n + 1 > cap) return false;
    memcpy(field, key, n);
    field[n] = 0;
    *log_d = -1;
    if (colon) {
        char *end = nullptr;
        const long v = strtol(colon + 1, &end, 10);
        if (end == colon + 1 || *end != 0 || v < 1 || v > b64::MAX_DIM) return false;
        *log_d = (int)v;
    }
    return true;
}

// Splits "ntt.<knob>" / "batch.<knob>" into the table it addresses (single-job
// or batch-plan schedules) and the knob name.
static bool tuning_table(const char *field, bool *batch, const char **knob) {
    if (!strncmp(field, "ntt.", 4)) {
        *batch = false;
        *knob = field + 4;
        return true;
    }
    if (!strncmp(field, "batch.", 6)) {
        *batch = true;
        *knob = field + 6;
        return true;
    }
    return false;
}

int b64_config_set(const char *key, int64_t value) {
    if (!key) return B64_ERR_INVALID_ARGS;
    if (!strcmp(key, "host.pclmul")) {
        b64::host::g_pclmul_allowed.store(value != 0);
        return B64_OK;
    }
    if (!strcmp(key, "ntt.clear_sizes")) {
        (void)b64::ntt_tuning_for(0);  // initialises the table
        for (auto &t : b64::g_ntt_tuning_size) t.variant = -1;
        return B64_OK;
    }
    if (!strcmp(key, "batch.clear_sizes")) {
        (void)b64::batch_tuning_for(0);
        for (auto &t : b64::g_batch_tuning_size) t.variant = -1;
        return B64_OK;
    }
    char field[32];
    int log_d;
    bool batch;
    const char *knob;
    if (!parse_tuning_key(key, field, sizeof field, &log_d) || !tuning_table(field, &batch, &knob))
        return B64_ERR_INVALID_ARGS;
    b64::NttTuning t = batch ? b64::batch_tuning_for(log_d) : b64::ntt_tuning_for(log_d);
    if (!strcmp(knob, "variant")) {
        if (value < 0 || value > 4 || (batch && value == 2)) return B64_ERR_INVALID_ARGS;
        t.variant = (int)value;
    } else if (!strcmp(knob, "max_k")) {
        if (value < 1 || value > 7) return B64_ERR_INVALID_ARGS;
        t.max_k = (int)value;
    } else if (!strcmp(knob, "tile_log")) {
        if (value < 1 || value > 11) return B64_ERR_INVALID_ARGS;
        t.tile_log = (int)value;
    } else if (!strcmp(knob, "tab_min")) {
        if (value < 1 || value > (1 << 30)) return B64_ERR_INVALID_ARGS;
        t.tab_min = (int)value;
    } else if (!strcmp(knob, "tab_group")) {
        if (value != 1 && value != 2 && value != 4 && value != 8) return B64_ERR_INVALID_ARGS;
        t.tab_group = (int)value;
    } else if (!strcmp(knob, "waves")) {
        if (value != 0 && value != 4) return B64_ERR_INVALID_ARGS;
        t.waves = (int)value;
    } else {
        return B64_ERR_INVALID_ARGS;
    }
    if (log_d < 0)
        (batch ? b64::g_batch_tuning : b64::g_ntt_tuning) = t;
    else
        (batch ? b64::g_batch_tuning_size : b64::g_ntt_tuning_size)[log_d] = t;
    return B64_OK;
}

int b64_config_get(const char *key, int64_t *value) {
    if (!key || !value) return B64_ERR_INVALID_ARGS;
    if (!strcmp(key, "host.pclmul")) {
        *value = b64::host::use_pclmul() ? 1 : 0;
        return B64_OK;
    }
    char field[32];
    int log_d;
    bool batch;
    const char *knob;
    if (!parse_tuning_key(key, field, sizeof field, &log_d) || !tuning_table(field, &batch, &knob))
        return B64_ERR_INVALID_ARGS;
    const b64::NttTuning &t = batch ? b64::batch_tuning_for(log_d) : b64::ntt_tuning_for(log_d);
    if (!strcmp(knob, "variant")) *value = t.variant;
    else if (!strcmp(knob, "max_k")) *value = t.max_k;
    else if (!strcmp(knob, "tile_log")) *value = t.tile_log;
    else if (!strcmp(knob, "tab_min")) *value = t.tab_min;
    else if (!strcmp(knob, "tab_group")) *value = t.tab_group;
    else if (!strcmp(knob, "waves")) *value = t.waves;
    else return B64_ERR_INVALID_ARGS;
    return B64_OK;
}

void b64_gf_mul(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]) {
    st128(out, b64::host::gf_mul(ld128(a), ld128(b)));
}

void b64_gf_inv(const uint64_t a[2], uint64_t out[2]) { st128(out, b64::host::gf_inv(ld128(a))); }

int b64_gf_mul_device(const void *d_a, const void *d_b, void *d_out, uint64_t n, void *stream) {
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    HIP_TRY(b64::launch_gf_mul((const uint4 *)d_a, (const uint4 *)d_b, (uint4 *)d_out, n, as_stream(stream)));
    return B64_OK;
}

int b64_gf_mul_const_device(const void *d_a, const uint64_t c[2], void *d_out, uint64_t n, void *stream) {
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    HIP_TRY(b64::launch_gf_mul_const((const uint4 *)d_a, to_u4(ld128(c)), (uint4 *)d_out, n, as_stream(stream)));
    return B64_OK;
}

// generate_evals_from_subspace (ntt/domain_context.rs:19-46); the W table holds
// each layer's twiddle basis (GenericPreExpanded's expansion basis, :114-132).
int b64_domain_create(const uint64_t *basis, uint32_t dim, b64_domain **out) {
    if (!out || dim > (uint32_t)b64::MAX_DIM || (dim && !basis)) return B64_ERR_INVALID_ARGS;
    b64_domain *d = new (std::nothrow) b64_domain();
    if (!d) return B64_ERR_OUT_OF_MEMORY;
    d->dim = (int)dim;
    d->evals.assign((size_t)dim * dim, 0);
    for (uint32_t k = 0; k < dim; k++) d->evals[k] = ld128(basis + 2 * k);
    for (int i = 1; i < (int)dim; i++) {
        const int prev_len = (int)dim - (i - 1);
        for (int k = 1; k < prev_len; k++) {
            const u128 pk = d->eval(i - 1, k), p0 = d->eval(i - 1, 0);
            d->evals[(size_t)i * dim + (k - 1)] = b64::host::gf_mul(pk, pk ^ p0);
        }
    }
    for (int i = 0; i < (int)dim; i++) {
        const u128 inv = b64::host::gf_inv(d->eval(i, 0));
        for (int k = 0; k < (int)dim - i; k++)
            d->evals[(size_t)i * dim + k] = b64::host::gf_mul(d->eval(i, k), inv);
    }
    d->W.assign((size_t)(dim ? dim : 1) * b64::W_STRIDE, make_uint4(0, 0, 0, 0));
    for (int layer = 0; layer < (int)dim; layer++)
        for (int b = 0; b < layer; b++)
            d->W[(size_t)layer * b64::W_STRIDE + b] = to_u4(d->eval((int)dim - 1 - layer, 1 + b));
    *out = d;
    return B64_OK;
}

int b64_domain_create_standard(uint32_t dim, b64_domain **out) {
    std::vector<uint64_t> basis(2 * (size_t)(dim ? dim : 1), 0);
    for (uint32_t i = 0; i < dim; i++) st128(&basis[2 * i], (u128)1 << i);
    return b64_domain_create(basis.data(), dim, out);
}

void b64_domain_destroy(b64_domain *d) {
    if (!d) return;
    if (d->d_W) (void)hipFree(d->d_W);
    if (d->d_X) (void)hipFree(d->d_X);
    delete d;
}

uint32_t b64_domain_log_size(const b64_domain *d) { return d ? (uint32_t)d->dim : 0; }

int b64_domain_subspace(const b64_domain *d, uint32_t i, uint64_t *basis_out) {
    if (!d || i > (uint32_t)d->dim) return B64_ERR_INVALID_ARGS;
    for (uint32_t k = 0; k < i; k++) st128(basis_out + 2 * k, d->eval(d->dim - (int)i, (int)k));
    return B64_OK;
}

int b64_domain_twiddle(const b64_domain *d, uint32_t layer, uint64_t block, uint64_t out[2]) {
    if (!d || layer >= (uint32_t)d->dim || (layer < 64 && (block >> layer) != 0)) return B64_ERR_INVALID_ARGS;
    st128(out, d->twiddle((int)layer, block));
    return B64_OK;
}

// ===========================================================================
// NTT / RS encode
// ===========================================================================
static int ntt_common(const b64_domain *dc, void *d_data, uint32_t log_len, uint32_t se, uint32_t sl,
                      bool inverse, void *stream) {
    b64_domain *d = const_cast<b64_domain *>(dc);
    if (!d || !d_data) return B64_ERR_INVALID_ARGS;
    // input_check (ntt/reference.rs:89-100)
    if (se + sl > log_len || log_len - sl > (uint32_t)d->dim || log_len > 40) return B64_ERR_INVALID_ARGS;
    b64::DeviceDomain dd;
    B64_TRY(d->device_view(dd));
    const hipError_t e =
        b64::launch_ntt(dd, (uint4 *)d_data, (int)log_len, (int)se, (int)sl, inverse, as_stream(stream));
    if (e == hipErrorInvalidConfiguration) return B64_ERR_INVALID_ARGS;  // a pass grid too large
    HIP_TRY(e);
    return B64_OK;
}

int b64_ntt_plan(uint32_t log_len, uint32_t se, uint32_t sl, int inverse, int encode, uint32_t n_cus,
                 b64_ntt_pass_info *passes, uint32_t cap, uint32_t *n_passes) {
    if (!n_passes || (cap && !passes) || se + sl > log_len || log_len > 40 || (encode && inverse))
        return B64_ERR_INVALID_ARGS;
    int starts[64], ks[64], cs[64], es[64];
    const b64::NttTuning &tu = b64::ntt_tuning_for((int)log_len);
    const int n = b64::ntt_pass_plan(tu, (int)log_len, (int)se, (int)sl, encode != 0, starts, ks, cs, es);
    *n_passes = (uint32_t)n;
    if ((uint32_t)n > cap) return cap ? B64_ERR_BUFFER_TOO_SMALL : B64_OK;
    const uint32_t cus = n_cus ? n_cus : b64::ntt_device_cus();
    for (int p = 0; p < n; p++) {
        b64_ntt_pass_info &o = passes[p];
        o.layer0 = (uint32_t)starts[p];
        o.n_layers = (uint32_t)ks[p];
        o.low_bits = (uint32_t)cs[p];
        o.extra_bits = (uint32_t)es[p];
        o.kernel = (uint32_t)b64::ntt_pass_kernel(tu, (int)log_len, starts[p], ks[p], cs[p], es[p], inverse != 0);
        o.workitems = b64::ntt_pass_workitems((int)o.kernel, (int)log_len, starts[p], ks[p], cs[p], es[p], cus);
    }
    return B64_OK;
}

int b64_ntt_forward(const b64_domain *d, void *d_data, uint32_t log_len, uint32_t se, uint32_t sl, void *stream) {
    return ntt_common(d, d_data, log_len, se, sl, false, stream);
}

int b64_ntt_inverse(const b64_domain *d, void *d_data, uint32_t log_len, uint32_t se, uint32_t sl, void *stream) {
    return ntt_common(d, d_data, log_len, se, sl, true, stream);
}

int b64_rs_encode_batch(const b64_domain *dc, uint32_t log_dim, uint32_t lir, const void *d_msg,
                        uint32_t log_batch, void *d_out, void *stream) {
    b64_domain *d = const_cast<b64_domain *>(dc);
    if (!d || !d_msg || !d_out) return B64_ERR_INVALID_ARGS;
    if (log_dim + lir > (uint32_t)d->dim) return B64_ERR_SUBSPACE_DIMENSION_MISMATCH;
    const uint32_t m = log_dim + log_batch;
    if (m + lir > 40 || lir > 8) return B64_ERR_INVALID_ARGS;
    b64::DeviceDomain dd;
    B64_TRY(d->device_view(dd));
    hipStream_t s = as_stream(stream);
    if (log_dim == 0) {
        // no butterfly layer: the codeword is the replicated bit-reversed message
        HIP_TRY(b64::launch_bitrev_replicate((const uint4 *)d_msg, (uint4 *)d_out, (int)m, (int)lir, s));
        return B64_OK;
    }
    // the first NTT pass gathers the bit-reversed, replicated message itself
    const hipError_t e = b64::launch_ntt(dd, (uint4 *)d_out, (int)(m + lir), (int)lir, (int)log_batch, false, s,
                                         (const uint4 *)d_msg, (int)m);
    if (e == hipErrorInvalidConfiguration) return B64_ERR_INVALID_ARGS;  // a pass grid too large
    HIP_TRY(e);
    return B64_OK;
}

// ===========================================================================
// Merkle
// ===========================================================================
static int merkle_commit_impl(const uint4 *elems, uint64_t n_elems, uint64_t batch, uint4 *tree, hipStream_t s) {
    if (batch == 0 || n_elems % batch) return B64_ERR_INCORRECT_BATCH_SIZE;
    const uint64_t n_leaves = n_elems / batch;
    if (n_leaves == 0 || (n_leaves & (n_leaves - 1))) return B64_ERR_POWER_OF_TWO_LENGTH_REQUIRED;
    HIP_TRY(b64::launch_leaf_hash(elems, n_leaves, 0, batch, tree, s));
    HIP_TRY(b64::launch_merkle_levels(tree, n_leaves, compress_iv(), s));
    return B64_OK;
}

int b64_merkle_commit(const void *d_elems, uint64_t n_elems, uint64_t batch, void *d_tree, uint8_t root_out[32],
                      void *stream) {
    if (!d_elems || !d_tree) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    hipStream_t s = as_stream(stream);
    B64_TRY(merkle_commit_impl((const uint4 *)d_elems, n_elems, batch, (uint4 *)d_tree, s));
    if (root_out) {
        const uint64_t n_leaves = n_elems / batch;
        const uint8_t *root = (const uint8_t *)d_tree + 32 * (2 * n_leaves - 2);
        HIP_TRY(hipMemcpyAsync(root_out, root, 32, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return B64_OK;
}

int b64_merkle_layer(const void *d_tree, uint32_t log_len, uint32_t depth, uint8_t *out, void *stream) {
    if (!d_tree || !out) return B64_ERR_INVALID_ARGS;
    if (depth > log_len) return B64_ERR_INCORRECT_LAYER_DEPTH;
    const uint64_t total = (2ull << log_len) - 1;
    const uint64_t start = total + 1 - (2ull << depth);
    hipStream_t s = as_stream(stream);
    HIP_TRY(hipMemcpyAsync(out, (const uint8_t *)d_tree + 32 * start, 32ull << depth, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return B64_OK;
}

int b64_merkle_branch(const void *d_tree, uint32_t log_len, uint64_t index, uint32_t depth, uint8_t *out,
                      void *stream) {
    if (!d_tree || !out) return B64_ERR_INVALID_ARGS;
    if (index >= (1ull << log_len) || depth > log_len) return B64_ERR_INDEX_OUT_OF_RANGE;
    hipStream_t s = as_stream(stream);
    for (uint32_t j = 0; j < log_len - depth; j++) {
        const uint64_t node = (((1ull << j) - 1) << (log_len + 1 - j)) | ((index >> j) ^ 1);
        HIP_TRY(hipMemcpyAsync(out + 32 * j, (const uint8_t *)d_tree + 32 * node, 32, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return B64_OK;
}

// ===========================================================================
// Folds
// ===========================================================================
int b64_fri_fold_interleaved(const void *d_cw, uint32_t log_len, uint32_t log_batch, const uint64_t *chal,
                             void *d_out, void *stream) {
    if (!d_cw || !d_out || (log_batch && !chal) || log_batch > 6 || log_len > 40) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    uint4 ch[16];
    for (uint32_t t = 0; t < log_batch; t++) ch[t] = to_u4(ld128(chal + 2 * t));
    HIP_TRY(b64::launch_fold_interleaved((const uint4 *)d_cw, (uint4 *)d_out, 1ull << log_len, (int)log_batch, ch,
                                         as_stream(stream)));
    return B64_OK;
}

int b64_fri_fold_codeword(const b64_domain *dc, const void *d_cw, uint32_t log_len, const uint64_t *chal,
                          uint32_t n_chal, void *d_out, void *stream) {
    b64_domain *d = const_cast<b64_domain *>(dc);
    if (!d || !d_cw || !d_out || (n_chal && !chal) || n_chal > log_len || n_chal > 8) return B64_ERR_INVALID_ARGS;
    if (log_len > (uint32_t)d->dim) return B64_ERR_INVALID_ARGS;
    b64::DeviceDomain dd;
    B64_TRY(d->device_view(dd));
    uint4 ch[16];
    for (uint32_t t = 0; t < n_chal; t++) ch[t] = to_u4(ld128(chal + 2 * t));
    HIP_TRY(b64::launch_fold_codeword(dd, (const uint4 *)d_cw, (uint4 *)d_out, (int)log_len, ch, (int)n_chal,
                                      as_stream(stream)));
    return B64_OK;
}

// ===========================================================================
// Ring switch (prover/prover/src/ring_switch.rs)
// ===========================================================================
int b64_ring_switch_fold_1b_rows(const void *d_mat, const void *d_vec, uint32_t log_len, void *d_out,
                                 void *stream) {
    if (!d_mat || !d_vec || !d_out || log_len > 40) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    const uint64_t n = 1ull << log_len;
    hipStream_t s = as_stream(stream);
    // per-call partial sums from the stream-ordered pool, so concurrent calls on
    // different streams never share them
    void *partial = nullptr;
    b64::abi::keep_pool_memory();
    HIP_TRY(hipMallocAsync(&partial, (size_t)b64::rs_fold_1b_rows_blocks(n) * 128 * sizeof(uint4), s));
    const hipError_t e = b64::launch_rs_fold_1b_rows((const uint4 *)d_mat, (const uint4 *)d_vec, n,
                                                     (uint4 *)partial, (uint4 *)d_out, s);
    HIP_TRY(hipFreeAsync(partial, s));
    HIP_TRY(e);
    return B64_OK;
}

int b64_ring_switch_fold_elems(void *d_elems, uint32_t log_len, const uint64_t *vec, void *stream) {
    if (!d_elems || !vec || log_len > 40) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    b64::RsVec v;
    for (int b = 0; b < 128; b++) v.v[b] = to_u4(ld128(vec + 2 * b));
    HIP_TRY(b64::launch_rs_fold_elems((uint4 *)d_elems, 1ull << log_len, v, as_stream(stream)));
    return B64_OK;
}

int b64_pack_witness(const void *d_words, uint64_t n_words, uint32_t log_elems, void *d_out, void *stream) {
    if (!d_words || !d_out || log_elems > 40 || n_words > (2ull << log_elems)) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    HIP_TRY(b64::launch_pack_witness((const uint64_t *)d_words, n_words, (uint4 *)d_out, 1ull << log_elems,
                                     as_stream(stream)));
    return B64_OK;
}

// ===========================================================================
// FRI parameters (verifier/verifier/src/fri/common.rs)
// ===========================================================================
uint32_t b64_calculate_n_test_queries(uint32_t security_bits, uint32_t lir) {
    const double rate = pow(2.0, -(double)lir);
    const double per_query_err = 0.5 * (1.0 + rate);
    return (uint32_t)ceil((double)security_bits / -log2(per_query_err));
}

uint32_t b64_estimate_optimal_arity(uint32_t L, uint32_t digest_size, uint32_t field_size) {
    bool have = false;
    uint64_t old = 0;
    uint32_t last = 1;
    for (uint32_t arity = 1; arity <= L; arity++) {
        const uint64_t est = (((uint64_t)(L / 2) * digest_size + (1ull << arity) * field_size) * (L - arity)) / arity;
        if (have && est > old) break;
        have = true;
        old = est;
        last = arity;
    }
    return last;
}

int b64_fri_params_new(uint32_t log_dim, uint32_t lir, uint32_t lbs, const uint32_t *arities, uint32_t n_ar,
                       uint32_t n_queries, b64_fri_params *out) {
    if (!out || n_ar > B64_MAX_FOLD_ARITIES || (n_ar && !arities)) return B64_ERR_INVALID_ARGS;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < n_ar; i++) sum += arities[i];
    if (sum > log_dim) return B64_ERR_INVALID_FOLD_ARITY_SEQUENCE;
    memset(out, 0, sizeof(*out));
    out->log_dim = log_dim;
    out->log_inv_rate = lir;
    out->log_batch_size = lbs;
    out->n_test_queries = n_queries;
    out->n_fold_arities = n_ar;
    for (uint32_t i = 0; i < n_ar; i++) out->fold_arities[i] = arities[i];
    out->log_terminal_dim = log_dim - sum;
    return B64_OK;
}

int b64_fri_params_choose(const b64_domain *d, uint32_t log_msg_len, uint32_t security_bits, uint32_t lir,
                          uint32_t arity, b64_fri_params *out) {
    if (!d || !out || arity == 0) return B64_ERR_INVALID_ARGS;
    const uint32_t lbs = log_msg_len < arity ? log_msg_len : arity;
    const uint32_t log_dim = log_msg_len - lbs;
    const uint32_t rs_log_len = log_dim + lir;
    if (rs_log_len > (uint32_t)d->dim) return B64_ERR_SUBSPACE_DIMENSION_MISMATCH;
    const uint32_t nq = b64_calculate_n_test_queries(security_bits, lir);
    const uint32_t cap = (uint32_t)log2_ceil_u64(nq);
    uint32_t term = cap < lir ? lir : cap;
    if (term > rs_log_len) term = rs_log_len;
    const uint32_t q = (rs_log_len - term) / arity, rem = (rs_log_len - term) % arity;
    // the reference's list has no length limit; one that does not fit the
    // fixed-size array would be a different protocol, so it is refused
    if (q + (rem != 0) > B64_MAX_FOLD_ARITIES) return B64_ERR_INVALID_ARGS;
    uint32_t ar[B64_MAX_FOLD_ARITIES];
    uint32_t n = 0;
    for (uint32_t i = 0; i < q; i++) ar[n++] = arity;
    if (rem) ar[n++] = rem;
    return b64_fri_params_new(log_dim, lir, lbs, ar, n, nq, out);
}

int b64_setup_commit_params(uint32_t log_witness_elems, uint32_t lir, uint32_t security_bits,
                            b64_fri_params *params_out, b64_domain **domain_out) {
    if (!params_out || !domain_out) return B64_ERR_INVALID_ARGS;
    // Verifier::setup (verify.rs:76-96): digest 32 B, field element 16 B
    const uint32_t log_code_len = log_witness_elems + lir;
    if (log_code_len > (uint32_t)b64::MAX_DIM) return B64_ERR_INVALID_ARGS;
    const uint32_t arity = b64_estimate_optimal_arity(log_code_len, 32, 16);
    b64_domain *std_dom = nullptr;
    B64_TRY(b64_domain_create_standard(log_code_len, &std_dom));
    int rc = b64_fri_params_choose(std_dom, log_witness_elems, security_bits, lir, arity, params_out);
    if (rc == B64_OK) {
        // Prover::setup (prove.rs:98-99): GenericPreExpanded over rs_code.subspace()
        const uint32_t rs_log_len = params_out->log_dim + lir;
        std::vector<uint64_t> basis(2 * (size_t)(rs_log_len ? rs_log_len : 1));
        rc = b64_domain_subspace(std_dom, rs_log_len, basis.data());
        if (rc == B64_OK) rc = b64_domain_create(basis.data(), rs_log_len, domain_out);
    }
    b64_domain_destroy(std_dom);
    return rc;
}

int b64_fri_commit_interleaved_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash,
                                  const void *d_msg, void *d_codeword, void *d_tree, uint8_t root_out[32],
                                  void *stream) {
    if (!p || !d || !d_msg || !d_codeword || !d_tree) return B64_ERR_INVALID_ARGS;
    if (merkle_hash != B64_MERKLE_SHA256 && merkle_hash != B64_MERKLE_VISION4 && merkle_hash != B64_MERKLE_VISION6)
        return B64_ERR_INVALID_ARGS;
    // encode_batch requires ntt.subspace(rs_log_len) == rs_code.subspace(); the
    // code built over `d` satisfies it by construction (reed_solomon.rs:54-65).
    B64_TRY(b64_rs_encode_batch(d, p->log_dim, p->log_inv_rate, d_msg, p->log_batch_size, d_codeword, stream));
    const uint64_t n = 1ull << (p->log_dim + p->log_inv_rate + p->log_batch_size);
    if (merkle_hash == B64_MERKLE_SHA256)
        return b64_merkle_commit(d_codeword, n, 1ull << p->log_batch_size, d_tree, root_out, stream);
    return b64_merkle_commit_vision(merkle_hash, d_codeword, n, 1ull << p->log_batch_size, d_tree, root_out, stream);
}

int b64_fri_commit_interleaved(const b64_fri_params *p, const b64_domain *d, const void *d_msg, void *d_codeword,
                               void *d_tree, uint8_t root_out[32], void *stream) {
    return b64_fri_commit_interleaved_ex(p, d, B64_MERKLE_SHA256, d_msg, d_codeword, d_tree, root_out, stream);
}

int b64_fri_commit_fold(const b64_fri_params *p, const b64_domain *d, const void *d_msg, void *d_codeword,
                        void *d_tree, const uint64_t *challenges, void *d_folded, uint8_t root_out[32],
                        void *stream) {
    if (!p || !d || !d_msg || !d_codeword || !d_tree || !d_folded || (p->log_batch_size && !challenges))
        return B64_ERR_INVALID_ARGS;
    const uint32_t lbs = p->log_batch_size, rs_log_len = p->log_dim + p->log_inv_rate;
    if (lbs > 6) return B64_ERR_INVALID_ARGS;
    hipStream_t s = as_stream(stream);
    B64_TRY(b64_rs_encode_batch(d, p->log_dim, p->log_inv_rate, d_msg, lbs, d_codeword, stream));
    uint4 ch[8];
    for (uint32_t t = 0; t < lbs; t++) ch[t] = to_u4(ld128(challenges + 2 * t));
    const uint64_t n_leaves = 1ull << rs_log_len;
    if (lbs == 4) {
        // one pass over the codeword: leaf digests and the folded oracle
        HIP_TRY(b64::launch_leaf16_fold((const uint4 *)d_codeword, n_leaves, (uint4 *)d_tree, (uint4 *)d_folded, ch, s));
        HIP_TRY(b64::launch_merkle_levels((uint4 *)d_tree, n_leaves, compress_iv(), s));
    } else {
        B64_TRY(merkle_commit_impl((const uint4 *)d_codeword, n_leaves << lbs, 1ull << lbs, (uint4 *)d_tree, s));
        HIP_TRY(b64::launch_fold_interleaved((const uint4 *)d_codeword, (uint4 *)d_folded, n_leaves, (int)lbs, ch, s));
    }
    if (root_out) {
        HIP_TRY(hipMemcpyAsync(root_out, (const uint8_t *)d_tree + 32 * (2 * n_leaves - 2), 32, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return B64_OK;
}

}  // extern "C"

// ===========================================================================
// Transcript: HasherChallenger<Sha256> + proof tape
// (verifier/transcript/src/fiat_shamir/hasher_challenger.rs, transcript.rs)
// ===========================================================================
struct b64_transcript {
    b64::host::Sha256 hasher;
    bool sampler = true;
    uint8_t buffer[32];
    size_t index = 0;
    std::vector<uint8_t> tape;

    b64_transcript() {
        uint8_t e[32];
        b64::host::Sha256::digest(nullptr, 0, e);
        hasher.update(e, 32);
        memcpy(buffer, e, 32);
    }
    void observe(const uint8_t *b, size_t n) {
        if (sampler) {  // Sampler -> Observer: absorb the LE usize read index
            uint8_t le[8];
            for (int i = 0; i < 8; i++) le[i] = (uint8_t)((uint64_t)index >> (8 * i));
            hasher.update(le, 8);
            sampler = false;
        }
        hasher.update(b, n);
    }
    void sample(uint8_t *out, size_t n) {
        if (!sampler) {  // Observer -> Sampler: buffer exhausted, refill on first read
            sampler = true;
            index = 32;
        }
        while (n) {
            if (index == 32) {
                uint8_t d[32];
                hasher.finalize(d);
                hasher.reset();
                hasher.update(d, 32);
                memcpy(buffer, d, 32);
                index = 0;
            }
            size_t take = 32 - index;
            if (take > n) take = n;
            memcpy(out, buffer + index, take);
            index += take;
            out += take;
            n -= take;
        }
    }
};

extern "C" {

int b64_transcript_create(b64_transcript **out) {
    if (!out) return B64_ERR_INVALID_ARGS;
    *out = new (std::nothrow) b64_transcript();
    return *out ? B64_OK : B64_ERR_OUT_OF_MEMORY;
}
void b64_transcript_destroy(b64_transcript *t) { delete t; }
void b64_transcript_message(b64_transcript *t, const uint8_t *b, size_t n) {
    t->tape.insert(t->tape.end(), b, b + n);
    t->observe(b, n);
}
void b64_transcript_observe(b64_transcript *t, const uint8_t *b, size_t n) { t->observe(b, n); }
void b64_transcript_decommit(b64_transcript *t, const uint8_t *b, size_t n) { t->tape.insert(t->tape.end(), b, b + n); }
void b64_transcript_sample(b64_transcript *t, uint8_t *out, size_t n) { t->sample(out, n); }
uint32_t b64_transcript_sample_bits(b64_transcript *t, uint32_t bits) {
    if (bits > 32) bits = 32;
    uint8_t b[4];
    t->sample(b, 4);
    const uint32_t v = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    return bits >= 32 ? v : (v & ((1u << bits) - 1));
}
size_t b64_transcript_tape(const b64_transcript *t, uint8_t *out, size_t cap) {
    if (out && cap >= t->tape.size()) memcpy(out, t->tape.data(), t->tape.size());
    return t->tape.size();
}

}  // extern "C"

// ===========================================================================
// Proof wire format (verifier/core/src/constraint_system.rs:1014-1100): the
// transcript bytes wrapped with a version and the challenger's name, using the
// SerializeBytes rules of verifier/utils/src/serialization.rs (u32 LE; usize
// as u32; str and Vec<u8> as a length then the bytes).
// ===========================================================================
namespace {

void put_u32(uint8_t *p, uint32_t v) {
    for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get_u32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// Well-formed UTF-8 as String::from_utf8 accepts it: no overlong forms, no
// surrogates, nothing above U+10FFFF.
bool valid_utf8(const uint8_t *s, uint64_t n) {
    uint64_t i = 0;
    while (i < n) {
        const uint8_t c = s[i];
        if (c < 0x80) {
            i++;
            continue;
        }
        int len;
        uint8_t lo = 0x80, hi = 0xBF;  // range of the second byte
        if (c >= 0xC2 && c <= 0xDF) {
            len = 2;
        } else if (c >= 0xE0 && c <= 0xEF) {
            len = 3;
            if (c == 0xE0) lo = 0xA0;
            if (c == 0xED) hi = 0x9F;
        } else if (c >= 0xF0 && c <= 0xF4) {
            len = 4;
            if (c == 0xF0) lo = 0x90;
            if (c == 0xF4) hi = 0x8F;
        } else {
            return false;
        }
        if (n - i < (uint64_t)len) return false;
        if (s[i + 1] < lo || s[i + 1] > hi) return false;
        for (int k = 2; k < len; k++)
            if ((s[i + k] & 0xC0) != 0x80) return false;
        i += (uint64_t)len;
    }
    return true;
}

}  // namespace

extern "C" {

int b64_proof_serialize(const uint8_t *data, uint64_t data_len, const char *challenger_type, uint8_t *out,
                        uint64_t out_cap, uint64_t *out_len) {
    if (!challenger_type || (!data && data_len) || !out_len) return B64_ERR_INVALID_ARGS;
    const uint64_t ct_len = strlen(challenger_type);
    const uint64_t need = 4 + 4 + ct_len + 4 + data_len;
    *out_len = need;
    if (ct_len > 0xFFFFFFFFull || data_len > 0xFFFFFFFFull) return B64_ERR_USIZE_TOO_LARGE;
    if (!out || out_cap < need) return B64_ERR_BUFFER_TOO_SMALL;
    uint8_t *p = out;
    put_u32(p, B64_PROOF_SERIALIZATION_VERSION);
    put_u32(p + 4, (uint32_t)ct_len);
    if (ct_len) memcpy(p + 8, challenger_type, ct_len);
    put_u32(p + 8 + ct_len, (uint32_t)data_len);
    if (data_len) memcpy(p + 12 + ct_len, data, data_len);
    return B64_OK;
}

int b64_proof_deserialize(const uint8_t *buf, uint64_t buf_len, uint64_t *challenger_off, uint64_t *challenger_len,
                          uint64_t *data_off, uint64_t *data_len) {
    if ((!buf && buf_len) || !challenger_off || !challenger_len || !data_off || !data_len)
        return B64_ERR_INVALID_ARGS;
    uint64_t at = 0;
    if (buf_len - at < 4) return B64_ERR_NOT_ENOUGH_BYTES;
    if (get_u32(buf) != B64_PROOF_SERIALIZATION_VERSION) return B64_ERR_INVALID_CONSTRUCTION;
    at += 4;
    if (buf_len - at < 4) return B64_ERR_NOT_ENOUGH_BYTES;
    const uint64_t ct = get_u32(buf + at);
    at += 4;
    if (buf_len - at < ct) return B64_ERR_NOT_ENOUGH_BYTES;
    if (!valid_utf8(buf + at, ct)) return B64_ERR_FROM_UTF8;
    *challenger_off = at;
    *challenger_len = ct;
    at += ct;
    if (buf_len - at < 4) return B64_ERR_NOT_ENOUGH_BYTES;
    const uint64_t dl = get_u32(buf + at);
    at += 4;
    if (buf_len - at < dl) return B64_ERR_NOT_ENOUGH_BYTES;
    *data_off = at;
    *data_len = dl;
    return B64_OK;
}

}  // extern "C"

// ===========================================================================
// FRI fold / query prover (prover/prover/src/fri/fold.rs, query.rs)
// ===========================================================================
struct b64_fri_prover {
    struct Oracle {
        uint4 *cw = nullptr;    // device codeword
        uint4 *tree = nullptr;  // device tree (2*n_leaves-1 digests)
        int log_len = 0;        // log2 codeword length
        int tree_log_len = 0;   // log2 number of leaves
    };
    b64_fri_params p;
    b64_domain *d = nullptr;
    Oracle first;  // the committed interleaved codeword (not owned)
    std::vector<Oracle> committed;
    hipStream_t s = nullptr;
    uint32_t merkle_hash = B64_MERKLE_SHA256;  // scheme of the folded oracles' trees
    uint32_t curr_round = 0;
    bool has_next = true;
    uint32_t next_commit = 0;
    std::vector<u128> unprocessed;
    uint8_t *h_root = nullptr;  // pinned: the last committed root, copied by the fold round

    ~b64_fri_prover() {
        for (auto &o : committed) {
            if (o.cw) (void)hipFreeAsync(o.cw, s);
            if (o.tree) (void)hipFreeAsync(o.tree, s);
        }
        if (h_root) {
            (void)hipStreamSynchronize(s);  // a root copy may still be in flight
            b64::abi::pinned_small_put(h_root);
        }
    }
    uint32_t n_rounds() const { return p.log_dim + p.log_batch_size; }
    uint32_t rs_log_len() const { return p.log_dim + p.log_inv_rate; }
};

int b64::FriQuery::prepare(const b64_fri_params &params, const FriOracleView &first, const FriOracleView *committed,
                           size_t n_committed) {
    p = params;
    // terminate codeword: the last committed oracle, or the codeword itself
    term = n_committed ? committed[n_committed - 1] : first;
    // optimal layer depths (verifier fri/common.rs:154-168, scheme.rs:70-72)
    const int cap = log2_ceil_u64(p.n_test_queries);
    depths.clear();
    int log_n_cosets = (int)(p.log_dim + p.log_inv_rate + p.log_batch_size);
    for (uint32_t i = 0; i <= p.n_fold_arities; i++) {
        log_n_cosets -= (int)(i == 0 ? p.log_batch_size : p.fold_arities[i - 1]);
        depths.push_back(cap < log_n_cosets ? cap : log_n_cosets);
    }
    // the first oracle and every committed oracle but the terminal one
    opened.clear();
    opened.push_back(first);
    for (size_t i = 0; i + 1 < n_committed; i++) opened.push_back(committed[i]);
    if (opened.size() > depths.size()) opened.resize(depths.size());
    return B64_OK;
}

int b64::FriQuery::sample(b64_transcript *t) {
    // queries (fold.rs:230-243): each index is sampled, then its opening written;
    // decommitments are not observed, so the indices can all come first
    const uint32_t rs_log_len = p.log_dim + p.log_inv_rate;
    idx.resize(p.n_test_queries);
    for (uint32_t q = 0; q < p.n_test_queries; q++) idx[q] = b64_transcript_sample_bits(t, rs_log_len);
    for (uint64_t v : idx) {
        uint64_t cur = v;
        for (size_t i = 0; i < opened.size(); i++) {
            if (i > 0) cur >>= p.fold_arities[i - 1];
            if (cur >> opened[i].tree_log_len) return B64_ERR_INDEX_OUT_OF_RANGE;
        }
    }
    return B64_OK;
}

int b64::FriQuery::args(QueryArgs *a) const {
    *a = QueryArgs();
    if (opened.size() > (size_t)MAX_OPENED || term.log_len > 40) return B64_ERR_INVALID_ARGS;
    a->term = term.cw;
    a->term_gran = 1ull << term.log_len;
    a->nq = (uint32_t)idx.size();
    a->n_opened = (int)opened.size();
    uint64_t per_query = 0;
    for (size_t i = 0; i < opened.size(); i++) {
        const FriOracleView &v = opened[i];
        QueryOracle &o = a->o[i];
        o.cw = v.cw;
        o.tree = v.tree;
        o.tree_log_len = v.tree_log_len;
        o.log_coset = (int)(i == 0 ? p.log_batch_size : p.fold_arities[i - 1]);
        o.depth = depths[i];
        // the kernel reads cw[< 2^log_len] and tree nodes < 2^(tree_log_len+1) only under these
        if (v.tree_log_len + o.log_coset != v.log_len || o.depth < 0 || o.depth > v.tree_log_len ||
            v.log_len > 40)
            return B64_ERR_INVALID_ARGS;
        o.n_gran = (1u << o.log_coset) + 2u * (uint32_t)(v.tree_log_len - o.depth);
        a->layers_gran += 2ull << o.depth;
        per_query += o.n_gran;
    }
    if (per_query * a->nq >= (1ull << 31)) return B64_ERR_INVALID_ARGS;  // 32-bit offsets in the kernel
    a->per_query = (uint32_t)per_query;
    a->n_gran = a->term_gran + a->layers_gran + per_query * a->nq;
    if (a->n_gran >= (1ull << 39)) return B64_ERR_INVALID_ARGS;  // 2^31 blocks at most
    return B64_OK;
}

extern "C" {

int b64_fri_prover_create(const b64_fri_params *p, const b64_domain *d, const void *d_codeword, const void *d_tree,
                          void *stream, b64_fri_prover **out) {
    return b64_fri_prover_create_ex(p, d, B64_MERKLE_SHA256, d_codeword, d_tree, stream, out);
}

int b64_fri_prover_create_ex(const b64_fri_params *p, const b64_domain *d, uint32_t merkle_hash, const void *d_codeword,
                             const void *d_tree, void *stream, b64_fri_prover **out) {
    if (!p || !d || !d_codeword || !d_tree || !out) return B64_ERR_INVALID_ARGS;
    if (merkle_hash != B64_MERKLE_SHA256 && merkle_hash != B64_MERKLE_VISION4 && merkle_hash != B64_MERKLE_VISION6)
        return B64_ERR_INVALID_ARGS;
    if (p->log_dim + p->log_inv_rate > (uint32_t)d->dim) return B64_ERR_INVALID_ARGS;
    b64_fri_prover *fp = new (std::nothrow) b64_fri_prover();
    if (!fp) return B64_ERR_OUT_OF_MEMORY;
    fp->p = *p;
    fp->d = const_cast<b64_domain *>(d);
    fp->first.cw = (uint4 *)d_codeword;
    fp->first.tree = (uint4 *)d_tree;
    fp->first.log_len = (int)(p->log_dim + p->log_inv_rate + p->log_batch_size);
    fp->first.tree_log_len = (int)(p->log_dim + p->log_inv_rate);
    fp->s = as_stream(stream);
    fp->next_commit = p->log_batch_size;  // FRIFoldProver::new (fold.rs:70)
    fp->merkle_hash = merkle_hash;
    fp->h_root = (uint8_t *)b64::abi::pinned_small_get();
    if (!fp->h_root) {
        delete fp;
        return B64_ERR_OUT_OF_MEMORY;
    }
    b64::abi::keep_pool_memory();
    *out = fp;
    return B64_OK;
}

void b64_fri_prover_destroy(b64_fri_prover *fp) { delete fp; }
uint32_t b64_fri_prover_n_rounds(const b64_fri_prover *fp) { return fp->n_rounds(); }
uint32_t b64_fri_prover_curr_round(const b64_fri_prover *fp) { return fp->curr_round; }

int b64_fri_prover_receive_challenge(b64_fri_prover *fp, const uint64_t ch[2]) {
    if (!fp || !ch) return B64_ERR_INVALID_ARGS;
    fp->unprocessed.push_back(ld128(ch));
    fp->curr_round++;
    return B64_OK;
}

}  // extern "C"

// The fold round's device work, enqueued on the prover's stream; a committed
// round also enqueues the copy of its root into fp->h_root.  The caller
// synchronises the stream before reading it.
int b64::fri_fold_round_enqueue(b64_fri_prover *fp, int *committed, const uint8_t **root_host) {
    if (!fp || !committed) return B64_ERR_INVALID_ARGS;
    *committed = 0;
    if (!(fp->has_next && fp->next_commit == fp->curr_round)) return B64_OK;
    b64::DeviceDomain dd;
    B64_TRY(fp->d->device_view(dd));
    b64_fri_prover::Oracle o;
    std::vector<uint4> ch(fp->unprocessed.size() + 1);
    for (size_t i = 0; i < fp->unprocessed.size(); i++) ch[i] = to_u4(fp->unprocessed[i]);
    const size_t n_ch = fp->unprocessed.size();
    if (fp->committed.empty()) {
        // fold_interleaved (fold.rs:260-282)
        o.log_len = (int)fp->rs_log_len();
        if (n_ch != fp->p.log_batch_size || n_ch > 6) return B64_ERR_INVALID_ARGS;
        HIP_TRY(hipMallocAsync((void **)&o.cw, 16ull << o.log_len, fp->s));
        hipError_t e = b64::launch_fold_interleaved(fp->first.cw, o.cw, 1ull << o.log_len, (int)n_ch, ch.data(), fp->s);
        if (e != hipSuccess) {
            (void)hipFreeAsync(o.cw, fp->s);
            return B64_ERR_HIP;
        }
    } else {
        // fold_codeword (fold.rs:297-318)
        const b64_fri_prover::Oracle &prev = fp->committed.back();
        if ((int)n_ch > prev.log_len || n_ch > 8) return B64_ERR_INVALID_ARGS;
        o.log_len = prev.log_len - (int)n_ch;
        HIP_TRY(hipMallocAsync((void **)&o.cw, 16ull << o.log_len, fp->s));
        hipError_t e = b64::launch_fold_codeword(dd, prev.cw, o.cw, prev.log_len, ch.data(), (int)n_ch, fp->s);
        if (e != hipSuccess) {
            (void)hipFreeAsync(o.cw, fp->s);
            return B64_ERR_HIP;
        }
    }
    fp->unprocessed.clear();
    const size_t k = fp->committed.size();
    const bool has_arity = k < fp->p.n_fold_arities;
    const int log_coset = has_arity ? (int)fp->p.fold_arities[k] : (int)fp->p.log_terminal_dim;
    if (log_coset > o.log_len) {
        (void)hipFreeAsync(o.cw, fp->s);
        return B64_ERR_INCORRECT_BATCH_SIZE;
    }
    o.tree_log_len = o.log_len - log_coset;
    if (hipMallocAsync((void **)&o.tree, 32ull * ((2ull << o.tree_log_len) - 1), fp->s) != hipSuccess) {
        (void)hipFreeAsync(o.cw, fp->s);
        return B64_ERR_OUT_OF_MEMORY;
    }
    fp->committed.push_back(o);  // owned from here on
    if (fp->merkle_hash == B64_MERKLE_SHA256)
        B64_TRY(merkle_commit_impl(o.cw, 1ull << o.log_len, 1ull << log_coset, o.tree, fp->s));
    else
        B64_TRY(b64::vision_merkle_commit((int)fp->merkle_hash, o.cw, 1ull << o.log_len, 1ull << log_coset, o.tree,
                                          fp->s));
    if (has_arity)
        fp->next_commit += fp->p.fold_arities[k];
    else
        fp->has_next = false;
    HIP_TRY(hipMemcpyAsync(fp->h_root, (const uint8_t *)o.tree + 32 * ((2ull << o.tree_log_len) - 2), 32,
                           hipMemcpyDeviceToHost, fp->s));
    if (root_host) *root_host = fp->h_root;
    *committed = 1;
    return B64_OK;
}

extern "C" {

int b64_fri_prover_execute_fold_round(b64_fri_prover *fp, uint8_t root_out[32], int *committed) {
    const uint8_t *root = nullptr;
    B64_TRY(b64::fri_fold_round_enqueue(fp, committed, &root));
    if (*committed && root_out) {
        HIP_TRY(hipStreamSynchronize(fp->s));
        memcpy(root_out, root, 32);
    }
    return B64_OK;
}

int b64_fri_prover_oracle(const b64_fri_prover *fp, uint32_t k, const void **d_cw, uint32_t *log_len,
                          const void **d_tree) {
    if (!fp || k >= fp->committed.size()) return B64_ERR_INDEX_OUT_OF_RANGE;
    if (d_cw) *d_cw = fp->committed[k].cw;
    if (log_len) *log_len = (uint32_t)fp->committed[k].log_len;
    if (d_tree) *d_tree = fp->committed[k].tree;
    return B64_OK;
}

// finish_proof (fold.rs:221-244) with vcs_optimal_layers (query.rs:93-109) and
// prove_query (query.rs:51-91).
int b64_fri_prover_finish_proof(b64_fri_prover *fp, b64_transcript *t) {
    if (!fp || !t) return B64_ERR_INVALID_ARGS;
    if (fp->curr_round != fp->n_rounds()) return B64_ERR_EARLY_PROVER_FINISH;  // finalize (:190-192)
    std::vector<b64::FriOracleView> committed;
    for (const auto &o : fp->committed) committed.push_back({o.cw, o.tree, o.log_len, o.tree_log_len});
    const b64::FriOracleView first = {fp->first.cw, fp->first.tree, fp->first.log_len, fp->first.tree_log_len};
    b64::FriQuery q;
    B64_TRY(q.prepare(fp->p, first, committed.data(), committed.size()));
    B64_TRY(q.sample(t));
    b64::QueryArgs a;
    B64_TRY(q.args(&a));
    // one device buffer: [args][first block][indices][tape]
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t idx_off = up(sizeof(a) + sizeof(uint32_t)), out_off = up(idx_off + 8 * q.idx.size());
    std::vector<uint8_t> host(idx_off + 8 * q.idx.size());
    memcpy(host.data(), &a, sizeof(a));
    memset(host.data() + sizeof(a), 0, sizeof(uint32_t));
    if (!q.idx.empty()) memcpy(host.data() + idx_off, q.idx.data(), 8 * q.idx.size());
    std::vector<uint8_t> out(16 * a.n_gran);
    uint8_t *d = nullptr;
    HIP_TRY(hipMallocAsync((void **)&d, out_off + out.size(), fp->s));
    int rc = B64_OK;
    if (hipMemcpyAsync(d, host.data(), host.size(), hipMemcpyHostToDevice, fp->s) != hipSuccess ||
        b64::launch_fri_query((const b64::QueryArgs *)d, (const uint32_t *)(d + sizeof(a)), 1, b64::fri_query_blocks(a),
                              (const uint64_t *)(d + idx_off), (uint4 *)(d + out_off), fp->s) != hipSuccess ||
        hipMemcpyAsync(out.data(), d + out_off, out.size(), hipMemcpyDeviceToHost, fp->s) != hipSuccess)
        rc = B64_ERR_HIP;
    (void)hipFreeAsync(d, fp->s);
    HIP_TRY(hipStreamSynchronize(fp->s));
    B64_TRY(rc);
    b64_transcript_decommit(t, out.data(), out.size());
    return B64_OK;
}

}  // extern "C"

// ===========================================================================
// Ring-switch fold without allocation.  The partial sums of
// k_rs_fold_1b_rows live in `scratch`, a file-scope device array; host code
// reaches it only through the per-device address hipGetSymbolAddress returns
// (the array's name on the host is a stand-in, not a device address).
// ===========================================================================
namespace {

static __device__ __align__(256) uint4 scratch[(size_t)B64_RS_SCRATCH_BLOCKS * 128];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint4 *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up once per device with that
// device current.
int scratch_for_current_device(uint4 **out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return B64_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return B64_ERR_HIP;
        g_scratch_ptr[dev] = (uint4 *)p;
    }
    *out = g_scratch_ptr[dev];
    return B64_OK;
}

}  // namespace

extern "C" {

int b64_ring_switch_fold_1b_rows_noalloc(const void *d_mat, const void *d_vec, uint32_t log_len, void *d_out,
                                         void *stream) {
    if (!d_mat || !d_vec || !d_out || log_len > 40) return B64_ERR_INVALID_ARGS;
    if (!device_ok()) return B64_ERR_NO_DEVICE;
    const uint64_t n = 1ull << log_len;
    if (b64::rs_fold_1b_rows_blocks(n) > B64_RS_SCRATCH_BLOCKS) return B64_ERR_BUFFER_TOO_SMALL;
    uint4 *partial = nullptr;
    B64_TRY(scratch_for_current_device(&partial));
    HIP_TRY(b64::launch_rs_fold_1b_rows((const uint4 *)d_mat, (const uint4 *)d_vec, n, partial, (uint4 *)d_out,
                                        as_stream(stream)));
    return B64_OK;
}

}  // extern "C"
