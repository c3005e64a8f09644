/*
 * CPU oracle for the Binius64 B128 RS-encode + Merkle-commit + FRI-fold path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See b64_oracle.h.  Each function
 * cites the reference file:line whose algorithm it restates.  The heavy
 * functions also take a thread count: with nthreads > 1 they follow the
 * reference's rayon structure and serve as the CPU baseline ("port").
 */
#define _GNU_SOURCE
#include "b64_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <wmmintrin.h>
#include <emmintrin.h>
#include <smmintrin.h>

typedef unsigned __int128 u128;

static inline u128 ld(const uint64_t *p) { return ((u128)p[1] << 64) | p[0]; }
static inline void st(uint64_t *p, u128 v) { p[0] = (uint64_t)v; p[1] = (uint64_t)(v >> 64); }

/* ===================================================================== */
/* Field: GF(2^128) mod x^128 + x^7 + x^2 + x + 1, bit i <-> x^i          */
/* verifier/field/src/ghash.rs:1-4,55                                      */
/* ===================================================================== */

/* ghash.rs:69-80 */
static inline u128 gf_mul_x(u128 v) {
    u128 mask = (u128)0 - (v >> 127);
    return (v << 1) ^ ((u128)0x87 & mask);
}

/* ghash.rs:82-97 */
static inline u128 gf_mul_inv_x(u128 v) {
    u128 mask = (u128)0 - (v & 1);
    return (v >> 1) ^ ((((u128)1 << 127) | 0x43) & mask);
}

/* arch/shared/ghash.rs:19-45 with gf2_128_reduce :71-84 (POLY = 0x87 :68) */
static inline __m128i clmul_reduce(__m128i t0, __m128i t1, __m128i poly) {
    t0 = _mm_xor_si128(t0, _mm_slli_si128(t1, 8));
    t0 = _mm_xor_si128(t0, _mm_clmulepi64_si128(t1, poly, 0x01));
    return t0;
}

static inline u128 gf_mul_clmul(u128 a, u128 b) {
    __m128i x = _mm_set_epi64x((long long)(uint64_t)(a >> 64), (long long)(uint64_t)a);
    __m128i y = _mm_set_epi64x((long long)(uint64_t)(b >> 64), (long long)(uint64_t)b);
    __m128i poly = _mm_set_epi64x(0, 0x87);
    __m128i t1a = _mm_clmulepi64_si128(x, y, 0x01);
    __m128i t1b = _mm_clmulepi64_si128(x, y, 0x10);
    __m128i t1 = _mm_xor_si128(t1a, t1b);
    __m128i t2 = _mm_clmulepi64_si128(x, y, 0x11);
    t1 = clmul_reduce(t1, t2, poly);
    __m128i t0 = _mm_clmulepi64_si128(x, y, 0x00);
    t0 = clmul_reduce(t0, t1, poly);
    uint64_t lo = (uint64_t)_mm_cvtsi128_si64(t0);
    uint64_t hi = (uint64_t)_mm_extract_epi64(t0, 1);
    return ((u128)hi << 64) | lo;
}

/* arch/portable/univariate_mul_utils_128.rs:120-141 (bmul64 with holes) */
static inline uint64_t bmul64(uint64_t x, uint64_t y) {
    const uint64_t m0 = 0x1111111111111111ull, m1 = 0x2222222222222222ull;
    const uint64_t m2 = 0x4444444444444444ull, m3 = 0x8888888888888888ull;
    uint64_t x0 = x & m0, x1 = x & m1, x2 = x & m2, x3 = x & m3;
    uint64_t y0 = y & m0, y1 = y & m1, y2 = y & m2, y3 = y & m3;
    uint64_t z0 = (x0 * y0) ^ (x1 * y3) ^ (x2 * y2) ^ (x3 * y1);
    uint64_t z1 = (x0 * y1) ^ (x1 * y0) ^ (x2 * y3) ^ (x3 * y2);
    uint64_t z2 = (x0 * y2) ^ (x1 * y1) ^ (x2 * y0) ^ (x3 * y3);
    uint64_t z3 = (x0 * y3) ^ (x1 * y2) ^ (x2 * y1) ^ (x3 * y0);
    return (z0 & m0) | (z1 & m1) | (z2 & m2) | (z3 & m3);
}

static inline uint64_t rev64(uint64_t x) {
    x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return __builtin_bswap64(x);
}

/* arch/portable/packed_ghash_128.rs:33-69 (ghash_mul) and :84-97 (reduce_64) */
static u128 gf_mul_portable(u128 x, u128 y) {
    uint64_t x1 = (uint64_t)(x >> 64), x0 = (uint64_t)x;
    uint64_t y1 = (uint64_t)(y >> 64), y0 = (uint64_t)y;
    uint64_t x0r = rev64(x0), x1r = rev64(x1), x2 = x0 ^ x1, x2r = x0r ^ x1r;
    uint64_t y0r = rev64(y0), y1r = rev64(y1), y2 = y0 ^ y1, y2r = y0r ^ y1r;
    uint64_t z0 = bmul64(y0, x0), z1 = bmul64(y1, x1), z2 = bmul64(y2, x2);
    uint64_t z0h = bmul64(y0r, x0r), z1h = bmul64(y1r, x1r), z2h = bmul64(y2r, x2r);
    z2 ^= z0 ^ z1;
    z2h ^= z0h ^ z1h;
    z0h = rev64(z0h) >> 1;
    z1h = rev64(z1h) >> 1;
    z2h = rev64(z2h) >> 1;
    uint64_t v0 = z0, v1 = z0h ^ z2, v2 = z1 ^ z2h, v3 = z1h;
    v1 ^= v3 ^ (v3 << 1) ^ (v3 << 2) ^ (v3 << 7);
    v2 ^= (v3 >> 63) ^ (v3 >> 62) ^ (v3 >> 57);
    v0 ^= v2 ^ (v2 << 1) ^ (v2 << 2) ^ (v2 << 7);
    v1 ^= (v2 >> 63) ^ (v2 >> 62) ^ (v2 >> 57);
    return ((u128)v1 << 64) | v0;
}

/* Textbook shift-and-add using the definition of mul_x (ghash.rs:69-80). */
static u128 gf_mul_bitserial(u128 a, u128 b) {
    u128 acc = 0;
    for (int i = 0; i < 128; i++) {
        if ((b >> i) & 1) acc ^= a;
        a = gf_mul_x(a);
    }
    return acc;
}

static inline u128 gf_mul(u128 a, u128 b) { return gf_mul_clmul(a, b); }

static u128 gf_pow(u128 a, u128 e) {
    u128 r = 1;
    while (e) {
        if (e & 1) r = gf_mul(r, a);
        a = gf_mul(a, a);
        e >>= 1;
    }
    return r;
}

/* invert_or_zero (arch/portable/packed_ghash_128.rs:117-128): a^(2^128-2), 0 -> 0 */
static u128 gf_inv(u128 a) {
    u128 e = ~(u128)0 - 1;
    return gf_pow(a, e);
}

void bo_gf_mul_many(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n, int variant) {
    for (size_t i = 0; i < n; i++) {
        u128 x = ld(a + 2 * i), y = ld(b + 2 * i), z;
        if (variant == 0) z = gf_mul_clmul(x, y);
        else if (variant == 1) z = gf_mul_portable(x, y);
        else z = gf_mul_bitserial(x, y);
        st(out + 2 * i, z);
    }
}
void bo_gf_square_many(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) { u128 x = ld(a + 2 * i); st(out + 2 * i, gf_mul(x, x)); }
}
void bo_gf_inv_many(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 2 * i, gf_inv(ld(a + 2 * i)));
}
void bo_gf_mul_x_many(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 2 * i, gf_mul_x(ld(a + 2 * i)));
}
void bo_gf_mul_inv_x_many(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 2 * i, gf_mul_inv_x(ld(a + 2 * i)));
}
void bo_gf_pow(const uint64_t a[2], const uint64_t e[2], uint64_t out[2]) {
    st(out, gf_pow(ld(a), ld(e)));
}

/* ===================================================================== */
/* Domain: verifier/math/src/ntt/domain_context.rs                         */
/* ===================================================================== */
#define MAX_EXPANDED_LAYER 22

struct bo_domain {
    int dim;
    u128 *evals;     /* evals[i][k] at evals[i*dim + k], k < dim - i */
    u128 **expanded; /* expanded[layer] has 2^layer entries */
};

/* generate_evals_from_subspace, domain_context.rs:19-46 */
bo_domain *bo_domain_new(const uint64_t *basis, int dim) {
    if (dim < 0 || dim > 40) return NULL;
    bo_domain *d = (bo_domain *)calloc(1, sizeof(bo_domain));
    d->dim = dim;
    d->evals = (u128 *)calloc((size_t)(dim > 0 ? dim * dim : 1), sizeof(u128));
    for (int k = 0; k < dim; k++) d->evals[k] = ld(basis + 2 * k);
    for (int i = 1; i < dim; i++) {
        u128 *prev = d->evals + (size_t)(i - 1) * dim;
        u128 *cur = d->evals + (size_t)i * dim;
        int prev_len = dim - (i - 1);
        for (int k = 1; k < prev_len; k++) cur[k - 1] = gf_mul(prev[k], prev[k] ^ prev[0]);
    }
    for (int i = 0; i < dim; i++) {
        u128 *row = d->evals + (size_t)i * dim;
        u128 inv = gf_inv(row[0]);
        for (int k = 0; k < dim - i; k++) row[k] = gf_mul(row[k], inv);
    }
    /* GenericPreExpanded::generate_from_subspace, domain_context.rs:114-132 */
    d->expanded = (u128 **)calloc((size_t)(dim > 0 ? dim : 1), sizeof(u128 *));
    /* layers above MAX_EXPANDED_LAYER are served on the fly (GenericOnTheFly,
       domain_context.rs:91-94); the values are identical. */
    for (int layer = 0; layer < dim && layer <= MAX_EXPANDED_LAYER; layer++) {
        const u128 *b = d->evals + (size_t)(dim - 1 - layer) * dim; /* layer + 1 entries */
        size_t n = (size_t)1 << layer;
        u128 *e = (u128 *)malloc(n * sizeof(u128));
        e[0] = 0;
        size_t len = 1;
        for (int i = 1; i <= layer; i++) {
            for (size_t j = 0; j < len; j++) e[len + j] = e[j] ^ b[i];
            len <<= 1;
        }
        d->expanded[layer] = e;
    }
    return d;
}

/* BinarySubspace::with_dim (binary_subspace.rs:81-86): basis 1, x, x^2, ... */
bo_domain *bo_domain_new_standard(int dim) {
    uint64_t *basis = (uint64_t *)calloc((size_t)(dim > 0 ? dim : 1) * 2, sizeof(uint64_t));
    for (int i = 0; i < dim; i++) st(basis + 2 * i, (u128)1 << i);
    bo_domain *d = bo_domain_new(basis, dim);
    free(basis);
    return d;
}

void bo_domain_free(bo_domain *d) {
    if (!d) return;
    for (int l = 0; l < d->dim; l++) free(d->expanded[l]);
    free(d->expanded);
    free(d->evals);
    free(d);
}

int bo_domain_dim(const bo_domain *d) { return d->dim; }

void bo_domain_evals(const bo_domain *d, uint64_t *out) {
    for (int i = 0; i < d->dim; i++)
        for (int k = 0; k < d->dim; k++)
            st(out + 2 * ((size_t)i * d->dim + k), k < d->dim - i ? d->evals[(size_t)i * d->dim + k] : 0);
}

static inline u128 dom_twiddle(const bo_domain *d, int layer, uint64_t block) {
    if (d->expanded[layer]) return d->expanded[layer][block];
    const u128 *b = d->evals + (size_t)(d->dim - 1 - layer) * d->dim;
    u128 t = 0;
    for (int i = 0; i < layer; i++)
        if ((block >> i) & 1) t ^= b[1 + i];
    return t;
}

void bo_domain_twiddle(const bo_domain *d, int layer, uint64_t block, uint64_t out[2]) {
    st(out, dom_twiddle(d, layer, block));
}

/* subspace(i) = evals[dim - i] (i > 0), domain_context.rs:142-147 */
void bo_domain_subspace(const bo_domain *d, int i, uint64_t *out) {
    if (i <= 0) return;
    const u128 *row = d->evals + (size_t)(d->dim - i) * d->dim;
    for (int k = 0; k < i; k++) st(out + 2 * k, row[k]);
}

/* ===================================================================== */
/* NTT                                                                     */
/* ===================================================================== */

/* input_check, ntt/reference.rs:89-100 */
static int ntt_check(const bo_domain *d, int log_d, int se, int sl) {
    if (se < 0 || sl < 0 || se + sl > log_d) return 1;
    if (log_d - sl > d->dim) return 1;
    return 0;
}

/* NeighborsLastReference::forward_transform, ntt/reference.rs:17-47 */
int bo_ntt_forward_ref(const bo_domain *d, uint64_t *data, int log_d, int se, int sl) {
    if (ntt_check(d, log_d, se, sl)) return 1;
    for (int layer = se; layer < log_d - sl; layer++) {
        size_t nblocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t block = 0; block < nblocks; block++) {
            u128 t = dom_twiddle(d, layer, block);
            size_t start = block << (log_d - layer);
            for (size_t i0 = start; i0 < start + half; i0++) {
                size_t i1 = half | i0;
                u128 u = ld(data + 2 * i0), v = ld(data + 2 * i1);
                u ^= gf_mul(v, t);
                v ^= u;
                st(data + 2 * i0, u);
                st(data + 2 * i1, v);
            }
        }
    }
    return 0;
}

/* NeighborsLastReference::inverse_transform, ntt/reference.rs:49-76 */
int bo_ntt_inverse_ref(const bo_domain *d, uint64_t *data, int log_d, int se, int sl) {
    if (ntt_check(d, log_d, se, sl)) return 1;
    for (int layer = log_d - sl - 1; layer >= se; layer--) {
        size_t nblocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t block = 0; block < nblocks; block++) {
            u128 t = dom_twiddle(d, layer, block);
            size_t start = block << (log_d - layer);
            for (size_t i0 = start; i0 < start + half; i0++) {
                size_t i1 = half | i0;
                u128 u = ld(data + 2 * i0), v = ld(data + 2 * i1);
                v ^= u;
                u ^= gf_mul(v, t);
                st(data + 2 * i0, u);
                st(data + 2 * i1, v);
            }
        }
    }
    return 0;
}

/* ---- NeighborsLastMultiThread restatement (ntt/neighbors_last.rs) ------ */
#define LOG_BASE_LEN 10 /* neighbors_last.rs:25 */

/* forward_breadth_first with a scalar packing (neighbors_last.rs:131-222) */
static void nl_breadth_first(const bo_domain *d, u128 *data, int log_d, int base_layer,
                             size_t base_block, int lr_start, int lr_end) {
    int log_n = log_d + base_layer;
    for (int layer = lr_start; layer < lr_end; layer++) {
        int log_block = log_n - layer;
        size_t half = (size_t)1 << (log_block - 1);
        int log_blocks = layer - base_layer;
        size_t nblocks = (size_t)1 << log_blocks;
        for (size_t b = 0; b < nblocks; b++) {
            u128 t = dom_twiddle(d, layer, (base_block << log_blocks) + b);
            u128 *b0 = data + (b << log_block), *b1 = b0 + half;
            for (size_t i = 0; i < half; i++) {
                u128 u = b0[i] ^ gf_mul(b1[i], t);
                b0[i] = u;
                b1[i] ^= u;
            }
        }
    }
}

/* forward_depth_first (neighbors_last.rs:57-120) */
static void nl_depth_first(const bo_domain *d, u128 *data, int log_d, int layer, size_t block,
                           int lr_start, int lr_end) {
    if (layer >= lr_end) return;
    if (log_d <= LOG_BASE_LEN) {
        nl_breadth_first(d, data, log_d, layer, block, lr_start, lr_end);
        return;
    }
    size_t half = (size_t)1 << (log_d - 1);
    if (layer >= lr_start) {
        u128 t = dom_twiddle(d, layer, block);
        for (size_t i = 0; i < half; i++) {
            u128 u = data[i] ^ gf_mul(data[half + i], t);
            data[i] = u;
            data[half + i] ^= u;
        }
        lr_start++;
    }
    nl_depth_first(d, data, log_d - 1, layer + 1, block << 1, lr_start, lr_end);
    nl_depth_first(d, data + half, log_d - 1, layer + 1, (block << 1) + 1, lr_start, lr_end);
}

typedef struct {
    void (*fn)(void *, size_t);
    void *arg;
    size_t n_tasks;
    size_t next;
    pthread_mutex_t mu;
} task_pool;

static void *pool_worker(void *p) {
    task_pool *tp = (task_pool *)p;
    for (;;) {
        pthread_mutex_lock(&tp->mu);
        size_t i = tp->next++;
        pthread_mutex_unlock(&tp->mu);
        if (i >= tp->n_tasks) break;
        tp->fn(tp->arg, i);
    }
    return NULL;
}

/* Runs fn(arg, i) for i < n_tasks on nthreads threads (the rayon stand-in). */
static void par_for(int nthreads, size_t n_tasks, void (*fn)(void *, size_t), void *arg) {
    if (nthreads <= 1 || n_tasks <= 1) {
        for (size_t i = 0; i < n_tasks; i++) fn(arg, i);
        return;
    }
    task_pool tp = {fn, arg, n_tasks, 0, PTHREAD_MUTEX_INITIALIZER};
    int nt = nthreads;
    if ((size_t)nt > n_tasks) nt = (int)n_tasks;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, pool_worker, &tp);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
}

/* with_middle_bit, neighbors_last.rs:288-300 */
static void with_middle_bit(size_t k, int shift, size_t *k0, size_t *k1) {
    size_t ms = k >> (shift - 1);
    size_t ls = k & (((size_t)1 << shift) - 1);
    *k0 = ls | ((ms & ~(size_t)1) << shift);
    *k1 = ls | ((ms | 1) << shift);
}

typedef struct {
    const bo_domain *d;
    u128 *data;
    int log_d, layer, log_num_shares;
    int lr_start, lr_end, log_d_chunk;
} nl_ctx;

/* forward_shared_layer task body (neighbors_last.rs:235-279) */
static void nl_shared_task(void *p, size_t k) {
    nl_ctx *c = (nl_ctx *)p;
    int log_num_chunks = c->log_num_shares + 1;
    int log_chunk_len = c->log_d - log_num_chunks;
    int shift = c->log_num_shares - c->layer;
    size_t c0, c1;
    with_middle_bit(k, shift, &c0, &c1);
    size_t block = c0 >> (log_num_chunks - c->layer);
    u128 t = dom_twiddle(c->d, c->layer, block);
    u128 *p0 = c->data + (c0 << log_chunk_len), *p1 = c->data + (c1 << log_chunk_len);
    size_t n = (size_t)1 << log_chunk_len;
    for (size_t i = 0; i < n; i++) {
        u128 u = p0[i] ^ gf_mul(p1[i], t);
        p0[i] = u;
        p1[i] ^= u;
    }
}

static void nl_chunk_task(void *p, size_t block) {
    nl_ctx *c = (nl_ctx *)p;
    nl_depth_first(c->d, c->data + (block << c->log_d_chunk), c->log_d_chunk, c->layer, block,
                   c->lr_start, c->lr_end);
}

/* NeighborsLastMultiThread::forward_transform (neighbors_last.rs:464-526) with scalar packing */
int bo_ntt_forward_mt(const bo_domain *d, uint64_t *data64, int log_d, int se, int sl,
                      int log_num_shares, int nthreads) {
    if (ntt_check(d, log_d, se, sl)) return 1;
    u128 *data = (u128 *)data64;
    if (log_d <= 0) return bo_ntt_forward_ref(d, data64, log_d, se, sl);
    int max_shares = log_d - 1;
    int actual = log_num_shares < max_shares ? log_num_shares : max_shares;
    int last_layer = log_d - sl;
    int shared_end = actual < last_layer ? actual : last_layer;
    int ind_start = actual > se ? actual : se;
    nl_ctx c = {d, data, log_d, 0, actual, 0, 0, 0};
    for (int layer = se; layer < shared_end; layer++) {
        c.layer = layer;
        par_for(nthreads, (size_t)1 << actual, nl_shared_task, &c);
    }
    int layer = ind_start < max_shares ? ind_start : max_shares;
    c.layer = layer;
    c.log_d_chunk = log_d - layer;
    c.lr_start = ind_start;
    c.lr_end = last_layer;
    par_for(nthreads, (size_t)1 << layer, nl_chunk_task, &c);
    return 0;
}

/* ===================================================================== */
/* RS encode: verifier/math/src/reed_solomon.rs:128-201                    */
/* ===================================================================== */
static inline size_t reverse_bits(size_t x, int bits) {
    /* bit_reverse.rs:19-21 */
    if (bits == 0) return 0;
    uint64_t r = rev64((uint64_t)x);
    return (size_t)(r >> (64 - bits));
}

int bo_encode_batch(const bo_domain *d, const uint64_t *msg, int log_dim, int lir, int lbs,
                    uint64_t *out, int nthreads) {
    /* encode_batch requires ntt.subspace(log_len) == code subspace; the caller
       passes the domain whose subspace(log_dim+lir) is the code's. */
    if (log_dim + lir > d->dim) return 3; /* SubspaceDimensionMismatch */
    int m = log_dim + lbs;
    size_t n = (size_t)1 << m;
    u128 *o = (u128 *)out;
    const u128 *in = (const u128 *)msg;
    for (size_t i = 0; i < n; i++) o[i] = in[reverse_bits(i, m)];
    for (size_t r = 1; r < ((size_t)1 << lir); r++) memcpy(o + r * n, o, n * sizeof(u128));
    int log_out = m + lir;
    if (nthreads > 1) {
        int lns = 0;
        while ((1 << (lns + 1)) <= nthreads) lns++;
        return bo_ntt_forward_mt(d, out, log_out, lir, lbs, lns, nthreads);
    }
    return bo_ntt_forward_ref(d, out, log_out, lir, lbs);
}

/* ===================================================================== */
/* SHA-256 (FIPS 180-4), the third-party sha2 crate's algorithm            */
/* ===================================================================== */
static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
static const uint32_t IV256[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                  0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

#define ROR(x, n) (((x) >> (n)) | ((x) << (32 - (n))))

static void sha256_block(uint32_t st8[8], const uint8_t *blk) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++)
        w[i] = ((uint32_t)blk[4 * i] << 24) | ((uint32_t)blk[4 * i + 1] << 16) |
               ((uint32_t)blk[4 * i + 2] << 8) | blk[4 * i + 3];
    for (int i = 16; i < 64; i++) {
        uint32_t s0 = ROR(w[i - 15], 7) ^ ROR(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = ROR(w[i - 2], 17) ^ ROR(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = st8[0], b = st8[1], c = st8[2], d = st8[3], e = st8[4], f = st8[5], g = st8[6], h = st8[7];
    for (int i = 0; i < 64; i++) {
        uint32_t S1 = ROR(e, 6) ^ ROR(e, 11) ^ ROR(e, 25);
        uint32_t ch = (e & f) ^ (~e & g);
        uint32_t t1 = h + S1 + ch + K256[i] + w[i];
        uint32_t S0 = ROR(a, 2) ^ ROR(a, 13) ^ ROR(a, 22);
        uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
        uint32_t t2 = S0 + mj;
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    st8[0] += a; st8[1] += b; st8[2] += c; st8[3] += d;
    st8[4] += e; st8[5] += f; st8[6] += g; st8[7] += h;
}

typedef struct {
    uint32_t h[8];
    uint8_t buf[64];
    size_t buflen;
    uint64_t total;
} sha_ctx;

static void sha_init(sha_ctx *c) {
    memcpy(c->h, IV256, sizeof(IV256));
    c->buflen = 0;
    c->total = 0;
}
static void sha_update(sha_ctx *c, const uint8_t *p, size_t n) {
    c->total += n;
    while (n) {
        size_t take = 64 - c->buflen;
        if (take > n) take = n;
        memcpy(c->buf + c->buflen, p, take);
        c->buflen += take;
        p += take;
        n -= take;
        if (c->buflen == 64) {
            sha256_block(c->h, c->buf);
            c->buflen = 0;
        }
    }
}
static void sha_final(sha_ctx *c, uint8_t out[32]) {
    uint64_t bits = c->total * 8;
    uint8_t pad = 0x80;
    sha_update(c, &pad, 1);
    uint8_t z = 0;
    while (c->buflen != 56) sha_update(c, &z, 1);
    uint8_t len[8];
    for (int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (56 - 8 * i));
    sha_update(c, len, 8);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(c->h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(c->h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(c->h[i] >> 8);
        out[4 * i + 3] = (uint8_t)c->h[i];
    }
}

void bo_sha256(const uint8_t *data, size_t len, uint8_t out[32]) {
    sha_ctx c;
    sha_init(&c);
    sha_update(&c, data, len);
    sha_final(&c, out);
}

void bo_sha256_compress_block(const uint32_t iv[8], const uint8_t block[64], uint32_t state_out[8]) {
    memcpy(state_out, iv, 32);
    sha256_block(state_out, block);
}

/* Sha256Compression (verifier/verifier/src/hash/compress.rs:33-57): the IV is
 * SHA-256("BINIUS SHA-256 COMPRESS") read as native little-endian u32 words;
 * compress256 runs once on left||right (no padding); the output is the state
 * words written little-endian. */
static uint32_t binius_iv[8];
static pthread_once_t binius_iv_once = PTHREAD_ONCE_INIT;
static void binius_iv_init(void) {
    static const char tag[] = "BINIUS SHA-256 COMPRESS";
    uint8_t dg[32];
    bo_sha256((const uint8_t *)tag, sizeof(tag) - 1, dg);
    for (int k = 0; k < 8; k++)
        binius_iv[k] = (uint32_t)dg[4 * k] | ((uint32_t)dg[4 * k + 1] << 8) |
                       ((uint32_t)dg[4 * k + 2] << 16) | ((uint32_t)dg[4 * k + 3] << 24);
}

void bo_binius_compress(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    pthread_once(&binius_iv_once, binius_iv_init);
    uint8_t blk[64];
    memcpy(blk, left, 32);
    memcpy(blk + 32, right, 32);
    uint32_t s[8];
    memcpy(s, binius_iv, 32);
    sha256_block(s, blk);
    for (int k = 0; k < 8; k++) {
        out[4 * k] = (uint8_t)s[k];
        out[4 * k + 1] = (uint8_t)(s[k] >> 8);
        out[4 * k + 2] = (uint8_t)(s[k] >> 16);
        out[4 * k + 3] = (uint8_t)(s[k] >> 24);
    }
}

/* ===================================================================== */
/* Merkle: prover/prover/src/merkle_tree/binary_merkle_tree.rs             */
/* ===================================================================== */
typedef struct {
    const uint8_t *elems; /* LE16 serialized elements, contiguous */
    size_t batch;
    uint8_t *tree;
    const uint8_t *prev;
    uint8_t *next;
} mk_ctx;

/* hash_leaves (binary_merkle_tree.rs:178-205) via the blanket ParallelDigest
 * impl (hash/parallel_digest.rs:145-173): leaf = SHA-256(LE16(e_0)||...). */
static void mk_leaf_task(void *p, size_t i) {
    mk_ctx *c = (mk_ctx *)p;
    bo_sha256(c->elems + i * c->batch * 16, c->batch * 16, c->tree + 32 * i);
}

/* ParallelCompressionAdaptor::parallel_compress (hash/parallel_compression.rs:79-88) */
static void mk_node_task(void *p, size_t i) {
    mk_ctx *c = (mk_ctx *)p;
    bo_binius_compress(c->prev + 64 * i, c->prev + 64 * i + 32, c->next + 32 * i);
}

/* build + build_from_iterator (binary_merkle_tree.rs:32-122) */
int bo_merkle_build(const uint64_t *elems, size_t n_elems, size_t batch, uint8_t *tree, int nthreads) {
    if (batch == 0 || n_elems % batch != 0) return 4;          /* IncorrectBatchSize */
    size_t n_leaves = n_elems / batch;
    if (n_leaves == 0 || (n_leaves & (n_leaves - 1))) return 5; /* PowerOfTwoLengthRequired */
    mk_ctx c = {(const uint8_t *)elems, batch, tree, NULL, NULL};
    par_for(nthreads, n_leaves, mk_leaf_task, &c);
    uint8_t *prev = tree;
    size_t len = n_leaves;
    while (len > 1) {
        c.prev = prev;
        c.next = prev + 32 * len;
        par_for(len / 2 >= 1024 ? nthreads : 1, len / 2, mk_node_task, &c);
        prev = c.next;
        len /= 2;
    }
    return 0;
}

/* BinaryMerkleTree::branch (binary_merkle_tree.rs:154-169) */
int bo_merkle_branch(const uint8_t *tree, int log_len, size_t index, int layer_depth, uint8_t *out) {
    if (index >= ((size_t)1 << log_len) || layer_depth > log_len) return 8;
    for (int j = 0; j < log_len - layer_depth; j++) {
        size_t node = ((((size_t)1 << j) - 1) << (log_len + 1 - j)) | ((index >> j) ^ 1);
        memcpy(out + 32 * j, tree + 32 * node, 32);
    }
    return 0;
}

/* ===================================================================== */
/* FRI: verifier/verifier/src/fri/{common,fold}.rs and prover/prover/src/fri/ */
/* ===================================================================== */
static int log2_ceil(size_t x) {
    int r = 0;
    while (((size_t)1 << r) < x) r++;
    return r;
}

/* calculate_n_test_queries, fri/common.rs:182-186 */
size_t bo_n_test_queries(int security_bits, int lir) {
    double rate = pow(2.0, -(double)lir);
    double per_query_err = 0.5 * (1.0 + rate);
    return (size_t)ceil((double)security_bits / -log2(per_query_err));
}

/* estimate_optimal_arity, fri/common.rs:191-219 */
int bo_estimate_optimal_arity(int L, int digest_size, int field_size) {
    int have_old = 0, last = 1;
    size_t old = 0;
    for (int arity = 1; arity <= L; arity++) {
        size_t est = (((size_t)(L / 2) * (size_t)digest_size + ((size_t)1 << arity) * (size_t)field_size) *
                      (size_t)(L - arity)) / (size_t)arity;
        if (have_old && est > old) break;
        have_old = 1;
        old = est;
        last = arity;
    }
    return last;
}

/* Verifier::setup (verify.rs:61-106, LOG_WORDS_PER_ELEM handled by caller)
 * -> FRIParams::choose_with_constant_fold_arity (fri/common.rs:64-115). */
int bo_fri_params_setup(int log_msg_len, int lir, int security_bits, int *out) {
    int log_code_len = log_msg_len + lir;
    int arity = bo_estimate_optimal_arity(log_code_len, 32, 16);
    if (arity <= 0) return 1;
    int lbs = log_msg_len < arity ? log_msg_len : arity;
    int log_dim = log_msg_len - lbs;
    int rs_log_len = log_dim + lir;
    size_t nq = bo_n_test_queries(security_bits, lir);
    int cap = log2_ceil(nq);
    int term = cap;
    if (term < lir) term = lir;
    if (term > rs_log_len) term = rs_log_len;
    int q = (rs_log_len - term) / arity, rem = (rs_log_len - term) % arity;
    int n = 0;
    for (int i = 0; i < q; i++) out[5 + n++] = arity;
    if (rem) out[5 + n++] = rem;
    int sum = 0;
    for (int i = 0; i < n; i++) sum += out[5 + i];
    if (sum > log_dim) return 2; /* InvalidFoldAritySequence */
    out[0] = lbs;
    out[1] = log_dim;
    out[2] = rs_log_len;
    out[3] = (int)nq;
    out[4] = n;
    return 0;
}

/* eq_ind_partial_eval (multilinear/eq.rs:123-131, tensor_prod_eq_ind :39-71):
 * bit i of the index <-> point[i]. */
static void eq_ind(const u128 *r, int n, u128 *out) {
    out[0] = 1;
    size_t len = 1;
    for (int i = 0; i < n; i++) {
        for (size_t j = 0; j < len; j++) {
            u128 prod = gf_mul(out[j], r[i]);
            out[len + j] = prod;
            out[j] ^= prod;
        }
        len <<= 1;
    }
}
void bo_eq_ind(const uint64_t *point, int n, uint64_t *out) {
    eq_ind((const u128 *)point, n, (u128 *)out);
}

typedef struct {
    const u128 *cw;
    u128 *out;
    const u128 *tensor;
    int log_batch;
    const bo_domain *d;
    int log_len, n_chal;
    const u128 *chal;
} fold_ctx;

#define FOLD_GRAIN 4096

/* fold_interleaved (prover fri/fold.rs:260-282) = per-chunk inner_product_packed
 * (inner_product.rs:73-91) with the eq tensor. */
static void fold_il_task(void *p, size_t g) {
    fold_ctx *c = (fold_ctx *)p;
    size_t bs = (size_t)1 << c->log_batch;
    for (size_t i = g * FOLD_GRAIN; i < (g + 1) * FOLD_GRAIN; i++) {
        u128 acc = 0;
        for (size_t j = 0; j < bs; j++) acc ^= gf_mul(c->cw[i * bs + j], c->tensor[j]);
        c->out[i] = acc;
    }
}

int bo_fold_interleaved(const uint64_t *cw, int log_len, int log_batch, const uint64_t *chal,
                        uint64_t *out, int nthreads) {
    u128 tensor[1 << 12];
    if (log_batch > 12) return 1;
    eq_ind((const u128 *)chal, log_batch, tensor);
    fold_ctx c = {(const u128 *)cw, (u128 *)out, tensor, log_batch, NULL, 0, 0, NULL};
    size_t n = (size_t)1 << log_len;
    if (n < FOLD_GRAIN) {
        size_t bs = (size_t)1 << log_batch;
        for (size_t i = 0; i < n; i++) {
            u128 acc = 0;
            for (size_t j = 0; j < bs; j++) acc ^= gf_mul(c.cw[i * bs + j], tensor[j]);
            c.out[i] = acc;
        }
        return 0;
    }
    par_for(nthreads, n / FOLD_GRAIN, fold_il_task, &c);
    return 0;
}

/* fold_pair (verifier fri/fold.rs:21-33): inverse butterfly, then
 * extrapolate_line_packed (line.rs:24-29) x0 + (x1 - x0) * z. */
static inline u128 fold_pair(const bo_domain *d, int round, size_t index, u128 u, u128 v, u128 r) {
    u128 t = dom_twiddle(d, round - 1, index);
    v ^= u;
    u ^= gf_mul(v, t);
    return u ^ gf_mul(v ^ u, r);
}

/* fold_chunk (verifier fri/fold.rs:61-99) */
static u128 fold_chunk(const bo_domain *d, int log_len, size_t chunk_index, u128 *vals,
                       const u128 *chal, int n_chal) {
    int log_size = n_chal;
    for (int k = 0; k < n_chal; k++) {
        for (size_t off = 0; off < ((size_t)1 << (log_size - 1)); off++) {
            vals[off] = fold_pair(d, log_len, (chunk_index << (log_size - 1)) | off, vals[off << 1],
                                  vals[(off << 1) | 1], chal[k]);
        }
        log_len--;
        log_size--;
    }
    return vals[0];
}

void bo_fold_chunk(const bo_domain *d, int log_len, uint64_t chunk_index, const uint64_t *vals,
                   const uint64_t *chal, int n_chal, uint64_t out[2]) {
    u128 scratch[1 << 12];
    memcpy(scratch, vals, ((size_t)1 << n_chal) * sizeof(u128));
    st(out, fold_chunk(d, log_len, chunk_index, scratch, (const u128 *)chal, n_chal));
}

static void fold_cw_task(void *p, size_t g) {
    fold_ctx *c = (fold_ctx *)p;
    size_t cs = (size_t)1 << c->n_chal;
    size_t n_out = (size_t)1 << (c->log_len - c->n_chal);
    size_t lo = g * FOLD_GRAIN, hi = lo + FOLD_GRAIN;
    if (hi > n_out) hi = n_out;
    u128 scratch[1 << 12];
    for (size_t i = lo; i < hi; i++) {
        memcpy(scratch, c->cw + i * cs, cs * sizeof(u128));
        c->out[i] = fold_chunk(c->d, c->log_len, i, scratch, c->chal, c->n_chal);
    }
}

/* fold_codeword (prover fri/fold.rs:297-318) */
int bo_fold_codeword(const bo_domain *d, const uint64_t *cw, int log_len, const uint64_t *chal,
                     int n_chal, uint64_t *out, int nthreads) {
    if (n_chal > log_len || n_chal > 12 || log_len > d->dim) return 1;
    fold_ctx c = {(const u128 *)cw, (u128 *)out, NULL, 0, d, log_len, n_chal, (const u128 *)chal};
    size_t n_out = (size_t)1 << (log_len - n_chal);
    par_for(nthreads, (n_out + FOLD_GRAIN - 1) / FOLD_GRAIN, fold_cw_task, &c);
    return 0;
}

/* evaluate (multilinear/evaluate.rs:29-...): sum_i evals[i] * eq(i, point) */
void bo_evaluate_mle(const uint64_t *evals, int log_n, const uint64_t *point, uint64_t out[2]) {
    size_t n = (size_t)1 << log_n;
    u128 *t = (u128 *)malloc(n * sizeof(u128));
    eq_ind((const u128 *)point, log_n, t);
    u128 acc = 0;
    for (size_t i = 0; i < n; i++) acc ^= gf_mul(((const u128 *)evals)[i], t[i]);
    free(t);
    st(out, acc);
}

/* pack_witness (prover/prover/src/prove.rs:286-327): B128 = w[2i+1] << 64 | w[2i] */
void bo_pack_witness(const uint64_t *words, size_t n_elems, uint64_t *out) {
    for (size_t i = 0; i < n_elems; i++) {
        out[2 * i] = words[2 * i];
        out[2 * i + 1] = words[2 * i + 1];
    }
}

/* ===================================================================== */
/* Ring switch (prover/prover/src/ring_switch.rs)                          */
/* ===================================================================== */

/* expand_subset_sums_array (verifier/math/src/span.rs:43-57): out[m] = sum of
 * elems[i] over the set bits i of m. */
static void subset_sums(const u128 *elems, int n, u128 *out) {
    out[0] = 0;
    for (int i = 0; i < n; i++)
        for (size_t m = 0; m < ((size_t)1 << i); m++) out[m | ((size_t)1 << i)] = out[m] ^ elems[i];
}

/* square_transpose_const_size::<_, 2, 4> (ring_switch.rs:267-320) on four
 * elements viewed as 128 B1 scalars: within each nibble q, bit q*4+r of element j
 * becomes bit q*4+j of element r. */
static void transpose4_nibbles(u128 m[4]) {
    u128 t[4] = {0, 0, 0, 0};
    for (int j = 0; j < 4; j++)
        for (int q = 0; q < 32; q++)
            for (int r = 0; r < 4; r++)
                if ((m[j] >> (4 * q + r)) & 1) t[r] |= (u128)1 << (4 * q + j);
    for (int j = 0; j < 4; j++) m[j] = t[j];
}

typedef struct {
    const u128 *mat, *vec;
    size_t n, chunk;
    u128 *partial; /* per task: 128 accumulators */
    int variant;
} rs1b_ctx;

static void rs1b_task(void *p, size_t task) {
    rs1b_ctx *c = (rs1b_ctx *)p;
    u128 *acc = c->partial + 128 * task;
    for (int b = 0; b < 128; b++) acc[b] = 0;
    const size_t r0 = task * c->chunk, r1 = r0 + c->chunk < c->n ? r0 + c->chunk : c->n;
    if (c->variant == 1) {
        for (size_t r = r0; r < r1; r++)
            for (int b = 0; b < 128; b++)
                if ((c->mat[r] >> b) & 1) acc[b] ^= c->vec[r];
        return;
    }
    /* ring_switch.rs:216-257: chunks of 4 rows, subset sums of the vec chunk,
     * transpose the mat chunk, then per byte i of transposed element j the low and
     * high nibbles index the sums for output bits 8i+j and 8i+4+j. */
    for (size_t r = r0; r < r1; r += 4) {
        u128 v4[4] = {0, 0, 0, 0}, m4[4] = {0, 0, 0, 0}, look[16];
        for (int t = 0; t < 4 && r + t < r1; t++) {
            v4[t] = c->vec[r + t];
            m4[t] = c->mat[r + t];
        }
        subset_sums(v4, 4, look);
        transpose4_nibbles(m4);
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 16; i++) {
                const unsigned byte = (unsigned)(m4[j] >> (8 * i)) & 0xFF;
                acc[(i << 3) | j] ^= look[byte & 0x0F];
                acc[(i << 3) | 4 | j] ^= look[byte >> 4];
            }
    }
}

void bo_rs_fold_1b_rows(const uint64_t *mat, const uint64_t *vec, int log_len, uint64_t *out, int variant,
                        int nthreads) {
    const size_t n = (size_t)1 << log_len;
    const size_t chunk = n < 4096 ? n : 4096;
    const size_t tasks = (n + chunk - 1) / chunk;
    u128 *partial = (u128 *)malloc(sizeof(u128) * 128 * tasks);
    rs1b_ctx c = {(const u128 *)mat, (const u128 *)vec, n, chunk, partial, variant};
    par_for(nthreads, tasks, rs1b_task, &c);
    u128 *o = (u128 *)out;
    for (int b = 0; b < 128; b++) {
        u128 a = 0;
        for (size_t t = 0; t < tasks; t++) a ^= partial[128 * t + b];
        o[b] = a;
    }
    free(partial);
}

typedef struct {
    u128 *elems;
    size_t n, chunk;
    const u128 *vec;
    const u128 *tables; /* 16 x 256 */
    int variant;
} rsel_ctx;

static void rsel_task(void *p, size_t task) {
    rsel_ctx *c = (rsel_ctx *)p;
    const size_t x0 = task * c->chunk, x1 = x0 + c->chunk < c->n ? x0 + c->chunk : c->n;
    for (size_t x = x0; x < x1; x++) {
        const u128 e = c->elems[x];
        u128 acc = 0;
        if (c->variant == 1) {
            for (int b = 0; b < 128; b++)
                if ((e >> b) & 1) acc ^= c->vec[b];
        } else {
            for (int i = 0; i < 16; i++) acc ^= c->tables[256 * i + ((unsigned)(e >> (8 * i)) & 0xFF)];
        }
        c->elems[x] = acc;
    }
}

void bo_rs_fold_elems(uint64_t *elems, size_t n, const uint64_t *vec, int variant, int nthreads) {
    const u128 *v = (const u128 *)vec;
    u128 *tables = (u128 *)malloc(sizeof(u128) * 16 * 256);
    for (int i = 0; i < 16; i++) subset_sums(v + 8 * i, 8, tables + 256 * i); /* ring_switch.rs:144-157 */
    const size_t chunk = 1 << 14;
    rsel_ctx c = {(u128 *)elems, n, chunk, v, tables, variant};
    par_for(nthreads, (n + chunk - 1) / chunk, rsel_task, &c);
    free(tables);
}

/* ===================================================================== */
/* Transcript: HasherChallenger<Sha256> (fiat_shamir/hasher_challenger.rs) */
/* + ProverTranscript tape (transcript.rs:205-420)                         */
/* ===================================================================== */
struct bo_transcript {
    sha_ctx hasher;
    int sampler; /* 1 = Sampler, 0 = Observer */
    uint8_t buffer[32];
    size_t index;
    uint8_t *tape;
    size_t tape_len, tape_cap;
};

/* HasherChallenger::default -> new(SHA256("")) (hasher_challenger.rs:42-62) */
bo_transcript *bo_transcript_new(void) {
    bo_transcript *t = (bo_transcript *)calloc(1, sizeof(bo_transcript));
    uint8_t e[32];
    bo_sha256(NULL, 0, e);
    sha_init(&t->hasher);
    sha_update(&t->hasher, e, 32);
    t->sampler = 1;
    memcpy(t->buffer, e, 32);
    t->index = 0;
    return t;
}
void bo_transcript_free(bo_transcript *t) {
    if (!t) return;
    free(t->tape);
    free(t);
}

static void tape_append(bo_transcript *t, const uint8_t *b, size_t n) {
    if (n == 0) return;
    if (t->tape_len + n > t->tape_cap) {
        size_t cap = t->tape_cap ? t->tape_cap : 1024;
        while (cap < t->tape_len + n) cap *= 2;
        t->tape = (uint8_t *)realloc(t->tape, cap);
        t->tape_cap = cap;
    }
    memcpy(t->tape + t->tape_len, b, n);
    t->tape_len += n;
}

/* Challenger::observer(): Sampler -> Observer absorbs LE64(index) (:95-104) */
void bo_transcript_observe(bo_transcript *t, const uint8_t *bytes, size_t n) {
    if (t->sampler) {
        uint8_t le[8];
        for (int i = 0; i < 8; i++) le[i] = (uint8_t)((uint64_t)t->index >> (8 * i));
        sha_update(&t->hasher, le, 8);
        t->sampler = 0;
    }
    sha_update(&t->hasher, bytes, n);
}

void bo_transcript_message(bo_transcript *t, const uint8_t *bytes, size_t n) {
    tape_append(t, bytes, n);
    bo_transcript_observe(t, bytes, n);
}

void bo_transcript_decommit(bo_transcript *t, const uint8_t *bytes, size_t n) { tape_append(t, bytes, n); }

/* Observer -> Sampler sets index = 32 (:124-133); reads refill with
 * d = finalize_reset(); hasher.update(d) (:111-121, :146-165). */
void bo_transcript_sample(bo_transcript *t, uint8_t *out, size_t n) {
    if (!t->sampler) {
        t->sampler = 1;
        t->index = 32;
    }
    while (n) {
        if (t->index == 32) {
            uint8_t dg[32];
            sha_ctx tmp = t->hasher;
            sha_final(&tmp, dg);
            sha_init(&t->hasher);
            sha_update(&t->hasher, dg, 32);
            memcpy(t->buffer, dg, 32);
            t->index = 0;
        }
        size_t take = 32 - t->index;
        if (take > n) take = n;
        memcpy(out, t->buffer + t->index, take);
        t->index += take;
        out += take;
        n -= take;
    }
}

/* sample_bits_reader (fiat_shamir/sampling.rs:34-50) */
uint32_t bo_transcript_sample_bits(bo_transcript *t, int bits) {
    if (bits > 32) bits = 32;
    uint8_t b[4];
    bo_transcript_sample(t, b, 4);
    uint32_t v = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    uint32_t mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
    return v & mask;
}

size_t bo_transcript_tape(const bo_transcript *t, uint8_t *out, size_t cap) {
    if (out && t->tape_len && cap >= t->tape_len) memcpy(out, t->tape, t->tape_len);
    return t->tape_len;
}

/* ===================================================================== */
/* FRI-only prover flow                                                    */
/* ===================================================================== */
typedef struct {
    u128 *cw;
    int log_len;
    uint8_t *tree;
    int tree_log_len;
} committed;

/* The Merkle scheme of the commitments: SHA-256 leaves with Sha256Compression
 * nodes (both NULL), or Vision leaves with VisionCompression nodes
 * (prover/examples/src/lib.rs setup_vision4). */
typedef struct bo_vision bo_vision;
typedef struct {
    const bo_vision *leaf, *node;
} mk_scheme;
int bo_merkle_build_vision(const bo_vision *leaf, const bo_vision *node4, const uint64_t *elems, size_t n_elems,
                           size_t batch, uint8_t *tree, int nthreads);

static int commit_vec(const u128 *v, int log_len, int log_batch, committed *o, int nthreads, mk_scheme sc) {
    o->log_len = log_len;
    o->tree_log_len = log_len - log_batch;
    size_t n_leaves = (size_t)1 << o->tree_log_len;
    o->tree = (uint8_t *)malloc(32 * (2 * n_leaves - 1));
    if (sc.leaf)
        return bo_merkle_build_vision(sc.leaf, sc.node, (const uint64_t *)v, (size_t)1 << log_len,
                                      (size_t)1 << log_batch, o->tree, nthreads);
    return bo_merkle_build((const uint64_t *)v, (size_t)1 << log_len, (size_t)1 << log_batch, o->tree,
                           nthreads);
}

static const uint8_t *tree_root(const committed *c) {
    size_t n = (size_t)1 << c->tree_log_len;
    return c->tree + 32 * (2 * n - 2);
}

/* BinaryMerkleTree::layer (binary_merkle_tree.rs:142-149) */
static const uint8_t *tree_layer(const committed *c, int depth) {
    size_t total = 2 * ((size_t)1 << c->tree_log_len) - 1;
    size_t start = total + 1 - ((size_t)1 << (depth + 1));
    return c->tree + 32 * start;
}

static void write_scalars(bo_transcript *t, const u128 *v, size_t n) {
    bo_transcript_decommit(t, (const uint8_t *)v, n * 16);
}

/* prove_coset_opening (fri/query.rs:112-134) + BinaryMerkleTreeProver::prove_opening
 * (merkle_tree/prover.rs:77-90): scalars, empty salt, branch digests. */
static void prove_coset_opening(bo_transcript *t, const committed *c, size_t coset_index,
                                int log_coset, int depth) {
    write_scalars(t, c->cw + (coset_index << log_coset), (size_t)1 << log_coset);
    uint8_t br[32 * 64];
    bo_merkle_branch(c->tree, c->tree_log_len, coset_index, depth, br);
    bo_transcript_decommit(t, br, 32 * (size_t)(c->tree_log_len - depth));
}

/* FRIFoldProver state (prover fri/fold.rs:34-82): the committed codeword c0
 * (borrowed), the folded oracles committed so far, unprocessed challenges. */
typedef struct {
    const bo_domain *d;
    int log_dim, lir, lbs, n_ar, n_queries, nthreads;
    int arities[64];
    int rs_log_len, log_len, n_final, n_rounds;
    committed c0;
    committed rc[65];
    int n_rc;
    u128 unproc[64];
    int n_unproc, curr_round, next_commit, has_next;
    mk_scheme sc;
} fri_st;

/* FRIFoldProver::new (fold.rs:57-82); c0 is borrowed. */
static int fri_init(fri_st *f, const bo_domain *d, int log_dim, int lir, int lbs, const int *arities, int n_ar,
                    int n_queries, const committed *c0, int nthreads, mk_scheme sc) {
    memset(f, 0, sizeof(*f));
    if (n_ar > 63) return 1;
    int sum = 0;
    for (int i = 0; i < n_ar; i++) sum += arities[i];
    if (sum > log_dim) return 2;
    f->d = d;
    f->log_dim = log_dim;
    f->lir = lir;
    f->lbs = lbs;
    f->n_ar = n_ar;
    memcpy(f->arities, arities, sizeof(int) * (size_t)n_ar);
    f->n_queries = n_queries;
    f->nthreads = nthreads;
    f->rs_log_len = log_dim + lir;
    f->log_len = f->rs_log_len + lbs;
    f->n_final = log_dim - sum;
    f->n_rounds = log_dim + lbs;
    f->c0 = *c0;
    f->next_commit = lbs;
    f->has_next = 1;
    f->sc = sc;
    return 0;
}

static void fri_free(fri_st *f) {
    for (int i = 0; i < f->n_rc; i++) {
        free(f->rc[i].cw);
        free(f->rc[i].tree);
    }
    f->n_rc = 0;
}

/* receive_challenge (fold.rs:107-110) */
static void fri_receive(fri_st *f, u128 ch) {
    f->unproc[f->n_unproc++] = ch;
    f->curr_round++;
}

/* execute_fold_round (fold.rs:117-175): returns 1 with root_out set when the
 * round commits a folded oracle, 0 when it does not, < 0 on error. */
static int fri_execute(fri_st *f, uint8_t root_out[32]) {
    if (!(f->has_next && f->next_commit == f->curr_round)) return 0;
    committed *nc = &f->rc[f->n_rc];
    if (f->n_rc == 0) {
        nc->log_len = f->rs_log_len;
        nc->cw = (u128 *)malloc(((size_t)1 << f->rs_log_len) * sizeof(u128));
        bo_fold_interleaved((const uint64_t *)f->c0.cw, f->rs_log_len, f->lbs, (const uint64_t *)f->unproc,
                            (uint64_t *)nc->cw, f->nthreads);
    } else {
        committed *prev = &f->rc[f->n_rc - 1];
        nc->log_len = prev->log_len - f->n_unproc;
        nc->cw = (u128 *)malloc(((size_t)1 << nc->log_len) * sizeof(u128));
        bo_fold_codeword(f->d, (const uint64_t *)prev->cw, prev->log_len, (const uint64_t *)f->unproc,
                         f->n_unproc, (uint64_t *)nc->cw, f->nthreads);
    }
    f->n_unproc = 0;
    int has_arity = f->n_rc < f->n_ar;
    int log_coset = has_arity ? f->arities[f->n_rc] : f->n_final;
    if (commit_vec(nc->cw, nc->log_len, log_coset, nc, f->nthreads, f->sc)) return -1;
    if (has_arity) f->next_commit += f->arities[f->n_rc];
    else f->has_next = 0;
    f->n_rc++;
    memcpy(root_out, tree_root(nc), 32);
    return 1;
}

/* finish_proof (fold.rs:221-244): terminal codeword, optimal layers, queries. */
static int fri_finish(fri_st *f, bo_transcript *t, uint64_t *terminal_out) {
    if (f->curr_round != f->n_rounds) return 6; /* EarlyProverFinish */
    const committed *term = f->n_rc ? &f->rc[f->n_rc - 1] : &f->c0;
    size_t term_len = (size_t)1 << term->log_len;
    write_scalars(t, term->cw, term_len);
    if (terminal_out) memcpy(terminal_out, term->cw, term_len * 16);

    /* vcs_optimal_layers (fri/query.rs:93-109) with
       vcs_optimal_layers_depths_iter (verifier fri/common.rs:154-168) */
    int cap = log2_ceil((size_t)f->n_queries);
    int depths[65];
    int log_n_cosets = f->log_len;
    for (int i = 0; i <= f->n_ar; i++) {
        int a = i == 0 ? f->lbs : f->arities[i - 1];
        log_n_cosets -= a;
        depths[i] = cap < log_n_cosets ? cap : log_n_cosets;
    }
    for (int i = 0; i <= f->n_ar; i++) {
        const committed *cc = i == 0 ? &f->c0 : &f->rc[i - 1];
        bo_transcript_decommit(t, tree_layer(cc, depths[i]), 32 * ((size_t)1 << depths[i]));
    }
    for (int q = 0; q < f->n_queries; q++) {
        size_t index = bo_transcript_sample_bits(t, f->rs_log_len);
        prove_coset_opening(t, &f->c0, index, f->lbs, depths[0]);
        for (int i = 0; i < f->n_ar; i++) {
            index >>= f->arities[i];
            prove_coset_opening(t, &f->rc[i], index, f->arities[i], depths[i + 1]);
        }
    }
    return 0;
}

int bo_fri_prove_ex(const bo_domain *d, const uint64_t *msg, int log_dim, int lir, int lbs,
                    const int *arities, int n_ar, int n_queries, const uint64_t *chal_in, uint8_t *tape,
                    size_t tape_cap, size_t *tape_len, uint8_t *roots_out, uint64_t *chal_out,
                    uint64_t *terminal_out, int nthreads, const bo_vision *leaf, const bo_vision *node) {
    const mk_scheme sc = {leaf, node};
    int log_len = log_dim + lir + lbs;
    committed c0;
    c0.cw = (u128 *)malloc(((size_t)1 << log_len) * sizeof(u128));
    int rc = bo_encode_batch(d, msg, log_dim, lir, lbs, (uint64_t *)c0.cw, nthreads);
    if (rc) return rc;
    rc = commit_vec(c0.cw, log_len, lbs, &c0, nthreads, sc);
    if (rc) return rc;
    fri_st f;
    rc = fri_init(&f, d, log_dim, lir, lbs, arities, n_ar, n_queries, &c0, nthreads, sc);
    if (rc) return rc;

    bo_transcript *t = bo_transcript_new();
    bo_transcript_message(t, tree_root(&c0), 32);
    memcpy(roots_out, tree_root(&c0), 32);
    int n_roots = 1;
    for (int step = 0; step <= f.n_rounds; step++) {
        if (step > 0) {
            u128 ch;
            if (chal_in) ch = ld(chal_in + 2 * (step - 1));
            else bo_transcript_sample(t, (uint8_t *)&ch, 16);
            st(chal_out + 2 * (step - 1), ch);
            fri_receive(&f, ch);
        }
        uint8_t root[32];
        int c = fri_execute(&f, root);
        if (c < 0) return 5;
        if (c) {
            bo_transcript_message(t, root, 32);
            memcpy(roots_out + 32 * (size_t)n_roots++, root, 32);
        }
    }
    if (f.n_rc != n_ar + 1) return 6; /* EarlyProverFinish */
    rc = fri_finish(&f, t, terminal_out);
    if (rc) return rc;

    *tape_len = t->tape_len;
    int ret = 0;
    if (tape) {
        if (tape_cap >= t->tape_len) memcpy(tape, t->tape, t->tape_len);
        else ret = 9;
    }
    bo_transcript_free(t);
    free(c0.cw);
    free(c0.tree);
    fri_free(&f);
    return ret;
}

int bo_fri_prove(const bo_domain *d, const uint64_t *msg, int log_dim, int lir, int lbs, const int *arities, int n_ar,
                 int n_queries, const uint64_t *chal_in, uint8_t *tape, size_t tape_cap, size_t *tape_len,
                 uint8_t *roots_out, uint64_t *chal_out, uint64_t *terminal_out, int nthreads) {
    return bo_fri_prove_ex(d, msg, log_dim, lir, lbs, arities, n_ar, n_queries, chal_in, tape, tape_cap, tape_len,
                           roots_out, chal_out, terminal_out, nthreads, NULL, NULL);
}

/* ===================================================================== */
/* Bivariate product sumcheck, BaseFold and the one-bit PCS opening        */
/* ===================================================================== */

/* BivariateProductSumcheckProver (prover protocols/sumcheck/bivariate_product.rs):
 * two multilinears folded in place, binding the highest variable first. */
typedef struct {
    u128 *a, *b;
    int n_vars, nthreads;
    int has_coeffs; /* RoundCoeffsOrSum: 1 = Coeffs, 0 = Sum */
    u128 sum, coeffs[3];
} bp_st;

typedef struct {
    const u128 *a, *b;
    size_t half;
    u128 *y1, *yinf; /* per task */
    u128 r;
    u128 *wa, *wb;
} bp_ctx;

#define BP_GRAIN 8192

static void bp_eval_task(void *p, size_t g) {
    bp_ctx *c = (bp_ctx *)p;
    size_t lo = g * BP_GRAIN, hi = lo + BP_GRAIN;
    if (hi > c->half) hi = c->half;
    u128 y1 = 0, yinf = 0;
    for (size_t i = lo; i < hi; i++) {
        u128 a0 = c->a[i], a1 = c->a[c->half + i], b0 = c->b[i], b1 = c->b[c->half + i];
        y1 ^= gf_mul(a1, b1);
        yinf ^= gf_mul(a0 ^ a1, b0 ^ b1);
    }
    c->y1[g] = y1;
    c->yinf[g] = yinf;
}

/* execute (bivariate_product.rs:43-82): F(1) and F(inf) summed over the low
 * half, then RoundEvals2::interpolate (round_evals.rs:73-78) with
 * calculate_round_coeffs_from_evals_2 (:126-140). */
static int bp_execute(bp_st *s, u128 coeffs[3]) {
    if (s->has_coeffs) return 1; /* ExpectedFold */
    if (s->n_vars == 0) return 2;
    size_t half = (size_t)1 << (s->n_vars - 1);
    size_t n_tasks = (half + BP_GRAIN - 1) / BP_GRAIN;
    u128 *y1 = (u128 *)calloc(n_tasks, sizeof(u128)), *yinf = (u128 *)calloc(n_tasks, sizeof(u128));
    bp_ctx c = {s->a, s->b, half, y1, yinf, 0, NULL, NULL};
    if (n_tasks == 1) bp_eval_task(&c, 0);
    else par_for(s->nthreads, n_tasks, bp_eval_task, &c);
    u128 s1 = 0, sinf = 0;
    for (size_t g = 0; g < n_tasks; g++) {
        s1 ^= y1[g];
        sinf ^= yinf[g];
    }
    free(y1);
    free(yinf);
    u128 y0 = s->sum ^ s1;
    s->coeffs[0] = y0;
    s->coeffs[2] = sinf;
    s->coeffs[1] = s1 ^ y0 ^ sinf;
    s->has_coeffs = 1;
    memcpy(coeffs, s->coeffs, sizeof(s->coeffs));
    return 0;
}

static void bp_fold_task(void *p, size_t g) {
    bp_ctx *c = (bp_ctx *)p;
    size_t lo = g * BP_GRAIN, hi = lo + BP_GRAIN;
    if (hi > c->half) hi = c->half;
    for (size_t i = lo; i < hi; i++) {
        c->wa[i] ^= gf_mul(c->r, c->wa[i] ^ c->wa[c->half + i]);
        c->wb[i] ^= gf_mul(c->r, c->wb[i] ^ c->wb[c->half + i]);
    }
}

/* evaluate_univariate (Horner) for RoundCoeffs::evaluate (verifier sumcheck/common.rs:22-24) */
static u128 poly_eval(const u128 *c, int n, u128 x) {
    u128 acc = 0;
    for (int i = n - 1; i >= 0; i--) acc = gf_mul(acc, x) ^ c[i];
    return acc;
}

/* fold (bivariate_product.rs:84-98) with fold_highest_var_inplace
 * (multilinear/fold.rs:14-29): zero += r * (one - zero). */
static int bp_fold(bp_st *s, u128 r) {
    if (!s->has_coeffs) return 1; /* ExpectedExecute */
    size_t half = (size_t)1 << (s->n_vars - 1);
    bp_ctx c = {NULL, NULL, half, NULL, NULL, r, s->a, s->b};
    size_t n_tasks = (half + BP_GRAIN - 1) / BP_GRAIN;
    if (n_tasks == 1) bp_fold_task(&c, 0);
    else par_for(s->nthreads, n_tasks, bp_fold_task, &c);
    s->n_vars--;
    s->sum = poly_eval(s->coeffs, 3, r);
    s->has_coeffs = 0;
    return 0;
}

/* prove_single (prover protocols/sumcheck/prove.rs:17-60) over a bivariate
 * product prover; a and b are folded in place (the prover owns them). */
int bo_sumcheck_prove(uint64_t *a, uint64_t *b, int n_vars, const uint64_t sum[2], bo_transcript *t,
                      uint64_t *challenges_out, uint64_t evals_out[4], int nthreads) {
    bp_st s = {(u128 *)a, (u128 *)b, n_vars, nthreads, 0, ld(sum), {0, 0, 0}};
    for (int i = 0; i < n_vars; i++) {
        u128 co[3];
        if (bp_execute(&s, co)) return 1;
        bo_transcript_message(t, (const uint8_t *)co, 32); /* truncate(): c0, c1 */
        u128 r;
        bo_transcript_sample(t, (uint8_t *)&r, 16);
        if (challenges_out) st(challenges_out + 2 * i, r);
        if (bp_fold(&s, r)) return 1;
    }
    st(evals_out, s.a[0]);
    st(evals_out + 2, s.b[0]);
    return 0;
}

/* BaseFoldProver::prove (prover protocols/basefold.rs:131-151) and finish
 * (:158-166).  multilinear / transparent (2^(log_dim+lbs) each) are folded in
 * place; codeword / tree are the committed interleaved codeword and its tree. */
int bo_basefold_prove_ex(const bo_domain *d, uint64_t *multilinear, uint64_t *transparent, const uint64_t claim[2],
                         const uint64_t *codeword, const uint8_t *tree, int log_dim, int lir, int lbs,
                         const int *arities, int n_ar, int n_queries, bo_transcript *t, uint64_t *challenges_out,
                         int nthreads, const bo_vision *leaf, const bo_vision *node) {
    const mk_scheme sc = {leaf, node};
    committed c0 = {(u128 *)codeword, log_dim + lir + lbs, (uint8_t *)tree, log_dim + lir};
    fri_st f;
    int rc = fri_init(&f, d, log_dim, lir, lbs, arities, n_ar, n_queries, &c0, nthreads, sc);
    if (rc) return rc;
    int n_vars = log_dim + lbs;
    bp_st s = {(u128 *)multilinear, (u128 *)transparent, n_vars, nthreads, 0, ld(claim), {0, 0, 0}};
    for (int i = 0; i < n_vars; i++) {
        /* execute (:99-113): sumcheck round, then FRI fold round */
        u128 co[3];
        if (bp_execute(&s, co)) return 1;
        uint8_t root[32];
        int c = fri_execute(&f, root);
        if (c < 0) return 5;
        bo_transcript_message(t, (const uint8_t *)co, 32);
        if (c) bo_transcript_message(t, root, 32);
        u128 r;
        bo_transcript_sample(t, (uint8_t *)&r, 16);
        if (challenges_out) st(challenges_out + 2 * i, r);
        /* fold (:119-123) */
        if (bp_fold(&s, r)) return 1;
        fri_receive(&f, r);
    }
    uint8_t root[32];
    int c = fri_execute(&f, root);
    if (c < 0) return 5;
    if (c) bo_transcript_message(t, root, 32);
    rc = fri_finish(&f, t, NULL);
    fri_free(&f);
    return rc;
}

int bo_basefold_prove(const bo_domain *d, uint64_t *multilinear, uint64_t *transparent, const uint64_t claim[2],
                      const uint64_t *codeword, const uint8_t *tree, int log_dim, int lir, int lbs,
                      const int *arities, int n_ar, int n_queries, bo_transcript *t, uint64_t *challenges_out,
                      int nthreads) {
    return bo_basefold_prove_ex(d, multilinear, transparent, claim, codeword, tree, log_dim, lir, lbs, arities, n_ar,
                                n_queries, t, challenges_out, nthreads, NULL, NULL);
}

/* TensorAlgebra::transpose for B1 x B128 (verifier/math/src/tensor_algebra.rs:137-141 ->
 * square_transforms_extension_field, verifier/field/src/transpose.rs:30-75):
 * the 128 x 128 bit matrix of the 128 elements transposed, out[i] bit j = in[j] bit i. */
static void transpose128(const u128 *in, u128 *out) {
    for (int i = 0; i < 128; i++) {
        u128 v = 0;
        for (int j = 0; j < 128; j++) v |= ((in[j] >> i) & 1) << j;
        out[i] = v;
    }
}
void bo_tensor_transpose(const uint64_t *in, uint64_t *out) { transpose128((const u128 *)in, (u128 *)out); }

/* OneBitPCSProver::prove (prover/prover/src/pcs.rs:107-170): ring-switch
 * reduction to a B128 sumcheck claim, then BaseFold.  packed (2^log_len
 * elements, log_len = log_dim + lbs) is consumed; eval_point has log_len + 7
 * coordinates. */
int bo_pcs_prove_ex(const bo_domain *d, uint64_t *packed, const uint64_t *eval_point, const uint64_t *codeword,
                    const uint8_t *tree, int log_dim, int lir, int lbs, const int *arities, int n_ar, int n_queries,
                    bo_transcript *t, int nthreads, const bo_vision *leaf, const bo_vision *node) {
    int log_len = log_dim + lbs;
    size_t n = (size_t)1 << log_len;
    u128 *suffix = (u128 *)malloc(n * sizeof(u128));
    eq_ind((const u128 *)eval_point + 7, log_len, suffix);
    u128 s_hat_v[128], s_hat_u[128], rpp[7], eq_rpp[128];
    bo_rs_fold_1b_rows(packed, (const uint64_t *)suffix, log_len, (uint64_t *)s_hat_v, 0, nthreads);
    bo_transcript_message(t, (const uint8_t *)s_hat_v, sizeof(s_hat_v));
    transpose128(s_hat_v, s_hat_u);
    bo_transcript_sample(t, (uint8_t *)rpp, sizeof(rpp));
    eq_ind(rpp, 7, eq_rpp);
    u128 claim = 0;
    for (int i = 0; i < 128; i++) claim ^= gf_mul(s_hat_u[i], eq_rpp[i]);
    bo_rs_fold_elems((uint64_t *)suffix, n, (const uint64_t *)eq_rpp, 0, nthreads);
    uint64_t cl[2];
    st(cl, claim);
    int rc = bo_basefold_prove_ex(d, packed, (uint64_t *)suffix, cl, codeword, tree, log_dim, lir, lbs, arities,
                                  n_ar, n_queries, t, NULL, nthreads, leaf, node);
    free(suffix);
    return rc;
}

int bo_pcs_prove(const bo_domain *d, uint64_t *packed, const uint64_t *eval_point, const uint64_t *codeword,
                 const uint8_t *tree, int log_dim, int lir, int lbs, const int *arities, int n_ar, int n_queries,
                 bo_transcript *t, int nthreads) {
    return bo_pcs_prove_ex(d, packed, eval_point, codeword, tree, log_dim, lir, lbs, arities, n_ar, n_queries, t,
                           nthreads, NULL, NULL);
}

/* ===================================================================== */
/* Vision-M hash (verifier/verifier/src/hash/vision_{4,6})                 */
/* ===================================================================== */
#define VIS_ROUNDS 8
#define VIS_MAXM 6

struct bo_vision {
    int m;
    u128 b_fwd[4];                /* B(x) = c0 + c1 x + c2 x^2 + c3 x^4 (constants.rs B_FWD_COEFFS) */
    u128 b_inv_cols[128], b_inv0; /* B^-1(y) = sum_j bit_j(y) cols[j] + b_inv0 */
    u128 rc[1 + 2 * VIS_ROUNDS][VIS_MAXM];
    u128 mds[VIS_MAXM][VIS_MAXM];
};

static inline u128 gf_sqr(u128 a) { return gf_mul(a, a); }

static u128 gf_sqr_n(u128 a, int n) {
    while (n--) a = gf_sqr(a);
    return a;
}

/* a^(2^128 - 2) by the Itoh-Tsujii chain: b_k = a^(2^k - 1) with
 * b_(i+j) = b_i^(2^j) * b_j over k = 1, 2, 3, 6, 7, 14, 15, 30, 31, 62, 63, 126, 127,
 * then one more squaring. */
static u128 gf_inv_chain(u128 a) {
    u128 b1 = a;
    u128 b2 = gf_mul(gf_sqr_n(b1, 1), b1);
    u128 b3 = gf_mul(gf_sqr_n(b2, 1), b1);
    u128 b6 = gf_mul(gf_sqr_n(b3, 3), b3);
    u128 b7 = gf_mul(gf_sqr_n(b6, 1), b1);
    u128 b14 = gf_mul(gf_sqr_n(b7, 7), b7);
    u128 b15 = gf_mul(gf_sqr_n(b14, 1), b1);
    u128 b30 = gf_mul(gf_sqr_n(b15, 15), b15);
    u128 b31 = gf_mul(gf_sqr_n(b30, 1), b1);
    u128 b62 = gf_mul(gf_sqr_n(b31, 31), b31);
    u128 b63 = gf_mul(gf_sqr_n(b62, 1), b1);
    u128 b126 = gf_mul(gf_sqr_n(b63, 63), b63);
    u128 b127 = gf_mul(gf_sqr_n(b126, 1), b1);
    return gf_sqr(b127);
}

/* batch_invert (verifier/math/src/batch_invert.rs): invert_or_zero of every
 * element with one field inversion (zeros stand in as ones and stay zero). */
static void batch_inv(u128 *x, int n) {
    u128 pre[VIS_MAXM], acc = 1;
    for (int i = 0; i < n; i++) {
        pre[i] = acc;
        acc = gf_mul(acc, x[i] ? x[i] : 1);
    }
    u128 inv = gf_inv_chain(acc);
    for (int i = n - 1; i >= 0; i--) {
        const u128 xi = x[i] ? x[i] : 1;
        const u128 out = gf_mul(inv, pre[i]);
        inv = gf_mul(inv, xi);
        x[i] = x[i] ? out : 0;
    }
}

void bo_vision_batch_invert(uint64_t *x, int n) {
    u128 t[VIS_MAXM];
    if (n > VIS_MAXM) return;
    for (int i = 0; i < n; i++) t[i] = ld(x + 2 * i);
    batch_inv(t, n);
    for (int i = 0; i < n; i++) st(x + 2 * i, t[i]);
}

static u128 b_fwd_linear(const bo_vision *v, u128 x) {
    u128 x2 = gf_sqr(x), x4 = gf_sqr(x2);
    return gf_mul(v->b_fwd[1], x) ^ gf_mul(v->b_fwd[2], x2) ^ gf_mul(v->b_fwd[3], x4);
}

static u128 b_fwd(const bo_vision *v, u128 x) { return v->b_fwd[0] ^ b_fwd_linear(v, x); }

static u128 b_inv(const bo_vision *v, u128 y) {
    u128 acc = v->b_inv0;
    for (int j = 0; j < 128; j++)
        if ((y >> j) & 1) acc ^= v->b_inv_cols[j];
    return acc;
}

/* The inverse of B's linear part, column by column: Gauss-Jordan elimination
 * over GF(2) on [L | I] with L's columns L(x^j).  B^-1(y) = L^-1(y) + L^-1(c0)
 * (the affine form that constants.rs B_INV_COEFFS encodes as a linearized
 * polynomial).  Returns nonzero if L is singular. */
static int invert_b_linear(bo_vision *v) {
    /* rows r of the augmented matrix: bits of row r of L (as u128 over columns j) and of I */
    u128 a[128], e[128];
    for (int r = 0; r < 128; r++) {
        a[r] = 0;
        e[r] = (u128)1 << r;
    }
    for (int j = 0; j < 128; j++) {
        u128 col = b_fwd_linear(v, (u128)1 << j);
        for (int r = 0; r < 128; r++)
            if ((col >> r) & 1) a[r] |= (u128)1 << j;
    }
    for (int c = 0; c < 128; c++) {
        int p = -1;
        for (int r = c; r < 128; r++)
            if ((a[r] >> c) & 1) {
                p = r;
                break;
            }
        if (p < 0) return 1;
        u128 t = a[p]; a[p] = a[c]; a[c] = t;
        t = e[p]; e[p] = e[c]; e[c] = t;
        for (int r = 0; r < 128; r++)
            if (r != c && ((a[r] >> c) & 1)) {
                a[r] ^= a[c];
                e[r] ^= e[c];
            }
    }
    /* now e = L^-1 as rows: bit j of row r = entry (r, j); column j = L^-1(x^j) */
    for (int j = 0; j < 128; j++) {
        u128 col = 0;
        for (int r = 0; r < 128; r++)
            if ((e[r] >> j) & 1) col |= (u128)1 << r;
        v->b_inv_cols[j] = col;
    }
    v->b_inv0 = 0;
    v->b_inv0 = b_inv(v, v->b_fwd[0]);
    return 0;
}

/* sbox (permutation.rs): batch inversion of the state, then B or B^-1. */
static void vis_sbox(const bo_vision *v, u128 *s, int inverse) {
    batch_inv(s, v->m);
    for (int i = 0; i < v->m; i++) s[i] = inverse ? b_inv(v, s[i]) : b_fwd(v, s[i]);
}

/* mds_mul as the matrix it implements (permutation.rs test_mds for M = 4; the
 * circulant rows of vision_6/permutation.rs mds_mul for M = 6). */
static void vis_mds(const bo_vision *v, u128 *s) {
    u128 o[VIS_MAXM];
    for (int i = 0; i < v->m; i++) {
        o[i] = 0;
        for (int j = 0; j < v->m; j++) o[i] ^= gf_mul(v->mds[i][j], s[j]);
    }
    memcpy(s, o, sizeof(u128) * (size_t)v->m);
}

static void vis_add(const bo_vision *v, u128 *s, const u128 *c) {
    for (int i = 0; i < v->m; i++) s[i] ^= c[i];
}

/* permutation (permutation.rs): RC[0], then 8 rounds of
 * (S-box with B^-1, MDS, RC) and (S-box with B, MDS, RC). */
static void vis_permute(const bo_vision *v, u128 *s) {
    vis_add(v, s, v->rc[0]);
    for (int r = 0; r < VIS_ROUNDS; r++) {
        vis_sbox(v, s, 1);
        vis_mds(v, s);
        vis_add(v, s, v->rc[1 + 2 * r]);
        vis_sbox(v, s, 0);
        vis_mds(v, s);
        vis_add(v, s, v->rc[2 + 2 * r]);
    }
}

bo_vision *bo_vision_new(int m, const uint64_t *b_fwd_coeffs, const uint64_t *initial, const uint64_t *matrix,
                         const uint64_t *constant) {
    if (m != 4 && m != 6) return NULL;
    bo_vision *v = (bo_vision *)calloc(1, sizeof(bo_vision));
    v->m = m;
    for (int i = 0; i < 4; i++) v->b_fwd[i] = ld(b_fwd_coeffs + 2 * i);
    if (invert_b_linear(v)) {
        free(v);
        return NULL;
    }
    const u128 two = 2, three = 3, inv_two = gf_inv(2);
    if (m == 4) {
        const u128 rows[4][4] = {{two, three, 1, 1}, {1, two, three, 1}, {1, 1, two, three}, {three, 1, 1, two}};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) v->mds[i][j] = rows[i][j];
    } else {
        const u128 row0[6] = {inv_two, inv_two, two, 1, two, three};
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) v->mds[i][j] = row0[(j - i + 6) % 6];
    }
    /* compute_round_constants (constants.rs test module): a key schedule run
     * through the permutation's own S-boxes and MDS. */
    u128 key_state[VIS_MAXM], key_inj[VIS_MAXM], cmat[VIS_MAXM * VIS_MAXM], cconst[VIS_MAXM];
    for (int i = 0; i < m; i++) {
        key_state[i] = key_inj[i] = v->rc[0][i] = ld(initial + 2 * i);
        cconst[i] = ld(constant + 2 * i);
    }
    for (int i = 0; i < m * m; i++) cmat[i] = ld(matrix + 2 * i);
    int idx = 1;
    for (int r = 0; r < VIS_ROUNDS; r++)
        for (int half = 0; half < 2; half++) {
            u128 t[VIS_MAXM];
            for (int i = 0; i < m; i++) {
                t[i] = cconst[i];
                for (int j = 0; j < m; j++) t[i] ^= gf_mul(cmat[i * m + j], key_inj[j]);
            }
            memcpy(key_inj, t, sizeof(u128) * (size_t)m);
            vis_sbox(v, key_state, half == 0);
            vis_mds(v, key_state);
            vis_add(v, key_state, key_inj);
            memcpy(v->rc[idx++], key_state, sizeof(u128) * (size_t)m);
        }
    return v;
}

void bo_vision_free(bo_vision *v) { free(v); }

void bo_vision_round_constants(const bo_vision *v, uint64_t *out) {
    for (int r = 0; r < 1 + 2 * VIS_ROUNDS; r++)
        for (int i = 0; i < v->m; i++) st(out + 2 * (r * v->m + i), v->rc[r][i]);
}

void bo_vision_b_maps(const bo_vision *v, const uint64_t *x, uint64_t *fwd, uint64_t *inv, size_t n) {
    for (size_t i = 0; i < n; i++) {
        st(fwd + 2 * i, b_fwd(v, ld(x + 2 * i)));
        st(inv + 2 * i, b_inv(v, ld(x + 2 * i)));
    }
}

void bo_vision_permutation(const bo_vision *v, uint64_t *states, size_t n_states) {
    for (size_t k = 0; k < n_states; k++) {
        u128 s[VIS_MAXM];
        for (int i = 0; i < v->m; i++) s[i] = ld(states + 2 * (k * v->m + i));
        vis_permute(v, s);
        for (int i = 0; i < v->m; i++) st(states + 2 * (k * v->m + i), s[i]);
    }
}

/* VisionHasherDigest (digest.rs): sponge with rate M - 2 elements; each full
 * block overwrites state[0 .. rate) and permutes; the last partial block is
 * padded 0x80 .. 0x01 (or a whole padding block is absorbed); the digest is
 * state[0] || state[1]. */
void bo_vision_digest(const bo_vision *v, const uint8_t *data, size_t len, uint8_t out[32]) {
    const int rate = v->m - 2;
    const size_t rb = 16 * (size_t)rate;
    u128 s[VIS_MAXM] = {0};
    uint8_t blk[64];
    size_t off = 0;
    for (; off + rb <= len; off += rb) {
        for (int i = 0; i < rate; i++) memcpy(&s[i], data + off + 16 * i, 16);
        vis_permute(v, s);
    }
    size_t rem = len - off;
    memset(blk, 0, rb);
    memcpy(blk, data + off, rem);
    blk[rem] |= 0x80;
    blk[rb - 1] |= 0x01;
    for (int i = 0; i < rate; i++) memcpy(&s[i], blk + 16 * i, 16);
    vis_permute(v, s);
    memcpy(out, &s[0], 16);
    memcpy(out + 16, &s[1], 16);
}

/* VisionCompression (vision_4/compression.rs): permute [l0, l1, r0, r1],
 * add the inputs' first two elements back, keep the first two. */
void bo_vision_compress(const bo_vision *v4, const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    u128 s[4];
    memcpy(&s[0], left, 16);
    memcpy(&s[1], left + 16, 16);
    memcpy(&s[2], right, 16);
    memcpy(&s[3], right + 16, 16);
    u128 o0 = s[0], o1 = s[1];
    vis_permute(v4, s);
    s[0] ^= o0;
    s[1] ^= o1;
    memcpy(out, &s[0], 16);
    memcpy(out + 16, &s[1], 16);
}

typedef struct {
    const bo_vision *leaf, *node;
    const u128 *elems;
    size_t batch;
    uint8_t *tree;
    size_t in_off, out_off;
} vmk_ctx;

static void vmk_leaf_task(void *p, size_t i) {
    vmk_ctx *c = (vmk_ctx *)p;
    bo_vision_digest(c->leaf, (const uint8_t *)(c->elems + i * c->batch), 16 * c->batch, c->tree + 32 * i);
}
static void vmk_node_task(void *p, size_t i) {
    vmk_ctx *c = (vmk_ctx *)p;
    bo_vision_compress(c->node, c->tree + 32 * (c->in_off + 2 * i), c->tree + 32 * (c->in_off + 2 * i + 1),
                       c->tree + 32 * (c->out_off + i));
}

/* BinaryMerkleTreeProver with a Vision leaf digest (Vision-4 or Vision-6
 * sponge, prover/prover/benches/binary_merkle_tree.rs) and VisionCompression
 * for the nodes; same flattened layout as bo_merkle_build. */
int bo_merkle_build_vision(const bo_vision *leaf, const bo_vision *node4, const uint64_t *elems, size_t n_elems,
                           size_t batch, uint8_t *tree, int nthreads) {
    if (!batch || n_elems % batch) return 4;
    size_t n_leaves = n_elems / batch;
    if (!n_leaves || (n_leaves & (n_leaves - 1))) return 5;
    if (node4->m != 4) return 1;
    vmk_ctx c = {leaf, node4, (const u128 *)elems, batch, tree, 0, 0};
    par_for(nthreads, n_leaves, vmk_leaf_task, &c);
    size_t in_off = 0, n = n_leaves;
    while (n > 1) {
        c.in_off = in_off;
        c.out_off = in_off + n;
        par_for(nthreads, n / 2, vmk_node_task, &c);
        in_off += n;
        n /= 2;
    }
    return 0;
}

/* ===================================================================== */
/* SplitMix64 input generator                                              */
/* ===================================================================== */
void bo_splitmix_fill(uint64_t seed, uint64_t *out, size_t n) {
    uint64_t x = seed;
    for (size_t i = 0; i < n; i++) {
        x += 0x9e3779b97f4a7c15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        out[i] = z ^ (z >> 31);
    }
}
