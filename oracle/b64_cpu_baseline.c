/*
 * CPU baseline for the commit step: the reference prover's x86-64 path as it is
 * built with RUSTFLAGS="-C target-cpu=native" (README.md:28) on an AVX-512 host.
 *
 * BENCHMARK / TEST INFRASTRUCTURE ONLY.  Only bench.py's cpu_baseline leg and
 * tests/ load it (never the product library); every output is checked against
 * the scalar oracle (b64_oracle.c) byte for byte.
 *
 * What it restates:
 *  - PackedBinaryGhash4x128b (verifier/field/src/arch/x86_64/packed_ghash_512.rs:38-64):
 *    four B128 elements per 512-bit vector, multiplied with VPCLMULQDQ by the
 *    mul_clmul sequence of arch/shared/ghash.rs:19-94; the packed type the
 *    reference selects on AVX-512 hosts (arch/arch_optimal.rs:6-11).  Without
 *    AVX-512 VPCLMULQDQ at run time, the 1x128 PCLMULQDQ form.
 *  - NeighborsLastMultiThread::forward_transform with that packing
 *    (verifier/math/src/ntt/neighbors_last.rs:57-279, 464-526): shared layers
 *    split into 2^log_num_shares shares, then independent chunks depth-first with a
 *    2^10-element breadth-first base case; twiddles from GenericPreExpanded tables
 *    (ntt/domain_context.rs:99-159).
 *  - ReedSolomonCode::encode_batch (reed_solomon.rs:128-201): parallel bit-reversed
 *    copy, replication, NTT.
 *  - BinaryMerkleTreeProver (prover/prover/src/merkle_tree/binary_merkle_tree.rs:65-122,
 *    178-205): leaf = SHA-256 of the LE16 leaf bytes, nodes = Sha256Compression
 *    (verifier/verifier/src/hash/compress.rs:33-57), each level in parallel; SHA-256
 *    rounds with the SHA extensions, as the sha2 crate selects at run time, scalar
 *    otherwise.
 *  - fold_interleaved (prover/prover/src/fri/fold.rs:260-282): eq tensor of the
 *    challenges, packed inner product per leaf.
 *  - A persistent pool of nthreads workers (the rayon global pool), with
 *    log_num_shares = floor(log2 nthreads) (prover/prover/src/prove.rs:98-104).
 */
#define _GNU_SOURCE
#include <immintrin.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ===================================================================== */
/* Field                                                                   */
/* ===================================================================== */
/* arch/shared/ghash.rs:19-45, 1x128 */
__attribute__((target("pclmul,sse4.1"))) static inline __m128i mul1(__m128i x, __m128i y) {
    const __m128i poly = _mm_set_epi64x(0, 0x87);
    __m128i t1 = _mm_xor_si128(_mm_clmulepi64_si128(x, y, 0x01), _mm_clmulepi64_si128(x, y, 0x10));
    __m128i t2 = _mm_clmulepi64_si128(x, y, 0x11);
    t1 = _mm_xor_si128(t1, _mm_slli_si128(t2, 8));
    t1 = _mm_xor_si128(t1, _mm_clmulepi64_si128(t2, poly, 0x01));
    __m128i t0 = _mm_clmulepi64_si128(x, y, 0x00);
    t0 = _mm_xor_si128(t0, _mm_slli_si128(t1, 8));
    return _mm_xor_si128(t0, _mm_clmulepi64_si128(t1, poly, 0x01));
}

/* the same sequence on four 128-bit lanes (PackedBinaryGhash4x128b) */
#define T512 __attribute__((target("avx512f,avx512bw,avx512vl,vpclmulqdq,pclmul,sse4.1")))
T512 static inline __m512i mul4(__m512i x, __m512i y) {
    const __m512i poly = _mm512_set_epi64(0, 0x87, 0, 0x87, 0, 0x87, 0, 0x87);
    const __m512i z = _mm512_setzero_si512();
    __m512i t1 = _mm512_xor_si512(_mm512_clmulepi64_epi128(x, y, 0x01), _mm512_clmulepi64_epi128(x, y, 0x10));
    __m512i t2 = _mm512_clmulepi64_epi128(x, y, 0x11);
    t1 = _mm512_xor_si512(t1, _mm512_unpacklo_epi64(z, t2));
    t1 = _mm512_xor_si512(t1, _mm512_clmulepi64_epi128(t2, poly, 0x01));
    __m512i t0 = _mm512_clmulepi64_epi128(x, y, 0x00);
    t0 = _mm512_xor_si512(t0, _mm512_unpacklo_epi64(z, t1));
    return _mm512_xor_si512(t0, _mm512_clmulepi64_epi128(t1, poly, 0x01));
}

static inline u128 ldu(const void *p) {
    u128 v;
    memcpy(&v, p, 16);
    return v;
}
__attribute__((target("pclmul,sse4.1"))) static u128 gmul(u128 a, u128 b) {
    __m128i r = mul1(_mm_loadu_si128((const __m128i *)&a), _mm_loadu_si128((const __m128i *)&b));
    u128 o;
    _mm_storeu_si128((__m128i *)&o, r);
    return o;
}
static u128 ginv(u128 a) { /* a^(2^128-2) */
    u128 r = 1, e = ~(u128)0 - 1;
    while (e) {
        if (e & 1) r = gmul(r, a);
        a = gmul(a, a);
        e >>= 1;
    }
    return r;
}

static int g_wide = -1, g_shani = -1;
static void detect(void) {
    if (g_wide < 0) {
        __builtin_cpu_init();
        g_wide = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
                 __builtin_cpu_supports("vpclmulqdq");
        g_shani = __builtin_cpu_supports("sha") && __builtin_cpu_supports("sse4.1");
    }
}

/* ===================================================================== */
/* Persistent thread pool (the rayon global pool)                          */
/* ===================================================================== */
typedef void (*task_fn)(void *, size_t);
typedef struct {
    int n;
    pthread_t *th;
    pthread_mutex_t mu;
    pthread_cond_t cv;
    uint64_t gen;
    int stop;
    task_fn fn;
    void *arg;
    size_t n_tasks;
    atomic_size_t next;
    atomic_int busy;
} pool_t;

static void pool_drain(pool_t *p) {
    for (;;) {
        size_t i = atomic_fetch_add_explicit(&p->next, 1, memory_order_relaxed);
        if (i >= p->n_tasks) break;
        p->fn(p->arg, i);
    }
}

static void *pool_main(void *a) {
    pool_t *p = (pool_t *)a;
    uint64_t seen = 0;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        while (!p->stop && p->gen == seen) pthread_cond_wait(&p->cv, &p->mu);
        if (p->stop) {
            pthread_mutex_unlock(&p->mu);
            return NULL;
        }
        seen = p->gen;
        pthread_mutex_unlock(&p->mu);
        pool_drain(p);
        atomic_fetch_sub_explicit(&p->busy, 1, memory_order_release);
    }
}

static pool_t *pool_new(int n) {
    pool_t *p = (pool_t *)calloc(1, sizeof(pool_t));
    p->n = n;
    pthread_mutex_init(&p->mu, NULL);
    pthread_cond_init(&p->cv, NULL);
    p->th = (pthread_t *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(pthread_t));
    for (int i = 0; i < n - 1; i++) pthread_create(&p->th[i], NULL, pool_main, p);
    return p;
}

static void pool_free(pool_t *p) {
    pthread_mutex_lock(&p->mu);
    p->stop = 1;
    pthread_cond_broadcast(&p->cv);
    pthread_mutex_unlock(&p->mu);
    for (int i = 0; i < p->n - 1; i++) pthread_join(p->th[i], NULL);
    free(p->th);
    free(p);
}

/* fn(arg, i) for i < n_tasks on every worker and the caller */
static void par_for(pool_t *p, size_t n_tasks, task_fn fn, void *arg) {
    if (p->n <= 1 || n_tasks <= 1) {
        for (size_t i = 0; i < n_tasks; i++) fn(arg, i);
        return;
    }
    pthread_mutex_lock(&p->mu);
    p->fn = fn;
    p->arg = arg;
    p->n_tasks = n_tasks;
    atomic_store(&p->next, 0);
    atomic_store(&p->busy, p->n - 1);
    p->gen++;
    pthread_cond_broadcast(&p->cv);
    pthread_mutex_unlock(&p->mu);
    pool_drain(p);
    while (atomic_load_explicit(&p->busy, memory_order_acquire) > 0) __builtin_ia32_pause();
}

/* ===================================================================== */
/* Domain: GenericPreExpanded (ntt/domain_context.rs:19-46, 99-159)        */
/* ===================================================================== */
typedef struct {
    int dim;
    u128 **expanded; /* expanded[layer][block], 2^layer entries */
    pool_t *pool;
    int log_num_shares;
} cb_ctx;

cb_ctx *cb_create(const uint64_t *basis, int dim, int nthreads) {
    detect();
    if (dim < 1 || dim > 40 || nthreads < 1) return NULL;
    cb_ctx *c = (cb_ctx *)calloc(1, sizeof(cb_ctx));
    c->dim = dim;
    /* generate_evals_from_subspace: evals[0] = basis; evals[i][k-1] =
       evals[i-1][k] * (evals[i-1][k] + evals[i-1][0]); each row scaled by its
       first entry's inverse */
    u128 *ev = (u128 *)calloc((size_t)dim * dim, sizeof(u128));
    for (int k = 0; k < dim; k++) ev[k] = ldu(basis + 2 * k);
    for (int i = 1; i < dim; i++) {
        const u128 *prev = ev + (size_t)(i - 1) * dim;
        u128 *row = ev + (size_t)i * dim;
        for (int k = 1; k < dim - i + 1; k++) row[k - 1] = gmul(prev[k], prev[k] ^ prev[0]);
    }
    for (int i = 0; i < dim; i++) {
        u128 *row = ev + (size_t)i * dim;
        u128 inv = ginv(row[0]);
        for (int k = 0; k < dim - i; k++) row[k] = gmul(row[k], inv);
    }
    /* expanded[layer][block] = subset sum of evals[dim-1-layer][1..] by the bits of block */
    c->expanded = (u128 **)calloc((size_t)dim, sizeof(u128 *));
    for (int layer = 0; layer < dim; layer++) {
        const u128 *row = ev + (size_t)(dim - 1 - layer) * dim;
        /* aligned_alloc takes a size that is a multiple of the alignment */
        const size_t bytes = (sizeof(u128) * ((size_t)1 << layer) + 63) & ~(size_t)63;
        u128 *e = (u128 *)aligned_alloc(64, bytes);
        e[0] = 0;
        for (int b = 0; b < layer; b++)
            for (size_t j = 0; j < ((size_t)1 << b); j++) e[((size_t)1 << b) + j] = e[j] ^ row[1 + b];
        c->expanded[layer] = e;
    }
    free(ev);
    c->pool = pool_new(nthreads);
    int l = 0;
    while ((2 << l) <= nthreads) l++;
    c->log_num_shares = l; /* ilog2(current_num_threads()) */
    return c;
}

void cb_destroy(cb_ctx *c) {
    if (!c) return;
    for (int l = 0; l < c->dim; l++) free(c->expanded[l]);
    free(c->expanded);
    pool_free(c->pool);
    free(c);
}

/* bit 0: 4-wide VPCLMULQDQ multiply in use, bit 1: SHA extensions in use */
int cb_features(void) {
    detect();
    return (g_wide ? 1 : 0) | (g_shani ? 2 : 0);
}

/* Restricts the paths to those in `mask` (bits as in cb_features) and what the
   host supports; the tests run every path against the oracle this way. */
void cb_restrict(int mask) {
    g_wide = -1;
    detect();
    g_wide = g_wide && (mask & 1);
    g_shani = g_shani && (mask & 2);
}

static inline u128 twiddle(const cb_ctx *c, int layer, size_t block) { return c->expanded[layer][block]; }

/* ===================================================================== */
/* Butterflies on packed elements                                          */
/* ===================================================================== */
/* u[i] += v[i] * t; v[i] += u[i] for n elements (n a multiple of 4 when wide) */
T512 static void bfly_run4(u128 *u, u128 *v, size_t n, u128 t) {
    const __m512i tw = _mm512_broadcast_i32x4(_mm_loadu_si128((const __m128i *)&t));
    for (size_t i = 0; i < n; i += 4) {
        __m512i a = _mm512_loadu_si512((const void *)(u + i)), b = _mm512_loadu_si512((const void *)(v + i));
        a = _mm512_xor_si512(a, mul4(b, tw));
        b = _mm512_xor_si512(b, a);
        _mm512_storeu_si512((void *)(u + i), a);
        _mm512_storeu_si512((void *)(v + i), b);
    }
}
__attribute__((target("pclmul,sse4.1"))) static void bfly_run1(u128 *u, u128 *v, size_t n, u128 t) {
    const __m128i tw = _mm_loadu_si128((const __m128i *)&t);
    for (size_t i = 0; i < n; i++) {
        __m128i a = _mm_loadu_si128((const __m128i *)(u + i)), b = _mm_loadu_si128((const __m128i *)(v + i));
        a = _mm_xor_si128(a, mul1(b, tw));
        b = _mm_xor_si128(b, a);
        _mm_storeu_si128((__m128i *)(u + i), a);
        _mm_storeu_si128((__m128i *)(v + i), b);
    }
}
static inline void bfly_run(u128 *u, u128 *v, size_t n, u128 t) {
    if (g_wide && (n & 3) == 0)
        bfly_run4(u, v, n, t);
    else
        bfly_run1(u, v, n, t);
}

/* ===================================================================== */
/* NeighborsLastMultiThread (ntt/neighbors_last.rs) with 4-element packing  */
/* ===================================================================== */
#define LOG_BASE_LEN 10 /* DEFAULT_LOG_BASE_LEN */
#define LOG_WIDTH 2     /* PackedBinaryGhash4x128b */

/* forward_breadth_first (neighbors_last.rs:131-222): the layers whose butterflies
   pair whole packed elements broadcast one twiddle per block; the last LOG_WIDTH
   layers pair lanes of one packed element, done here lane by lane. */
static void nl_breadth_first(const cb_ctx *c, u128 *data, int log_d, int base_layer, size_t base_block,
                             int lr_start, int lr_end) {
    const int log_n = log_d + base_layer;
    for (int layer = lr_start; layer < lr_end; layer++) {
        const int log_block = log_n - layer;
        const size_t half = (size_t)1 << (log_block - 1);
        const int log_blocks = layer - base_layer;
        const u128 *tws = c->expanded[layer] + (base_block << log_blocks);
        for (size_t b = 0; b < ((size_t)1 << log_blocks); b++) {
            u128 *b0 = data + (b << log_block);
            bfly_run(b0, b0 + half, half, tws[b]);
        }
    }
}

/* forward_depth_first (neighbors_last.rs:57-120) */
static void nl_depth_first(const cb_ctx *c, u128 *data, int log_d, int layer, size_t block, int lr_start,
                           int lr_end) {
    if (layer >= lr_end) return;
    if (log_d <= LOG_BASE_LEN) {
        nl_breadth_first(c, data, log_d, layer, block, lr_start, lr_end);
        return;
    }
    const size_t half = (size_t)1 << (log_d - 1);
    if (layer >= lr_start) {
        bfly_run(data, data + half, half, twiddle(c, layer, block));
        lr_start++;
    }
    nl_depth_first(c, data, log_d - 1, layer + 1, block << 1, lr_start, lr_end);
    nl_depth_first(c, data + half, log_d - 1, layer + 1, (block << 1) + 1, lr_start, lr_end);
}

typedef struct {
    const cb_ctx *c;
    u128 *data;
    int log_d, layer, log_num_shares, lr_start, lr_end, log_d_chunk;
} nl_job;

/* forward_shared_layer task (neighbors_last.rs:235-279) with with_middle_bit (:288-300) */
static void nl_shared_task(void *p, size_t k) {
    const nl_job *j = (const nl_job *)p;
    const int log_num_chunks = j->log_num_shares + 1;
    const int log_chunk_len = j->log_d - log_num_chunks;
    const int shift = j->log_num_shares - j->layer;
    const size_t ms = k >> (shift - 1), ls = k & (((size_t)1 << shift) - 1);
    const size_t c0 = ls | ((ms & ~(size_t)1) << shift), c1 = ls | ((ms | 1) << shift);
    const size_t block = c0 >> (log_num_chunks - j->layer);
    bfly_run(j->data + (c0 << log_chunk_len), j->data + (c1 << log_chunk_len), (size_t)1 << log_chunk_len,
             twiddle(j->c, j->layer, block));
}

static void nl_chunk_task(void *p, size_t block) {
    const nl_job *j = (const nl_job *)p;
    nl_depth_first(j->c, j->data + (block << j->log_d_chunk), j->log_d_chunk, j->layer, block, j->lr_start,
                   j->lr_end);
}

/* NeighborsLastMultiThread::forward_transform (neighbors_last.rs:464-526) */
int cb_ntt_forward(cb_ctx *c, uint64_t *data64, int log_d, int se, int sl) {
    if (se + sl > log_d || log_d - sl > c->dim) return 1;
    u128 *data = (u128 *)data64;
    if (log_d <= LOG_WIDTH) {
        nl_breadth_first(c, data, log_d, 0, 0, se, log_d - sl);
        return 0;
    }
    const int max_shares = log_d - LOG_WIDTH - 1;
    const int actual = c->log_num_shares < max_shares ? c->log_num_shares : max_shares;
    const int last_layer = log_d - sl;
    const int shared_end = actual < last_layer ? actual : last_layer;
    const int ind_start = actual > se ? actual : se;
    nl_job j = {c, data, log_d, 0, actual, 0, 0, 0};
    for (int layer = se; layer < shared_end; layer++) {
        j.layer = layer;
        par_for(c->pool, (size_t)1 << actual, nl_shared_task, &j);
    }
    const int layer = ind_start < max_shares ? ind_start : max_shares;
    j.layer = layer;
    j.log_d_chunk = log_d - layer;
    j.lr_start = ind_start;
    j.lr_end = last_layer;
    par_for(c->pool, (size_t)1 << layer, nl_chunk_task, &j);
    return 0;
}

/* ===================================================================== */
/* encode_batch (reed_solomon.rs:128-201)                                   */
/* ===================================================================== */
typedef struct {
    const u128 *msg;
    u128 *out;
    int m, chunk_log;
} br_job;

static inline uint64_t rev64(uint64_t x) {
    x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return __builtin_bswap64(x);
}

static void br_task(void *p, size_t t) {
    const br_job *j = (const br_job *)p;
    const size_t n = (size_t)1 << j->chunk_log;
    for (size_t i = t * n; i < (t + 1) * n; i++) j->out[i] = j->msg[j->m ? (size_t)(rev64(i) >> (64 - j->m)) : 0];
}

int cb_encode_batch(cb_ctx *c, const uint64_t *msg, int log_dim, int lir, int lbs, uint64_t *out) {
    if (log_dim + lir > c->dim) return 3;
    const int m = log_dim + lbs;
    br_job j = {(const u128 *)msg, (u128 *)out, m, m > 12 ? m - 12 : 0};
    par_for(c->pool, (size_t)1 << (m - j.chunk_log), br_task, &j);
    const size_t n = (size_t)1 << m;
    for (size_t r = 1; r < ((size_t)1 << lir); r++) memcpy((u128 *)out + r * n, out, n * sizeof(u128));
    return cb_ntt_forward(c, out, m + lir, lir, lbs);
}

/* ===================================================================== */
/* SHA-256: SHA extensions, or scalar (FIPS 180-4)                          */
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

#define ROR(x, n) (((x) >> (n)) | ((x) << (32 - (n))))
static void sha_blocks_scalar(uint32_t s[8], const uint8_t *p, size_t nb) {
    for (; nb; nb--, p += 64) {
        uint32_t w[64];
        for (int i = 0; i < 16; i++)
            w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) | ((uint32_t)p[4 * i + 2] << 8) |
                   p[4 * i + 3];
        for (int i = 16; i < 64; i++) {
            uint32_t s0 = ROR(w[i - 15], 7) ^ ROR(w[i - 15], 18) ^ (w[i - 15] >> 3);
            uint32_t s1 = ROR(w[i - 2], 17) ^ ROR(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
        for (int i = 0; i < 64; i++) {
            uint32_t t1 = h + (ROR(e, 6) ^ ROR(e, 11) ^ ROR(e, 25)) + ((e & f) ^ (~e & g)) + K256[i] + w[i];
            uint32_t t2 = (ROR(a, 2) ^ ROR(a, 13) ^ ROR(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        s[0] += a; s[1] += b; s[2] += c; s[3] += d; s[4] += e; s[5] += f; s[6] += g; s[7] += h;
    }
}

/* The SHA extensions keep the state as (A,B,E,F) and (C,D,G,H) and run two
   rounds per sha256rnds2; the schedule of words 16..63 comes from sha256msg1 /
   sha256msg2 four words at a time. */
__attribute__((target("sha,sse4.1,ssse3"))) static void sha_blocks_ni(uint32_t s[8], const uint8_t *p, size_t nb) {
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
    __m128i tmp = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i *)&s[0]), 0xB1); /* C D A B */
    __m128i st1 = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i *)&s[4]), 0x1B); /* E F G H */
    __m128i st0 = _mm_alignr_epi8(tmp, st1, 8);                                    /* A B E F */
    st1 = _mm_blend_epi16(st1, tmp, 0xF0);                                         /* C D G H */
    for (; nb; nb--, p += 64) {
        const __m128i save0 = st0, save1 = st1;
        __m128i w[4];
        for (int g = 0; g < 16; g++) {
            __m128i x;
            if (g < 4) {
                x = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i *)(p + 16 * g)), bswap);
            } else {
                x = _mm_sha256msg1_epu32(w[g & 3], w[(g + 1) & 3]);
                x = _mm_add_epi32(x, _mm_alignr_epi8(w[(g + 3) & 3], w[(g + 2) & 3], 4));
                x = _mm_sha256msg2_epu32(x, w[(g + 3) & 3]);
            }
            w[g & 3] = x;
            const __m128i k = _mm_loadu_si128((const __m128i *)&K256[4 * g]);
            const __m128i m = _mm_add_epi32(x, k);
            st1 = _mm_sha256rnds2_epu32(st1, st0, m);
            st0 = _mm_sha256rnds2_epu32(st0, st1, _mm_shuffle_epi32(m, 0x0E));
        }
        st0 = _mm_add_epi32(st0, save0);
        st1 = _mm_add_epi32(st1, save1);
    }
    tmp = _mm_shuffle_epi32(st0, 0x1B);     /* F E B A */
    st1 = _mm_shuffle_epi32(st1, 0xB1);     /* D C H G */
    st0 = _mm_blend_epi16(tmp, st1, 0xF0);  /* D C B A */
    st1 = _mm_alignr_epi8(st1, tmp, 8);     /* H G F E */
    _mm_storeu_si128((__m128i *)&s[0], st0);
    _mm_storeu_si128((__m128i *)&s[4], st1);
}

static inline void sha_blocks(uint32_t s[8], const uint8_t *p, size_t nb) {
    if (g_shani)
        sha_blocks_ni(s, p, nb);
    else
        sha_blocks_scalar(s, p, nb);
}

static const uint32_t IV256[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                  0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

/* SHA-256 of a message of nb whole blocks plus one padding block */
static void sha256_whole_blocks(const uint8_t *p, size_t nb, uint8_t out[32]) {
    uint32_t s[8];
    memcpy(s, IV256, 32);
    sha_blocks(s, p, nb);
    uint8_t pad[64] = {0x80};
    const uint64_t bits = (uint64_t)nb * 512;
    for (int i = 0; i < 8; i++) pad[63 - i] = (uint8_t)(bits >> (8 * i));
    sha_blocks(s, pad, 1);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(s[i] >> 24);
        out[4 * i + 1] = (uint8_t)(s[i] >> 16);
        out[4 * i + 2] = (uint8_t)(s[i] >> 8);
        out[4 * i + 3] = (uint8_t)s[i];
    }
}

/* Sha256Compression IV: SHA-256("BINIUS SHA-256 COMPRESS") as little-endian words */
static uint32_t g_iv[8];
static pthread_once_t g_iv_once = PTHREAD_ONCE_INIT;
static void iv_init(void) {
    static const char tag[] = "BINIUS SHA-256 COMPRESS";
    uint32_t s[8];
    memcpy(s, IV256, 32);
    uint8_t blk[64] = {0};
    memcpy(blk, tag, sizeof(tag) - 1);
    blk[sizeof(tag) - 1] = 0x80;
    blk[63] = (uint8_t)((sizeof(tag) - 1) * 8);
    sha_blocks_scalar(s, blk, 1);
    for (int k = 0; k < 8; k++) g_iv[k] = __builtin_bswap32(s[k]); /* digest bytes read as LE words */
}

static inline void compress(const uint8_t *lr, uint8_t *out) {
    uint32_t s[8];
    memcpy(s, g_iv, 32);
    sha_blocks(s, lr, 1);
    memcpy(out, s, 32); /* state words written little-endian */
}

/* ===================================================================== */
/* Merkle tree (binary_merkle_tree.rs:65-122, 178-205)                       */
/* ===================================================================== */
typedef struct {
    const uint8_t *elems;
    size_t leaf_bytes;
    uint8_t *prev, *next;
    size_t per_task;
    size_t n;
} mk_job;

static void leaf_task(void *p, size_t t) {
    const mk_job *j = (const mk_job *)p;
    for (size_t i = t * j->per_task; i < (t + 1) * j->per_task && i < j->n; i++)
        sha256_whole_blocks(j->elems + i * j->leaf_bytes, j->leaf_bytes / 64, j->prev + 32 * i);
}

static void node_task(void *p, size_t t) {
    const mk_job *j = (const mk_job *)p;
    for (size_t i = t * j->per_task; i < (t + 1) * j->per_task && i < j->n; i++) compress(j->prev + 64 * i, j->next + 32 * i);
}

/* leaves of batch elements each (batch * 16 a multiple of 64 bytes) */
int cb_merkle_build(cb_ctx *c, const uint64_t *elems, size_t n_elems, size_t batch, uint8_t *tree) {
    pthread_once(&g_iv_once, iv_init);
    if (batch == 0 || (batch * 16) % 64 || n_elems % batch) return 4;
    const size_t n_leaves = n_elems / batch;
    if (n_leaves == 0 || (n_leaves & (n_leaves - 1))) return 5;
    mk_job j = {(const uint8_t *)elems, batch * 16, tree, NULL, 64, n_leaves};
    par_for(c->pool, (n_leaves + j.per_task - 1) / j.per_task, leaf_task, &j);
    uint8_t *prev = tree;
    for (size_t len = n_leaves; len > 1; len /= 2) {
        j.prev = prev;
        j.next = prev + 32 * len;
        j.n = len / 2;
        par_for(c->pool, (j.n + j.per_task - 1) / j.per_task, node_task, &j);
        prev = j.next;
    }
    return 0;
}

/* ===================================================================== */
/* fold_interleaved (prover fri/fold.rs:260-282) for a batch of 16          */
/* ===================================================================== */
typedef struct {
    const u128 *cw;
    u128 *out;
    u128 eq[16];
    size_t n;
} fold_job;

T512 static void fold_task4(void *p, size_t t) {
    const fold_job *j = (const fold_job *)p;
    const __m512i e0 = _mm512_loadu_si512((const void *)&j->eq[0]), e1 = _mm512_loadu_si512((const void *)&j->eq[4]);
    const __m512i e2 = _mm512_loadu_si512((const void *)&j->eq[8]), e3 = _mm512_loadu_si512((const void *)&j->eq[12]);
    const size_t per = 4096;
    for (size_t i = t * per; i < (t + 1) * per && i < j->n; i++) {
        const u128 *c = j->cw + 16 * i;
        __m512i acc = mul4(_mm512_loadu_si512((const void *)c), e0);
        acc = _mm512_xor_si512(acc, mul4(_mm512_loadu_si512((const void *)(c + 4)), e1));
        acc = _mm512_xor_si512(acc, mul4(_mm512_loadu_si512((const void *)(c + 8)), e2));
        acc = _mm512_xor_si512(acc, mul4(_mm512_loadu_si512((const void *)(c + 12)), e3));
        __m256i h = _mm256_xor_si256(_mm512_castsi512_si256(acc), _mm512_extracti64x4_epi64(acc, 1));
        __m128i r = _mm_xor_si128(_mm256_castsi256_si128(h), _mm256_extracti128_si256(h, 1));
        _mm_storeu_si128((__m128i *)(j->out + i), r);
    }
}
__attribute__((target("pclmul,sse4.1"))) static void fold_task1(void *p, size_t t) {
    const fold_job *j = (const fold_job *)p;
    const size_t per = 4096;
    for (size_t i = t * per; i < (t + 1) * per && i < j->n; i++) {
        __m128i acc = _mm_setzero_si128();
        for (int k = 0; k < 16; k++)
            acc = _mm_xor_si128(acc, mul1(_mm_loadu_si128((const __m128i *)(j->cw + 16 * i + k)),
                                          _mm_loadu_si128((const __m128i *)&j->eq[k])));
        _mm_storeu_si128((__m128i *)(j->out + i), acc);
    }
}

/* eq_ind_partial_eval (multilinear/eq.rs:123-131) of 4 challenges, then the
   per-leaf inner product of 16 codeword elements */
int cb_fold_interleaved16(cb_ctx *c, const uint64_t *cw, size_t n_out, const uint64_t *chal, uint64_t *out) {
    fold_job j;
    j.cw = (const u128 *)cw;
    j.out = (u128 *)out;
    j.n = n_out;
    j.eq[0] = 1;
    for (int k = 0; k < 4; k++) {
        const u128 r = ldu(chal + 2 * k);
        for (int i = 0; i < (1 << k); i++) {
            const u128 hi = gmul(j.eq[i], r);
            j.eq[i + (1 << k)] = hi;
            j.eq[i] ^= hi;
        }
    }
    par_for(c->pool, (n_out + 4095) / 4096, g_wide ? fold_task4 : fold_task1, &j);
    return 0;
}

/* One commit step: encode (rate 1/2^lir, batch 2^lbs), Merkle tree with leaves of
   2^lbs elements, first (interleaved) fold with lbs = 4 challenges. */
int cb_commit_step(cb_ctx *c, const uint64_t *msg, int log_dim, int lir, int lbs, uint64_t *cw, uint8_t *tree,
                   const uint64_t *chal, uint64_t *folded) {
    if (lbs != 4) return 1;
    int rc = cb_encode_batch(c, msg, log_dim, lir, lbs, cw);
    if (rc) return rc;
    const size_t n = (size_t)1 << (log_dim + lir + lbs);
    rc = cb_merkle_build(c, cw, n, 16, tree);
    if (rc) return rc;
    return cb_fold_interleaved16(c, cw, n / 16, chal, folded);
}
