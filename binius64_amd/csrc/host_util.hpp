// Host-side arithmetic for the library: GF(2^128) (setup of twiddle bases,
// fold challenges) and SHA-256 (Fiat-Shamir transcript, compression IV).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#if !defined(__HIP_DEVICE_COMPILE__)
#include <cpuid.h>
#include <immintrin.h>
#endif

namespace b64 {
namespace host {

typedef unsigned __int128 u128;

static inline u128 clmul64(uint64_t a, uint64_t b) {
    u128 r = 0;
    for (int i = 0; i < 64; i++)
        if ((b >> i) & 1) r ^= (u128)a << i;
    return r;
}

// Reduction of the 256-bit product (lo, mid at x^64, hi at x^128) by
// x^128 = x^7 + x^2 + x + 1.
static inline u128 gf_reduce(u128 lo, u128 mid, u128 hi) {
    uint64_t v0 = (uint64_t)lo, v1 = (uint64_t)(lo >> 64) ^ (uint64_t)mid;
    uint64_t v2 = (uint64_t)hi ^ (uint64_t)(mid >> 64), v3 = (uint64_t)(hi >> 64);
    v1 ^= v3 ^ (v3 << 1) ^ (v3 << 2) ^ (v3 << 7);
    v2 ^= (v3 >> 63) ^ (v3 >> 62) ^ (v3 >> 57);
    v0 ^= v2 ^ (v2 << 1) ^ (v2 << 2) ^ (v2 << 7);
    v1 ^= (v2 >> 63) ^ (v2 >> 62) ^ (v2 >> 57);
    return ((u128)v1 << 64) | v0;
}

// The x86 carry-less multiply (PCLMULQDQ), detected once; the config key
// "host.pclmul" = 0 turns its use off (the loop below then runs, as on a CPU
// without it).
inline std::atomic<bool> g_pclmul_allowed{true};
#if !defined(__HIP_DEVICE_COMPILE__)
static inline bool have_pclmul() {
    static const bool have = [] {
        unsigned a, b, c, d;
        return __get_cpuid(1, &a, &b, &c, &d) && (c & bit_PCLMUL) != 0;
    }();
    return have;
}
#endif
static inline bool use_pclmul() {
#if !defined(__HIP_DEVICE_COMPILE__)
    return have_pclmul() && g_pclmul_allowed.load(std::memory_order_relaxed);
#else
    return false;
#endif
}
#if !defined(__HIP_DEVICE_COMPILE__)
__attribute__((target("pclmul,sse2"))) static inline u128 gf_mul_pclmul(u128 a, u128 b) {
    const __m128i x = _mm_set_epi64x((long long)(uint64_t)(a >> 64), (long long)(uint64_t)a);
    const __m128i y = _mm_set_epi64x((long long)(uint64_t)(b >> 64), (long long)(uint64_t)b);
    const __m128i l = _mm_clmulepi64_si128(x, y, 0x00), h = _mm_clmulepi64_si128(x, y, 0x11);
    const __m128i m = _mm_xor_si128(_mm_clmulepi64_si128(x, y, 0x10), _mm_clmulepi64_si128(x, y, 0x01));
    u128 lo, mid, hi;
    memcpy(&lo, &l, 16);
    memcpy(&mid, &m, 16);
    memcpy(&hi, &h, 16);
    return gf_reduce(lo, mid, hi);
}
#endif

// GHASH product (verifier/field/src/ghash.rs:188-195): schoolbook carry-less
// product, then reduction by x^128 = x^7 + x^2 + x + 1.  PCLMULQDQ where the
// CPU has it (as the reference's x86 backend), else a shift-and-XOR loop.
static inline u128 gf_mul(u128 a, u128 b) {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (use_pclmul()) return gf_mul_pclmul(a, b);
#endif
    const uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64);
    const uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    return gf_reduce(clmul64(a0, b0), clmul64(a0, b1) ^ clmul64(a1, b0), clmul64(a1, b1));
}

// a^(2^128 - 2): the inverse, and 0 for 0 (invert_or_zero).
static inline u128 gf_inv(u128 a) {
    u128 r = 1, base = a;
    u128 e = ~(u128)0 - 1;
    while (e) {
        if (e & 1) r = gf_mul(r, base);
        base = gf_mul(base, base);
        e >>= 1;
    }
    return r;
}

class Sha256 {
   public:
    Sha256() { reset(); }
    void reset() {
        static const uint32_t iv[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                       0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
        memcpy(h_, iv, sizeof(iv));
        len_ = 0;
        total_ = 0;
    }
    void update(const uint8_t *p, size_t n) {
        total_ += n;
        while (n) {
            size_t take = 64 - len_;
            if (take > n) take = n;
            memcpy(buf_ + len_, p, take);
            len_ += take;
            p += take;
            n -= take;
            if (len_ == 64) {
                block(h_, buf_);
                len_ = 0;
            }
        }
    }
    // The digest of everything absorbed so far (the state is left as it was):
    // the pending bytes, 0x80, zeros and the 64-bit bit length, in one block or two.
    void finalize(uint8_t out[32]) const {
        uint32_t h[8];
        memcpy(h, h_, sizeof(h));
        uint8_t blk[128];
        memcpy(blk, buf_, len_);
        size_t n = len_;
        blk[n++] = 0x80;
        const size_t end = n <= 56 ? 64 : 128;
        memset(blk + n, 0, end - 8 - n);
        const uint64_t bits = total_ * 8;
        for (int i = 0; i < 8; i++) blk[end - 8 + i] = (uint8_t)(bits >> (56 - 8 * i));
        block(h, blk);
        if (end == 128) block(h, blk + 64);
        for (int i = 0; i < 8; i++) {
            out[4 * i] = (uint8_t)(h[i] >> 24);
            out[4 * i + 1] = (uint8_t)(h[i] >> 16);
            out[4 * i + 2] = (uint8_t)(h[i] >> 8);
            out[4 * i + 3] = (uint8_t)h[i];
        }
    }
    static void digest(const uint8_t *p, size_t n, uint8_t out[32]) {
        Sha256 s;
        s.update(p, n);
        s.finalize(out);
    }

   private:
    static inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
    static void block(uint32_t h[8], const uint8_t *b) {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (sha_ni()) {
            block_ni(h, b);
            return;
        }
#endif
        block_scalar(h, b);
    }
#if !defined(__HIP_DEVICE_COMPILE__)
    // The x86 SHA extensions (with SSSE3 and SSE4.1), detected once.
    static bool sha_ni() {
        static const bool have = [] {
            unsigned a, b, c, d;
            if (!__get_cpuid(1, &a, &b, &c, &d) || !(c & bit_SSSE3) || !(c & bit_SSE4_1)) return false;
            if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return false;
            return (b & bit_SHA) != 0;
        }();
        return have;
    }
    // One block with sha256rnds2 (two rounds each) on the state split as
    // (A,B,E,F) / (C,D,G,H), the schedule four words at a time from
    // sha256msg1 / sha256msg2 (Intel's SHA extensions reference flow).
    __attribute__((target("sha,sse4.1,ssse3"))) static void block_ni(uint32_t h[8], const uint8_t *b) {
        static const uint32_t K[64] = {
            0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
            0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
            0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
            0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
            0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
            0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
            0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
            0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
        const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
        const __m128i cdab = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i *)&h[0]), 0xB1);
        const __m128i efgh = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i *)&h[4]), 0x1B);
        __m128i abef = _mm_alignr_epi8(cdab, efgh, 8);
        __m128i cdgh = _mm_blend_epi16(efgh, cdab, 0xF0);
        const __m128i abef0 = abef, cdgh0 = cdgh;
        __m128i w[4];
        for (int g = 0; g < 16; g++) {
            __m128i x;
            if (g < 4) {
                x = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i *)(b + 16 * g)), bswap);
            } else {
                x = _mm_sha256msg1_epu32(w[g & 3], w[(g + 1) & 3]);
                x = _mm_add_epi32(x, _mm_alignr_epi8(w[(g + 3) & 3], w[(g + 2) & 3], 4));
                x = _mm_sha256msg2_epu32(x, w[(g + 3) & 3]);
            }
            w[g & 3] = x;
            const __m128i m = _mm_add_epi32(x, _mm_loadu_si128((const __m128i *)&K[4 * g]));
            cdgh = _mm_sha256rnds2_epu32(cdgh, abef, m);
            abef = _mm_sha256rnds2_epu32(abef, cdgh, _mm_shuffle_epi32(m, 0x0E));
        }
        abef = _mm_add_epi32(abef, abef0);
        cdgh = _mm_add_epi32(cdgh, cdgh0);
        const __m128i feba = _mm_shuffle_epi32(abef, 0x1B);
        const __m128i dchg = _mm_shuffle_epi32(cdgh, 0xB1);
        _mm_storeu_si128((__m128i *)&h[0], _mm_blend_epi16(feba, dchg, 0xF0));
        _mm_storeu_si128((__m128i *)&h[4], _mm_alignr_epi8(dchg, feba, 8));
    }
#endif
    static void block_scalar(uint32_t h[8], const uint8_t *b) {
        static const uint32_t K[64] = {
            0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
            0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
            0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
            0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
            0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
            0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
            0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
            0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
        uint32_t w[64];
        for (int i = 0; i < 16; i++)
            w[i] = ((uint32_t)b[4 * i] << 24) | ((uint32_t)b[4 * i + 1] << 16) | ((uint32_t)b[4 * i + 2] << 8) |
                   b[4 * i + 3];
        for (int i = 16; i < 64; i++) {
            const uint32_t s0 = ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], bb = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; i++) {
            const uint32_t t1 = hh + (ror(e, 6) ^ ror(e, 11) ^ ror(e, 25)) + ((e & f) ^ (~e & g)) + K[i] + w[i];
            const uint32_t t2 = (ror(a, 2) ^ ror(a, 13) ^ ror(a, 22)) + ((a & bb) ^ (a & c) ^ (bb & c));
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = bb;
            bb = a;
            a = t1 + t2;
        }
        h[0] += a;
        h[1] += bb;
        h[2] += c;
        h[3] += d;
        h[4] += e;
        h[5] += f;
        h[6] += g;
        h[7] += hh;
    }
    uint32_t h_[8];
    uint8_t buf_[64];
    size_t len_;
    uint64_t total_;
};

}  // namespace host
}  // namespace b64
