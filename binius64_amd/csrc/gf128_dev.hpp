// GF(2^128) arithmetic for gfx950 kernels: the GHASH field of
// verifier/field/src/ghash.rs (modulus x^128 + x^7 + x^2 + x + 1, bit i of the
// little-endian u128 is the coefficient of x^i).  An element lives in a uint4:
// .x = bits 0..31, .y = 32..63, .z = 64..95, .w = 96..127, which is exactly the
// 16-byte HBM image, so elements move with single 16-byte loads and stores.
//
// Two multipliers:
//  * gf_mul      -- general product.  CDNA4 has no carry-less multiply, so the
//                   32x32 carry-less products come from integer multiplies on
//                   operands with 3-bit "holes" (the method of the reference's
//                   portable backend, arch/portable/univariate_mul_utils_128.rs:120-141,
//                   re-cut for 32-bit lanes), combined by two Karatsuba levels and a
//                   word-wise reduction (packed_ghash_128.rs:84-97).
//  * WinTable    -- product by a constant c shared by many lanes: a 4-bit window
//                   table T[k][n] = c * n * x^(4k) (32 x 16 entries, 8 KiB) in LDS;
//                   v*c is the XOR of 32 table rows selected by v's nibbles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bitop_dev.hpp"

namespace b64 {

__device__ __forceinline__ uint4 gxor(uint4 a, uint4 b) {
    return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}
__device__ __forceinline__ bool gzero(uint4 a) { return (a.x | a.y | a.z | a.w) == 0u; }

__device__ __forceinline__ uint64_t xor3_64(uint64_t a, uint64_t b, uint64_t c) {
    const uint32_t lo = xor3((uint32_t)a, (uint32_t)b, (uint32_t)c);
    const uint32_t hi = xor3((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t xor4_64(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
    const uint32_t lo = xor3((uint32_t)a, (uint32_t)b, (uint32_t)c) ^ (uint32_t)d;
    const uint32_t hi = xor3((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32)) ^ (uint32_t)(d >> 32);
    return ((uint64_t)hi << 32) | lo;
}

// 32x32 -> 64 carry-less product: operands split into four masks with 3-bit
// holes; each integer product sums at most 8 one-bit terms per kept position,
// so carries stay inside the holes.  The four partial sums z_t are exact at the
// positions = t (mod 4) and are merged with two levels of bit selects.
// Two products may also be ADDED (one v_mad_u64_u32) when their term counts at
// every kept position sum to at most 15.  The count of a_s b_t at position
// 4k + (s + t mod 4) rises by one per step of k to a peak of 8 and falls again;
// the pairs taken here peak at neighbouring k (a0 b0 with a1 b3; a0 b1 with
// a2 b3 and a1 b0 with a3 b2; a0 b2 with a3 b3), so their sums stay at most 15
// at every position, while any other pairing reaches 16.  That saves six XORs
// per call.
__device__ __forceinline__ uint64_t clmul32(uint32_t a, uint32_t b) {
    const uint32_t m0 = 0x11111111u, m1 = 0x22222222u, m2 = 0x44444444u, m3 = 0x88888888u;
    const uint32_t a0 = a & m0, a1 = a & m1, a2 = a & m2, a3 = a & m3;
    const uint32_t b0 = b & m0, b1 = b & m1, b2 = b & m2, b3 = b & m3;
    const uint64_t z0 = xor3_64((uint64_t)a0 * b0 + (uint64_t)a1 * b3, (uint64_t)a2 * b2, (uint64_t)a3 * b1);
    const uint64_t z1 = ((uint64_t)a0 * b1 + (uint64_t)a2 * b3) ^ ((uint64_t)a1 * b0 + (uint64_t)a3 * b2);
    const uint64_t z2 = xor3_64((uint64_t)a0 * b2 + (uint64_t)a3 * b3, (uint64_t)a1 * b1, (uint64_t)a2 * b0);
    const uint64_t z3 = xor4_64((uint64_t)a0 * b3, (uint64_t)a1 * b2, (uint64_t)a2 * b1, (uint64_t)a3 * b0);
    // positions 0,1 (mod 4) from z0/z1, positions 2,3 from z2/z3
    const uint32_t lo = bsel(bsel((uint32_t)z0, (uint32_t)z1, m0), bsel((uint32_t)z2, (uint32_t)z3, m2), m0 | m1);
    const uint32_t hi = bsel(bsel((uint32_t)(z0 >> 32), (uint32_t)(z1 >> 32), m0),
                             bsel((uint32_t)(z2 >> 32), (uint32_t)(z3 >> 32), m2), m0 | m1);
    return ((uint64_t)hi << 32) | lo;
}

// 64x64 -> 128 carry-less product (one Karatsuba level over clmul32).
__device__ __forceinline__ void clmul64(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p0 = clmul32(a0, b0);
    const uint64_t p2 = clmul32(a1, b1);
    const uint64_t p1 = xor3_64(clmul32(a0 ^ a1, b0 ^ b1), p0, p2);
    lo = p0 ^ (p1 << 32);
    hi = p2 ^ (p1 >> 32);
}

// Reduce the 256-bit product (v0..v3, 64-bit words, v0 lowest) modulo
// x^128 + x^7 + x^2 + x + 1.
__device__ __forceinline__ uint4 gf_reduce(uint64_t v0, uint64_t v1, uint64_t v2, uint64_t v3) {
    v1 = xor3_64(xor3_64(v1, v3, v3 << 1), v3 << 2, v3 << 7);
    v2 = xor4_64(v2, v3 >> 63, v3 >> 62, v3 >> 57);
    v0 = xor3_64(xor3_64(v0, v2, v2 << 1), v2 << 2, v2 << 7);
    v1 = xor4_64(v1, v2 >> 63, v2 >> 62, v2 >> 57);
    return make_uint4((uint32_t)v0, (uint32_t)(v0 >> 32), (uint32_t)v1, (uint32_t)(v1 >> 32));
}

__device__ __forceinline__ uint4 gf_mul(uint4 a, uint4 b) {
    const uint64_t a0 = ((uint64_t)a.y << 32) | a.x, a1 = ((uint64_t)a.w << 32) | a.z;
    const uint64_t b0 = ((uint64_t)b.y << 32) | b.x, b1 = ((uint64_t)b.w << 32) | b.z;
    uint64_t l0, h0, l2, h2, l1, h1;
    clmul64(a0, b0, l0, h0);
    clmul64(a1, b1, l2, h2);
    clmul64(a0 ^ a1, b0 ^ b1, l1, h1);
    // middle term l1 ^ l0 ^ l2 lands on words 1 and 2
    return gf_reduce(l0, xor4_64(h0, l1, l0, l2), xor4_64(l2, h1, h0, h2), h2);
}

// v * x (ghash.rs:69-80)
__device__ __forceinline__ uint4 gf_mul_x(uint4 v) {
    const uint32_t carry = (v.w >> 31) ? 0x87u : 0u;
    return make_uint4((v.x << 1) ^ carry, (v.y << 1) | (v.x >> 31), (v.z << 1) | (v.y >> 31),
                      (v.w << 1) | (v.z >> 31));
}

// ---- 4-bit window tables ---------------------------------------------------
constexpr int WIN_ROWS = 32;
constexpr int WIN_ENTRIES = WIN_ROWS * 16;  // uint4 entries per table (8 KiB)

// Entry n of row r is stored at slot n ^ (r & 15) of the row.  Row builders
// (one lane per row) then write 16 distinct 16-byte slots across any 16
// consecutive rows at each step, so the stores are free of LDS bank conflicts;
// lookups fold the same XOR into the nibble (win_swz).
__device__ __forceinline__ constexpr uint32_t win_swz(int row0) {
    // nibble j of a word whose rows are row0 .. row0+7 (row0 a multiple of 8)
    return (row0 & 8) ? 0xFEDCBA98u : 0x76543210u;
}

// v * x^(4r) for 0 <= r < 32: the 256-bit shift, then the general reduction.
__device__ __forceinline__ uint4 gf_mul_xpow4(uint4 v, uint32_t r) {
    const uint64_t lo = ((uint64_t)v.y << 32) | v.x, hi = ((uint64_t)v.w << 32) | v.z;
    const uint32_t s = 4 * r, t = s & 63;
    const bool big = s >= 64;
    const uint64_t a0 = big ? 0 : lo, a1 = big ? lo : hi, a2 = big ? hi : 0;
    const uint64_t w0 = a0 << t;
    const uint64_t w1 = (a1 << t) | (t ? a0 >> (64 - t) : 0);
    const uint64_t w2 = (a2 << t) | (t ? a1 >> (64 - t) : 0);
    const uint64_t w3 = t ? a2 >> (64 - t) : 0;
    return gf_reduce(w0, w1, w2, w3);
}

// Builds row `row` (0..31) of the window table of constant c: entries
// c * n * x^(4*row) for n = 0..15.  32 lanes build a whole table.
__device__ inline void win_build_row(uint4 *__restrict__ tab, uint4 c, int row) {
    const uint4 p0 = gf_mul_xpow4(c, (uint32_t)row);
    const uint4 p1 = gf_mul_x(p0);
    const uint4 p2 = gf_mul_x(p1);
    const uint4 p3 = gf_mul_x(p2);
    uint4 *r = tab + row * 16;
    const int sw = row & 15;
    const uint4 p01 = gxor(p0, p1), p32 = gxor(p3, p2);
    r[0 ^ sw] = make_uint4(0, 0, 0, 0);
    r[1 ^ sw] = p0;
    r[2 ^ sw] = p1;
    r[3 ^ sw] = p01;
    r[4 ^ sw] = p2;
    r[5 ^ sw] = gxor(p2, p0);
    r[6 ^ sw] = gxor(p2, p1);
    r[7 ^ sw] = gxor(p2, p01);
    r[8 ^ sw] = p3;
    r[9 ^ sw] = gxor(p3, p0);
    r[10 ^ sw] = gxor(p3, p1);
    r[11 ^ sw] = gxor(p3, p01);
    r[12 ^ sw] = p32;
    r[13 ^ sw] = gxor(p32, p0);
    r[14 ^ sw] = gxor(p32, p1);
    r[15 ^ sw] = gxor(p32, p01);
}

__device__ __forceinline__ uint4 win_word(const uint4 *__restrict__ tab, uint32_t w, int row0, uint4 acc) {
    w ^= win_swz(row0);
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
        const uint4 e0 = tab[(row0 + j) * 16 + ((w >> (4 * j)) & 15u)];
        const uint4 e1 = tab[(row0 + j + 1) * 16 + ((w >> (4 * j + 4)) & 15u)];
        acc = make_uint4(xor3(acc.x, e0.x, e1.x), xor3(acc.y, e0.y, e1.y), xor3(acc.z, e0.z, e1.z),
                         xor3(acc.w, e0.w, e1.w));
    }
    return acc;
}

// Same product with the lookup addresses built by byte permutes: the even and
// odd nibbles of a word, pre-scaled by 16 (the entry size), sit in the bytes of
// two masked words; v_perm_b32 drops one of those bytes into the low byte of the
// 256-byte-aligned table address, and the row offset (row * 256) rides in the
// ds_read immediate.  One VALU op per lookup instead of two.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_cu32x4;
typedef __attribute__((address_space(3))) const uint4 lds_cuint4;

__device__ __forceinline__ uint4 gxor3v(uint4 a, u32x4 b, u32x4 c) {
    return make_uint4(xor3(a.x, b.x, c.x), xor3(a.y, b.y, c.y), xor3(a.z, b.z, c.z), xor3(a.w, b.w, c.w));
}

// The lookups go out in groups of G (2, 4 or 8) before their XORs, so a wave
// keeps G table reads in flight; G trades LDS latency against VGPRs.
// (a & b) ^ c
__device__ __forceinline__ uint32_t and_xor(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x6A);
}

// The nibbles of w ^ swz, pre-scaled by 16, in two masked words: lo holds
// nibbles 0,2,4,6 and hi nibbles 1,3,5,7 (one byte each).  The swizzle is
// folded into the masking (three VALU ops per word).
__device__ __forceinline__ void nib_split(uint32_t w, uint32_t swz, uint32_t &lo, uint32_t &hi) {
    lo = and_xor(w << 4, 0xF0F0F0F0u, (swz << 4) & 0xF0F0F0F0u);
    hi = and_xor(w, 0xF0F0F0F0u, swz & 0xF0F0F0F0u);
}

template <int G = 8>
__device__ __forceinline__ uint4 win_word_perm(uint32_t base, uint32_t w, int row0, uint4 acc) {
    uint32_t lo, hi;  // nibbles 0,2,4,6 and 1,3,5,7 (after the row swizzle) * 16
    nib_split(w, win_swz(row0), lo, hi);
#pragma unroll
    for (int g = 0; g < 8; g += G) {
        u32x4 e[G];
#pragma unroll
        for (int k = g / 2; k < (g + G) / 2; k++) {
            const uint32_t sel = 0x07060500u | (uint32_t)k;  // byte k of src1, bytes 1..3 of base
            const uint32_t a0 = __builtin_amdgcn_perm(base, lo, sel);
            const uint32_t a1 = __builtin_amdgcn_perm(base, hi, sel);
            const lds_cu32x4 *p0 = (const lds_cu32x4 *)(uintptr_t)a0;
            const lds_cu32x4 *p1 = (const lds_cu32x4 *)(uintptr_t)a1;
            e[2 * k - g] = p0[(row0 + 2 * k) * 16];
            e[2 * k + 1 - g] = p1[(row0 + 2 * k + 1) * 16];
        }
#pragma unroll
        for (int k = 0; k < G / 2; k++) acc = gxor3v(acc, e[2 * k], e[2 * k + 1]);
    }
    return acc;
}

// tab must be 256-byte aligned in LDS.
__device__ __forceinline__ uint32_t lds_addr(const uint4 *p) {
    return (uint32_t)(uintptr_t)(const lds_cuint4 *)p;
}

// acc0 + v * c (acc0 = 0 for the plain product; an NTT butterfly passes u).
template <int G = 8>
__device__ __forceinline__ uint4 win_mul_perm(uint32_t tab_base, uint4 v, uint4 acc0 = make_uint4(0, 0, 0, 0)) {
    uint4 acc = acc0;
    acc = win_word_perm<G>(tab_base, v.x, 0, acc);
    acc = win_word_perm<G>(tab_base, v.y, 8, acc);
    acc = win_word_perm<G>(tab_base, v.z, 16, acc);
    acc = win_word_perm<G>(tab_base, v.w, 24, acc);
    return acc;
}

// ---- 2 KiB window tables with pre-shifted rows -------------------------------
// Row s (0..7) of the table of a constant t holds t * n * x^(4s) mod P for the 16
// nibbles n, entry n at slot n ^ s (8 rows x 16 entries x 16 B = 2 KiB).  A 32-bit
// word w of v is the sum of its nibbles times x^(4s), so w * t is the XOR of 8
// row entries, and word q of v contributes that sum times x^(32q): word-aligned,
// so the 32 lookups of a product accumulate into 7 unreduced words and one
// reduction of the 96-bit high part finishes it (x^128 = 1 + x + x^2 + x^7 and
// the folded high part stays below x^103, so a single fold suffices).  A table is
// a quarter of the 32-row form's size and is built from one shift and three
// doublings per row, so it pays for constants shared by as few as 16 products.
constexpr int W8_ROWS = 8;
constexpr int W8_ENTRIES = W8_ROWS * 16;  // uint4 entries per table (2 KiB)

// v * x^k for 0 <= k < 32.
__device__ __forceinline__ uint4 gf_mul_xk(uint4 v, uint32_t k) {
    const uint32_t r = 32u - k;  // 1..32; alignbit takes r & 31, so k = 0 is selected below
    const uint32_t o = v.w >> (r & 31u);  // the k bits shifted out (garbage for k = 0)
    uint32_t w0 = v.x << k, w1 = __builtin_amdgcn_alignbit(v.y, v.x, r);
    uint32_t w2 = __builtin_amdgcn_alignbit(v.z, v.y, r), w3 = __builtin_amdgcn_alignbit(v.w, v.z, r);
    // o * (1 + x + x^2 + x^7): at most 38 bits
    w0 ^= xor3(o, o << 1, o << 2) ^ (o << 7);
    w1 ^= xor3(o >> 31, o >> 30, o >> 25);
    return k ? make_uint4(w0, w1, w2, w3) : v;
}

// Builds entries 8 * half .. 8 * half + 7 of row `row` of t's table (16 lanes,
// two per row, build a whole table).
__device__ __forceinline__ void w8_build_half(uint4 *__restrict__ tab, uint4 t, int row, int half) {
    const uint4 p0 = gf_mul_xk(t, 4u * (uint32_t)row);
    const uint4 p1 = gf_mul_x(p0);
    const uint4 p2 = gf_mul_x(p1);
    const uint4 p3 = gf_mul_x(p2);
    const uint4 z = make_uint4(0, 0, 0, 0);
    const uint4 b = half ? p3 : z;  // entries 8..15 add the x^3 term
    const uint4 e1 = gxor(b, p0), e2 = gxor(b, p1), e4 = gxor(b, p2);
    const uint4 e3 = gxor(e1, p1), e5 = gxor(e4, p0), e6 = gxor(e4, p1), e7 = gxor(e5, p1);
    uint4 *r = tab + row * 16 + 8 * half;
    r[0 ^ row] = b;
    r[1 ^ row] = e1;
    r[2 ^ row] = e2;
    r[3 ^ row] = e3;
    r[4 ^ row] = e4;
    r[5 ^ row] = e5;
    r[6 ^ row] = e6;
    r[7 ^ row] = e7;
}

// Builds row `row` of t's table (all 16 entries, one lane per row).
__device__ __forceinline__ void w8_build_row(uint4 *__restrict__ tab, uint4 t, int row) {
    const uint4 p0 = gf_mul_xk(t, 4u * (uint32_t)row);
    const uint4 p1 = gf_mul_x(p0);
    const uint4 p2 = gf_mul_x(p1);
    const uint4 p3 = gf_mul_x(p2);
    uint4 e[16];
    e[0] = make_uint4(0, 0, 0, 0);
    e[1] = p0;
    e[2] = p1;
    e[3] = gxor(p0, p1);
    e[4] = p2;
#pragma unroll
    for (int n = 5; n < 8; n++) e[n] = gxor(e[n - 4], p2);
    e[8] = p3;
#pragma unroll
    for (int n = 9; n < 16; n++) e[n] = gxor(e[n - 8], p3);
    uint4 *r = tab + row * 16;
#pragma unroll
    for (int n = 0; n < 16; n++) r[n ^ row] = e[n];
}

// v * t through t's 2 KiB table at LDS address tab_base (256-byte aligned); the
// lookups go out in groups of G (as win_word_perm).
template <int G = 8>
__device__ __forceinline__ uint4 w8_mul(uint32_t tab_base, uint4 v, uint4 acc0 = make_uint4(0, 0, 0, 0)) {
    uint32_t acc[7] = {acc0.x, acc0.y, acc0.z, acc0.w, 0, 0, 0};
    const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t lo, hi;  // nibble s -> slot nib ^ s of row s
        nib_split(vw[q], 0x76543210u, lo, hi);
#pragma unroll
        for (int g = 0; g < 8; g += G) {
            u32x4 e[G];
#pragma unroll
            for (int k = g / 2; k < (g + G) / 2; k++) {
                const uint32_t sel = 0x07060500u | (uint32_t)k;
                const lds_cu32x4 *p0 = (const lds_cu32x4 *)(uintptr_t)__builtin_amdgcn_perm(tab_base, lo, sel);
                const lds_cu32x4 *p1 = (const lds_cu32x4 *)(uintptr_t)__builtin_amdgcn_perm(tab_base, hi, sel);
                e[2 * k - g] = p0[(2 * k) * 16];
                e[2 * k + 1 - g] = p1[(2 * k + 1) * 16];
            }
#pragma unroll
            for (int k = 0; k < G / 2; k++) {
                acc[q + 0] = xor3(acc[q + 0], e[2 * k].x, e[2 * k + 1].x);
                acc[q + 1] = xor3(acc[q + 1], e[2 * k].y, e[2 * k + 1].y);
                acc[q + 2] = xor3(acc[q + 2], e[2 * k].z, e[2 * k + 1].z);
                acc[q + 3] = xor3(acc[q + 3], e[2 * k].w, e[2 * k + 1].w);
            }
        }
    }
    // fold the high words h = acc[4..6] (below x^96): r = lo ^ h * (1 + x + x^2 + x^7)
    const uint32_t h0 = acc[4], h1 = acc[5], h2 = acc[6];
    const uint32_t r0 = xor3(xor3(acc[0], h0, h0 << 1), h0 << 2, h0 << 7);
    const uint32_t r1 = xor3(xor3(acc[1], h1, __builtin_amdgcn_alignbit(h1, h0, 31)),
                             __builtin_amdgcn_alignbit(h1, h0, 30), __builtin_amdgcn_alignbit(h1, h0, 25));
    const uint32_t r2 = xor3(xor3(acc[2], h2, __builtin_amdgcn_alignbit(h2, h1, 31)),
                             __builtin_amdgcn_alignbit(h2, h1, 30), __builtin_amdgcn_alignbit(h2, h1, 25));
    const uint32_t r3 = xor3(xor3(acc[3], h2 >> 31, h2 >> 30), h2 >> 25, 0u);
    return make_uint4(r0, r1, r2, r3);
}

// ---- 1 KiB window tables: the first four rows of the 2 KiB form -------------
// Rows 0..3 (t * n * x^(4s), same layout, built by w8_build_half with row < 4).
// Nibbles 0..3 of word q of v index rows 0..3 at word offset q, nibbles 4..7 the
// same rows 16 bits further: two word-aligned accumulators, one 16-bit funnel
// to merge them, and a full reduction of the 255-bit product.
constexpr int W4_ROWS = 4;
constexpr int W4_ENTRIES = W4_ROWS * 16;  // uint4 entries per table (1 KiB)

__device__ __forceinline__ uint4 w4_mul(uint32_t tab_base, uint4 v, uint4 acc0 = make_uint4(0, 0, 0, 0)) {
    uint32_t A[7] = {acc0.x, acc0.y, acc0.z, acc0.w, 0, 0, 0}, B[7] = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        // nibble j -> slot nib ^ (j & 3); lo: nibbles 0,2,4,6 (rows 0,2,0,2), hi: 1,3,5,7 (rows 1,3,1,3)
        uint32_t lo, hi;
        nib_split(vw[q], 0x32103210u, lo, hi);
        u32x4 e[8];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t sel = 0x07060500u | (uint32_t)k;
            const lds_cu32x4 *p0 = (const lds_cu32x4 *)(uintptr_t)__builtin_amdgcn_perm(tab_base, lo, sel);
            const lds_cu32x4 *p1 = (const lds_cu32x4 *)(uintptr_t)__builtin_amdgcn_perm(tab_base, hi, sel);
            e[2 * k] = p0[(2 * (k & 1)) * 16];
            e[2 * k + 1] = p1[(2 * (k & 1) + 1) * 16];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t *acc = k < 2 ? A : B;
            acc[q + 0] = xor3(acc[q + 0], e[2 * k].x, e[2 * k + 1].x);
            acc[q + 1] = xor3(acc[q + 1], e[2 * k].y, e[2 * k + 1].y);
            acc[q + 2] = xor3(acc[q + 2], e[2 * k].z, e[2 * k + 1].z);
            acc[q + 3] = xor3(acc[q + 3], e[2 * k].w, e[2 * k + 1].w);
        }
    }
    // u = A + B * x^16 (8 words, degree <= 254)
    uint32_t u[8];
    u[0] = A[0] ^ (B[0] << 16);
#pragma unroll
    for (int i = 1; i < 7; i++) u[i] = A[i] ^ __builtin_amdgcn_alignbit(B[i], B[i - 1], 16);
    u[7] = B[6] >> 16;
    // fold h = u[4..7] (below x^127): h * (1 + x + x^2 + x^7), whose bits at
    // x^128 and up (g, below x^7) fold once more
    const uint32_t h0 = u[4], h1 = u[5], h2 = u[6], h3 = u[7];
    const uint32_t g = xor3(h3 >> 31, h3 >> 30, h3 >> 25);
    const uint32_t r0 = xor3(xor3(u[0], h0, h0 << 1), xor3(h0 << 2, h0 << 7, g), xor3(g << 1, g << 2, g << 7));
    const uint32_t r1 = xor3(xor3(u[1], h1, __builtin_amdgcn_alignbit(h1, h0, 31)),
                             __builtin_amdgcn_alignbit(h1, h0, 30), __builtin_amdgcn_alignbit(h1, h0, 25));
    const uint32_t r2 = xor3(xor3(u[2], h2, __builtin_amdgcn_alignbit(h2, h1, 31)),
                             __builtin_amdgcn_alignbit(h2, h1, 30), __builtin_amdgcn_alignbit(h2, h1, 25));
    const uint32_t r3 = xor3(xor3(u[3], h3, __builtin_amdgcn_alignbit(h3, h2, 31)),
                             __builtin_amdgcn_alignbit(h3, h2, 30), __builtin_amdgcn_alignbit(h3, h2, 25));
    return make_uint4(r0, r1, r2, r3);
}

// v * c using c's window table.
__device__ __forceinline__ uint4 win_mul(const uint4 *__restrict__ tab, uint4 v) {
    uint4 acc = make_uint4(0, 0, 0, 0);
    acc = win_word(tab, v.x, 0, acc);
    acc = win_word(tab, v.y, 8, acc);
    acc = win_word(tab, v.z, 16, acc);
    acc = win_word(tab, v.w, 24, acc);
    return acc;
}

}  // namespace b64
